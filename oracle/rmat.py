"""Numpy twin of the counter-based RMAT generator (oracle/oracle.c orc_rmat_edges).

TEST INFRASTRUCTURE ONLY.  An independent restatement in vectorised numpy so the
C generator, this twin and the device generator can be compared edge for edge.

draw(e, j) = mix64(mix64(seed) ^ (e*32 + j)); level l uses the low (even l) or
high (odd l) 32 bits of draw(e, l // 2) and decides bit (scale-1-l) of the
(src, dst) pair: u < t0 -> (0,0), u < t1 -> dst bit, u < t2 -> src bit, else both.
"""
import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def mix64(x):
    x = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = x + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def thresholds(a=0.57, b=0.19, c=0.19):
    return np.array([np.floor(p * 4294967296.0) for p in (a, a + b, a + b + c)],
                    dtype=np.uint64).astype(np.uint32)


def edges(scale, edge_factor, seed, a=0.57, b=0.19, c=0.19, e0=0, n=None):
    E = edge_factor << scale
    if n is None:
        n = E - e0
    t = thresholds(a, b, c).astype(np.uint64)
    sk = mix64(np.uint64(seed))
    e = np.arange(e0, e0 + n, dtype=np.uint64)
    src = np.zeros(n, np.uint32)
    dst = np.zeros(n, np.uint32)
    word = None
    for l in range(scale):
        if l % 2 == 0:
            with np.errstate(over="ignore"):
                word = mix64(sk ^ (e * np.uint64(32) + np.uint64(l // 2)))
        u = (word >> np.uint64(32)) if l % 2 else (word & np.uint64(0xFFFFFFFF))
        bit = np.uint32(1 << (scale - 1 - l))
        dbit = (u >= t[0]) & ((u < t[1]) | (u >= t[2]))
        sbit = (u >= t[1])
        dst |= np.where(dbit, bit, np.uint32(0)).astype(np.uint32)
        src |= np.where(sbit, bit, np.uint32(0)).astype(np.uint32)
    return src, dst


def weights(seed, e0, n):
    sk = mix64(np.uint64(seed))
    e = np.arange(e0, e0 + n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        r = mix64(sk ^ (e * np.uint64(32) + np.uint64(31)))
    return ((r >> np.uint64(40)) + np.uint64(1)).astype(np.float32) * np.float32(1.0 / 16777216.0)
