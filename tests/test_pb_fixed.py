"""CPU: the binned pull's fixed-point terms (plato_amd/csrc/pb_fixed.hpp, the same
code the kernels compile) against a 128-bit integer reference -- the single-word terms
the kernels use exhaustively over every FP32 value in range at F = 33 and F = 60, FP64
weighted products, the choice of F from the mass bound and 64-bit sums; and the
two-word 96-bit form over every non-negative FP32 value below 2^32."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_fixed_point_terms_exhaustive(tmp_path):
    exe = tmp_path / "test_pb_fixed"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "plato_amd", "csrc"),
                    os.path.join(ROOT, "tests", "cpp", "test_pb_fixed.cc"), "-o", str(exe)],
                   check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "fp32 exhaustive: 0 mismatches" in r.stdout
    assert "single-word terms: 0 mismatches" in r.stdout
    assert "scale and single-word sums: 0 mismatches" in r.stdout
