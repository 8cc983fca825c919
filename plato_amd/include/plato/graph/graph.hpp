// Umbrella header, as plato/graph/graph.hpp.
#pragma once

#include "plato/graph/base.hpp"
#include "plato/graph/message_passing.hpp"
#include "plato/graph/state/dense_state.hpp"
#include "plato/graph/structure.hpp"
#include "plato/util/bitmap.hpp"
#include "plato/util/foutput.h"
#include "plato/util/perf.hpp"
