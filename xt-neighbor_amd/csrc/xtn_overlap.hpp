// xtn_overlap.hpp -- repertoire-overlap kernels (filled in below the NN path).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "xtn_prims.hpp"
