// sphx_hydro_exact.hip - the std-SPH loops compiled with -ffp-contract=off
// (FP64 results round operation by operation like the reference CPU loops).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>

#include "sphx_internal.hpp"

#define SPHX_HYDRO_NS exact
#define SPHX_HYDRO_EXACT 1
#include "sphx_hydro_kernels.inc"
#include "sphx_ve_kernels.inc"
