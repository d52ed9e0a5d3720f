// plh_angles.hpp -- cos and sin of a rotation angle, bit for bit as the
// reference's own build computes them.
//
// evolve (include/PauliSumOperations.h:113-114) and apply_RX/RY/RZ (:30-31)
// take std::cos and std::sin of one angle.  The reference's compilers merge
// that pair into one call of the C library's sincos (GCC at any
// optimisation level, Clang under the reference's -ffast-math), whose
// results can differ in the last bit from separate cos and sin calls: with
// glibc 2.35, theta = 1.7152103698830625 gives cos = -0x1.26bba5d0ad6b6p-3
// from sincos but -0x1.26bba5d0ad6b5p-3 from cos, which the depth-20
// brickwork propagation meets at its 1213th rotation (program op 1274).  The engine therefore
// calls sincos itself instead of leaving the choice to its compiler.
#pragma once
#include <math.h>

namespace plh {

inline void angle_cos_sin(double theta, double* c, double* s) { ::sincos(theta, s, c); }

}  // namespace plh
