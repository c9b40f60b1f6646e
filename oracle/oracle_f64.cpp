// oracle_f64.cpp — the oracle's differentiable Path estimator evaluated in
// float64 (TEST INFRASTRUCTURE).
//
// The same restated sources (oracle_core.cpp, oracle_trace.cpp,
// oracle_adjoint.cpp) compiled with the keyword `float` widened to double and
// the float libm calls mapped to their double versions, so that central
// differences of the estimator with h = 1e-3 are not swamped by float32
// rounding (about 1e-7 of a pixel, i.e. 5e-5 of it after dividing by 2h).
// The wire-format structures of lt.h keep their float32 fields (the header
// is included before the widening), the RNG keeps its 32-bit float steps and
// the random numbers are the float32 ones, so a float64 sample follows the
// float32 sample's path except where a decision sits within a rounding
// error of its threshold.  Built into _build/liboracle_f64.so by the Makefile;
// only tests/ load it.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <thread>
#include <utility>
#include <vector>

#include "../include/lt.h"

#define ORA_F64 1
#define ora ora64  // its own C++ symbols, distinct from the float32 library's
#define float double
#define sqrtf sqrt
#define fabsf fabs
#define fmaxf fmax
#define fminf fmin
#define cosf cos
#define sinf sin
#define powf pow
#define floorf floor
#define copysignf copysign
#define ldexpf ldexp
#define fmaf fma
#define expf exp
#define logf log
#define tanf tan
#define acosf acos
#define atan2f atan2

#include "oracle_core.cpp"
#include "oracle_trace.cpp"
#include "oracle_adjoint.cpp"

extern "C" {
// the float64 library's own host BVH (its triangles hold doubles)
void* ora64_accel_build(const lt_scene_desc* s) { return ora64::accel_build(*s); }
void ora64_accel_free(void* a) { ora64::accel_free((ora64::Accel*)a); }
}
