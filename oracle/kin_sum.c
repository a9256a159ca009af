/* CPU oracle for the kinematic scattering sum -- TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this library, and only as the checker: the product path never calls it.
 *
 * Exports:
 *   oracle_kin_sum_f64  -- double precision, term for term the reference's
 *                          compute_scattering_cffi (Beam.py:663-816);
 *   oracle_kin_sum_f80  -- the same with x87 long double (64-bit
 *                          significand), the high-precision truth that the
 *                          float64 path and the GPU path are both measured
 *                          against.
 * Both accumulate into double out_r/out_i and return 0 (2 on bad sizes).
 * See kin_sum_body.h for the arithmetic.
 */
#include <math.h>
#include <stdlib.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define REAL double
#define SUFFIX _f64
#define PI_R 3.14159265358979323846
#define SQRT_R sqrt
#define EXP_R exp
#define SIN_R sin
#define COS_R cos
#define FMOD_R fmod
#include "kin_sum_body.h"
#undef REAL
#undef SUFFIX
#undef PI_R
#undef SQRT_R
#undef EXP_R
#undef SIN_R
#undef COS_R
#undef FMOD_R

#define REAL long double
#define SUFFIX _f80
#define PI_R 3.141592653589793238462643383279502884L
#define SQRT_R sqrtl
#define EXP_R expl
#define SIN_R sinl
#define COS_R cosl
#define FMOD_R fmodl
#include "kin_sum_body.h"
