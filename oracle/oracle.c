/* CPU oracle: both precisions of oracle_impl.h.  TEST INFRASTRUCTURE ONLY:
 * loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg. */
#include <math.h>
#include <stdlib.h>
#include "oracle.h"

#define PI_D 3.14159265358979323846

/* Subject stride of the fluid/wall neighbour loops: 1 computes every
 * subject.  k > 1 computes every k-th subject only -- a bounded sample used
 * to time the oracle on large workloads, never to produce results. */
static int ora_stride = 1;
void ora_set_stride(int k) { ora_stride = k > 0 ? k : 1; }

/* FP32 pow: evaluated in double and rounded once to float (the correctly
 * rounded value up to a double-rounding tie).  Warp's float pow is an
 * implementation of the same function with a few ulp of error; this fixes
 * one value so the oracle and the HIP path can be compared bit for bit. */
static float powf_cr(float x, float y) { return (float)pow((double)x, (double)y); }

#define REAL float
#define SFX _f32
#define SQRT sqrtf
#define POW powf_cr
#define FABS fabsf
#include "oracle_impl.h"
#undef REAL
#undef SFX
#undef SQRT
#undef POW
#undef FABS

#define REAL double
#define SFX _f64
#define SQRT sqrt
#define POW pow
#define FABS fabs
#include "oracle_impl.h"

int ora_num_threads(void) {
#ifdef _OPENMP
    extern int omp_get_max_threads(void);
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* Threads of the OpenMP loops from here on (the baseline's 1-thread leg). */
void ora_set_num_threads(int n) {
#ifdef _OPENMP
    extern void omp_set_num_threads(int);
    omp_set_num_threads(n > 0 ? n : 1);
#else
    (void)n;
#endif
}
