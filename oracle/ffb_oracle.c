/*
 * ffb_oracle.c -- CPU oracle for the fused D2Q9 pull-stream-collide BGK path.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker.  Never linked into, called by
 * or substituted for the HIP product path.
 *
 * Instantiates ffb_oracle_impl.h twice:
 *   ffbo_d_*  REAL = double  (the reference's FFB_FP64 build, settings.f90:7-9)
 *   ffbo_s_*  REAL = float   (the reference's default build, settings.f90:11-12)
 *
 * Build: gcc -O3 -ffp-contract=off -fopenmp -shared -fPIC (oracle/Makefile).
 * -ffp-contract=off keeps every product and sum separately rounded, as the
 * reference kernels are when compiled without contraction.  Without
 * -ffast-math gcc never reassociates, so the OpenMP build is bitwise equal to
 * a serial -O0 build.
 */
#include <math.h>
#include <stddef.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define REAL double
#define NAME(x) ffbo_d_##x
#define SIN_FN sin
#include "ffb_oracle_impl.h"
#undef REAL
#undef NAME
#undef SIN_FN

#define REAL float
#define NAME(x) ffbo_s_##x
#define SIN_FN sinf
#include "ffb_oracle_impl.h"
#undef REAL
#undef NAME
#undef SIN_FN

/* Thread count for the kernel loop (1 = the reference's one thread per image). */
void ffbo_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int ffbo_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
