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

/* ---- 16-bit scaled-deviation storage (the precision form; no reference
 * counterpart).  The GPU stores d_i = (f_i - w_i) * 2^s rounded to FP16 or
 * BF16 (round to nearest, ties to even) and decodes f_i = w_i + d_i * 2^-s in
 * FP32 when it loads.  The oracle runs the FP32 path and, after every step,
 * rounds each owned population through the same encode/decode, which gives
 * the values the GPU decodes at its next load. ---- */

/* IEEE binary16 from binary32, round to nearest even (subnormals, overflow to inf). */
unsigned short ffbo_float_to_half(float f)
{
    unsigned int x;
    memcpy(&x, &f, 4);
    unsigned int sign = (x >> 16) & 0x8000u;
    unsigned int ax = x & 0x7fffffffu;
    if (ax > 0x7f800000u) return (unsigned short)(sign | 0x7e00u);      /* NaN */
    if (ax >= 0x477ff000u) return (unsigned short)(sign | 0x7c00u);     /* >= 65520: inf */
    if (ax < 0x38800000u) {                                             /* below 2^-14: subnormal */
        if (ax < 0x33000000u) return (unsigned short)sign;              /* < 2^-25: rounds to 0 */
        unsigned int e = ax >> 23;                                      /* 102..112 */
        unsigned int m = (ax & 0x7fffffu) | 0x800000u;
        unsigned int shift = 126 - e;                                   /* 14..24 */
        unsigned int q = m >> shift;
        unsigned int rem = m & ((1u << shift) - 1u);
        unsigned int half = 1u << (shift - 1);
        if (rem > half || (rem == half && (q & 1u))) ++q;
        return (unsigned short)(sign | q);
    }
    unsigned int q = ((ax - 0x38000000u) >> 13);                        /* rebias exponent */
    unsigned int rem = ax & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (q & 1u))) ++q;
    return (unsigned short)(sign | q);
}

float ffbo_half_to_float(unsigned short h)
{
    unsigned int sign = ((unsigned int)h & 0x8000u) << 16;
    unsigned int e = ((unsigned int)h >> 10) & 0x1fu;
    unsigned int m = (unsigned int)h & 0x3ffu;
    unsigned int x;
    if (e == 0) {
        if (m == 0) {
            x = sign;
        } else {                                                        /* subnormal: normalise */
            int k = -1;
            do { ++k; m <<= 1; } while (!(m & 0x400u));
            x = sign | ((unsigned int)(112 - k) << 23) | ((m & 0x3ffu) << 13);
        }
    } else if (e == 31) {
        x = sign | 0x7f800000u | (m << 13);
    } else {
        x = sign | ((e + 112u) << 23) | (m << 13);
    }
    float f;
    memcpy(&f, &x, 4);
    return f;
}

unsigned short ffbo_float_to_bf16(float f)
{
    unsigned int u;
    memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (unsigned short)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

float ffbo_bf16_to_float(unsigned short b)
{
    unsigned int u = (unsigned int)b << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

/* Rounds the owned populations of a float halo array through the stored
 * form: kind 1 = FP16, 2 = BF16 deviations with scale 2^scale_log2.  The
 * float operations are the GPU's: (f - w) * 2^s on store, w + d * 2^-s on load. */
void ffbo_s_quantize(int kind, int scale_log2, int nx, int ny, float *f)
{
    const size_t pitch = (size_t)(nx + 2), plane = pitch * (size_t)(ny + 2);
    const float scale = ldexpf(1.0f, scale_log2), inv = ldexpf(1.0f, -scale_log2);
    for (int i = 1; i <= 9; ++i) {
        const float w = ffbo_s_w(i);
        for (int y = 1; y <= ny; ++y) {
            float *row = f + (size_t)(i - 1) * plane + (size_t)y * pitch;
            for (int x = 1; x <= nx; ++x) {
                const float d = (row[x] - w) * scale;
                const float back = kind == 1 ? ffbo_half_to_float(ffbo_float_to_half(d))
                                             : ffbo_bf16_to_float(ffbo_float_to_bf16(d));
                row[x] = w + back * inv;
            }
        }
    }
}

/* n_steps of the FP32 loop with the 16-bit storage rounding after every step
 * (kind as in ffbo_s_quantize).  Final state in f_a. */
void ffbo_s_run_quantized(int kind, int scale_log2, int mode, int nx, int ny, int n_steps,
                          float omega, float rho_0, float u_wall, float u_lid, float rho_in,
                          float rho_out, float *f_a, float *f_b, float *pf_state,
                          int write_macro_last, float *rho, float *u_x, float *u_y, float force)
{
    float *cur = f_a, *nxt = f_b;
    for (int step = 1; step <= n_steps; ++step) {
        int wm = write_macro_last && step == n_steps;
        ffbo_s_step(mode, nx, ny, omega, rho_0, u_wall, u_lid, rho_in, rho_out, cur, nxt,
                    pf_state, wm, rho, u_x, u_y, force);
        ffbo_s_quantize(kind, scale_log2, nx, ny, nxt);
        float *t = cur;
        cur = nxt;
        nxt = t;
    }
    if (cur != f_a) memcpy(f_a, cur, sizeof(float) * (size_t)(nx + 2) * (size_t)(ny + 2) * 9);
}

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
