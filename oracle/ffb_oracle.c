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

/* ---- 16-bit conversions of the deviation forms (below). ---- */

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

/* ---- Deviation forms (the precision form; no reference counterpart) ----
 * The GPU stores D_i = (f_i - w_i) * 2^s -- as FP16 or BF16 (kind 1, 2;
 * round to nearest even) or unrounded FP32 (kind 0) -- and works on D
 * throughout a step: the halo rules act on D (copies and bounce-back are the
 * same, since w_i = w_opp(i); the moving-wall and pressure increments are
 * scaled by 2^s), and the collision runs in deviation space with the operation
 * order below.  This restates that algorithm on the halo arrays of the
 * reference form; the state array holds D (float).  Kind 3, the 1-3-12 form,
 * works on D * 2^-120 (scale_log2 = s - 120 here), whose small values are
 * FP32 subnormals: the oracle is built without flush-to-zero, as the GPU
 * kernels run with FP32 denormals. */

/* 1-3-12 form (kind 3): the stored 16 bits are the sign, a 3-bit exponent and
 * a 12-bit significand of the FP32 value v = D * 2^-120 the step works on
 * (the oracle's state holds v): FP32 bits 11..25 and 31, i.e. magnitudes
 * k * 2^-138 (k < 2^12, subnormal) and (1 + k 2^-12) 2^(e-127), e = 1..7.
 * Encoding clamps |v| to the largest magnitude (bits 0x03fff800), then adds
 * `add` to the FP32 bits and cuts the low 11 (the sign bit is untouched: the
 * clamp leaves room for the carry).  add = 0x3ff + (bit 11) rounds to nearest
 * even; (r << 4) | 8 with 7 random bits r rounds stochastically. */
static unsigned short q_e3m12_bits(float v, unsigned int add)
{
    const unsigned int max_bits = 0x03fff800u;
    float m;
    memcpy(&m, &max_bits, 4);
    if (v > m) v = m;
    if (v < -m) v = -m;
    unsigned int u;
    memcpy(&u, &v, 4);
    if (add == 0xffffffffu) add = 0x3ffu + ((u >> 11) & 1u);
    u += add;
    return (unsigned short)(((u >> 11) & 0x7fffu) | ((u >> 16) & 0x8000u));
}

float ffbo_e3m12_to_float(unsigned short h)
{
    unsigned int u = ((unsigned int)(h & 0x7fffu) << 11) | ((unsigned int)(h & 0x8000u) << 16);
    float f;
    memcpy(&f, &u, 4);
    return f;
}

/* Stored bits of v: nearest even (sr = 0) or stochastic with random bits r. */
unsigned short ffbo_float_to_e3m12(float v, int sr, unsigned int r)
{
    return q_e3m12_bits(v, sr ? (((r & 127u) << 4) | 8u) : 0xffffffffu);
}

/* Exponent offset of direction d (0-based) in the 1-3-12 form: the stored
 * value is v * 2^off (rest -1, axis 0, diagonal +1), read back as
 * stored * 2^-off; FP32 multiplications, exact except on subnormal products. */
static int q_off(int kind, int d) { return kind == 3 ? (d == 0 ? -1 : (d <= 4 ? 0 : 1)) : 0; }

static float q_e3m12_round(float v, int d, unsigned int add)
{
    const int o = q_off(3, d);
    const float x = o == 0 ? v : v * ldexpf(1.0f, o);
    const float y = ffbo_e3m12_to_float(q_e3m12_bits(x, add));
    return o == 0 ? y : y * ldexpf(1.0f, -o);
}

static float q_round(int kind, float v, int d)
{
    if (kind == 3) return q_e3m12_round(v, d, 0xffffffffu);
    if (kind == 1) return ffbo_half_to_float(ffbo_float_to_half(v));
    if (kind == 2) return ffbo_bf16_to_float(ffbo_float_to_bf16(v));
    return v;
}

/* Stochastic rounding of the 16-bit forms (the GPU's out_sr): 7 random bits
 * per stored value from the MurmurHash3 32-bit finaliser (Appleby, public
 * domain) of cell ^ (step * 0x9e3779b9), cell = y * nx + x over the owned cells
 * (0-based), step = index of the step being computed (1, 2, ...).  The GPU
 * rounds with gfx950's v_cvt_sr_f16_f32 / v_cvt_sr_bf16_f32 and the random
 * operand R = (r << 25) | (1 << 24); restated here as measured on MI355X for
 * 2^20 random (value, R) pairs: BF16 = (bits + (R >> 16)) >> 16; FP16 normal
 * = magnitude bits + (R >> 19) with the low 13 bits cleared; FP16 below 2^-14
 * (k = 126 - exponent discarded bits of the significand) = (significand +
 * R aligned to the top of the k bits) >> k. */
static unsigned int q_fmix32(unsigned int h)
{
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}

/* Random bits of population d (0-based): d = 0, 1, 2 from the first word,
 * d = 5, 6 from the second, 7 bits each; the opposite directions 3, 4, 7, 8
 * take 127 - r of their partner 1, 2, 5, 6 (antithetic pairs). */
static unsigned int q_pop_bits(unsigned int h0, unsigned int h1, int d)
{
    static const int partner[9] = {0, 1, 2, 1, 2, 5, 6, 5, 6};
    if (partner[d] != d) return 127u - q_pop_bits(h0, h1, partner[d]);
    if (d < 4) return (h0 >> (7 * d)) & 127u;
    return (h1 >> (7 * (d - 4))) & 127u;
}

static float q_round_sr(int kind, float v, unsigned int r, int d)
{
    if (kind == 3) return q_e3m12_round(v, d, (r << 4) | 8u);
    unsigned int u;
    memcpy(&u, &v, 4);
    if (kind == 2) { /* BF16: add r to the 16 bits below the mantissa, truncate */
        unsigned int b = (u + ((r << 9) | 256u)) & 0xffff0000u;
        float o;
        memcpy(&o, &b, 4);
        return o;
    }
    /* FP16 */
    unsigned int sgn = u & 0x80000000u, a = u & 0x7fffffffu;
    unsigned int R = (r << 25) | (1u << 24);
    if (a >= 0x38800000u) {
        unsigned int b = sgn | ((a + (R >> 19)) & ~0x1fffu);
        float q;
        memcpy(&q, &b, 4);
        return ffbo_half_to_float(ffbo_float_to_half(q)); /* exact (or overflow to inf) */
    }
    int e = (int)(a >> 23);
    int k = 126 - e;                                     /* discarded bits, >= 14 */
    unsigned long long m = (a & 0x7fffffu) | (e ? 0x800000u : 0u);
    unsigned int q;
    if (k <= 32) q = (unsigned int)((m + (R >> (32 - k))) >> k);
    else if (k < 56) q = (unsigned int)((m + ((unsigned long long)R << (k - 32))) >> k);
    else q = 0;
    return ffbo_half_to_float((unsigned short)((sgn >> 16) | q));
}

/* The stochastic rounding of one value with random bits r (0..127): the
 * stored value decoded to float (kind 1 FP16, 2 BF16, 3 the 1-3-12 form). */
float ffbo_round_sr(int kind, float v, unsigned int r) { return q_round_sr(kind, v, r & 127u, 1); }

/* Stored form of the populations of a float halo array (initial state):
 * D = round((f - w) * 2^s), every entry. */
void ffbo_q_encode(int kind, int scale_log2, int nx, int ny, const float *f, float *D)
{
    const size_t plane = (size_t)(nx + 2) * (size_t)(ny + 2);
    const float scale = ldexpf(1.0f, scale_log2);
    for (int i = 1; i <= 9; ++i) {
        const float w = ffbo_s_w(i);
        for (size_t k = 0; k < plane; ++k) {
            const size_t j = (size_t)(i - 1) * plane + k;
            D[j] = q_round(kind, (f[j] - w) * scale, i - 1);
        }
    }
}

/* Populations the GPU returns for stored D: f = w + D * 2^-s in FP32. */
void ffbo_q_decode(int scale_log2, int nx, int ny, const float *D, float *f)
{
    const size_t plane = (size_t)(nx + 2) * (size_t)(ny + 2);
    const float inv = ldexpf(1.0f, -scale_log2);
    for (int i = 1; i <= 9; ++i) {
        const float w = ffbo_s_w(i);
        for (size_t k = 0; k < plane; ++k) {
            const size_t j = (size_t)(i - 1) * plane + k;
            f[j] = w + D[j] * inv;
        }
    }
}

#define FQ(f, x, y, i) (f)[(size_t)((i) - 1) * plane + (size_t)(y) * (size_t)pitch + (size_t)(x)]

/* The pressure increment of pressure_periodic_distribution
 * (modes/poiseuille_flow.f90:189-194) times 2^s, added to a stored D. */
static float q_ppd(int i, float D_src, float rho_src, float ux_src, float uy_src, float rho_b, float scale)
{
    static const int cx[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
    static const int cy[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
    float usq = ux_src * ux_src + uy_src * uy_src;
    float cu = (float)cx[i - 1] * ux_src + (float)cy[i - 1] * uy_src;
    float pf = ((1.0f + 3.0f * cu) + 4.5f * cu * cu) - 1.5f * usq;
    return D_src + (ffbo_s_w(i) * (rho_b - rho_src) * pf) * scale;
}

/* prepare_poiseuille_flow_halos_PF (modes/poiseuille_flow.f90:9-89) on D. */
static void q_halos_pf(int nx, int ny, float rho_in, float rho_out, const float *ml, const float *mr,
                       float scale, float *f)
{
    const long pitch = nx + 2;
    const size_t plane = (size_t)(nx + 2) * (size_t)(ny + 2);
#define QL(y, k) ml[(size_t)((k) - 1) * (size_t)ny + (size_t)((y) - 1)]
#define QR(y, k) mr[(size_t)((k) - 1) * (size_t)ny + (size_t)((y) - 1)]
    ffbo_s_bb_bottom(nx, ny, f);
    ffbo_s_bb_top(nx, ny, 0, 0, 0, f);
    for (int y = 1; y <= ny; ++y) {
        FQ(f, 0, y, 2) = q_ppd(2, FQ(f, 0, y, 2), QL(y, 1), QL(y, 2), QL(y, 3), rho_in, scale);
        FQ(f, 0, y, 6) = q_ppd(6, FQ(f, 0, y, 6), QL(y, 1), QL(y, 2), QL(y, 3), rho_in, scale);
        FQ(f, 0, y, 9) = q_ppd(9, FQ(f, 0, y, 9), QL(y, 1), QL(y, 2), QL(y, 3), rho_in, scale);
    }
    FQ(f, 0, 0, 6) = q_ppd(6, FQ(f, 0, 0, 6), QL(1, 1), QL(1, 2), QL(1, 3), rho_in, scale);
    FQ(f, 0, ny + 1, 9) = q_ppd(9, FQ(f, 0, ny + 1, 9), QL(ny, 1), QL(ny, 2), QL(ny, 3), rho_in, scale);
    for (int y = 1; y <= ny; ++y) {
        FQ(f, nx + 1, y, 4) = q_ppd(4, FQ(f, nx + 1, y, 4), QR(y, 1), QR(y, 2), QR(y, 3), rho_out, scale);
        FQ(f, nx + 1, y, 7) = q_ppd(7, FQ(f, nx + 1, y, 7), QR(y, 1), QR(y, 2), QR(y, 3), rho_out, scale);
        FQ(f, nx + 1, y, 8) = q_ppd(8, FQ(f, nx + 1, y, 8), QR(y, 1), QR(y, 2), QR(y, 3), rho_out, scale);
    }
    FQ(f, nx + 1, 0, 7) = q_ppd(7, FQ(f, nx + 1, 0, 7), QR(1, 1), QR(1, 2), QR(1, 3), rho_out, scale);
    FQ(f, nx + 1, ny + 1, 8) = q_ppd(8, FQ(f, nx + 1, ny + 1, 8), QR(ny, 1), QR(ny, 2), QR(ny, 3), rho_out, scale);
#undef QL
#undef QR
}

/* Pulled D of cell (x, y), v[d] = D_{d+1}(x - c). */
static void q_pull(const float *f, long pitch, size_t plane, int x, int y, float v[9])
{
    v[0] = FQ(f, x, y, 1);
    v[1] = FQ(f, x - 1, y, 2);
    v[2] = FQ(f, x, y - 1, 3);
    v[3] = FQ(f, x + 1, y, 4);
    v[4] = FQ(f, x, y + 1, 5);
    v[5] = FQ(f, x - 1, y - 1, 6);
    v[6] = FQ(f, x + 1, y - 1, 7);
    v[7] = FQ(f, x + 1, y + 1, 8);
    v[8] = FQ(f, x - 1, y + 1, 9);
}

/* update_poiseuille_flow_macro_strips_PF (modes/poiseuille_flow.f90:92-167)
 * on D: rho = 1 + 2^-s Sum D, u = 2^-s Sum c D / rho, with the pair sums of
 * the deviation collision. */
static void q_macro_strip(int nx, int ny, int x, const float *f, float inv, float *strip)
{
    const long pitch = nx + 2;
    const size_t plane = (size_t)(nx + 2) * (size_t)(ny + 2);
    for (int y = 1; y <= ny; ++y) {
        float v[9];
        q_pull(f, pitch, plane, x, y, v);
        float a13 = v[1] + v[3], a24 = v[2] + v[4], a57 = v[5] + v[7], a68 = v[6] + v[8];
        float b13 = v[1] - v[3], b24 = v[2] - v[4], b57 = v[5] - v[7], b68 = v[6] - v[8];
        float rho = 1.0f + (((v[0] + a13) + a24) + (a57 + a68)) * inv;
        strip[(size_t)(y - 1)] = rho;
        strip[(size_t)ny + (size_t)(y - 1)] = (((b13 + b57) - b68) * inv) / rho;
        strip[2 * (size_t)ny + (size_t)(y - 1)] = (((b24 + b57) + b68) * inv) / rho;
    }
}

/* Collision in deviation space (the factor omega w_i 2^s carried as a float
 * pair kw + kw_lo from the exact weight).  With Sum w = 1 and Sum w c = 0:
 * rho - 1 = 2^-s Sum D, rho u = 2^-s Sum c D (+ F/2), and
 * 2^s (f_eq_i - w_i) = 2^s w_i ((rho - 1) + rho (3 c.u + 4.5 (c.u)^2 - 1.5 u^2)),
 * so D'_i = (1 - omega) D_i + (omega w_i 2^s) inner_i, computed pairwise for
 * opposite directions; body force: plus the Guo source
 * 2^s w_i (1 - omega/2) [3 (c_i - u) + 9 (c_i.u) c_i] . F. */
static void q_kernel(int kind, int stochastic, unsigned int step, int guo, int nx, int ny, int write_macro,
                     float omega, float force, float scale, float inv, const float *f, float *f_next,
                     float *rho_o, float *ux_o, float *uy_o)
{
    const long pitch = nx + 2;
    const size_t plane = (size_t)(nx + 2) * (size_t)(ny + 2);
    float kw[9], kw_lo[9], kf[9];
    const float om1 = 1.0f - omega;
    const float pre = (1.0f - 0.5f * omega) * force;
    const float half_force = 0.5f * force;
    for (int d = 0; d < 9; ++d) {
        /* omega w_d 2^s with the exact weight, as a float pair from a double product */
        const double wd = d == 0 ? 4.0 / 9.0 : (d <= 4 ? 1.0 / 9.0 : 1.0 / 36.0);
        const double kd = (double)omega * wd * (double)scale;
        kw[d] = (float)kd;
        kw_lo[d] = (float)(kd - (double)kw[d]);
        kf[d] = (ffbo_s_w(d + 1) * pre) * scale;
    }
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int y = 1; y <= ny; ++y) {
        for (int x = 1; x <= nx; ++x) {
            float D[9];
            q_pull(f, pitch, plane, x, y, D);
            float a13 = D[1] + D[3], a24 = D[2] + D[4], a57 = D[5] + D[7], a68 = D[6] + D[8];
            float b13 = D[1] - D[3], b24 = D[2] - D[4], b57 = D[5] - D[7], b68 = D[6] - D[8];
            float m0 = ((D[0] + a13) + a24) + (a57 + a68);
            float jx = (b13 + b57) - b68;
            float jy = (b24 + b57) + b68;
            float drho = m0 * inv;
            float rho = 1.0f + drho;
            float r = 1.0f / rho;
            float ux = (guo ? fmaf(jx, inv, half_force) : jx * inv) * r;
            float uy = (jy * inv) * r;
            if (write_macro) {
                size_t c = (size_t)(y - 1) * (size_t)nx + (size_t)(x - 1);
                rho_o[c] = rho;
                ux_o[c] = ux;
                uy_o[c] = uy;
            }
            float usq = fmaf(ux, ux, uy * uy);
            float base = fmaf(-(1.5f * usq), rho, drho);
            float hr = 0.5f * rho;
            float h1 = 3.0f * ux, h2 = 3.0f * uy;
            float h5 = h1 + h2, h8 = h1 - h2;
            static const int pd[4] = {1, 2, 5, 8}, pe[4] = {3, 4, 7, 6};
            float hs[4] = {h1, h2, h5, h8};
            D[0] = fmaf(om1, D[0], fmaf(kw[0], base, kw_lo[0] * base));
            for (int k = 0; k < 4; ++k) {
                const int d = pd[k], e = pe[k];
                float p = fmaf(hr, hs[k] * hs[k], base);
                float q = rho * hs[k];
                const float pq = p + q, pm = p - q;
                D[d] = fmaf(om1, D[d], fmaf(kw[d], pq, kw_lo[d] * pq));
                D[e] = fmaf(om1, D[e], fmaf(kw[e], pm, kw_lo[e] * pm));
            }
            if (guo) {
                float g = 3.0f * ux;
                D[0] = fmaf(kf[0], -g, D[0]);
                D[2] = fmaf(kf[2], -g, D[2]);
                D[4] = fmaf(kf[4], -g, D[4]);
                D[1] = fmaf(kf[1], (3.0f * h1 + 3.0f) - g, D[1]);
                D[3] = fmaf(kf[3], (3.0f * h1 - 3.0f) - g, D[3]);
                D[5] = fmaf(kf[5], (3.0f * h5 + 3.0f) - g, D[5]);
                D[7] = fmaf(kf[7], (3.0f * h5 - 3.0f) - g, D[7]);
                D[8] = fmaf(kf[8], (3.0f * h8 + 3.0f) - g, D[8]);
                D[6] = fmaf(kf[6], (3.0f * h8 - 3.0f) - g, D[6]);
            }
            if (stochastic && kind >= 1) {
                unsigned long long cell = (unsigned long long)(y - 1) * (unsigned long long)nx + (unsigned)(x - 1);
                unsigned int c = (unsigned int)cell ^ (unsigned int)(cell >> 32);
                unsigned int h0 = q_fmix32(c ^ (step * 0x9e3779b9u));
                unsigned int h1 = q_fmix32(h0 + 0x632be5abu);
                for (int d = 0; d < 9; ++d) FQ(f_next, x, y, d + 1) = q_round_sr(kind, D[d], q_pop_bits(h0, h1, d), d);
            } else {
                for (int d = 0; d < 9; ++d) FQ(f_next, x, y, d + 1) = q_round(kind, D[d], d);
            }
        }
    }
}

/* One step of the deviation form: the driver order of ffbo_s_step (exchange,
 * PF pre-pass and strips, mode halo prep, kernel) on D. */
void ffbo_q_step(int kind, int scale_log2, int stochastic, unsigned int step, int mode, int nx, int ny,
                 float omega, float rho_0, float u_wall, float u_lid, float rho_in, float rho_out, float *f,
                 float *f_next, float *pf_state, int write_macro, float *rho, float *u_x, float *u_y, float force)
{
    const float scale = ldexpf(1.0f, scale_log2), inv = ldexpf(1.0f, -scale_log2);
    ffbo_s_exchange_1domain(mode, nx, ny, f);
    if (mode == 3) {
        size_t s = (size_t)ny * 3;
        float *recv_l = pf_state, *recv_r = pf_state + s;
        float *win_l = pf_state + 2 * s, *win_r = pf_state + 3 * s;
        q_halos_pf(nx, ny, rho_in, rho_out, recv_l, recv_r, scale, f);
        q_macro_strip(nx, ny, 1, f, inv, win_l);
        q_macro_strip(nx, ny, nx, f, inv, win_r);
        memcpy(recv_l, win_r, s * sizeof(float));
        memcpy(recv_r, win_l, s * sizeof(float));
    }
    switch (mode) {
    case 1: ffbo_s_halos_sw(nx, ny, f); break;
    /* ((6 w_6) rho_0)(u 2^s) = ((6 w_6) rho_0 u) 2^s exactly: scaled moving-wall term */
    case 2: ffbo_s_halos_cf(nx, ny, rho_0, u_wall * scale, f); break;
    case 3: q_halos_pf(nx, ny, rho_in, rho_out, pf_state, pf_state + (size_t)ny * 3, scale, f); break;
    case 5:
        ffbo_s_bb_bottom(nx, ny, f);
        ffbo_s_bb_top(nx, ny, 0, 0, 0, f);
        break;
    default: ffbo_s_halos_sl(nx, ny, rho_0, u_lid * scale, f); break;
    }
    q_kernel(kind, stochastic, step, mode == 5, nx, ny, write_macro, omega, force, scale, inv, f, f_next, rho,
             u_x, u_y);
}

/* n_steps of the deviation form with ping-pong buffers; final D in f_a.
 * steps_done = steps taken before (the stochastic-rounding step index). */
void ffbo_q_run(int kind, int scale_log2, int stochastic, int steps_done, int mode, int nx, int ny,
                int n_steps, float omega, float rho_0, float u_wall, float u_lid, float rho_in,
                float rho_out, float *f_a, float *f_b, float *pf_state, int write_macro_last, float *rho,
                float *u_x, float *u_y, float force)
{
    float *cur = f_a, *nxt = f_b;
    for (int step = 1; step <= n_steps; ++step) {
        int wm = write_macro_last && step == n_steps;
        ffbo_q_step(kind, scale_log2, stochastic, (unsigned int)(steps_done + step), mode, nx, ny, omega, rho_0,
                    u_wall, u_lid, rho_in, rho_out, cur, nxt, pf_state, wm, rho, u_x, u_y, force);
        float *t = cur;
        cur = nxt;
        nxt = t;
    }
    if (cur != f_a) memcpy(f_a, cur, sizeof(float) * (size_t)(nx + 2) * (size_t)(ny + 2) * 9);
}

#undef FQ

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
