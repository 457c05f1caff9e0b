// d2q9_kernels.h -- device code of the fused D2Q9 pull-stream-collide BGK step
// for gfx950 (MI355X), shared by the single-domain library (ffb_api.hip) and
// the batched launcher.
//
// Data layout in HBM (one domain): f_d(x, y) at d*plane + y*pitch + x, x
// fastest, no halo; two such sets ping-pong between steps.  Row-blocked
// (default): plane = n_x rounded up to 64 elements and pitch = 9*plane, so
// row y of all nine directions is one contiguous block.  Planes: pitch = n_x
// rounded up to 64 elements, plane = pitch*n_y (nine separate planes).
// Direction index d = 0..8 is the reference's i = 1..9 minus one
// (source/settings.f90:96-113): 0 rest, 1 (1,0), 2 (0,1), 3 (-1,0), 4 (0,-1),
// 5 (1,1), 6 (-1,1), 7 (-1,-1), 8 (1,-1).
//
// The reference writes boundary values into a one-cell halo before every
// step (source/layouts/xyd/modes/*.f90, exchange.f90:213-291) and then pulls
// f_i(x - c_i).  Each halo entry is read by exactly one (cell, direction)
// pair, so here the value that halo entry would hold is computed where it is
// read (pull_edge below); the arithmetic is the reference's, in its order.
//
// Floating-point: every expression keeps the reference's left-to-right order
// and the file is compiled without contraction, so FP64/FP32 results are
// bitwise those of the reference kernels compiled without FMA contraction.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <stdint.h>

#include <type_traits>

#pragma clang fp contract(off)

namespace ffb {

// SW, CF, PF, SL: the reference's four sim modes (source/settings.f90:56-59).
// BF: body-force channel (periodic in x, halfway bounce-back walls in y, a
// uniform body force along x with Guo forcing); no reference counterpart.
enum Mode { SW = 1, CF = 2, PF = 3, SL = 4, BF = 5 };

__host__ __device__ constexpr int cx_of(int d) {
    return (d == 1 || d == 5 || d == 8) ? 1 : ((d == 3 || d == 6 || d == 7) ? -1 : 0);
}
__host__ __device__ constexpr int cy_of(int d) {
    return (d == 2 || d == 5 || d == 6) ? 1 : ((d == 4 || d == 7 || d == 8) ? -1 : 0);
}
__host__ __device__ constexpr int opp_of(int d) {
    return d == 0 ? 0 : (d <= 4 ? ((d + 1) % 4) + 1 : ((d - 3) % 4) + 5);
}
static_assert(opp_of(1) == 3 && opp_of(2) == 4 && opp_of(3) == 1 && opp_of(4) == 2, "opp axis");
static_assert(opp_of(5) == 7 && opp_of(6) == 8 && opp_of(7) == 5 && opp_of(8) == 6, "opp diag");

// ---------------------------------------------------------------------------
// storage of one population value.  FP64 / FP32 store f itself and the step
// works on f.  The deviation forms store D = (f - w) * 2^s (FP16, BF16, FP32
// or the 1-3-12 layout); the step reads D as FP32, works on D throughout
// (boundary rules and collision in deviation space, collide_dev) and rounds
// the new D to the storage type.  `in` gives the value
// the step works on, `out` the stored form of it, `from_f` turns a population
// f into that value (initial state).
// ---------------------------------------------------------------------------
template <typename S> struct Store;

// Stochastic rounding of the 16-bit forms.  Each stored value takes 7 random
// bits from a hash of (cell, step): the MurmurHash3 32-bit finaliser
// (Appleby, public domain) of cell ^ (step * 0x9e3779b9) gives 32 bits, a
// second finaliser call 32 more; population d = 0, 1, 2 uses bits 7d.. of the
// first word, d = 5, 6 bits 7(d-4).. of the second, and the opposite
// directions 3, 4, 7, 8 take the antithetic value 127 - r of their partner
// 1, 2, 5, 6 (still uniform, so every value stays unbiased; rounding errors
// of opposite populations then tend to cancel, measured 13-14 % less velocity
// error against FP64 for the FP16 and 1-3-12 forms on a 1024^2 cavity).  The
// same (cell, step) always gives the same bits, so runs are reproducible and
// the oracle restates them exactly.
__host__ __device__ __forceinline__ unsigned fmix32(unsigned h) {
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}
struct CellBits {
    unsigned h0, h1;
};
__device__ __forceinline__ CellBits cell_bits(unsigned long long cell, unsigned step) {
    const unsigned c = (unsigned)cell ^ (unsigned)(cell >> 32);
    const unsigned h0 = fmix32(c ^ (step * 0x9e3779b9u));
    return {h0, fmix32(h0 + 0x632be5abu)};
}
__device__ __forceinline__ unsigned pop_bits(const CellBits &b, int d) {
    const bool anti = (d == 3) | (d == 4) | (d == 7) | (d == 8);
    const int e = anti ? opp_of(d) : d;  // 0, 1, 2, 5 or 6
    const unsigned r = e < 4 ? (b.h0 >> (7 * e)) & 127u : (b.h1 >> (7 * (e - 4))) & 127u;
    return anti ? r ^ 127u : r;  // 127 - r
}

template <> struct Store<double> {
    using C = double;  // compute type
    using R = double;  // raw stored type
    static constexpr bool dev = false;
    static __device__ __forceinline__ C in(R r) { return r; }
    static __device__ __forceinline__ R out(C v) { return v; }
    static __device__ __forceinline__ C from_f(C f, C, C) { return f; }
};
template <> struct Store<float> {
    using C = float;
    using R = float;
    static constexpr bool dev = false;
    static __device__ __forceinline__ C in(R r) { return r; }
    static __device__ __forceinline__ R out(C v) { return v; }
    static __device__ __forceinline__ C from_f(C f, C, C) { return f; }
};
// FP32 deviations: D stored unrounded in FP32 (FFB_PREC_FP32_DEV).
struct f32dev {};
template <> struct Store<f32dev> {
    using C = float;
    using R = float;
    static constexpr bool dev = true;
    static __device__ __forceinline__ C in(R r) { return r; }
    static __device__ __forceinline__ R out(C v) { return v; }
    static __device__ __forceinline__ C from_f(C f, C w, C scale) { return (f - w) * scale; }
};
template <> struct Store<__half> {
    using C = float;
    using R = unsigned short;
    static constexpr bool dev = true;
    static __device__ __forceinline__ C in(R r) { return __half2float(__ushort_as_half(r)); }
    static __device__ __forceinline__ R out(C v) {
        // v is the FP32 result of the collision; the empty asm makes it an
        // opaque register value so that the conversion rounds that FP32 value.
        // Without it the compiler fuses the collision's last FMA and the
        // conversion into v_fma_mixlo_f16, which rounds the exact result to
        // FP16 once: a different value in about 1 of 2^14 conversions.
        asm("" : "+v"(v));
        return __half_as_ushort(__float2half_rn(v));  // round to nearest even
    }
    // Stochastic rounding with 7 random bits r: gfx950's v_cvt_sr_f16_f32 with
    // the random operand (r << 25) | (1 << 24), i.e. r and a half in the top
    // bits of the bits the conversion discards (13 for FP16 normals, more
    // below 2^-14): the magnitude bits plus those are truncated.  Up with
    // probability (discarded fraction + 1/256) rounded to 1/128.
    static __device__ __forceinline__ R out_sr(C v, unsigned r) {
        typedef _Float16 h2 __attribute__((ext_vector_type(2)));
        const h2 h = __builtin_amdgcn_cvt_sr_f16_f32(h2{}, v, (r << 25) | (1u << 24), false);
        return __builtin_bit_cast(R, h[0]);
    }
    static __device__ __forceinline__ C from_f(C f, C w, C scale) { return (f - w) * scale; }
};
template <> struct Store<__hip_bfloat16> {
    using C = float;
    using R = unsigned short;
    static constexpr bool dev = true;
    static __device__ __forceinline__ C in(R r) { return __uint_as_float((unsigned)r << 16); }
    static __device__ __forceinline__ R out(C v) {
        // v_cvt_pk_bf16_f32 (gfx950): round to nearest even, as the FP32 value
        // v itself (the empty asm keeps the conversion on that value)
        asm("" : "+v"(v));
        return __builtin_bit_cast(R, (__bf16)v);
    }
    // Stochastic rounding with 7 random bits r: gfx950's v_cvt_sr_bf16_f32 with
    // the random operand (r << 25) | (1 << 24), which adds (r << 9) | 256 to the
    // 16 bits below the BF16 mantissa and truncates them.
    static __device__ __forceinline__ R out_sr(C v, unsigned r) {
        typedef __bf16 b2 __attribute__((ext_vector_type(2)));
        const b2 h = __builtin_amdgcn_cvt_sr_bf16_f32(b2{}, v, (r << 25) | (1u << 24), false);
        return __builtin_bit_cast(R, h[0]);
    }
    static __device__ __forceinline__ C from_f(C f, C w, C scale) { return (f - w) * scale; }
};

// 16-bit deviations in a 1-3-12 float layout (FFB_PREC_E3M12_DEV): sign bit 15,
// a 3-bit exponent and a 12-bit significand with subnormals, i.e. the FP32
// bit pattern of (D * 2^-120) with the exponent field limited to 0..7 and the
// low 11 significand bits dropped.  D * 2^-120 spans [2^-138, 2^-119): seven
// binades of 12-bit precision below the maximum (2 - 2^-12) * 2^-120 and a
// uniform step of 2^-138 under them; scaled by 2^(120-s) (and by the
// direction's offset 2^off, store_off) the stored deviation f - w covers
// |f - w| < 2^(1-s-off) with 12-bit precision down to 2^(-5-s-off).
// The step works on the raw FP32 value D * 2^-120 (the deviation form with
// scale 2^(s-120), so values under 2^-126 are FP32 subnormals): decoding is a
// shift and a mask, encoding clamps to the largest magnitude and drops 11
// bits (rounded).  Two more significand bits than FP16 where deviations are
// small and bounded, as in a flow with known largest speed.
struct e3m12 {};
template <> struct Store<e3m12> {
    using C = float;
    using R = unsigned short;
    static constexpr bool dev = true;
    static constexpr unsigned kMaxBits = 0x03fff800u;  // (2 - 2^-12) * 2^-120
    static __device__ __forceinline__ C in(R r) {
        // sign-extended: bits 11..25 get the magnitude, bit 31 the sign
        return __uint_as_float(((unsigned)(int)(short)r << 11) & 0x83ffffffu);
    }
    static __device__ __forceinline__ unsigned clamped(C v) {
        const float m = __uint_as_float(kMaxBits);
        return __float_as_uint(__builtin_amdgcn_fmed3f(v, -m, m));
    }
    // magnitude bits 11..25 -> 0..14, sign bit 31 -> 15
    static __device__ __forceinline__ R pack(unsigned u) {
        return (R)(((u >> 11) & 0x7fffu) | ((u >> 16) & 0x8000u));
    }
    // to nearest, ties to even (on the magnitude bits; the clamp leaves room
    // for the carry)
    static __device__ __forceinline__ R out(C v) {
        const unsigned u = clamped(v);
        return pack(u + 0x3ffu + ((u >> 11) & 1u));
    }
    // Stochastic rounding with 7 random bits r: (r << 4) | 8 is added to the
    // 11 dropped bits before they are cut, as the FP16/BF16 hardware
    // conversions add r and a half in the top discarded bits.
    static __device__ __forceinline__ R out_sr(C v, unsigned r) {
        return pack(clamped(v) + ((r << 4) | 8u));
    }
    static __device__ __forceinline__ C from_f(C f, C w, C scale) { return (f - w) * scale; }
};

// Exponent offset of direction d in the stored form: direction d stores the
// value the step works on times 2^off.  Only the 1-3-12 form, whose range is
// eight binades, uses offsets, one per direction class: the largest |f - w| of
// a class is its weight times (rho - 1) plus, for the moving directions,
// 3 |c.u| and smaller terms.  Rest (w = 4/9, no velocity term, but the density
// peak at a sliding lid's corners, rho - 1 ~ 0.14 for u = 0.1): -1, one binade
// more range.  Axis (1/9): 0.  Diagonal (1/36): +1, one bit more precision.
// Scaling by a power of two is exact except where the product is an FP32
// subnormal (then rounded as FP32 arithmetic rounds it).
template <typename S>
__host__ __device__ constexpr int store_off(int d) {
    return std::is_same_v<S, e3m12> ? (d == 0 ? -1 : (d <= 4 ? 0 : 1)) : 0;
}
__host__ __device__ constexpr float pow2f(int e) { return e >= 0 ? (float)(1u << e) : 1.0f / (float)(1u << -e); }
template <typename S>
__device__ __forceinline__ typename Store<S>::C pop_in(typename Store<S>::R r, int d) {
    const typename Store<S>::C v = Store<S>::in(r);
    return store_off<S>(d) == 0 ? v : v * pow2f(-store_off<S>(d));
}
template <typename S>
__device__ __forceinline__ typename Store<S>::R pop_out(typename Store<S>::C v, int d) {
    return Store<S>::out(store_off<S>(d) == 0 ? v : v * pow2f(store_off<S>(d)));
}
template <typename S>
__device__ __forceinline__ typename Store<S>::R pop_out_sr(typename Store<S>::C v, unsigned r, int d) {
    return Store<S>::out_sr(store_off<S>(d) == 0 ? v : v * pow2f(store_off<S>(d)), r);
}

template <typename S>
__device__ __forceinline__ typename Store<S>::C ld_pop(const typename Store<S>::R *p, size_t k, int d) {
    return pop_in<S>(p[k], d);
}
template <typename S>
__device__ __forceinline__ void st_pop(typename Store<S>::R *p, size_t k, typename Store<S>::C v, int d) {
    __builtin_nontemporal_store(pop_out<S>(v, d), p + k);
}

// Parameters of one step over one domain.
template <typename C> struct StepArgs {
    const void *src;       // populations read (pitch/plane layout above)
    void *dst;             // populations written
    int nx, ny, pitch;
    long long plane;
    C omega;
    C corr8, corr9;        // moving-wall terms ((6*w_6)*rho_0)*u_wall, couette_flow.f90:118-119
                           // (times 2^s in the deviation forms)
    C rho_in, rho_out;     // Poiseuille boundary densities
    const C *m_old;        // Poiseuille strips of the previous step [side][rho,ux,uy][ny]
    C *m_new;              // strips of this step (written by the strip pass, read by the step)
    C *rho_o, *ux_o, *uy_o;// macro fields (nx*ny, x fastest), written when write_macro
    int write_macro;
    C w[9];                // lattice weights in C precision
    C scale, inv_scale;    // 2^s and 2^-s of the deviation forms
    C force;               // BF: body-force density along x (lattice units)
    // deviation-space collision of the deviation forms (collide_dev):
    C om1;                 // 1 - omega
    C kw[9], kw_lo[9];     // omega * w_d * 2^s with the exact weights 4/9, 1/9, 1/36 as a
                           // pair of floats (kw + kw_lo, from a double product)
    C kf[9];               // BF: (w_d * ((1 - omega/2) * force)) * 2^s
    C half_force;          // BF: force / 2
    int n_tiles_x;         // tiles per row
    int stochastic;        // 16-bit forms: 1 = stochastic rounding of the stored values
    unsigned step_seed;    // step index for the stochastic-rounding hash
    int n_blocks;          // grid size (for the XCD-aware tile order)
    int xcd_group;         // G > 0: G consecutive tiles per XCD (tile_of_block)
    int check_rho;         // FFB_DENSITY_CHECKS: flag densities rho <= 0 (or NaN) in *status
    unsigned *status;      // bit 0: rho <= 0 met, bit 1: rho NaN met
    int march_rows;        // march2_kernel: output rows per workgroup (one chunk of a strip)
    int march_inner;       // march2_kernel: 1 = strips away from the x edges pull without edge rules
    int xcd_rows;          // step2_kernel: H > 1: blocks of xcd_group x H tiles per XCD
    int n_tiles_y;         // step2_kernel: tile rows
};

// FFB_DENSITY_CHECKS (simulation.f90:229-233, :133-137,
// modes/poiseuille_flow.f90:156-160): the reference stops at a density
// rho <= 0 before dividing by it; here the cell sets bit 0 of the status word
// (bit 1 for a NaN density, which the reference's test lets pass) and the
// host reports it (ffb_sync).  A branch on a kernel argument when off.  `act`
// false: a lane that computes a cell only to keep its wave's stores
// unconditional (march2_kernel), whose result is dropped.
template <typename C>
__device__ __forceinline__ void density_check(const StepArgs<C> &a, C rho, bool act = true) {
    if (a.check_rho && act) {
        if (rho <= (C)0) atomicOr(a.status, 1u);
        else if (rho != rho) atomicOr(a.status, 2u);
    }
}

// Workgroups are dispatched round-robin over the 8 XCDs of an MI355X
// (workgroup p to XCD p mod 8), each XCD with its own L2.  With G > 0 the
// workgroups p of one chunk of 8G map to tiles so that XCD j takes the G
// consecutive tiles [jG, (j+1)G) of the chunk: x-neighbouring tiles, which
// share the cache lines at their common edge, then hit one L2, while the eight
// XCDs still work on neighbouring tiles.  The last, partial chunk keeps the
// plain order.  G = 0: tile = workgroup index.
__device__ __forceinline__ int tile_of_block(int p, int n, int G) {
    if (G <= 0) return p;
    const int chunk = 8 * G;
    const int c = p / chunk;
    if ((c + 1) * chunk > n) return p;
    const int r = p - c * chunk;
    return c * chunk + (r & 7) * G + (r >> 3);
}

template <typename S, typename C>
__device__ __forceinline__ C LD(const StepArgs<C> &a, const typename Store<S>::R *f, int d, int x, int y) {
    return ld_pop<S>(f, (size_t)d * (size_t)a.plane + (size_t)y * (size_t)a.pitch + (size_t)x, d);
}

// Pressure-periodic increment of pressure_periodic_distribution
// (modes/poiseuille_flow.f90:189-194): (w_d * (rho_b - rho_src)) * factor.
template <typename C>
__device__ __forceinline__ C pf_delta(const StepArgs<C> &a, const C *m, int d, int side, int row) {
    const int ny = a.ny;
    const C rho_s = m[(side * 3 + 0) * ny + row];
    const C ux = m[(side * 3 + 1) * ny + row];
    const C uy = m[(side * 3 + 2) * ny + row];
    const C rho_b = side == 0 ? a.rho_in : a.rho_out;
    const C usq = ux * ux + uy * uy;
    const C cu = (C)cx_of(d) * ux + (C)cy_of(d) * uy;
    const C fac = (((C)1 + (C)3 * cu) + (C)4.5 * cu * cu) - (C)1.5 * usq;
    return a.w[d] * (rho_b - rho_s) * fac;
}
// The increment in the units the step works on (times 2^s in the deviation forms).
template <typename S, typename C>
__device__ __forceinline__ C pf_inc(const StepArgs<C> &a, const C *m, int d, int side, int row) {
    const C v = pf_delta<C>(a, m, d, side, row);
    return Store<S>::dev ? v * a.scale : v;
}

// Value of f_d pulled into cell (x, y) when the source (x - c_d) may lie
// outside the domain: the halo entry the reference would have written.
//   SW: periodic in x and y (shear_wave.f90:31-55)
//   CF: bounce-back + moving top wall in y, periodic in x; the wall rule wins
//       at the corners (couette_flow.f90:95-126)
//   SL: bounce-back on all sides, moving lid on top; the lid wins at the top
//       corners (sliding_lid.f90:167-203)
//   BF: bounce-back in y, periodic in x (resting walls)
//   PF: bounce-back in y; in x the periodic value plus the pressure
//       increments of the previous strips (first application, main.f90:216)
//       and of this step's strips (second application, simulation.f90:47-52);
//       the four corner entries are bounce-back plus the increment of the
//       application that wrote them last.
// PHASE 1 gives the halo after the first application only (the state the
// strip pass reads, poiseuille_flow.f90:140-165), PHASE 2 after both.
template <typename S, typename C, int MODE, int PHASE>
__device__ __forceinline__ C pull_edge(const StepArgs<C> &a, const typename Store<S>::R *f, int d, int x,
                                      int y) {
    const int nx = a.nx, ny = a.ny;
    int xs = x - cx_of(d), ys = y - cy_of(d);
    const bool xo = (xs < 0) | (xs >= nx);
    const bool yo = (ys < 0) | (ys >= ny);
    if (MODE == SW) {
        xs = xs < 0 ? xs + nx : (xs >= nx ? xs - nx : xs);
        ys = ys < 0 ? ys + ny : (ys >= ny ? ys - ny : ys);
        return LD<S, C>(a, f, d, xs, ys);
    }
    if ((MODE == SL) ? (xo | yo) : yo) {
        C v = LD<S, C>(a, f, opp_of(d), x, y);
        if (MODE == CF || MODE == SL) {
            if (ys >= ny) {
                if (d == 7) v = v - a.corr8;
                else if (d == 8) v = v + a.corr9;
            }
        }
        if (MODE == PF && xo) {
            const int side = xs < 0 ? 0 : 1;
            v = v + pf_inc<S, C>(a, PHASE == 1 ? a.m_old : a.m_new, d, side, y);
        }
        return v;
    }
    if (xo) {
        const int xw = xs < 0 ? xs + nx : xs - nx;
        C v = LD<S, C>(a, f, d, xw, ys);
        if (MODE == PF) {
            const int side = xs < 0 ? 0 : 1;
            v = v + pf_inc<S, C>(a, a.m_old, d, side, ys);
            if (PHASE == 2) v = v + pf_inc<S, C>(a, a.m_new, d, side, ys);
        }
        return v;
    }
    return LD<S, C>(a, f, d, xs, ys);
}

// The nine stored elements a domain-edge cell (x, y) pulls in the modes
// without pressure increments (SW, CF, SL, BF): the halo entries of
// pull_edge, with the source element of every direction chosen first
// (integer selects) and the nine loads then issued unconditionally, so that
// an edge cell waits for memory once, not once per direction.  A bounce-back
// source is the plane of opp(d), stored with the offset of d (same direction
// class).  The moving-wall terms of CF and SL are added after decoding
// (pull_pops).
template <typename S, typename C, int MODE, typename L>
__device__ __forceinline__ void pull_edge_cell(const StepArgs<C> &a, const typename Store<S>::R *f, int x, int y,
                                               L (&raw)[9]) {
    static_assert(MODE != PF, "PF edge cells add the strip increments (pull_edge)");
    static_assert(store_off<S>(opp_of(1)) == store_off<S>(1) && store_off<S>(opp_of(5)) == store_off<S>(5),
                  "opposite directions share a storage offset");
    const int nx = a.nx, ny = a.ny;
#pragma unroll
    for (int d = 0; d < 9; ++d) {
        int xs = x - cx_of(d), ys = y - cy_of(d), ds = d;
        const bool xo = (xs < 0) | (xs >= nx);
        const bool yo = (ys < 0) | (ys >= ny);
        if (MODE == SW) {
            xs = xs < 0 ? xs + nx : (xs >= nx ? xs - nx : xs);
            ys = ys < 0 ? ys + ny : (ys >= ny ? ys - ny : ys);
        } else if ((MODE == SL) ? (xo | yo) : yo) {
            ds = opp_of(d);
            xs = x;
            ys = y;
        } else if (xo) {
            xs = xs < 0 ? xs + nx : xs - nx;
        }
        raw[d] = f[(size_t)ds * (size_t)a.plane + (size_t)ys * (size_t)a.pitch + (size_t)xs];
    }
}

// Quotients a / b of FP32 operands sharing one denominator b.  The sequence
// is the compiler's IEEE-correct division on gfx950 -- an approximate
// reciprocal, one Newton step, two quotient corrections by FMA -- without the
// range scaling (v_div_scale) and special-case fixup (v_div_fixup) it wraps
// around every quotient: both are identities when b and a/b are normal
// numbers well inside the exponent range and a is 0 or above 2^-100 in
// magnitude, as the densities and momenta of a step are, so each quotient is
// bitwise the one a / b gives there, and the reciprocal's refinement is
// computed once per denominator (9 instructions less for the two velocity
// quotients of a cell).  FP64: the compiler's sequence is a reciprocal, two
// Newton steps and one quotient correction (11 instructions less per cell).
template <typename C>
struct Quot;
template <>
struct Quot<double> {
    double b, y;
    __device__ __forceinline__ explicit Quot(double den) : b(den) {
        const double y0 = __builtin_amdgcn_rcp(den);
        const double y1 = __builtin_fma(y0, __builtin_fma(-den, y0, 1.0), y0);
        y = __builtin_fma(y1, __builtin_fma(-den, y1, 1.0), y1);
    }
    __device__ __forceinline__ double operator()(double a) const {
        const double q = a * y;
        return __builtin_fma(__builtin_fma(-b, q, a), y, q);
    }
};
template <>
struct Quot<float> {
    float b, y;
    __device__ __forceinline__ explicit Quot(float den) : b(den) {
        const float y0 = __builtin_amdgcn_rcpf(den);
        y = __builtin_fmaf(__builtin_fmaf(-den, y0, 1.0f), y0, y0);
    }
    __device__ __forceinline__ float operator()(float a) const {
        const float q0 = a * y;
        const float q1 = __builtin_fmaf(__builtin_fmaf(-b, q0, a), y, q0);
        return __builtin_fmaf(__builtin_fmaf(-b, q1, a), y, q1);
    }
};

// BGK collision of fuzed_pull_streaming_collision_local_unrolled_universal
// (source/layouts/xyd/simulation.f90:224-293), same operation order.  f[]
// holds the pulled populations on entry and the post-collision ones on exit.
template <typename C>
__device__ __forceinline__ void collide(const StepArgs<C> &a, C (&f)[9], int x, int y, bool act = true) {
    const C f1 = f[0], f2 = f[1], f3 = f[2], f4 = f[3], f5 = f[4], f6 = f[5], f7 = f[6], f8 = f[7], f9 = f[8];
    C rho = f1 + f2 + f3 + f4 + f5 + f6 + f7 + f8 + f9;
    C ux = f2 - f4 + f6 - f7 - f8 + f9;
    C uy = f3 - f5 + f6 + f7 - f8 - f9;
    density_check(a, rho, act);
    const Quot<C> by_rho(rho);
    ux = by_rho(ux);
    uy = by_rho(uy);
    const C usq = ux * ux + uy * uy;
    if (a.write_macro && act) {
        const size_t m = (size_t)y * (size_t)a.nx + (size_t)x;
        a.rho_o[m] = rho;
        a.ux_o[m] = ux;
        a.uy_o[m] = uy;
    }
    const C om = a.omega;
    const C one = 1, three = 3, c45 = (C)4.5, c15 = (C)1.5;
    const C w0 = a.w[0], w1 = a.w[1], w2 = a.w[5];
    C s;
    f[0] = f1 + om * (w0 * rho * (one - c15 * usq) - f1);
    f[1] = f2 + om * (w1 * rho * (one + three * ux + c45 * ux * ux - c15 * usq) - f2);
    f[2] = f3 + om * (w1 * rho * (one + three * uy + c45 * uy * uy - c15 * usq) - f3);
    f[3] = f4 + om * (w1 * rho * (one - three * ux + c45 * ux * ux - c15 * usq) - f4);
    f[4] = f5 + om * (w1 * rho * (one - three * uy + c45 * uy * uy - c15 * usq) - f5);
    s = ux + uy;
    f[5] = f6 + om * (w2 * rho * (one + three * s + c45 * s * s - c15 * usq) - f6);
    s = -ux + uy;
    f[6] = f7 + om * (w2 * rho * (one + three * s + c45 * s * s - c15 * usq) - f7);
    s = ux + uy;
    f[7] = f8 + om * (w2 * rho * (one - three * s + c45 * s * s - c15 * usq) - f8);
    s = ux - uy;
    f[8] = f9 + om * (w2 * rho * (one + three * s + c45 * s * s - c15 * usq) - f9);
}

// BGK collision with Guo forcing for a force density F = (force, 0)
// (Guo, Zheng & Shi, Phys. Rev. E 65, 046308 (2002)): velocity
// u = (sum c f + F/2) / rho, f' = f + omega (f_eq(rho, u) - f) + S with
// S_i = (1 - omega/2) w_i [3 (c_i - u) + 9 (c_i . u) c_i] . F.  The macro
// fields written are rho and this u.  No reference counterpart (the
// reference's Poiseuille flow is pressure-driven).
template <typename C>
__device__ __forceinline__ void collide_guo(const StepArgs<C> &a, C (&f)[9], int x, int y, bool act = true) {
    const C f1 = f[0], f2 = f[1], f3 = f[2], f4 = f[3], f5 = f[4], f6 = f[5], f7 = f[6], f8 = f[7], f9 = f[8];
    const C F = a.force;
    C rho = f1 + f2 + f3 + f4 + f5 + f6 + f7 + f8 + f9;
    C ux = f2 - f4 + f6 - f7 - f8 + f9;
    C uy = f3 - f5 + f6 + f7 - f8 - f9;
    density_check(a, rho, act);
    const Quot<C> by_rho(rho);
    ux = by_rho(ux + (C)0.5 * F);
    uy = by_rho(uy);
    const C usq = ux * ux + uy * uy;
    if (a.write_macro && act) {
        const size_t m = (size_t)y * (size_t)a.nx + (size_t)x;
        a.rho_o[m] = rho;
        a.ux_o[m] = ux;
        a.uy_o[m] = uy;
    }
    const C om = a.omega;
    const C one = 1, three = 3, c45 = (C)4.5, c15 = (C)1.5, nine = 9;
    const C pre = (one - (C)0.5 * om) * F;  // (1 - omega/2) F
    // Per direction, with c = c_d:
    //   cu  = c_x u_x + c_y u_y,
    //   feq = w rho (((1 + 3 cu) + 4.5 cu cu) - 1.5 u^2),
    //   src = w pre (3 (c_x - u_x) + 9 cu c_x),
    //   f'  = f + omega (feq - f) + src.
    // Products by a zero velocity component are written out of these
    // expressions (x * 0 may not be dropped by the compiler): c_x = 0 gives
    // 3 (0 - u_x) + 9 cu 0 = 3 (-u_x) up to the sign of a zero, and a zero
    // component adds a signed zero to cu; products by +-1 the compiler folds
    // itself.  Every rounded operation that remains is the one of the full
    // expressions, so the values are theirs (a zero's sign can differ only in
    // a term whose sum with a non-zero population is unchanged by it).
#pragma unroll
    for (int d = 0; d < 9; ++d) {
        const int ix = cx_of(d), iy = cy_of(d);
        C cu;
        if (ix == 0 && iy == 0) cu = (C)0;
        else if (ix == 0) cu = (C)iy * uy;
        else if (iy == 0) cu = (C)ix * ux;
        else cu = (C)ix * ux + (C)iy * uy;
        const C w = a.w[d == 0 ? 0 : (d < 5 ? 1 : 5)];  // the weight of the class (equal values)
        const C feq = w * rho * (((one + three * cu) + c45 * cu * cu) - c15 * usq);
        const C g = ix == 0 ? three * -ux : three * ((C)ix - ux) + nine * cu * (C)ix;
        const C src = w * pre * g;
        f[d] = f[d] + om * (feq - f[d]) + src;
    }
}

// BGK collision of the deviation forms.  D[] holds the scaled deviations
// D_d = (f_d - w_d) * 2^s on entry and the new ones on exit (not yet rounded
// to the storage type).  With Sum w_d = 1 and Sum w_d c_d = 0,
//   rho - 1 = 2^-s Sum D_d,   rho u = 2^-s Sum c_d D_d (+ F/2 for BF),
// and the equilibrium deviation is
//   2^s (f_eq_d - w_d) = 2^s w_d ((rho - 1) + rho (3 c.u + 4.5 (c.u)^2 - 1.5 u^2)),
// so D'_d = (1 - omega) D_d + (omega w_d 2^s) * inner_d (+ Guo source for BF),
// the weight factor carried to double precision (kw + kw_lo).
// Working on deviations keeps the FP32 arithmetic on values of the size of
// the deviations instead of f ~ w (rounding errors ~1e3 times smaller for
// the cavity), and needs one division (1/rho) instead of two.  Opposite
// directions share the even part of the equilibrium:
//   inner_{d,opp} = (base + (rho/2) h^2) +- rho h,  h = 3 c_d.u,
//   base = (rho - 1) - 1.5 rho u^2.
// The operation order here is the definition of the deviation forms (the oracle
// restates it, ffbo_q_step).  No reference counterpart.
template <typename C, int MODE>
__device__ __forceinline__ void collide_dev(const StepArgs<C> &a, C (&D)[9], int x, int y, bool act = true) {
    const C inv = a.inv_scale;
    const C a13 = D[1] + D[3], a24 = D[2] + D[4], a57 = D[5] + D[7], a68 = D[6] + D[8];
    const C b13 = D[1] - D[3], b24 = D[2] - D[4], b57 = D[5] - D[7], b68 = D[6] - D[8];
    const C m0 = ((D[0] + a13) + a24) + (a57 + a68);
    const C jx = (b13 + b57) - b68;
    const C jy = (b24 + b57) + b68;
    const C drho = m0 * inv;
    const C rho = (C)1 + drho;
    density_check(a, rho, act);
    const C r = Quot<C>(rho)((C)1);
    const C ux = (MODE == BF ? __builtin_fmaf(jx, inv, a.half_force) : jx * inv) * r;
    const C uy = (jy * inv) * r;
    if (a.write_macro && act) {
        const size_t m = (size_t)y * (size_t)a.nx + (size_t)x;
        a.rho_o[m] = rho;
        a.ux_o[m] = ux;
        a.uy_o[m] = uy;
    }
    const C usq = __builtin_fmaf(ux, ux, uy * uy);
    const C base = __builtin_fmaf(-((C)1.5 * usq), rho, drho);
    const C hr = (C)0.5 * rho;
    const C om1 = a.om1;
    const C h1 = (C)3 * ux, h2 = (C)3 * uy;
    const C h5 = h1 + h2, h8 = h1 - h2;
    // (omega w_d 2^s) * v with the weight to double precision: the float
    // weights 1/9, 1/36 are off by ~1e-8, which would break
    // Sum w c_x^2 = 1/3 and grow the momentum by ~7e-9 per step (a
    // viscosity bias of ~1e-4 for slowly decaying flows)
    auto kmul = [&](int d, C v) { return __builtin_fmaf(a.kw[d], v, a.kw_lo[d] * v); };
    D[0] = __builtin_fmaf(om1, D[0], kmul(0, base));
    // pair (d, e) with c_e = -c_d and h = 3 c_d.u
    auto pair = [&](int d, int e, C h) {
        const C p = __builtin_fmaf(hr, h * h, base);
        const C q = rho * h;
        D[d] = __builtin_fmaf(om1, D[d], kmul(d, p + q));
        D[e] = __builtin_fmaf(om1, D[e], kmul(e, p - q));
    };
    pair(1, 3, h1);
    pair(2, 4, h2);
    pair(5, 7, h5);
    pair(8, 6, h8);
    if (MODE == BF) {
        // Guo source 2^s w_d (1 - omega/2) [3 (c_d - u) + 9 (c_d.u) c_d] . F,
        // F = (force, 0): kf_d * (9 (c_d.u) c_dx + 3 c_dx - 3 u_x)
        const C g = (C)3 * ux;
        D[0] = __builtin_fmaf(a.kf[0], -g, D[0]);
        D[2] = __builtin_fmaf(a.kf[2], -g, D[2]);
        D[4] = __builtin_fmaf(a.kf[4], -g, D[4]);
        // c_x = +1: 3 h + 3 - g; c_x = -1: 3 h - 3 - g, with h = 3 c_d.u of that direction
        D[1] = __builtin_fmaf(a.kf[1], ((C)3 * h1 + (C)3) - g, D[1]);
        D[3] = __builtin_fmaf(a.kf[3], ((C)3 * h1 - (C)3) - g, D[3]);
        D[5] = __builtin_fmaf(a.kf[5], ((C)3 * h5 + (C)3) - g, D[5]);
        D[7] = __builtin_fmaf(a.kf[7], ((C)3 * h5 - (C)3) - g, D[7]);
        D[8] = __builtin_fmaf(a.kf[8], ((C)3 * h8 + (C)3) - g, D[8]);
        D[6] = __builtin_fmaf(a.kf[6], ((C)3 * h8 - (C)3) - g, D[6]);
    }
}

template <typename S, typename C, int MODE>
__device__ __forceinline__ void collide_mode(const StepArgs<C> &a, C (&f)[9], int x, int y, bool act = true) {
    if constexpr (Store<S>::dev) collide_dev<C, MODE>(a, f, x, y, act);
    else if (MODE == BF) collide_guo<C>(a, f, x, y, act);
    else collide<C>(a, f, x, y, act);
}

// x offset and y offset of the pull source of direction d: f_d(x - c_d).
__host__ __device__ constexpr int pull_dx(int d) { return -cx_of(d); }
__host__ __device__ constexpr int pull_dy(int d) { return -cy_of(d); }

// Row-relative addressing.  A tile (or a batched row segment) lies in one row,
// so the address of that row in every direction set is the same for the whole
// wave (scalar registers) and each lane adds one 32-bit byte offset: the
// global loads and stores take the scalar-base form, with no 64-bit address
// arithmetic per access.  `row` may point before the array when the access
// itself stays inside it (a pull from row y-1 of an interior cell).
template <typename R>
__device__ __forceinline__ const R *at(const R *row, unsigned xb) {
    return reinterpret_cast<const R *>(reinterpret_cast<const char *>(row) + xb);
}
template <typename R>
__device__ __forceinline__ R *at(R *row, unsigned xb) {
    return reinterpret_cast<R *>(reinterpret_cast<char *>(row) + xb);
}

// Raw buffer access of one direction row: a buffer descriptor whose base is the
// row address (wave-uniform, built in scalar registers; word 3 0x00020000 is
// the raw 32-bit data format of gfx9 parts) and one 32-bit byte offset per
// lane.  The x shift of a pull (+-1 element) is folded into the base, so the
// accesses of a cell share one offset register.  num_records = 2^32-1
// disables the range check: every access made this way lies inside its array.
// Measured on MI355X (8192^2 cavity): buffer stores beat global stores for
// every element size; buffer loads beat global loads for 4- and 8-byte
// elements (FP32 192.5 against 193.6 us at 4096^2) but are 12 % slower for
// 2-byte ones (FP16 450 against 394 us per step), which keep global loads.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t row_rsrc(const void *row) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(row), 0, -1, 0x00020000);
}
template <typename R>
__device__ __forceinline__ R buf_ld(__amdgpu_buffer_rsrc_t r, unsigned xb) {
    if constexpr (sizeof(R) == 2) return __builtin_bit_cast(R, __builtin_amdgcn_raw_buffer_load_b16(r, xb, 0, 0));
    else if constexpr (sizeof(R) == 4) return __builtin_bit_cast(R, __builtin_amdgcn_raw_buffer_load_b32(r, xb, 0, 0));
    else return __builtin_bit_cast(R, __builtin_amdgcn_raw_buffer_load_b64(r, xb, 0, 0));
}
// Nontemporal store (cache policy bit 1, "nt").
template <typename R>
__device__ __forceinline__ void buf_st_nt(R v, __amdgpu_buffer_rsrc_t r, unsigned xb, unsigned so = 0) {
    if constexpr (sizeof(R) == 2) __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(unsigned short, v), r, xb, so, 2);
    else if constexpr (sizeof(R) == 4) __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, xb, so, 2);
    else {
        typedef unsigned u2 __attribute__((ext_vector_type(2)));
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), r, xb, so, 2);
    }
}
// Store with the default cache policy (partial cache lines that neighbouring
// workgroups on one XCD write merge in its L2).
template <typename R>
__device__ __forceinline__ void buf_st(R v, __amdgpu_buffer_rsrc_t r, unsigned xb) {
    if constexpr (sizeof(R) == 2) __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(unsigned short, v), r, xb, 0, 0);
    else if constexpr (sizeof(R) == 4) __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, xb, 0, 0);
    else {
        typedef unsigned u2 __attribute__((ext_vector_type(2)));
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), r, xb, 0, 0);
    }
}

// Row descriptor with range checking: num_records 2^31 - 1, so a lane whose
// byte offset is kDropOffset lies out of range and the hardware drops its store
// (the range check of raw buffer accesses: offset >= num_records).  Every
// offset of a real element is below 2^31 (n_x <= 2^24 elements of at most 8
// bytes).  Lets a wave store unconditionally with some lanes masked.
constexpr unsigned kDropOffset = 0x80000000u;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t row_rsrc_checked(const void *row) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(row), 0, 0x7fffffff, 0x00020000);
}

// Element xb/sizeof(R) of a row: raw buffer load for 4- and 8-byte elements,
// global load for 2-byte ones (faster there, measured).
template <typename R>
__device__ __forceinline__ R ld_row(const R *row, unsigned xb) {
    if constexpr (sizeof(R) >= 4) return buf_ld<R>(row_rsrc(row), xb);
    else return *at(row, xb);
}

// A stored population as loaded, before decoding (2-byte elements held
// zero-extended, one per register).
template <typename S>
using RawPop = std::conditional_t<sizeof(typename Store<S>::R) == 2, unsigned, typename Store<S>::R>;

// The loads of the nine stored elements cell (x, y) pulls, in the modes
// without pressure increments (SW, CF, SL, BF).  Interior cells pull through
// their row addresses, edge cells by the edge rule (which gives the plain
// pull for an interior cell too).  Both sides only load and decoding follows
// (raw_pops), so a caller can issue the loads early.  For 2-byte elements a
// wave holding an edge cell (a row's first or last segment, the first or last
// row) takes the edge rule in every lane: a uniform branch, one batch of loads
// and one wait per wave.  For 4- and 8-byte elements only the edge lanes do,
// and the interior lanes keep their buffer loads; measured faster there (FP32
// batch of 512 lattices 80,300 against 78,700 MLUPS), slower for the 16-bit
// forms (FP16 129,600 against 137,500).
// UNIFORM_EDGE: every lane of a wave holding an edge cell takes the edge rule
// (the default for 2-byte elements).
template <typename S, typename C, int MODE, bool UNIFORM_EDGE = sizeof(typename Store<S>::R) == 2>
__device__ __forceinline__ void pull_raw(const StepArgs<C> &a, const typename Store<S>::R *__restrict__ f, int x,
                                         int y, RawPop<S> (&raw)[9]) {
    using R = typename Store<S>::R;
    const ptrdiff_t P = (ptrdiff_t)a.plane, W = (ptrdiff_t)a.pitch;
    const R *fr = f + (ptrdiff_t)y * W;  // row y, direction 0 (wave-uniform)
    const unsigned xb = (unsigned)x * (unsigned)sizeof(R);
    const bool edge = (x == 0) | (x == a.nx - 1) | (y == 0) | (y == a.ny - 1);
    const bool edge_rule = UNIFORM_EDGE ? __builtin_amdgcn_ballot_w64(edge) != 0 : edge;
    if (!edge_rule) {
#pragma unroll
        for (int d = 0; d < 9; ++d) raw[d] = ld_row(fr + d * P + pull_dy(d) * W + pull_dx(d), xb);
    } else {
        pull_edge_cell<S, C, MODE>(a, f, x, y, raw);
    }
}

// Working values of the loaded elements of a cell of row y: decoded, plus
// the moving-wall terms of CF and SL on the top row (bounce-back of d = 7, 8
// from row ny, pull_edge).
template <typename S, typename C, int MODE>
__device__ __forceinline__ void raw_pops(const StepArgs<C> &a, const RawPop<S> (&raw)[9], int y, C (&v)[9]) {
    using R = typename Store<S>::R;
#pragma unroll
    for (int d = 0; d < 9; ++d) v[d] = pop_in<S>((R)raw[d], d);
    if (MODE == CF || MODE == SL) {
        if (y == a.ny - 1) {
            v[7] = v[7] - a.corr8;
            v[8] = v[8] + a.corr9;
        }
    }
}

// The nine populations cell (x, y) pulls, as working values: plain pull in
// the interior, boundary rule on the edge.
template <typename S, typename C, int MODE>
__device__ __forceinline__ void pull_pops(const StepArgs<C> &a, const typename Store<S>::R *__restrict__ f, int x,
                                          int y, C (&v)[9]) {
    using R = typename Store<S>::R;
    const ptrdiff_t P = (ptrdiff_t)a.plane, W = (ptrdiff_t)a.pitch;
    const R *fr = f + (ptrdiff_t)y * W;  // row y, direction 0 (wave-uniform)
    const unsigned xb = (unsigned)x * (unsigned)sizeof(R);
    const bool edge = (x == 0) | (x == a.nx - 1) | (y == 0) | (y == a.ny - 1);
    if constexpr (MODE == PF) {
        if (!edge) {
#pragma unroll
            for (int d = 0; d < 9; ++d) v[d] = pop_in<S>(ld_row(fr + d * P + pull_dy(d) * W + pull_dx(d), xb), d);
        } else {
            v[0] = LD<S, C>(a, f, 0, x, y);
#pragma unroll
            for (int d = 1; d < 9; ++d) v[d] = pull_edge<S, C, MODE, 2>(a, f, d, x, y);
        }
    } else {
        RawPop<S> raw[9];
        pull_raw<S, C, MODE>(a, f, x, y, raw);
        raw_pops<S, C, MODE>(a, raw, y, v);
    }
}

// Stored value of direction d of a cell: nontemporal buffer store at byte xb
// of the direction row d of row gr, through that row's own descriptor, or,
// with ONE_DESC, through one descriptor at direction 0 and a scalar offset
// (the same descriptor for the nine rows: fewer scalar registers; needs the
// nine rows within 4 GiB, which the caller guarantees).
template <typename R, bool ONE_DESC, bool NT = true>
__device__ __forceinline__ void store_pop(R v, R *gr, ptrdiff_t P, unsigned xb, int d) {
    if constexpr (!NT) buf_st<R>(v, row_rsrc(gr + d * P), xb);
    else if constexpr (ONE_DESC) buf_st_nt<R>(v, row_rsrc(gr), xb, (unsigned)d * (unsigned)P * (unsigned)sizeof(R));
    else buf_st_nt<R>(v, row_rsrc(gr + d * P), xb);
}

// Stored form of the post-collision values v of cell (x, y): the 16-bit
// forms round stochastically with the bits of (cell, a.step_seed) when asked.
template <typename S, typename C>
__device__ __forceinline__ void encode_cell(const StepArgs<C> &a, const C (&v)[9], int x, int y,
                                            typename Store<S>::R (&o)[9]) {
    if constexpr (Store<S>::dev && sizeof(typename Store<S>::R) == 2) {
        if (a.stochastic) {
            const CellBits b = cell_bits((unsigned long long)y * (unsigned long long)a.nx + (unsigned)x, a.step_seed);
#pragma unroll
            for (int d = 0; d < 9; ++d) o[d] = pop_out_sr<S>(v[d], pop_bits(b, d), d);
            return;
        }
    }
#pragma unroll
    for (int d = 0; d < 9; ++d) o[d] = pop_out<S>(v[d], d);
}

// Collision of the pulled values v of cell (x, y) and the stores of its new
// populations into g.
template <typename S, typename C, int MODE, bool ONE_DESC = false, bool NT = true>
__device__ __forceinline__ void collide_store(const StepArgs<C> &a, C (&v)[9], typename Store<S>::R *__restrict__ g,
                                              int x, int y) {
    using R = typename Store<S>::R;
    const ptrdiff_t P = (ptrdiff_t)a.plane, W = (ptrdiff_t)a.pitch;
    R *gr = g + (ptrdiff_t)y * W;  // row y, direction 0 (wave-uniform)
    const unsigned xb = (unsigned)x * (unsigned)sizeof(R);
    collide_mode<S, C, MODE>(a, v, x, y);
    if constexpr (Store<S>::dev && sizeof(R) == 2) {
        if (a.stochastic) {
            const CellBits b = cell_bits((unsigned long long)y * (unsigned long long)a.nx + (unsigned)x, a.step_seed);
#pragma unroll
            for (int d = 0; d < 9; ++d) store_pop<R, ONE_DESC, NT>(pop_out_sr<S>(v[d], pop_bits(b, d), d), gr, P, xb, d);
            return;
        }
    }
#pragma unroll
    for (int d = 0; d < 9; ++d) store_pop<R, ONE_DESC, NT>(pop_out<S>(v[d], d), gr, P, xb, d);
}

// collide_store for every lane of a wave, without a branch around the stores:
// lanes with act false compute on whatever they hold, flag no density and
// write no macro field, and their population stores go to kDropOffset, which
// the hardware drops.  The stores are then unconditional in the program, so
// the compiler's count of outstanding memory operations includes them on
// every path (a store under a branch is counted as possibly absent, and a
// later wait for loads issued before it then also waits for the stores).
// y is wave-uniform (the row addresses are scalar); it may lie outside the
// domain when no lane is active.
template <typename S, typename C, int MODE>
__device__ __forceinline__ void collide_store_masked(const StepArgs<C> &a, C (&v)[9], typename Store<S>::R *__restrict__ g,
                                                     int x, int y, bool act) {
    using R = typename Store<S>::R;
    const ptrdiff_t P = (ptrdiff_t)a.plane, W = (ptrdiff_t)a.pitch;
    R *gr = g + (ptrdiff_t)y * W;
    const unsigned xb = act ? (unsigned)x * (unsigned)sizeof(R) : kDropOffset;
    collide_mode<S, C, MODE>(a, v, x, y, act);
    R o[9];
    encode_cell<S, C>(a, v, x, y, o);
    // one descriptor for the row, direction d at the scalar offset d P
    // elements (the nine direction rows of a row lie within 9 plane elements,
    // below 2^31 bytes for the row-blocked layout; the plane layout keeps a
    // descriptor per direction)
    const bool one = a.pitch > a.plane;
    const __amdgpu_buffer_rsrc_t r0 = row_rsrc_checked(gr);
#pragma unroll
    for (int d = 0; d < 9; ++d) {
        const __amdgpu_buffer_rsrc_t r = one ? r0 : row_rsrc_checked(gr + d * P);
        const unsigned so = one ? (unsigned)(d * P * (ptrdiff_t)sizeof(R)) : 0u;
        if constexpr (sizeof(R) == 2) __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(unsigned short, o[d]), r, xb, so, 0);
        else if constexpr (sizeof(R) == 4) __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, o[d]), r, xb, so, 0);
        else {
            typedef unsigned u2 __attribute__((ext_vector_type(2)));
            __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, o[d]), r, xb, so, 0);
        }
    }
}

// One cell, any position: pull, collide, store.
template <typename S, typename C, int MODE, bool ONE_DESC = false>
__device__ __forceinline__ void step_cell(const StepArgs<C> &a, const typename Store<S>::R *__restrict__ f,
                                          typename Store<S>::R *__restrict__ g, int x, int y) {
    C v[9];
    pull_pops<S, C, MODE>(a, f, x, y, v);
    collide_store<S, C, MODE, ONE_DESC>(a, v, g, x, y);
}

// One fused pull-stream-collide step over one domain.  Each thread updates V
// consecutive cells of one row (V * element size = 16 bytes); a workgroup is a
// tile of TW*V cells.  Where all V cells are interior, every plane is read
// with one aligned 16-byte vector load at the thread's own x0; the x-1 and x+1
// planes add one scalar load of the element just outside the vector (a cache
// hit: the neighbouring lane's vector holds it) and shift it in, so every
// vector load is naturally aligned.  Each plane is written with one aligned nontemporal 16-byte
// store.  Cells on the domain edge take the boundary rule cell by cell.
template <typename S, typename C, int MODE, int TW, int V>
__global__ void __launch_bounds__(TW) step_kernel(StepArgs<C> a) {
    using R = typename Store<S>::R;
    typedef R RV __attribute__((ext_vector_type(V)));
    const int t = tile_of_block((int)blockIdx.x, a.n_blocks, a.xcd_group);
    const int y = t / a.n_tiles_x;
    const int x0 = ((t - y * a.n_tiles_x) * TW + (int)threadIdx.x) * V;
    if (x0 >= a.nx || y >= a.ny) return;
    const R *__restrict__ f = static_cast<const R *>(a.src);
    R *__restrict__ g = static_cast<R *>(a.dst);
    const bool interior = (V > 1) & (y > 0) & (y < a.ny - 1) & (x0 > 0) & (x0 + V < a.nx);
    if (!interior) {
#pragma unroll
        for (int k = 0; k < V; ++k)
            if (x0 + k < a.nx) step_cell<S, C, MODE>(a, f, g, x0 + k, y);
        return;
    }
    const ptrdiff_t P = (ptrdiff_t)a.plane, W = (ptrdiff_t)a.pitch;
    const R *fr = f + (ptrdiff_t)y * W;  // row y, direction 0 (wave-uniform)
    R *gr = g + (ptrdiff_t)y * W;
    const unsigned xb = (unsigned)x0 * (unsigned)sizeof(R);
    // neighbour elements first, then a compiler-only memory barrier (emits no
    // instruction): it keeps these element loads separate from the aligned
    // vector loads below, which the load combiner would otherwise fuse into
    // 4-byte-aligned 16-byte loads.
    R side[9];
#pragma unroll
    for (int d = 0; d < 9; ++d) {
        const R *row = at(fr + d * P + pull_dy(d) * W, xb);
        if (pull_dx(d) < 0) side[d] = row[-1];
        else if (pull_dx(d) > 0) side[d] = row[V];
    }
    asm volatile("" ::: "memory");
    RV in[9];
#pragma unroll
    for (int d = 0; d < 9; ++d) {
        const R *row = at(fr + d * P + pull_dy(d) * W, xb);  // 16-byte aligned
        const RV own = *reinterpret_cast<const RV *>(row);
        if (pull_dx(d) == 0) {
            in[d] = own;
        } else if (pull_dx(d) < 0) {  // f(x0-1 .. x0+V-2)
            RV t;
            t[0] = side[d];
#pragma unroll
            for (int k = 1; k < V; ++k) t[k] = own[k - 1];
            in[d] = t;
        } else {                       // f(x0+1 .. x0+V)
            RV t;
#pragma unroll
            for (int k = 0; k < V - 1; ++k) t[k] = own[k + 1];
            t[V - 1] = side[d];
            in[d] = t;
        }
    }
    RV out[9];
#pragma unroll
    for (int k = 0; k < V; ++k) {
        C v[9];
#pragma unroll
        for (int d = 0; d < 9; ++d) v[d] = pop_in<S>(in[d][k], d);
        collide_mode<S, C, MODE>(a, v, x0 + k, y);
        if constexpr (Store<S>::dev && sizeof(R) == 2) {
            if (a.stochastic) {
                const CellBits b =
                    cell_bits((unsigned long long)y * (unsigned long long)a.nx + (unsigned)(x0 + k), a.step_seed);
#pragma unroll
                for (int d = 0; d < 9; ++d) out[d][k] = pop_out_sr<S>(v[d], pop_bits(b, d), d);
                continue;
            }
        }
#pragma unroll
        for (int d = 0; d < 9; ++d) out[d][k] = pop_out<S>(v[d], d);
    }
#pragma unroll
    for (int d = 0; d < 9; ++d) __builtin_nontemporal_store(out[d], reinterpret_cast<RV *>(at(gr + d * P, xb)));
}

// Two steps in one pass (temporal blocking), the modes without strip
// passes (SW, CF, SL, BF).  A workgroup of WAVES waves owns an output tile of
// 62 x (WAVES RPT - 2) cells and computes step 1 on the tile plus a one-cell
// ring: a region of 64 x WAVES RPT cells, RPT rows per thread (wave w takes
// rows w, w + WAVES, ...).  On a periodic axis the ring cell is the wrapped
// cell, whose step-1 values the reference's halo copy would hold.  The
// region's step-1 results stay in LDS in the stored representation (so the
// 16-bit forms round exactly as two single steps do), and step 2 of the tile
// reads them there: pulled values from the neighbouring LDS cells, and, where
// the source lies beyond a wall, the cell's own opposite population plus the
// moving-wall term, as pull_edge.  Every cell update is the single step's
// arithmetic, so the result is bitwise that of two step_kernel launches.
// Step 1 runs with a.step_seed and no macro output, step 2 with
// a.step_seed + 1 and a.write_macro.  Tiles are numbered row by row
// (a.n_tiles_x per tile row) and mapped to workgroups XCD-grouped
// (tile_of_block).  HBM traffic per two steps: ~(1.2 + 1) x 9 x s bytes per
// cell against 36 x s.
// The pair of one tile (tx, ty) of domain a, LDS of the workgroup; every
// thread of the workgroup calls it with the same tile.
template <typename S, typename C, int MODE, int WAVES, int RPT, bool NT>
__device__ __forceinline__ void pair_tile(const StepArgs<C> &a, int tx, int ty,
                                          typename Store<S>::R (&lds)[9][WAVES * RPT][64]) {
    static_assert(MODE != PF, "Poiseuille flow takes its strip pass between steps");
    using R = typename Store<S>::R;
    constexpr int ROWS = WAVES * RPT;
    constexpr bool PX = MODE != SL, PY = MODE == SW;  // periodic axes
    // wave index made wave-uniform: the region row of a wave then lies in
    // scalar registers, and so do the buffer descriptors built from it (a
    // per-lane row would wrap every buffer access in a readfirstlane loop)
    const int lane = (int)threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    const int nx = a.nx, ny = a.ny;
    const int ux = tx * 62 - 1 + lane;  // unwrapped region column
    const int uy0 = ty * (ROWS - 2) - 1;  // unwrapped first region row
    const R *__restrict__ f = static_cast<const R *>(a.src);
    R *__restrict__ g = static_cast<R *>(a.dst);
    // step 1 on the region cells (needed where -1 <= ux <= nx, -1 <= uy <= ny)
    int gx = ux;
    bool xin = (ux >= -1) & (ux <= nx);
    if (PX) gx = ux < 0 ? ux + nx : (ux >= nx ? ux - nx : ux);
    else xin &= (ux >= 0) & (ux < nx);
    StepArgs<C> a1 = a;
    a1.write_macro = 0;
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        const int r = w + WAVES * k, uy = uy0 + r;
        int gy = uy;
        bool in1 = xin & (uy >= -1) & (uy <= ny);
        if (PY) gy = uy < 0 ? uy + ny : (uy >= ny ? uy - ny : uy);
        else in1 &= (uy >= 0) & (uy < ny);
        if (in1) {
            C v[9];
            pull_pops<S, C, MODE>(a1, f, gx, gy, v);
            collide_mode<S, C, MODE>(a1, v, gx, gy);
            R o[9];
            encode_cell<S, C>(a1, v, gx, gy, o);
#pragma unroll
            for (int d = 0; d < 9; ++d) lds[d][r][lane] = o[d];
        }
    }
    __syncthreads();
    // step 2 on the tile.  A tile whose output cells all have their eight
    // neighbours inside the domain (most tiles) pulls without the boundary
    // rules: a branch uniform over the workgroup.
    StepArgs<C> a2 = a;
    a2.step_seed = a.step_seed + 1u;
    const int x0 = tx * 62;
    const bool inner = (x0 >= 1) & (x0 + 62 <= nx - 1) & (uy0 >= 0) & (uy0 + ROWS <= ny);
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        const int r = w + WAVES * k, uy = uy0 + r;
        if ((lane >= 1) & (lane <= 62) & (r >= 1) & (r <= ROWS - 2) & (ux < nx) & (uy < ny)) {
            C v[9];
            if (inner) {
#pragma unroll
                for (int d = 0; d < 9; ++d) v[d] = pop_in<S>(lds[d][r - cy_of(d)][lane - cx_of(d)], d);
            } else {
#pragma unroll
                for (int d = 0; d < 9; ++d) {
                    const int sx = ux - cx_of(d), sy = uy - cy_of(d);
                    const bool xo = (sx < 0) | (sx >= nx), yo = (sy < 0) | (sy >= ny);
                    const bool bb = MODE == SW ? false : (MODE == SL ? (xo | yo) : yo);
                    const R raw = bb ? lds[opp_of(d)][r][lane] : lds[d][r - cy_of(d)][lane - cx_of(d)];
                    v[d] = pop_in<S>(raw, d);  // opp(d) stores with the offset of d
                }
                if (MODE == CF || MODE == SL) {  // moving top wall (pull_edge)
                    if (uy == ny - 1) {
                        v[7] = v[7] - a.corr8;
                        v[8] = v[8] + a.corr9;
                    }
                }
            }
            collide_store<S, C, MODE, false, NT>(a2, v, g, ux, uy);
        }
    }
}

// K steps in one pass (K = 3: step triples), the modes without strip
// passes.  A workgroup of WAVES waves owns an output tile of
// (64 - 2(K-1)) x (WAVES RPT - 2(K-1)) cells and computes step 1 on the tile
// plus a ring of K-1 cells (a region of 64 x WAVES RPT cells; wave w takes
// region rows w, w + WAVES, ...), step k on the tile plus a ring of K-k
// cells, and step K on the tile.  Step 1 pulls from HBM into LDS (stored
// representation, as pair_tile); every later step pulls from LDS, with the
// boundary rules of pull_edge for sources beyond a wall (the cell's own
// opposite population, plus the moving-wall term on the top row); on a
// periodic axis the ring cells are the wrapped cells, whose values at every
// step are those the reference's halo copies would hold.  An intermediate
// step keeps its results in registers until every wave has read the LDS
// (barrier), then writes them over its sources (barrier).  Step k runs with
// a.step_seed + k - 1; step K writes the macro fields when asked.  Every cell
// update is the single step's arithmetic, so the result is bitwise that of K
// step_kernel launches.
// A coordinate at most RING cells outside [0, n), wrapped into it (n >= 1 may
// be smaller than RING).
template <int RING>
__device__ __forceinline__ int mt_wrap(int u, int n) {
#pragma unroll
    for (int k = 0; k < RING; ++k) u = u < 0 ? u + n : (u >= n ? u - n : u);
    return u;
}

// Whether region cell (u, v) of a K-step tile takes part in step k: within
// K - k cells of the domain on a periodic axis, inside it on a wall axis.
template <int MODE, int K>
__device__ __forceinline__ bool mt_live(int u, int v, int k, int nx, int ny) {
    constexpr bool PX = MODE != SL, PY = MODE == SW;  // periodic axes
    const int e = K - k;
    bool in = PX ? (u >= -e) & (u <= nx - 1 + e) : (u >= 0) & (u < nx);
    in &= PY ? (v >= -e) & (v <= ny - 1 + e) : (v >= 0) & (v < ny);
    return in;
}

// The HBM loads of step 1 of K-step tile (tx, ty): the stored elements each
// region cell of this thread (region rows w, w + WAVES, ...) pulls, not yet
// decoded (pull_raw).  Issued one tile ahead by stepk_pipe_kernel.
template <typename S, typename C, int MODE, int WAVES, int RPT, int K>
__device__ __forceinline__ void multi_tile_loads(const StepArgs<C> &a, int tx, int ty, RawPop<S> (&raw)[RPT][9]) {
    using R = typename Store<S>::R;
    constexpr int ROWS = WAVES * RPT, RING = K - 1, OW = 64 - 2 * RING, OH = ROWS - 2 * RING;
    constexpr bool PX = MODE != SL, PY = MODE == SW;
    const int lane = (int)threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    const int ux = tx * OW - RING + lane, uy0 = ty * OH - RING;
    const int gx = PX ? mt_wrap<RING>(ux, a.nx) : ux;
    const R *__restrict__ f = static_cast<const R *>(a.src);
#pragma unroll
    for (int j = 0; j < RPT; ++j) {
        const int uy = uy0 + w + WAVES * j;
        if (mt_live<MODE, K>(ux, uy, 1, a.nx, a.ny)) {
            const int gy = PY ? mt_wrap<RING>(uy, a.ny) : uy;
            pull_raw<S, C, MODE>(a, f, gx, gy, raw[j]);
        }
    }
}

// PIPE: step 1 takes its loaded elements from raw (multi_tile_loads of this
// tile), and once every row of it has been computed the loads of tile
// (ntx, nty) go into raw (none for ntx < 0), so that they are in flight
// while steps 2 .. K run from the LDS.
template <typename S, typename C, int MODE, int WAVES, int RPT, int K, bool PIPE = false>
__device__ __forceinline__ void multi_tile(const StepArgs<C> &a, int tx, int ty,
                                           typename Store<S>::R (&lds)[9][WAVES * RPT][64],
                                           RawPop<S> (&raw)[RPT][9], int ntx = -1, int nty = 0) {
    static_assert(MODE != PF, "Poiseuille flow takes its strip pass between steps");
    static_assert(K >= 2 && K <= 5, "two to five steps per pass");
    using R = typename Store<S>::R;
    constexpr int ROWS = WAVES * RPT, RING = K - 1, OW = 64 - 2 * RING, OH = ROWS - 2 * RING;
    constexpr bool PX = MODE != SL, PY = MODE == SW;  // periodic axes
    const int lane = (int)threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    const int nx = a.nx, ny = a.ny;
    const int ux = tx * OW - RING + lane;  // unwrapped region column
    const int uy0 = ty * OH - RING;        // unwrapped first region row
    const typename Store<S>::R *__restrict__ f = static_cast<const R *>(a.src);
    R *__restrict__ g = static_cast<R *>(a.dst);
    auto wrap = [](int u, int n) { return mt_wrap<RING>(u, n); };
    auto live = [&](int u, int v, int k) { return mt_live<MODE, K>(u, v, k, nx, ny); };
    const int gx = PX ? wrap(ux, nx) : ux;
    // step 1 on the region
    {
        StepArgs<C> a1 = a;
        a1.write_macro = 0;
#pragma unroll
        for (int j = 0; j < RPT; ++j) {
            const int r = w + WAVES * j, uy = uy0 + r;
            if (live(ux, uy, 1)) {
                const int gy = PY ? wrap(uy, ny) : uy;
                C v[9];
                if constexpr (PIPE) raw_pops<S, C, MODE>(a1, raw[j], gy, v);
                else pull_pops<S, C, MODE>(a1, f, gx, gy, v);
                collide_mode<S, C, MODE>(a1, v, gx, gy);
                R o[9];
                encode_cell<S, C>(a1, v, gx, gy, o);
#pragma unroll
                for (int d = 0; d < 9; ++d) lds[d][r][lane] = o[d];
            }
        }
    }
    if constexpr (PIPE) {
        if (ntx >= 0) multi_tile_loads<S, C, MODE, WAVES, RPT, K>(a, ntx, nty, raw);
    }
    __syncthreads();
    // A region inside the domain on every wall axis (most tiles) has no
    // later-step pull that crosses a wall: every source of a step-k cell lies
    // in the region, and a cell on a wall row is on the region's edge, which
    // no later step computes.  Such a tile pulls without the boundary rules (a
    // branch uniform over the workgroup); on a periodic axis the ring cells
    // hold the wrapped values, so no rule is needed there either.
    const int x0 = tx * OW - RING;
    const bool inner = (PX || ((x0 >= 0) & (x0 + 63 <= nx - 1))) && (PY || ((uy0 >= 0) & (uy0 + ROWS - 1 <= ny - 1)));
    // the values a later step pulls into cell (lane, r) at (ux, uy) from the LDS
    auto pull_lds = [&](int r, int uy, C (&v)[9]) {
        if (inner) {
#pragma unroll
            for (int d = 0; d < 9; ++d) v[d] = pop_in<S>(lds[d][r - cy_of(d)][lane - cx_of(d)], d);
            return;
        }
#pragma unroll
        for (int d = 0; d < 9; ++d) {
            const int sx = ux - cx_of(d), sy = uy - cy_of(d);
            const bool xo = (sx < 0) | (sx >= nx), yo = (sy < 0) | (sy >= ny);
            const bool bb = MODE == SW ? false : (MODE == SL ? (xo | yo) : yo);
            const R raw = bb ? lds[opp_of(d)][r][lane] : lds[d][r - cy_of(d)][lane - cx_of(d)];
            v[d] = pop_in<S>(raw, d);  // opp(d) stores with the offset of d
        }
        if (MODE == CF || MODE == SL) {  // moving top wall (pull_edge)
            if (uy == ny - 1) {
                v[7] = v[7] - a.corr8;
                v[8] = v[8] + a.corr9;
            }
        }
    };
    // intermediate steps 2 .. K-1, in place in the LDS
#pragma unroll
    for (int k = 2; k < K; ++k) {
        StepArgs<C> ak = a;
        ak.write_macro = 0;
        ak.step_seed = a.step_seed + (unsigned)(k - 1);
        R o[RPT][9];
        bool on[RPT];
#pragma unroll
        for (int j = 0; j < RPT; ++j) {
            const int r = w + WAVES * j, uy = uy0 + r;
            on[j] = (lane >= k - 1) & (lane <= 64 - k) & (r >= k - 1) & (r <= ROWS - k) & live(ux, uy, k);
            if (on[j]) {
                const int gy = PY ? wrap(uy, ny) : uy;
                C v[9];
                pull_lds(r, uy, v);
                collide_mode<S, C, MODE>(ak, v, gx, gy);
                encode_cell<S, C>(ak, v, gx, gy, o[j]);
            }
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < RPT; ++j) {
            const int r = w + WAVES * j;
            if (on[j]) {
#pragma unroll
                for (int d = 0; d < 9; ++d) lds[d][r][lane] = o[j][d];
            }
        }
        __syncthreads();
    }
    // step K on the tile, to HBM
    StepArgs<C> aK = a;
    aK.step_seed = a.step_seed + (unsigned)(K - 1);
#pragma unroll
    for (int j = 0; j < RPT; ++j) {
        const int r = w + WAVES * j, uy = uy0 + r;
        if ((lane >= RING) & (lane <= 63 - RING) & (r >= RING) & (r <= ROWS - 1 - RING) & (ux < nx) & (uy < ny)) {
            C v[9];
            pull_lds(r, uy, v);
            collide_store<S, C, MODE, false, false>(aK, v, g, ux, uy);
        }
    }
}

// K steps per launch over tiles (multi_tile), XCD-grouped as step2_kernel.
template <typename S, typename C, int MODE, int WAVES, int RPT, int K>
__global__ void __launch_bounds__(64 * WAVES) stepk_kernel(StepArgs<C> a) {
    __shared__ typename Store<S>::R lds[9][WAVES * RPT][64];
    RawPop<S> raw[RPT][9];  // unused without PIPE
    const int t = tile_of_block((int)blockIdx.x, a.n_blocks, a.xcd_group);
    const int ty = t / a.n_tiles_x;
    multi_tile<S, C, MODE, WAVES, RPT, K>(a, t - ty * a.n_tiles_x, ty, lds, raw);
}

// K steps per launch over tiles by a persistent grid: workgroup p takes the
// tiles of blocks p, p + gridDim.x, ... (gridDim.x a multiple of 8, so every
// block it takes lies on its XCD in the grouped order of tile_of_block) and
// issues the HBM loads of its next tile's step 1 before steps 2 .. K of the
// current one (multi_tile, PIPE), so that a workgroup keeps loads in flight
// through its LDS-only steps.  Results are those of stepk_kernel bit for bit
// (same cell updates; only the time of the loads differs).
template <typename S, typename C, int MODE, int WAVES, int RPT, int K>
__global__ void __launch_bounds__(64 * WAVES) stepk_pipe_kernel(StepArgs<C> a) {
    __shared__ typename Store<S>::R lds[9][WAVES * RPT][64];
    RawPop<S> raw[RPT][9];
    int p = (int)blockIdx.x;
    if (p >= a.n_blocks) return;  // whole workgroup
    auto coords = [&](int q, int &tx, int &ty) {
        const int t = tile_of_block(q, a.n_blocks, a.xcd_group);
        ty = t / a.n_tiles_x;
        tx = t - ty * a.n_tiles_x;
    };
    int tx, ty;
    coords(p, tx, ty);
    multi_tile_loads<S, C, MODE, WAVES, RPT, K>(a, tx, ty, raw);
    for (;;) {
        const int q = p + (int)gridDim.x;
        int ntx = -1, nty = 0;
        if (q < a.n_blocks) coords(q, ntx, nty);
        multi_tile<S, C, MODE, WAVES, RPT, K, true>(a, tx, ty, lds, raw, ntx, nty);
        if (q >= a.n_blocks) break;
        __syncthreads();  // step K has read the LDS before the next step 1 writes it
        p = q;
        tx = ntx;
        ty = nty;
    }
}

// a.n_blocks tiles; a grid smaller than that walks them (persistent
// workgroups, grid a multiple of 8 so that every tile of a block lies on the
// block's XCD in the grouped order).
template <typename S, typename C, int MODE, int WAVES, int RPT, bool NT>
__global__ void __launch_bounds__(64 * WAVES) step2_kernel(StepArgs<C> a) {
    __shared__ typename Store<S>::R lds[9][WAVES * RPT][64];
    for (int p = (int)blockIdx.x; p < a.n_blocks; p += (int)gridDim.x) {
        int tx, ty;
        if (a.xcd_rows > 1) {
            // blocks of G x H tiles per XCD: workgroups p of one chunk of 8 G H
            // map to a band of 8 G tiles by H tile rows, XCD j taking the
            // G x H block j (row by row), so that the tiles above and below
            // a tile, which read the same ring rows, share its L2 as the
            // x-neighbours do; chunks cover the tile grid row-major, tiles
            // past its edge are skipped
            const int G = a.xcd_group, H = a.xcd_rows, per = 8 * G * H;
            const int c = p / per, r = p - c * per, k = r >> 3;
            const int cx = (a.n_tiles_x + 8 * G - 1) / (8 * G);
            tx = (c % cx) * 8 * G + (r & 7) * G + k % G;
            ty = (c / cx) * H + k / G;
            if (tx >= a.n_tiles_x || ty >= a.n_tiles_y) continue;
        } else {
            const int t = tile_of_block(p, a.n_blocks, a.xcd_group);
            ty = t / a.n_tiles_x;
            tx = t - ty * a.n_tiles_x;
        }
        pair_tile<S, C, MODE, WAVES, RPT, NT>(a, tx, ty, lds);
        if (p + (int)gridDim.x < a.n_blocks) __syncthreads();  // the next tile rewrites the LDS
    }
}

// Step-1 pulls of a wave of the column march (row-blocked layout) in a strip
// away from the x edges (every lane's cell has 1 <= x <= nx - 2): nine unconditional loads
// through row addresses chosen per wave -- the source row of each direction,
// wrapped on a periodic y axis, or, beyond a wall, the cell's own row of the
// opposite direction (bounce-back; the moving-wall terms follow in
// raw_pops) -- so no branch surrounds them (march_pull).  0 <= gy < ny,
// wave-uniform; xb = x * element size.
template <typename S, typename C, int MODE>
__device__ __forceinline__ void march_pull_inner(const StepArgs<C> &a, const typename Store<S>::R *__restrict__ f,
                                                 unsigned xb, int gy, RawPop<S> (&raw)[9]) {
    using R = typename Store<S>::R;
    const ptrdiff_t P = (ptrdiff_t)a.plane, W = (ptrdiff_t)a.pitch;
    const int ny = a.ny;
    if constexpr (sizeof(R) >= 4) {
        // one descriptor per source row (one element before its direction-0
        // row, so that every offset is >= 0), the direction and the x shift
        // of the pull in the scalar offset: (d P + 1 - c_x) elements, or, for
        // a bounce-back, (opp(d) P + 1) elements of the cell's own row
        __amdgpu_buffer_rsrc_t rs[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            int sy = gy + k - 1;
            if (MODE == SW) sy = sy < 0 ? sy + ny : (sy >= ny ? sy - ny : sy);
            else sy = sy < 0 ? 0 : (sy >= ny ? ny - 1 : sy);  // out of range: not read
            rs[k] = row_rsrc(f + (ptrdiff_t)sy * W - 1);
        }
#pragma unroll
        for (int d = 0; d < 9; ++d) {
            const int sy = gy - cy_of(d);
            const bool out = MODE != SW && ((sy < 0) | (sy >= ny));
            const unsigned so = (unsigned)((out ? (ptrdiff_t)opp_of(d) * P + 1 : (ptrdiff_t)d * P + 1 - cx_of(d)) *
                                           (ptrdiff_t)sizeof(R));
            const __amdgpu_buffer_rsrc_t r = out ? rs[1] : rs[1 - cy_of(d)];
            if constexpr (sizeof(R) == 4)
                raw[d] = __builtin_bit_cast(R, __builtin_amdgcn_raw_buffer_load_b32(r, xb, so, 0));
            else
                raw[d] = __builtin_bit_cast(R, __builtin_amdgcn_raw_buffer_load_b64(r, xb, so, 0));
        }
    } else {
#pragma unroll
        for (int d = 0; d < 9; ++d) {
            int sy = gy - cy_of(d);
            const R *row;
            if (MODE == SW) {
                sy = sy < 0 ? sy + ny : (sy >= ny ? sy - ny : sy);
                row = f + d * P + (ptrdiff_t)sy * W + pull_dx(d);
            } else {
                const bool out = (sy < 0) | (sy >= ny);
                row = out ? f + opp_of(d) * P + (ptrdiff_t)gy * W : f + d * P + (ptrdiff_t)sy * W + pull_dx(d);
            }
            raw[d] = ld_row(row, xb);
        }
    }
}

// The nine stored elements step 1 of the column march pulls into region cell
// (gx, gy), issued unconditionally: every lane of the wave loads nine elements
// on either side of one wave-uniform branch, so the compiler counts them on
// every path (a load under a branch that may be skipped is counted as
// possibly absent, and a later wait for older loads then also waits for
// everything issued since).  A wave holding a domain-edge cell, or whose row
// is not needed (row_act false), takes the edge rule in every lane
// (pull_edge_cell, which gives the plain pull for an interior cell); lanes
// that compute nothing (lane_act false) and rows not needed read cell
// (0, 0), whose values are not used.  Otherwise every lane pulls through the
// row addresses.  gy and row_act are wave-uniform.
template <typename S, typename C, int MODE>
__device__ __forceinline__ void march_pull(const StepArgs<C> &a, const typename Store<S>::R *__restrict__ f, int gx,
                                           int gy, bool lane_act, bool row_act, RawPop<S> (&raw)[9]) {
    using R = typename Store<S>::R;
    const bool edge = lane_act & ((gx == 0) | (gx == a.nx - 1) | (gy == 0) | (gy == a.ny - 1));
    if (!row_act | (__builtin_amdgcn_ballot_w64(edge) != 0)) {
        pull_edge_cell<S, C, MODE>(a, f, row_act & lane_act ? gx : 0, row_act ? gy : 0, raw);
    } else {
        const ptrdiff_t P = (ptrdiff_t)a.plane, W = (ptrdiff_t)a.pitch;
        const R *fr = f + (ptrdiff_t)gy * W;  // row gy, direction 0 (wave-uniform); 1 <= gy <= ny - 2 here
        const unsigned xb = lane_act ? (unsigned)gx * (unsigned)sizeof(R) : (unsigned)sizeof(R);
#pragma unroll
        for (int d = 0; d < 9; ++d) raw[d] = ld_row(fr + d * P + pull_dy(d) * W + pull_dx(d), xb);
    }
}

// Two steps in one pass by a column march (the modes without strip passes).
// A workgroup owns a strip of 62 output columns (64 region columns with the
// one-cell ring in x) and a chunk of a.march_rows output rows [Y0, Y1), and
// walks down it in bands of B = WAVES RPT rows.  Iteration `it` computes step
// 1 of the region rows Y0 - 1 + it B + [0, B) into a ring of NR = 2B + 2 LDS
// rows, then, after one barrier, step 2 of the rows one below them
// (Y0 - 2 + it B + [0, B), within [Y0, Y1)), whose three step-1 source rows
// are in the ring: each step-1 row is computed once per chunk (the tile
// pairs recompute a ring row above and below every tile), and the ring of
// 2B + 2 rows keeps the rows that step 2 of band `it` reads apart from those
// step 1 of band it + 1 writes, so one barrier per band suffices.  The HBM
// loads of band it + 2's step 1 are issued as soon as step 1 of band it has
// run (two register sets, the loop unrolled by two), so they are in flight
// while the next one and a half bands compute.
// Workgroup p takes strip (p / 8) mod n_strips and chunk 8 ((p / 8) /
// n_strips) + p mod 8: the eight XCDs (workgroup p runs on XCD p mod 8) each
// take one row chunk of every strip, so x-neighbouring strips, which read
// each other's edge columns, share an L2.  Boundary rules, random bits and
// arithmetic are those of pair_tile, so the result is bitwise that of two
// step_kernel launches.
// One-row bands are held to 64 VGPRs (8 waves per SIMD): two 16-wave
// workgroups per CU then fit for FP32 (2 x 78.3 KB of LDS).
template <typename S, typename C, int MODE, int WAVES, int RPT>
__global__ void __launch_bounds__(64 * WAVES) __attribute__((amdgpu_waves_per_eu(RPT == 1 ? 8 : 1)))
march2_kernel(StepArgs<C> a) {
    static_assert(MODE != PF, "Poiseuille flow takes its strip pass between steps");
    using R = typename Store<S>::R;
    constexpr int B = WAVES * RPT, NR = 2 * B + 2;
    constexpr bool PX = MODE != SL, PY = MODE == SW;  // periodic axes
    // the ring, flat: direction d, ring row k, region column c at
    // 1 + (d NR + k) 64 + c; one spare element at either end keeps the pulls
    // of lanes 0 and 63 (columns -1 and 64, results dropped) inside the array
    __shared__ R ring[9 * NR * 64 + 2];
    auto lds = [&](int d, int k, int c) -> R & { return ring[1 + (d * NR + k) * 64 + c]; };
    const int p = (int)blockIdx.x, q = p >> 3;
    const int sx = q % a.n_tiles_x;
    const int Y0 = ((q / a.n_tiles_x) * 8 + (p & 7)) * a.march_rows;
    if (Y0 >= a.ny) return;  // chunks past the last row (grid rounded up to 8 chunks per strip)
    const int Y1 = min(Y0 + a.march_rows, a.ny);
    const int lane = (int)threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    const int nx = a.nx, ny = a.ny;
    const int ux = sx * 62 - 1 + lane;  // unwrapped region column
    int gx = ux;
    bool xin = (ux >= -1) & (ux <= nx);
    if (PX) gx = ux < 0 ? ux + nx : (ux >= nx ? ux - nx : ux);
    else xin &= (ux >= 0) & (ux < nx);
    const bool out_x = (lane >= 1) & (lane <= 62) & (ux < nx);
    const bool inner_x = (sx * 62 >= 1) & (sx * 62 + 62 <= nx - 1);
    const R *__restrict__ f = static_cast<const R *>(a.src);
    R *__restrict__ g = static_cast<R *>(a.dst);
    StepArgs<C> a1 = a;
    a1.write_macro = 0;
    StepArgs<C> a2 = a;
    a2.step_seed = a.step_seed + 1u;
    const int last = Y1 - Y0 + 1;                  // step-1 rows l = 0 .. last (row Y0 - 1 + l)
    const int n_it = (last + 1 + B - 1) / B;
    // step-1 row of band `it`, band row j of this wave: wrapped row and whether the lane computes it
    auto row1 = [&](int it, int j, int &gy) {
        const int l = it * B + w + WAVES * j, uy = Y0 - 1 + l;
        bool act = xin & (l <= last) & (uy >= -1) & (uy <= ny);
        gy = uy;
        if (PY) gy = uy < 0 ? uy + ny : (uy >= ny ? uy - ny : uy);
        else act &= (uy >= 0) & (uy < ny);
        return act;
    };
    // whether any lane of the wave computes step-1 row (it, j) (wave-uniform)
    auto row_on = [&](int it, int j) {
        const int l = it * B + w + WAVES * j, uy = Y0 - 1 + l;
        return (l <= last) & (PY ? (uy >= -1) & (uy <= ny) : (uy >= 0) & (uy < ny));
    };
    // strips whose region lies in 1 <= x <= nx - 2 pull without edge rules in x
    // (march_pull_inner): one copy of the march per kind, chosen per workgroup
    const bool inner_strip = (sx >= 1) & (sx * 62 + 64 <= nx) & (a.march_inner != 0);
    const unsigned xb1 = (unsigned)gx * (unsigned)sizeof(R);
    auto run = [&](auto inner_tag) {
        constexpr bool INNER = decltype(inner_tag)::value;
        // two register sets of raw step-1 inputs, bands alternating between them:
        // a set is reloaded with band it + 2 as soon as step 1 of band it has
        // consumed it, so the loads are in flight for one and a half bands, and
        // the set is never free for the compiler to reuse (a load into registers
        // that held the data of a store still in flight waits for the store)
        RawPop<S> raw_a[RPT][9], raw_b[RPT][9];
        auto load_band = [&](int it, RawPop<S>(&dst)[RPT][9]) {
#pragma unroll
            for (int j = 0; j < RPT; ++j) {
                int gy;
                const bool act = row1(it, j, gy);
                const bool on = (it < n_it) & row_on(it, j);
                if constexpr (INNER) march_pull_inner<S, C, MODE>(a1, f, xb1, on ? gy : 0, dst[j]);
                else march_pull<S, C, MODE>(a1, f, gx, gy, act, on, dst[j]);
            }
        };
        // scheduling barriers keep the loads, the dropped stores and each band's
        // stores in program order: the counts above assume that order, and the
        // scheduler would otherwise move stores ahead of loads and loads across
        // the loop entry
        load_band(0, raw_a);
        __builtin_amdgcn_sched_barrier(0);
        load_band(1, raw_b);
        __builtin_amdgcn_sched_barrier(0);
        // Dropped stores after the first bands' loads, as many as step 2 issues
        // after every later band's loads: the compiler counts outstanding memory
        // operations at the loop head as the fewest over the paths into it, and
        // without these its wait for a band's loads would also wait for the
        // stores issued after them.  Each at its own out-of-range offset
        // (identical stores would be merged into one).
        {
            const __amdgpu_buffer_rsrc_t r = row_rsrc_checked(g);
#pragma unroll
            for (int k = 0; k < 9 * RPT; ++k) {
                const unsigned o = kDropOffset + 256u * (unsigned)k;
                if constexpr (sizeof(R) == 2) __builtin_amdgcn_raw_buffer_store_b16((unsigned short)0, r, o, 0, 0);
                else if constexpr (sizeof(R) == 4) __builtin_amdgcn_raw_buffer_store_b32(0u, r, o, 0, 0);
                else {
                    typedef unsigned u2 __attribute__((ext_vector_type(2)));
                    __builtin_amdgcn_raw_buffer_store_b64(u2{0u, 0u}, r, o, 0, 0);
                }
            }
        }
        __builtin_amdgcn_sched_barrier(0);
        // band `it`: step 1 from cur into the ring, loads of band it + 2 into cur,
        // barrier, step 2
        auto band = [&](int it, RawPop<S>(&cur)[RPT][9]) {
#pragma unroll
            for (int j = 0; j < RPT; ++j) {
                int gy;
                if (row1(it, j, gy)) {
                    C v[9];
                    raw_pops<S, C, MODE>(a1, cur[j], gy, v);
                    collide_mode<S, C, MODE>(a1, v, gx, gy);
                    R o[9];
                    encode_cell<S, C>(a1, v, gx, gy, o);
                    const int slot = (it * B + w + WAVES * j) % NR;
#pragma unroll
                    for (int d = 0; d < 9; ++d) lds(d, slot, lane) = o[d];
                }
            }
            __builtin_amdgcn_sched_barrier(0);
            load_band(it + 2, cur);
            __builtin_amdgcn_sched_barrier(0);
            __syncthreads();
            // step 2: every lane of every band row computes and stores (collide_store_masked);
            // the cells outside [Y0, Y1) x [0, nx) and lanes 0, 63 are dropped
#pragma unroll
            for (int j = 0; j < RPT; ++j) {
                const int l = it * B + w + WAVES * j - 1, uy = Y0 - 1 + l;  // step-2 row, ring row l >= -1
                const bool act = out_x & (uy >= Y0) & (uy < Y1);
                const int r0 = (l + NR) % NR, rm = (l + NR - 1) % NR, rp = (l + 1) % NR;
                auto src = [&](int d) { return cy_of(d) > 0 ? rm : (cy_of(d) < 0 ? rp : r0); };
                C v[9];
                if (inner_x & (uy >= 1) & (uy <= ny - 2)) {
#pragma unroll
                    for (int d = 0; d < 9; ++d) v[d] = pop_in<S>(lds(d, src(d), lane - cx_of(d)), d);
                } else {
#pragma unroll
                    for (int d = 0; d < 9; ++d) {
                        const int sxs = ux - cx_of(d), sys = uy - cy_of(d);
                        const bool xo = (sxs < 0) | (sxs >= nx), yo = (sys < 0) | (sys >= ny);
                        const bool bb = MODE == SW ? false : (MODE == SL ? (xo | yo) : yo);
                        const R rw = bb ? lds(opp_of(d), r0, lane) : lds(d, src(d), lane - cx_of(d));
                        v[d] = pop_in<S>(rw, d);  // opp(d) stores with the offset of d
                    }
                    if (MODE == CF || MODE == SL) {  // moving top wall (pull_edge)
                        if (uy == ny - 1) {
                            v[7] = v[7] - a.corr8;
                            v[8] = v[8] + a.corr9;
                        }
                    }
                }
                collide_store_masked<S, C, MODE>(a2, v, g, ux, uy, act);
            }
            __builtin_amdgcn_sched_barrier(0);
        };
        int it = 0;
        for (; it + 1 < n_it; it += 2) {
            band(it, raw_a);
            band(it + 1, raw_b);
        }
        if (it < n_it) band(it, raw_a);

    };
    if (inner_strip) run(std::true_type{});
    else run(std::false_type{});
}

// Poiseuille strip pass (update_poiseuille_flow_macro_strips_PF,
// modes/poiseuille_flow.f90:92-167, then the one-domain macro exchange of
// exchange.f90:416-431): side 0 (inlet strip, recv_macro_left) is the pulled
// (rho, u_x, u_y) of column nx-1, side 1 (outlet, recv_macro_right) that of
// column 0, both from the halo after the first application.
template <typename S, typename C>
__global__ void __launch_bounds__(256) pf_strip_kernel(StepArgs<C> a) {
    const int k = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (k >= 2 * a.ny) return;
    const int side = k / a.ny, y = k - side * a.ny;
    const int x = side == 0 ? a.nx - 1 : 0;
    const typename Store<S>::R *f = static_cast<const typename Store<S>::R *>(a.src);
    C v[9];
#pragma unroll
    for (int d = 0; d < 9; ++d) v[d] = pull_edge<S, C, PF, 1>(a, f, d, x, y);
    C *m = a.m_new;
    if constexpr (Store<S>::dev) {
        // deviation form: the sums of collide_dev
        const C inv = a.inv_scale;
        const C a13 = v[1] + v[3], a24 = v[2] + v[4], a57 = v[5] + v[7], a68 = v[6] + v[8];
        const C b13 = v[1] - v[3], b24 = v[2] - v[4], b57 = v[5] - v[7], b68 = v[6] - v[8];
        const C rho = (C)1 + (((v[0] + a13) + a24) + (a57 + a68)) * inv;
        density_check(a, rho);
        m[(side * 3 + 0) * a.ny + y] = rho;
        m[(side * 3 + 1) * a.ny + y] = (((b13 + b57) - b68) * inv) / rho;
        m[(side * 3 + 2) * a.ny + y] = (((b24 + b57) + b68) * inv) / rho;
    } else {
        const C f1 = v[0], f2 = v[1], f3 = v[2], f4 = v[3], f5 = v[4], f6 = v[5], f7 = v[6], f8 = v[7], f9 = v[8];
        const C rho = f1 + f2 + f3 + f4 + f5 + f6 + f7 + f8 + f9;
        const C ux = f2 - f4 + f6 - f7 - f8 + f9;
        const C uy = f3 - f5 + f6 + f7 - f8 - f9;
        density_check(a, rho);
        m[(side * 3 + 0) * a.ny + y] = rho;
        m[(side * 3 + 1) * a.ny + y] = ux / rho;
        m[(side * 3 + 2) * a.ny + y] = uy / rho;
    }
}

// Equilibrium initial state (initialization.f90:9-198): f_d = w_d * rho_0 *
// (1 + 3 c.u + 4.5 (c.u)^2 - 1.5 u^2) with u = (u_x(y), 0) from a per-row
// profile (shear wave) or u = 0.
template <typename S, typename C>
__global__ void __launch_bounds__(256) init_kernel(StepArgs<C> a, C rho_0, const C *ux_row) {
    const int tiles_x = (a.nx + 255) / 256;
    const int y = (int)blockIdx.x / tiles_x;
    const int x = ((int)blockIdx.x - y * tiles_x) * 256 + (int)threadIdx.x;
    if (x >= a.nx || y >= a.ny) return;
    typename Store<S>::R *g = static_cast<typename Store<S>::R *>(a.dst);
    const size_t c = (size_t)y * (size_t)a.pitch + (size_t)x;
    const C ux = ux_row ? ux_row[y] : (C)0;
    const C uy = 0;
    const C usq = ux * ux;
#pragma unroll
    for (int d = 0; d < 9; ++d) {
        C feq;
        if (ux_row) {
            const C cu = (C)cx_of(d) * ux + (C)cy_of(d) * uy;
            feq = a.w[d] * rho_0 * ((((C)1 + (C)3 * cu) + (C)4.5 * cu * cu) - (C)1.5 * usq);
        } else {
            feq = a.w[d] * rho_0;
        }
        st_pop<S>(g, (size_t)d * (size_t)a.plane + c, Store<S>::from_f(feq, a.w[d], a.scale), d);
    }
    if (a.rho_o) {
        const size_t m = (size_t)y * (size_t)a.nx + (size_t)x;
        a.rho_o[m] = rho_0;
        a.ux_o[m] = ux;
        a.uy_o[m] = uy;
    }
}

// FP64 sums of mass and momentum over the owned cells: one partial per block
// (deterministic order), reduced by moments_final_kernel.
template <typename S, typename C>
__global__ void __launch_bounds__(256) moments_partial_kernel(StepArgs<C> a, double *partial) {
    __shared__ double sh[3][256];
    double m = 0, mx = 0, my = 0;
    const long long n = (long long)a.nx * a.ny;
    const typename Store<S>::R *f = static_cast<const typename Store<S>::R *>(a.src);
    for (long long k = (long long)blockIdx.x * 256 + threadIdx.x; k < n; k += (long long)gridDim.x * 256) {
        const int y = (int)(k / a.nx), x = (int)(k - (long long)y * a.nx);
        const size_t c = (size_t)y * (size_t)a.pitch + (size_t)x;
#pragma unroll
        for (int d = 0; d < 9; ++d) {
            double v;
            if (Store<S>::dev)
                v = (double)a.w[d] + (double)ld_pop<S>(f, (size_t)d * (size_t)a.plane + c, d) * (double)a.inv_scale;
            else
                v = (double)ld_pop<S>(f, (size_t)d * (size_t)a.plane + c, d);
            m += v;
            mx += (double)cx_of(d) * v;
            my += (double)cy_of(d) * v;
        }
    }
    sh[0][threadIdx.x] = m;
    sh[1][threadIdx.x] = mx;
    sh[2][threadIdx.x] = my;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + s];
            sh[1][threadIdx.x] += sh[1][threadIdx.x + s];
            sh[2][threadIdx.x] += sh[2][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[3 * blockIdx.x + 0] = sh[0][0];
        partial[3 * blockIdx.x + 1] = sh[1][0];
        partial[3 * blockIdx.x + 2] = sh[2][0];
    }
}

// Largest |f - w| of the owned cells per direction class (0 rest, 1 axis,
// 2 diagonal), in FP64: one partial per block, reduced by extent_final_kernel.
// The deviation forms give |D| 2^-s of the stored value as the step reads it.
__host__ __device__ constexpr int dir_class(int d) { return d == 0 ? 0 : (d <= 4 ? 1 : 2); }
template <typename S, typename C>
__global__ void __launch_bounds__(256) extent_partial_kernel(StepArgs<C> a, double *partial) {
    __shared__ double sh[3][256];
    double mx[3] = {0.0, 0.0, 0.0};
    const long long n = (long long)a.nx * a.ny;
    const typename Store<S>::R *f = static_cast<const typename Store<S>::R *>(a.src);
    for (long long k = (long long)blockIdx.x * 256 + threadIdx.x; k < n; k += (long long)gridDim.x * 256) {
        const int y = (int)(k / a.nx), x = (int)(k - (long long)y * a.nx);
        const size_t c = (size_t)y * (size_t)a.pitch + (size_t)x;
#pragma unroll
        for (int d = 0; d < 9; ++d) {
            const double v = (double)ld_pop<S>(f, (size_t)d * (size_t)a.plane + c, d);
            const double dev = Store<S>::dev ? fabs(v * (double)a.inv_scale) : fabs(v - (double)a.w[d]);
            mx[dir_class(d)] = fmax(mx[dir_class(d)], dev);
        }
    }
#pragma unroll
    for (int j = 0; j < 3; ++j) sh[j][threadIdx.x] = mx[j];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
#pragma unroll
            for (int j = 0; j < 3; ++j) sh[j][threadIdx.x] = fmax(sh[j][threadIdx.x], sh[j][threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x < 3) partial[3 * blockIdx.x + threadIdx.x] = sh[threadIdx.x][0];
}

}  // namespace ffb
