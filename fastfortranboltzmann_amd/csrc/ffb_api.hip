// ffb_api.hip -- C ABI (include/ffb.h) of the single-domain D2Q9 solver:
// handle, device buffers, stream, launches and host conversions.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "d2q9_kernels.h"
#include "ffb.h"
#include "ffb_internal.h"

using namespace ffb;

namespace ffb {
thread_local std::string g_last_error;

int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}
}  // namespace ffb

#define HIPCHK(expr)                                                                           \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return ffb::fail(FFB_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));    \
    } while (0)



// cells per thread of the step kernel
static int cells_per_thread(const ffb_lattice *L) { return L->vec > 1 ? L->vec : 1; }

namespace {

// Second pass of the FP64 moment sums: reduces the per-block partials.
__global__ void __launch_bounds__(256) moments_final_kernel(const double *partial, int n, double *out) {
    __shared__ double sh[3][256];
    double m = 0, mx = 0, my = 0;
    for (int k = (int)threadIdx.x; k < n; k += 256) {
        m += partial[3 * k];
        mx += partial[3 * k + 1];
        my += partial[3 * k + 2];
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
        out[0] = sh[0][0];
        out[1] = sh[1][0];
        out[2] = sh[2][0];
    }
}


// D2Q9 weight to double precision (the deviation forms' collision factors)
inline double exact_weight(int d) { return d == 0 ? 4.0 / 9.0 : (d <= 4 ? 1.0 / 9.0 : 1.0 / 36.0); }

inline bool is_16bit_form(int p) {
    return p == FFB_PREC_FP16_DEV || p == FFB_PREC_BF16_DEV || p == FFB_PREC_E3M12_DEV;
}
inline bool is_deviation_form(int p) { return is_16bit_form(p) || p == FFB_PREC_FP32_DEV; }
// Exponent of the scale the step works in: 2^s, except the 1-3-12 form,
// whose kernels work on D * 2^-120 (the raw FP32 pattern of the stored value).
inline int work_scale_log2(const ffb_config_t &c) {
    return c.precision == FFB_PREC_E3M12_DEV ? c.storage_scale_log2 - 120 : c.storage_scale_log2;
}

template <typename C> StepArgs<C> make_args(const ffb_lattice *L, int write_macro) {
    StepArgs<C> a{};
    const ffb_config_t &c = L->cfg;
    a.src = L->buf[L->cur];
    a.dst = L->buf[1 - L->cur];
    a.nx = c.n_x;
    a.ny = c.n_y;
    a.pitch = L->pitch;
    a.plane = L->plane;
    a.omega = (C)c.omega;
    const C rho0 = (C)c.rho_0;
    const C w6 = (C)1 / (C)36;
    // ((6*w_6)*rho_0)*u, couette_flow.f90:118-119, sliding_lid.f90:195-196
    const C u = (C)(c.sim_mode == FFB_COUETTE_FLOW ? c.u_wall : c.u_lid);
    a.corr8 = (C)6 * w6 * rho0 * u;
    a.corr9 = (C)6 * w6 * rho0 * u;
    a.rho_in = (C)c.rho_in;
    a.rho_out = (C)c.rho_out;
    a.m_old = static_cast<const C *>(L->strips[L->strip_old]);
    a.m_new = static_cast<C *>(L->strips[1 - L->strip_old]);
    a.rho_o = static_cast<C *>(L->macro[0]);
    a.ux_o = static_cast<C *>(L->macro[1]);
    a.uy_o = static_cast<C *>(L->macro[2]);
    a.write_macro = write_macro;
    a.w[0] = (C)4 / (C)9;
    for (int d = 1; d <= 4; ++d) a.w[d] = (C)1 / (C)9;
    for (int d = 5; d <= 8; ++d) a.w[d] = (C)1 / (C)36;
    const bool dev = is_deviation_form(c.precision);
    const int s = dev ? work_scale_log2(c) : 0;
    a.scale = (C)std::ldexp(1.0, s);
    a.inv_scale = (C)std::ldexp(1.0, -s);
    const int cells_per_tile = L->tile_w * cells_per_thread(L);
    a.n_tiles_x = (c.n_x + cells_per_tile - 1) / cells_per_tile;
    a.force = (C)c.body_force;
    if (dev) {  // the deviation forms work on deviations times 2^s (collide_dev):
                // ((6*w_6)*rho_0)*(u*2^s), exact wherever u*2^s is an FP32 normal
        a.corr8 = (C)6 * w6 * rho0 * (u * a.scale);
        a.corr9 = a.corr8;
    }
    a.om1 = (C)1 - a.omega;
    const C pre = ((C)1 - (C)0.5 * a.omega) * a.force;
    for (int d = 0; d < 9; ++d) {
        const double kd = (double)a.omega * exact_weight(d) * (double)a.scale;
        a.kw[d] = (C)kd;
        a.kw_lo[d] = (C)(kd - (double)a.kw[d]);
        a.kf[d] = (a.w[d] * pre) * a.scale;
    }
    a.half_force = (C)0.5 * a.force;
    a.stochastic = is_16bit_form(c.precision) && c.storage_rounding == FFB_ROUND_STOCHASTIC;
    a.step_seed = (unsigned)(L->steps + 1);  // the step this launch computes
    a.n_blocks = a.n_tiles_x * c.n_y;
    a.xcd_group = L->xcd_group;
    a.check_rho = c.density_checks != 0;
    a.status = L->status;
    return a;
}

template <typename S, typename C, int MODE, int TW, int V>
void launch_v(ffb_lattice *L, const StepArgs<C> &a) {
    hipLaunchKernelGGL((step_kernel<S, C, MODE, TW, V>), dim3((unsigned)a.n_blocks), dim3(TW), 0, L->stream, a);
}

template <typename S, typename C, int MODE, int TW>
int launch_step_tw(ffb_lattice *L, const StepArgs<C> &a) {
    // cells per thread: 1, or 2/4/8 with one aligned vector load per plane
    // (at most 16 bytes); wider tiles only for the 256- and 512-thread shapes
    constexpr int RB = (int)sizeof(typename Store<S>::R);
    const int v = cells_per_thread(L);
    if constexpr (TW >= 256) {
        if (v == 2 && 2 * RB <= 16) { launch_v<S, C, MODE, TW, 2>(L, a); HIPCHK(hipGetLastError()); return FFB_OK; }
        if constexpr (4 * RB <= 16)
            if (v == 4) { launch_v<S, C, MODE, TW, 4>(L, a); HIPCHK(hipGetLastError()); return FFB_OK; }
        if constexpr (8 * RB <= 16)
            if (v == 8) { launch_v<S, C, MODE, TW, 8>(L, a); HIPCHK(hipGetLastError()); return FFB_OK; }
    }
    launch_v<S, C, MODE, TW, 1>(L, a);
    HIPCHK(hipGetLastError());
    return FFB_OK;
}

template <typename S, typename C, int MODE>
int launch_step_mode(ffb_lattice *L, const StepArgs<C> &a) {
    switch (L->tile_w) {
    case 64: return launch_step_tw<S, C, MODE, 64>(L, a);
    case 128: return launch_step_tw<S, C, MODE, 128>(L, a);
    case 512: return launch_step_tw<S, C, MODE, 512>(L, a);
    default: return launch_step_tw<S, C, MODE, 256>(L, a);
    }
}

template <typename S, typename C> int launch_step(ffb_lattice *L, int write_macro) {
    StepArgs<C> a = make_args<C>(L, write_macro);
    switch (L->cfg.sim_mode) {
    case FFB_SHEAR_WAVE: return launch_step_mode<S, C, SW>(L, a);
    case FFB_COUETTE_FLOW: return launch_step_mode<S, C, CF>(L, a);
    case FFB_BODY_FORCE_CHANNEL: return launch_step_mode<S, C, BF>(L, a);
    case FFB_POISEUILLE_FLOW: {
        const unsigned nb = (unsigned)((2 * a.ny + 255) / 256);
        hipLaunchKernelGGL((pf_strip_kernel<S, C>), dim3(nb), dim3(256), 0, L->stream, a);
        HIPCHK(hipGetLastError());
        return launch_step_mode<S, C, PF>(L, a);
    }
    default: return launch_step_mode<S, C, SL>(L, a);
    }
}

// One step from src to dst (not the handle's ping-pong buffers), any mode but
// Poiseuille flow (whose strips live in the handle).
template <typename S, typename C>
int launch_between(ffb_lattice *L, const void *src, void *dst, int write_macro) {
    StepArgs<C> a = make_args<C>(L, write_macro);
    a.src = src;
    a.dst = dst;
    switch (L->cfg.sim_mode) {
    case FFB_SHEAR_WAVE: return launch_step_mode<S, C, SW>(L, a);
    case FFB_COUETTE_FLOW: return launch_step_mode<S, C, CF>(L, a);
    case FFB_BODY_FORCE_CHANNEL: return launch_step_mode<S, C, BF>(L, a);
    case FFB_SLIDING_LID: return launch_step_mode<S, C, SL>(L, a);
    default: return fail(FFB_E_INVALID, "step between buffers: Poiseuille flow is not supported");
    }
}

// One K-step launch over a.n_blocks tiles: one workgroup per tile
// (stepk_kernel), or with L->multi_pipe a persistent grid that loads each
// workgroup's next tile during the LDS-only steps of the current one
// (stepk_pipe_kernel): L->multi_grid workgroups, or as many as fit on the
// device at once (occupancy x compute units, rounded up to a multiple of 8).
template <typename S, typename C, int MODE, int WAVES, int RPT, int K, bool PIPE_OK = true>
int launch_multi(ffb_lattice *L, const StepArgs<C> &a) {
    const dim3 block(64 * WAVES);
    if (PIPE_OK && L->multi_pipe) {
        static int per_cu = -1;  // resident workgroups per CU, one query per kernel
        if (per_cu < 0) {
            int n = 0;
            HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &n, reinterpret_cast<const void *>(&stepk_pipe_kernel<S, C, MODE, WAVES, RPT, K>), 64 * WAVES, 0));
            per_cu = std::max(1, n);
        }
        // an explicit L->multi_grid is taken as given (any size is correct; a
        // multiple of 8 keeps each workgroup's tiles on its XCD's group)
        const int g = std::min(a.n_blocks, L->multi_grid > 0 ? L->multi_grid : ((per_cu * L->n_cu + 7) / 8) * 8);
        hipLaunchKernelGGL((stepk_pipe_kernel<S, C, MODE, WAVES, RPT, K>), dim3((unsigned)g), block, L->lds_pad,
                           L->stream, a);
    } else {
        hipLaunchKernelGGL((stepk_kernel<S, C, MODE, WAVES, RPT, K>), dim3((unsigned)a.n_blocks), block, L->lds_pad,
                           L->stream, a);
    }
    HIPCHK(hipGetLastError());
    return FFB_OK;
}

// K steps in one launch over (64 - 2(K-1)) x (WAVES RPT - 2(K-1)) tiles
// (stepk_kernel); PIPE_OK: the persistent variant can be selected.
template <typename S, typename C, int WAVES, int RPT, int K, bool PIPE_OK = true>
int launch_multik_shape(ffb_lattice *L, StepArgs<C> a) {
    constexpr int OW = 64 - 2 * (K - 1), OH = WAVES * RPT - 2 * (K - 1);
    a.n_tiles_x = (a.nx + OW - 1) / OW;
    a.n_tiles_y = (a.ny + OH - 1) / OH;
    a.n_blocks = a.n_tiles_x * a.n_tiles_y;
    a.xcd_group = L->pair_xcd_group;
    switch (L->cfg.sim_mode) {
    case FFB_SHEAR_WAVE: return launch_multi<S, C, SW, WAVES, RPT, K, PIPE_OK>(L, a);
    case FFB_COUETTE_FLOW: return launch_multi<S, C, CF, WAVES, RPT, K, PIPE_OK>(L, a);
    case FFB_BODY_FORCE_CHANNEL: return launch_multi<S, C, BF, WAVES, RPT, K, PIPE_OK>(L, a);
    case FFB_SLIDING_LID: return launch_multi<S, C, SL, WAVES, RPT, K, PIPE_OK>(L, a);
    default: return fail(FFB_E_INVALID, "multi-step pass: Poiseuille flow is not supported");
    }
}

// Two steps in one launch (step2_kernel): tiles of 62 x (WAVES RPT - 2) cells.
template <typename S, typename C, int WAVES, int RPT, bool NT> int launch_step2_shape(ffb_lattice *L, StepArgs<C> a) {
    constexpr int ROWS = WAVES * RPT;
    a.n_tiles_x = (a.nx + 61) / 62;
    a.n_tiles_y = (a.ny + ROWS - 3) / (ROWS - 2);
    a.n_blocks = a.n_tiles_x * a.n_tiles_y;
    a.xcd_group = L->pair_xcd_group;
    a.xcd_rows = a.xcd_group > 0 ? L->pair_xcd_rows : 1;
    if (a.xcd_rows > 1) {  // whole chunks of 8 G x H tiles
        const long long G = a.xcd_group, H = a.xcd_rows;
        const long long n = ((a.n_tiles_x + 8 * G - 1) / (8 * G)) * ((a.n_tiles_y + H - 1) / H) * 8 * G * H;
        if (n > 0x7fffffffLL) return fail(FFB_E_INVALID, "two-step pass: too many tiles for FFB_PAIR_XCD_ROWS");
        a.n_blocks = (int)n;
    }
    // persistent grid of L->pair_grid workgroups (rounded up to a multiple of 8), or one per tile
    const int g = L->pair_grid > 0 ? std::min(a.n_blocks, ((L->pair_grid + 7) / 8) * 8) : a.n_blocks;
    const dim3 grid((unsigned)g), block(64 * WAVES);
    switch (L->cfg.sim_mode) {
    case FFB_SHEAR_WAVE: hipLaunchKernelGGL((step2_kernel<S, C, SW, WAVES, RPT, NT>), grid, block, 0, L->stream, a); break;
    case FFB_COUETTE_FLOW: hipLaunchKernelGGL((step2_kernel<S, C, CF, WAVES, RPT, NT>), grid, block, 0, L->stream, a); break;
    case FFB_BODY_FORCE_CHANNEL: hipLaunchKernelGGL((step2_kernel<S, C, BF, WAVES, RPT, NT>), grid, block, 0, L->stream, a); break;
    case FFB_SLIDING_LID: hipLaunchKernelGGL((step2_kernel<S, C, SL, WAVES, RPT, NT>), grid, block, 0, L->stream, a); break;
    default: return fail(FFB_E_INVALID, "two-step pass: Poiseuille flow is not supported");
    }
    HIPCHK(hipGetLastError());
    return FFB_OK;
}

// Two steps in one launch by column march (march2_kernel): strips of 62
// columns, chunks of L->march_rows rows, eight chunks of every strip per
// group of eight workgroups (one chunk per XCD).
template <typename S, typename C, int WAVES, int RPT> int launch_march_shape(ffb_lattice *L, StepArgs<C> a) {
    a.n_tiles_x = (a.nx + 61) / 62;
    a.march_rows = L->march_rows;
    a.march_inner = L->march_inner;
    const long long chunks = (a.ny + a.march_rows - 1) / a.march_rows;
    const long long g = (long long)a.n_tiles_x * (((chunks + 7) / 8) * 8);
    if (g > 0x7fffffffLL) return fail(FFB_E_INVALID, "column march: grid too large (raise FFB_MARCH_ROWS)");
    a.n_blocks = (int)g;
    const dim3 grid((unsigned)g), block(64 * WAVES);
    switch (L->cfg.sim_mode) {
    case FFB_SHEAR_WAVE: hipLaunchKernelGGL((march2_kernel<S, C, SW, WAVES, RPT>), grid, block, 0, L->stream, a); break;
    case FFB_COUETTE_FLOW: hipLaunchKernelGGL((march2_kernel<S, C, CF, WAVES, RPT>), grid, block, 0, L->stream, a); break;
    case FFB_BODY_FORCE_CHANNEL: hipLaunchKernelGGL((march2_kernel<S, C, BF, WAVES, RPT>), grid, block, 0, L->stream, a); break;
    case FFB_SLIDING_LID: hipLaunchKernelGGL((march2_kernel<S, C, SL, WAVES, RPT>), grid, block, 0, L->stream, a); break;
    default: return fail(FFB_E_INVALID, "two-step pass: Poiseuille flow is not supported");
    }
    HIPCHK(hipGetLastError());
    return FFB_OK;
}

template <typename S, typename C> int launch_march(ffb_lattice *L, const StepArgs<C> &a) {
    switch (L->march_shape) {
    case 2: return launch_march_shape<S, C, 8, 2>(L, a);
    case 3: return launch_march_shape<S, C, 16, 1>(L, a);
    case 4: return launch_march_shape<S, C, 4, 2>(L, a);
    default: return launch_march_shape<S, C, 8, 1>(L, a);
    }
}

// Region shapes, measured on the 4096^2 and 8192^2 cavities with 4 tiles per
// XCD group: 4-byte and smaller elements 16 waves x 2 rows (64 x 32 cells, LDS
// 73.7 KB for FP32; FP32 4096^2 150.5-152.1 us per step against 153.4-154.2
// with one row per thread, same box), FP64 8 waves x 1 row (64 x 8, 36.9 KB;
// 219-220 us per step, against 263 with 16 waves x 1 row on another box).
template <typename S, typename C> int launch_step2(ffb_lattice *L, int write_macro) {
    const StepArgs<C> a = make_args<C>(L, write_macro);
    if (L->pair_kernel == 1 && L->layout == 1) return launch_march<S, C>(L, a);  // row-blocked layout only
    const bool nt = L->pair_nt != 0;
    if constexpr (sizeof(typename Store<S>::R) <= 4) {
        // 8 waves x 4 rows (64 x 32, 62 x 30 tiles) by stepk_kernel
        if (L->pair_shape == 3) return launch_multik_shape<S, C, 8, 4, 2, false>(L, a);
        if (L->pair_shape == 2)
            return nt ? launch_step2_shape<S, C, 8, 2, true>(L, a) : launch_step2_shape<S, C, 8, 2, false>(L, a);
        return nt ? launch_step2_shape<S, C, 16, 2, true>(L, a) : launch_step2_shape<S, C, 16, 2, false>(L, a);
    } else {
        return nt ? launch_step2_shape<S, C, 8, 1, true>(L, a) : launch_step2_shape<S, C, 8, 1, false>(L, a);
    }
}

// Three steps in one launch (stepk_kernel): tiles of 60 x (WAVES RPT - 4) cells.
template <typename S, typename C, int WAVES, int RPT> int launch_step3_shape(ffb_lattice *L, StepArgs<C> a) {
    constexpr int ROWS = WAVES * RPT;
    a.n_tiles_x = (a.nx + 59) / 60;
    a.n_tiles_y = (a.ny + ROWS - 5) / (ROWS - 4);
    a.n_blocks = a.n_tiles_x * a.n_tiles_y;
    a.xcd_group = L->pair_xcd_group;
    switch (L->cfg.sim_mode) {
    case FFB_SHEAR_WAVE: return launch_multi<S, C, SW, WAVES, RPT, 3>(L, a);
    case FFB_COUETTE_FLOW: return launch_multi<S, C, CF, WAVES, RPT, 3>(L, a);
    case FFB_BODY_FORCE_CHANNEL: return launch_multi<S, C, BF, WAVES, RPT, 3>(L, a);
    case FFB_SLIDING_LID: return launch_multi<S, C, SL, WAVES, RPT, 3>(L, a);
    default: return fail(FFB_E_INVALID, "three-step pass: Poiseuille flow is not supported");
    }
}

// Region shapes of the triples: 2 = 8 waves x 2 rows (64 x 16 cells, 60 x 12
// outputs), 1 = 16 waves x 2 rows (64 x 32, 60 x 28; FP64 73.7 KB of LDS),
// 3 = 8 waves x 4 rows (64 x 32, 60 x 28; 4-byte and narrower elements).
template <typename S, typename C> int launch_step3(ffb_lattice *L, int write_macro) {
    const StepArgs<C> a = make_args<C>(L, write_macro);
    if (L->triple_shape == 1) return launch_step3_shape<S, C, 16, 2>(L, a);
    if constexpr (sizeof(typename Store<S>::R) <= 4) {
        if (L->triple_shape == 3) return launch_multik_shape<S, C, 8, 4, 3, false>(L, a);  // 64 x 32, 60 x 28 tiles
    }
    return launch_step3_shape<S, C, 8, 2>(L, a);
}

// Four steps in one launch (stepk_kernel): tiles of 58 x (WAVES RPT - 6) cells;
// regions 2 = 12 waves x 2 rows (64 x 24, 58 x 18 tiles, the default), 1 = 16
// waves x 2 rows, 3 = 8 waves x 2 rows.
template <typename S, typename C, int WAVES, int RPT> int launch_step4_shape(ffb_lattice *L, StepArgs<C> a) {
    constexpr int ROWS = WAVES * RPT;
    a.n_tiles_x = (a.nx + 57) / 58;
    a.n_tiles_y = (a.ny + ROWS - 7) / (ROWS - 6);
    a.n_blocks = a.n_tiles_x * a.n_tiles_y;
    a.xcd_group = L->pair_xcd_group;
    switch (L->cfg.sim_mode) {
    case FFB_SHEAR_WAVE: return launch_multi<S, C, SW, WAVES, RPT, 4>(L, a);
    case FFB_COUETTE_FLOW: return launch_multi<S, C, CF, WAVES, RPT, 4>(L, a);
    case FFB_BODY_FORCE_CHANNEL: return launch_multi<S, C, BF, WAVES, RPT, 4>(L, a);
    case FFB_SLIDING_LID: return launch_multi<S, C, SL, WAVES, RPT, 4>(L, a);
    default: return fail(FFB_E_INVALID, "four-step pass: Poiseuille flow is not supported");
    }
}

template <typename S, typename C> int launch_step4(ffb_lattice *L, int write_macro) {
    const StepArgs<C> a = make_args<C>(L, write_macro);
    if (L->quad_shape == 1) return launch_step4_shape<S, C, 16, 2>(L, a);
    if (L->quad_shape == 3) return launch_step4_shape<S, C, 8, 2>(L, a);
    if constexpr (sizeof(typename Store<S>::R) <= 4) {
        if (L->quad_shape == 4) return launch_multik_shape<S, C, 8, 4, 4, false>(L, a);  // 64 x 32, 58 x 26 tiles
    }
    return launch_step4_shape<S, C, 12, 2>(L, a);
}

// Five steps in one launch (4-byte elements): tiles of 56 x (WAVES RPT - 8)
// cells; regions 1 = 8 waves x 4 rows (64 x 32, 56 x 24 tiles, the default),
// 2 = 12 waves x 2 rows (64 x 24, 56 x 16).
template <typename S, typename C, int WAVES, int RPT> int launch_step5_shape(ffb_lattice *L, StepArgs<C> a) {
    constexpr int ROWS = WAVES * RPT;
    a.n_tiles_x = (a.nx + 55) / 56;
    a.n_tiles_y = (a.ny + ROWS - 9) / (ROWS - 8);
    a.n_blocks = a.n_tiles_x * a.n_tiles_y;
    a.xcd_group = L->pair_xcd_group;
    switch (L->cfg.sim_mode) {
    case FFB_SHEAR_WAVE: return launch_multi<S, C, SW, WAVES, RPT, 5>(L, a);
    case FFB_COUETTE_FLOW: return launch_multi<S, C, CF, WAVES, RPT, 5>(L, a);
    case FFB_BODY_FORCE_CHANNEL: return launch_multi<S, C, BF, WAVES, RPT, 5>(L, a);
    case FFB_SLIDING_LID: return launch_multi<S, C, SL, WAVES, RPT, 5>(L, a);
    default: return fail(FFB_E_INVALID, "five-step pass: Poiseuille flow is not supported");
    }
}

int dispatch_step5(ffb_lattice *L, int write_macro) {
    if (L->cfg.precision == FFB_PREC_FP32_DEV) {
        const StepArgs<float> a = make_args<float>(L, write_macro);
        return L->penta_shape == 2 ? launch_step5_shape<f32dev, float, 12, 2>(L, a)
                                   : launch_step5_shape<f32dev, float, 8, 4>(L, a);
    }
    const StepArgs<float> a = make_args<float>(L, write_macro);
    return L->penta_shape == 2 ? launch_step5_shape<float, float, 12, 2>(L, a)
                               : launch_step5_shape<float, float, 8, 4>(L, a);
}

int dispatch_step4(ffb_lattice *L, int write_macro) {
    switch (L->cfg.precision) {
    case FFB_PREC_FP64: return launch_step4<double, double>(L, write_macro);
    case FFB_PREC_FP16_DEV: return launch_step4<__half, float>(L, write_macro);
    case FFB_PREC_FP32_DEV: return launch_step4<f32dev, float>(L, write_macro);
    default: return launch_step4<float, float>(L, write_macro);
    }
}

int dispatch_step3(ffb_lattice *L, int write_macro) {
    switch (L->cfg.precision) {
    case FFB_PREC_FP64: return launch_step3<double, double>(L, write_macro);
    case FFB_PREC_FP16_DEV: return launch_step3<__half, float>(L, write_macro);
    case FFB_PREC_BF16_DEV: return launch_step3<__hip_bfloat16, float>(L, write_macro);
    case FFB_PREC_FP32_DEV: return launch_step3<f32dev, float>(L, write_macro);
    case FFB_PREC_E3M12_DEV: return launch_step3<e3m12, float>(L, write_macro);
    default: return launch_step3<float, float>(L, write_macro);
    }
}

int dispatch_step2(ffb_lattice *L, int write_macro) {
    switch (L->cfg.precision) {
    case FFB_PREC_FP64: return launch_step2<double, double>(L, write_macro);
    case FFB_PREC_FP16_DEV: return launch_step2<__half, float>(L, write_macro);
    case FFB_PREC_BF16_DEV: return launch_step2<__hip_bfloat16, float>(L, write_macro);
    case FFB_PREC_FP32_DEV: return launch_step2<f32dev, float>(L, write_macro);
    case FFB_PREC_E3M12_DEV: return launch_step2<e3m12, float>(L, write_macro);
    default: return launch_step2<float, float>(L, write_macro);
    }
}

int dispatch_step(ffb_lattice *L, int write_macro) {
    switch (L->cfg.precision) {
    case FFB_PREC_FP64: return launch_step<double, double>(L, write_macro);
    case FFB_PREC_FP16_DEV: return launch_step<__half, float>(L, write_macro);
    case FFB_PREC_BF16_DEV: return launch_step<__hip_bfloat16, float>(L, write_macro);
    case FFB_PREC_FP32_DEV: return launch_step<f32dev, float>(L, write_macro);
    case FFB_PREC_E3M12_DEV: return launch_step<e3m12, float>(L, write_macro);
    default: return launch_step<float, float>(L, write_macro);
    }
}

template <typename S, typename C> int launch_init(ffb_lattice *L) {
    StepArgs<C> a = make_args<C>(L, 0);
    a.dst = L->buf[0];
    const ffb_config_t &c = L->cfg;
    const C *ux_row = nullptr;
    if (c.sim_mode == FFB_SHEAR_WAVE) {
        // u_x(y) = U_MAX * sin((2*PI*N_SIN*(y-1)) / N_Y), initialization.f90:42,
        // evaluated on the host with the C library's sin in C precision.
        std::vector<C> row((size_t)c.n_y);
        const C pi = (C)3.141592653589793238462643383279502884197L;
        for (int y = 0; y < c.n_y; ++y) {
            const C arg = (((C)2 * pi) * (C)c.n_sin * (C)y) / (C)c.n_y;
            row[(size_t)y] = (C)c.u_max * std::sin(arg);
        }
        HIPCHK(hipMemcpyAsync(L->ux_row, row.data(), row.size() * sizeof(C), hipMemcpyHostToDevice, L->stream));
        HIPCHK(hipStreamSynchronize(L->stream));
        ux_row = static_cast<const C *>(L->ux_row);
    }
    const dim3 grid((unsigned)(((c.n_x + 255) / 256) * c.n_y));
    hipLaunchKernelGGL((init_kernel<S, C>), grid, dim3(256), 0, L->stream, a, (C)c.rho_0, ux_row);
    HIPCHK(hipGetLastError());
    // Poiseuille strip seeds (main.f90:139-155): rho_in / rho_out, zero velocity.
    if (c.sim_mode == FFB_POISEUILLE_FLOW) {
        std::vector<C> seed((size_t)6 * (size_t)c.n_y, (C)0);
        for (int y = 0; y < c.n_y; ++y) {
            seed[(size_t)y] = (C)c.rho_in;
            seed[(size_t)3 * c.n_y + (size_t)y] = (C)c.rho_out;
        }
        for (int k = 0; k < 2; ++k)
            HIPCHK(hipMemcpyAsync(L->strips[k], seed.data(), seed.size() * sizeof(C), hipMemcpyHostToDevice, L->stream));
        HIPCHK(hipStreamSynchronize(L->stream));
    }
    return FFB_OK;
}

int dispatch_init(ffb_lattice *L) {
    switch (L->cfg.precision) {
    case FFB_PREC_FP64: return launch_init<double, double>(L);
    case FFB_PREC_FP16_DEV: return launch_init<__half, float>(L);
    case FFB_PREC_BF16_DEV: return launch_init<__hip_bfloat16, float>(L);
    case FFB_PREC_FP32_DEV: return launch_init<f32dev, float>(L);
    case FFB_PREC_E3M12_DEV: return launch_init<e3m12, float>(L);
    default: return launch_init<float, float>(L);
    }
}

__global__ void __launch_bounds__(256) extent_final_kernel(const double *partial, int n, double *out) {
    __shared__ double sh[3][256];
    double m[3] = {0.0, 0.0, 0.0};
    for (int k = (int)threadIdx.x; k < n; k += 256)
        for (int j = 0; j < 3; ++j) m[j] = fmax(m[j], partial[3 * k + j]);
    for (int j = 0; j < 3; ++j) sh[j][threadIdx.x] = m[j];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int j = 0; j < 3; ++j) sh[j][threadIdx.x] = fmax(sh[j][threadIdx.x], sh[j][threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x < 3) out[threadIdx.x] = sh[threadIdx.x][0];
}

template <typename S, typename C> int launch_extent(ffb_lattice *L) {
    StepArgs<C> a = make_args<C>(L, 0);
    hipLaunchKernelGGL((extent_partial_kernel<S, C>), dim3((unsigned)L->mom_blocks), dim3(256), 0, L->stream, a,
                       L->mom_partial);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(extent_final_kernel, dim3(1), dim3(256), 0, L->stream, L->mom_partial, L->mom_blocks,
                       L->mom_out);
    HIPCHK(hipGetLastError());
    return FFB_OK;
}

template <typename S, typename C> int launch_moments(ffb_lattice *L) {
    StepArgs<C> a = make_args<C>(L, 0);
    hipLaunchKernelGGL((moments_partial_kernel<S, C>), dim3((unsigned)L->mom_blocks), dim3(256), 0, L->stream, a,
                       L->mom_partial);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(moments_final_kernel, dim3(1), dim3(256), 0, L->stream, L->mom_partial, L->mom_blocks,
                       L->mom_out);
    HIPCHK(hipGetLastError());
    return FFB_OK;
}

// host-side conversions between the stored form and the reference's values
float half_to_float(uint16_t h) {
    __half v;
    std::memcpy(&v, &h, 2);
    return __half2float(v);
}
uint16_t float_to_half(float f) {
    __half v = __float2half_rn(f);
    uint16_t h;
    std::memcpy(&h, &v, 2);
    return h;
}
float bf16_to_float(uint16_t b) {
    uint32_t u = (uint32_t)b << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}
// 1-3-12 form: stored bits <-> the FP32 value D * 2^-120 (as Store<e3m12>)
float e3m12_to_float(uint16_t h) {
    const uint32_t u = ((uint32_t)(int32_t)(int16_t)h << 11) & 0x83ffffffu;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}
uint16_t float_to_e3m12(float v) {  // to nearest, ties to even; clamped to the largest magnitude
    const uint32_t max_bits = 0x03fff800u;
    float m;
    std::memcpy(&m, &max_bits, 4);
    v = v > m ? m : (v < -m ? -m : v);
    uint32_t u;
    std::memcpy(&u, &v, 4);
    u += 0x3ffu + ((u >> 11) & 1u);
    return (uint16_t)(((u >> 11) & 0x7fffu) | ((u >> 16) & 0x8000u));
}
uint16_t float_to_bf16(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);  // quiet NaN
    const uint32_t lsb = (u >> 16) & 1u;
    u += 0x7fffu + lsb;  // round to nearest, ties to even
    return (uint16_t)(u >> 16);
}

typedef float v4f __attribute__((ext_vector_type(4)));

// Copy probes for the measured HBM ceiling.  VARIANT 0: grid-stride loop;
// 1: one 16-byte element per thread, plain store; 2: same with a nontemporal
// store; 3: four 16-byte elements per thread (strided by the block), nontemporal.
template <int VARIANT>
__global__ void __launch_bounds__(256) copy_kernel(const v4f *__restrict__ src, v4f *__restrict__ dst, long long n) {
    if (VARIANT == 0) {
        for (long long k = (long long)blockIdx.x * 256 + threadIdx.x; k < n; k += (long long)gridDim.x * 256)
            __builtin_nontemporal_store(src[k], dst + k);
    } else if (VARIANT == 3) {  // nontemporal load and store (fastest on MI355X)
        const long long k = (long long)blockIdx.x * 256 + threadIdx.x;
        if (k < n) __builtin_nontemporal_store(__builtin_nontemporal_load(src + k), dst + k);
    } else {
        const long long k = (long long)blockIdx.x * 256 + threadIdx.x;
        if (k < n) {
            if (VARIANT == 2) __builtin_nontemporal_store(src[k], dst + k);
            else dst[k] = src[k];
        }
    }
}

// Scratch of ffb_moments_noalloc: the FP64 partial sums of the moment kernel,
// three per workgroup for at most kMomentBlocks workgroups (48 KiB per
// device, held until the process ends).  In host code `scratch` names a
// host-side stand-in; kernels get the device address from scratch_on().
constexpr int kMomentBlocks = 2048;
static __device__ __align__(256) double scratch[3 * kMomentBlocks];

std::mutex g_scratch_mu;
double *g_scratch[64] = {};  // device address of `scratch` per device, resolved once

// Device address of `scratch` on `device` (current), resolved at the first call.
int scratch_on(int device, double **out) {
    if (device < 0 || device >= 64) return fail(FFB_E_INVALID, "ffb_moments_noalloc: device ordinal above 63");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(FFB_E_HIP, std::string("ffb_moments_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e));
        g_scratch[device] = static_cast<double *>(p);
    }
    *out = g_scratch[device];
    return FFB_OK;
}

template <typename S, typename C>
void launch_moments_into(const ffb_config_t &c, const void *src, long long pitch, long long plane, int blocks,
                         double *partial, double *out, hipStream_t stream) {
    StepArgs<C> a{};
    a.src = src;
    a.nx = c.n_x;
    a.ny = c.n_y;
    a.pitch = (int)pitch;
    a.plane = plane;
    a.w[0] = (C)4 / (C)9;
    for (int d = 1; d <= 4; ++d) a.w[d] = (C)1 / (C)9;
    for (int d = 5; d <= 8; ++d) a.w[d] = (C)1 / (C)36;
    a.inv_scale = (C)std::ldexp(1.0, -(is_deviation_form(c.precision) ? work_scale_log2(c) : 0));
    hipLaunchKernelGGL((moments_partial_kernel<S, C>), dim3((unsigned)blocks), dim3(256), 0, stream, a, partial);
    hipLaunchKernelGGL(moments_final_kernel, dim3(1), dim3(256), 0, stream, partial, blocks, out);
}

// lattice weight in FP32, folded like the device's (C)4 / (C)9 etc.
float weightf(int d) { return d == 0 ? 4.0f / 9.0f : (d <= 4 ? 1.0f / 9.0f : 1.0f / 36.0f); }

}  // namespace

namespace ffb {
int step_between(ffb_lattice *L, const void *src, void *dst, int write_macro) {
    switch (L->cfg.precision) {
    case FFB_PREC_FP64: return launch_between<double, double>(L, src, dst, write_macro);
    case FFB_PREC_FP32: return launch_between<float, float>(L, src, dst, write_macro);
    case FFB_PREC_FP32_DEV: return launch_between<f32dev, float>(L, src, dst, write_macro);
    default: return fail(FFB_E_INVALID, "step between buffers: FP32, FP32 deviation or FP64 storage only");
    }
}
}  // namespace ffb

extern "C" {

void ffb_config_default(ffb_config_t *cfg) {
    if (!cfg) return;
    std::memset(cfg, 0, sizeof(*cfg));
    // source/settings.f90:27-63
    cfg->n_x = 2520;
    cfg->n_y = 2520;
    cfg->sim_mode = FFB_SLIDING_LID;
    cfg->precision = FFB_PREC_FP32;
    cfg->omega = 1.1;
    cfg->rho_0 = 1.0;
    cfg->u_max = 0.1;
    cfg->n_sin = 2.0;
    cfg->u_wall = 0.1;
    cfg->u_lid = 0.1;
    cfg->rho_in = 1.00005;
    cfg->rho_out = 0.99995;
    cfg->storage_scale_log2 = 0;
    cfg->device = 0;
    cfg->body_force = 0.0;
    cfg->storage_rounding = FFB_ROUND_STOCHASTIC;
    cfg->density_checks = 0;
}

int ffb_create(const ffb_config_t *cfg, ffb_lattice_t **out) {
    if (!cfg || !out) return fail(FFB_E_INVALID, "ffb_create: null argument");
    *out = nullptr;
    if (cfg->n_x < 1 || cfg->n_y < 1) return fail(FFB_E_INVALID, "ffb_create: n_x and n_y must be >= 1");
    if ((long long)cfg->n_x * cfg->n_y > (1LL << 34))
        return fail(FFB_E_INVALID, "ffb_create: domain larger than 2^34 cells");
    // row stride 9 * n_x (rounded up) and byte offsets x * 8 are 32-bit values
    if (cfg->n_x > (1 << 24)) return fail(FFB_E_INVALID, "ffb_create: n_x larger than 2^24");
    if (cfg->sim_mode < FFB_SHEAR_WAVE || cfg->sim_mode > FFB_BODY_FORCE_CHANNEL)
        return fail(FFB_E_INVALID, "ffb_create: unknown sim mode " + std::to_string(cfg->sim_mode));
    if (cfg->precision < FFB_PREC_FP32 || cfg->precision > FFB_PREC_E3M12_DEV)
        return fail(FFB_E_INVALID, "ffb_create: unknown precision " + std::to_string(cfg->precision));
    if (is_deviation_form(cfg->precision) && (cfg->storage_scale_log2 < -30 || cfg->storage_scale_log2 > 30))
        return fail(FFB_E_INVALID, "ffb_create: storage_scale_log2 outside [-30, 30]");
    if (cfg->storage_rounding != FFB_ROUND_NEAREST && cfg->storage_rounding != FFB_ROUND_STOCHASTIC)
        return fail(FFB_E_INVALID, "ffb_create: storage_rounding must be FFB_ROUND_NEAREST or FFB_ROUND_STOCHASTIC");
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (cfg->device < 0 || cfg->device >= ndev)
        return fail(FFB_E_INVALID, "ffb_create: device " + std::to_string(cfg->device) + " not present");
    HIPCHK(hipSetDevice(cfg->device));

    ffb_lattice *L = new ffb_lattice();
    L->cfg = *cfg;
    const int px = ((cfg->n_x + 63) / 64) * 64;
    // Layout and launch shape of the step kernel, defaults measured on MI355X
    // (cavity, one launch per step; us per step):
    //   layout 1 (rows: the nine direction rows of each y next to each other)
    //     against layout 0 (nine planes): 18 streams of one step fall into a
    //     few neighbouring DRAM regions instead of 18 distant ones --
    //     FP32 4096^2 193 against 206, FP64 393 against 428, FP16 8192^2 424
    //     against 473, BF16 451 against 491;
    //   one cell per thread, 256-thread tiles (FP32 197 with 512; 16-bit forms
    //     440-489 with 2 or 4 cells per thread in 4- or 8-byte vectors);
    //   XCD groups (tile_of_block): consecutive tiles of two rows per XCD for
    //     the 16-bit forms (FP16 390 against 424 in dispatch order, BF16 411
    //     against 452: a wave reads one 128-byte line per direction, so the
    //     line shared with the neighbouring tile is a large share of the
    //     traffic when that tile runs on another XCD's L2), half a row for
    //     FP64 (389 against 393-411), dispatch order for FP32 (193 against
    //     194-199 with groups of 16-32 tiles).
    // FFB_LAYOUT, FFB_TILE_W, FFB_VEC (cells per thread), FFB_XCD_GROUP,
    // FFB_PLANE_PAD, FFB_BUF_PAD override them.
    auto env_ll = [](const char *name, long long dflt) {
        const char *v = std::getenv(name);
        return v ? std::atoll(v) : dflt;
    };
    L->plane_pad = ((env_ll("FFB_PLANE_PAD", 0) + 63) / 64) * 64;
    L->buf_pad = ((env_ll("FFB_BUF_PAD", 0) + 63) / 64) * 64;
    L->layout = (int)env_ll("FFB_LAYOUT", 1) == 0 ? 0 : 1;
    if (L->layout == 1) {
        L->plane = px + L->plane_pad;
        L->pitch = (int)(9 * L->plane);
        L->set_elems = (long long)L->pitch * cfg->n_y;
    } else {
        L->pitch = px;
        L->plane = (long long)px * cfg->n_y + L->plane_pad;
        L->set_elems = 9 * L->plane;
    }
    L->elem = cfg->precision == FFB_PREC_FP64
                  ? 8
                  : (cfg->precision == FFB_PREC_FP32 || cfg->precision == FFB_PREC_FP32_DEV ? 4 : 2);
    L->celem = cfg->precision == FFB_PREC_FP64 ? 8 : 4;
    const bool half = L->elem == 2;
    const bool rows = L->layout == 1;
    L->vec = (int)env_ll("FFB_VEC", half && !rows ? 4 : 1);
    L->tile_w = (int)env_ll("FFB_TILE_W", rows || cfg->precision == FFB_PREC_FP64 ? 256 : 512);
    if (L->tile_w != 64 && L->tile_w != 128 && L->tile_w != 256 && L->tile_w != 512) L->tile_w = 256;
    // cells per thread must be 1, 2, 4 or 8 with at most 16 bytes per vector and
    // a 256- or 512-thread tile; anything else falls back to one cell per thread
    if (!(L->vec == 2 || L->vec == 4 || L->vec == 8) || (size_t)L->vec * L->elem > 16 || L->tile_w < 256)
        L->vec = 1;
    {
        const int tiles_x = (cfg->n_x + L->tile_w * L->vec - 1) / (L->tile_w * L->vec);
        const int g = rows ? (half ? 2 * tiles_x : (cfg->precision == FFB_PREC_FP64 ? std::max(1, tiles_x / 2) : 0)) : 0;
        L->xcd_group = (int)std::max(0LL, env_ll("FFB_XCD_GROUP", g));
    }
    // step pairs in one pass (step2_kernel), measured faster for every form:
    // FP32 4096^2 cavity 192.5 -> 124 us per step, FP64 387.5 -> 218; FP16 and
    // BF16 deviations 8192^2 380 -> 341-342 (8-wave regions), the 1-3-12 form
    // 390 -> 371.  FFB_TWO_STEP=0/1 overrides.
    L->two_step = env_ll("FFB_TWO_STEP", 1) ? 1 : 0;
    // steps per pass where the mode allows: 4 or 3 (stepk_kernel) and fewer
    // for a remainder, 2 = pairs, 1 = one launch per step (FFB_STEPS_PER_PASS).  Triples against pairs, us per
    // three steps on one box: FP32 4096^2 264 against 363 (190,800 against
    // 138,700 MLUPS), FP64 4096^2 476 against 644, FP32 8192^2 1039 against
    // 1458, FP32 deviations 8192^2 997 against 1403, FP16 8192^2 1149
    // against 1199 (64 x 16 regions; 64 x 32 regions 9-19 % slower)
    // Four steps per pass (64 x 24 regions, 58 x 18 tiles) for FP32 and FP32
    // deviations: FP32 4096^2 205,200 against 191,900 MLUPS in triples, FP32
    // deviations 8192^2 223,900 against 199,100; FP64 110,600 against 107,900
    // with 64 x 16 regions (its 64 x 24 regions hold one workgroup per CU:
    // 92,200), so FP64 and the 16-bit forms keep triples.  Five steps (8 waves
    // x 4 rows, 64 x 32 regions, 56 x 24 tiles; 4-byte elements only), the
    // default for FP32 and FP32 deviations: FP32 4096^2 215,400 against 205,600
    // in fours, FP32 deviations 8192^2 249,900 against 224,300 (one box,
    // alternating; 12 x 2 regions 205,300 / 242,700, 16 x 2 157,000 / 174,700).
    // The 16-bit forms run 64 x 32 regions of 8 waves x 4 rows (below):
    // FP16 deviations 8192^2 230,400 MLUPS in triples against 196,400 with 64 x
    // 16 regions (223,300 in pairs, 224,000 in fours); the 1-3-12 form 206,200
    // in pairs against 200,500 in triples and 174,400 with 64 x 16 triples.
    // FP64: four steps in 64 x 16 regions (8 waves x 2 rows, 58 x 10 tiles)
    // 115,200 MLUPS at 4096^2 against 107,600 in triples (with the shared
    // reciprocal of the velocity quotients; 110,600 against 107,900 before).
    {
        const bool four = cfg->precision == FFB_PREC_FP32 || cfg->precision == FFB_PREC_FP32_DEV;
        const bool fp64 = cfg->precision == FFB_PREC_FP64;
        const int dflt = four ? 5 : (fp64 ? 4 : (cfg->precision == FFB_PREC_E3M12_DEV ? 2 : 3));
        L->steps_per_pass = (int)std::max(1LL, std::min(four ? 5LL : 4LL, env_ll("FFB_STEPS_PER_PASS", dflt)));
    }
    L->quad_shape = (int)env_ll("FFB_QUAD_SHAPE", cfg->precision == FFB_PREC_FP64 ? 3 : 2);
    L->penta_shape = (int)env_ll("FFB_PENTA_SHAPE", 1);
    // LDS bytes added to each multi-step workgroup (caps the workgroups per CU)
    L->lds_pad = (unsigned)std::max(0LL, std::min(65536LL, env_ll("FFB_MULTI_LDS_PAD", 0)));
    // multi-step launches (three to five steps, the default region shapes) by
    // a persistent grid that loads each workgroup's next tile while it runs the
    // current one's LDS steps (stepk_pipe_kernel), FFB_MULTI_PIPE; FFB_MULTI_GRID
    // workgroups (0: as many as fit at once).  Bitwise equal and measured
    // slower (FP32 4096^2 four steps 365.8 against 325.3 us per launch): the
    // multi-step launches keep the vector ALUs busy, so there is no idle time
    // for the early loads to fill.
    L->multi_pipe = (int)env_ll("FFB_MULTI_PIPE", 0) ? 1 : 0;
    L->multi_grid = (int)std::max(0LL, env_ll("FFB_MULTI_GRID", 0));
    {
        int n_cu = 0;  // a failed query leaves a grid of occupancy x 8 workgroups: slower, still correct
        if (hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, cfg->device) != hipSuccess) n_cu = 8;
        L->n_cu = std::max(1, n_cu);
    }
    if (L->steps_per_pass == 1) L->two_step = 0;
    L->triple_shape = (int)env_ll("FFB_TRIPLE_SHAPE", is_16bit_form(cfg->precision) ? 3 : 2);
    // pair tiles per XCD group, by the width of the domain in 62-cell tiles:
    // 8 up to 96 tiles per row, else 32.  Measured on one box, us per pair
    // launch: FP32 4096^2 (67 tiles per row) 241 at 8, 243 at 12, 247-249 at
    // 16 and 32, 250 at 6, 330 without groups; FP64 4096^2 431 at 8, 445-448
    // at 4, 16, 32; FP32 8192^2 (133 tiles per row) 969 at 32, 984 at 16, 996
    // at 12, 1018 at 8.  Blocks of G x H tiles per XCD (FFB_PAIR_XCD_ROWS = H)
    // were slower in every shape tried (FP32 4096^2 317-423 us for 2x4 .. 16x2).
    {
        const int tiles_x = (cfg->n_x + 61) / 62;
        L->pair_xcd_group = (int)std::max(0LL, env_ll("FFB_PAIR_XCD_GROUP", tiles_x <= 96 ? 8 : 32));
    }
    L->pair_xcd_rows = (int)std::max(1LL, std::min(64LL, env_ll("FFB_PAIR_XCD_ROWS", 1)));
    // pair region of 4-byte and narrower elements: 2 = 8 waves x 2 rows (64 x 16
    // cells, 62 x 14 outputs), 1 = 16 waves x 2 rows (64 x 32, 62 x 30).  Shape 2
    // for FP32 (4096^2: 248-254 against 259-268 us per pair; 8192^2 995 against
    // 1033) and the 16-bit forms (FP16 8192^2 683 against 771), shape 1 for FP32
    // deviations (8192^2 866 against 943).
    L->pair_shape = (int)env_ll("FFB_PAIR_SHAPE", is_16bit_form(cfg->precision)
                                                      ? 3
                                                      : (cfg->precision == FFB_PREC_FP32_DEV ? 1 : 2));
    L->pair_nt = (int)env_ll("FFB_PAIR_NT", 0);
    // pair workgroups: 0 = one per tile, else a persistent grid of this size
    L->pair_grid = (int)std::max(0LL, env_ll("FFB_PAIR_GRID", 0));
    // step pairs by tiles (step2_kernel, the default) or by column march
    // (march2_kernel, FFB_PAIR_KERNEL=1): the march computes no ring rows and
    // keeps the next bands' loads in flight, but its workgroups spread over many
    // more rows at once and it ran slower on MI355X (FP32 4096^2 292-373 us per
    // pair launch against 241-247 for the tiles, FP16 8192^2 723-889 against
    // 765-801); with the collision removed both ran as fast as with it (tiles
    // 246 against 244 us, march 296 against 292): the pairs are bound by their
    // memory access pattern, not by the arithmetic or its overlap with loads
    L->pair_kernel = (int)env_ll("FFB_PAIR_KERNEL", 0) == 1 ? 1 : 0;
    L->march_shape = (int)env_ll("FFB_MARCH_SHAPE", 1);
    L->march_rows = (int)std::max(1LL, std::min(1LL << 20, env_ll("FFB_MARCH_ROWS", 126)));
    L->march_inner = (int)env_ll("FFB_MARCH_INNER", 1) ? 1 : 0;

    auto cleanup_fail = [&](hipError_t e, const char *what) {
        ffb_destroy(L);
        return fail(FFB_E_HIP, std::string("ffb_create: ") + what + ": " + hipGetErrorString(e));
    };
    hipError_t e;
    const size_t set_elems = (size_t)L->set_elems + (size_t)L->buf_pad;
    if ((e = hipMalloc(&L->pop_alloc, 2 * set_elems * L->elem)) != hipSuccess)
        return cleanup_fail(e, "population buffers");
    L->buf[0] = L->pop_alloc;
    L->buf[1] = (unsigned char *)L->pop_alloc + set_elems * L->elem;
    const size_t mbytes = (size_t)cfg->n_x * (size_t)cfg->n_y * L->celem;
    for (int k = 0; k < 3; ++k)
        if ((e = hipMalloc(&L->macro[k], mbytes)) != hipSuccess) return cleanup_fail(e, "macro fields");
    const size_t sbytes = (size_t)6 * (size_t)cfg->n_y * L->celem;
    for (int k = 0; k < 2; ++k) {
        if ((e = hipMalloc(&L->strips[k], sbytes)) != hipSuccess) return cleanup_fail(e, "strips");
        if ((e = hipMemset(L->strips[k], 0, sbytes)) != hipSuccess) return cleanup_fail(e, "strips");
    }
    if ((e = hipMalloc(&L->ux_row, (size_t)cfg->n_y * L->celem)) != hipSuccess) return cleanup_fail(e, "profile");
    long long cells = (long long)cfg->n_x * cfg->n_y;
    L->mom_blocks = (int)std::min<long long>(kMomentBlocks, (cells + 255) / 256);
    if ((e = hipMalloc(&L->mom_partial, (size_t)L->mom_blocks * 3 * sizeof(double))) != hipSuccess)
        return cleanup_fail(e, "moment partials");
    if ((e = hipMalloc(&L->mom_out, 3 * sizeof(double))) != hipSuccess) return cleanup_fail(e, "moments");
    if ((e = hipMalloc(&L->status, sizeof(unsigned))) != hipSuccess) return cleanup_fail(e, "status word");
    if ((e = hipMemset(L->status, 0, sizeof(unsigned))) != hipSuccess) return cleanup_fail(e, "status word");
    if ((e = hipStreamCreateWithFlags(&L->stream, hipStreamNonBlocking)) != hipSuccess) return cleanup_fail(e, "stream");
    if ((e = hipEventCreate(&L->ev_start)) != hipSuccess) return cleanup_fail(e, "event");
    if ((e = hipEventCreate(&L->ev_stop)) != hipSuccess) return cleanup_fail(e, "event");
    *out = L;
    return FFB_OK;
}

int ffb_destroy(ffb_lattice_t *L) {
    if (!L) return FFB_OK;
    // best effort: a failing release must not stop the others
    (void)hipSetDevice(L->cfg.device);
    if (L->stream) (void)hipStreamSynchronize(L->stream);
    if (L->pop_alloc) (void)hipFree(L->pop_alloc);
    for (void *p : L->macro) if (p) (void)hipFree(p);
    for (void *p : L->strips) if (p) (void)hipFree(p);
    if (L->ux_row) (void)hipFree(L->ux_row);
    if (L->mom_partial) (void)hipFree(L->mom_partial);
    if (L->mom_out) (void)hipFree(L->mom_out);
    if (L->status) (void)hipFree(L->status);
    if (L->ev_start) (void)hipEventDestroy(L->ev_start);
    if (L->ev_stop) (void)hipEventDestroy(L->ev_stop);
    if (L->stream) (void)hipStreamDestroy(L->stream);
    delete L;
    return FFB_OK;
}

int ffb_init(ffb_lattice_t *L) {
    if (!L) return fail(FFB_E_INVALID, "ffb_init: null handle");
    HIPCHK(hipSetDevice(L->cfg.device));
    L->cur = 0;
    L->strip_old = 0;
    L->steps = 0;
    L->timing = false;
    HIPCHK(hipMemsetAsync(L->status, 0, sizeof(unsigned), L->stream));
    int rc = dispatch_init(L);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(L->stream));
    return FFB_OK;
}

int ffb_step(ffb_lattice_t *L, int32_t n_steps, int32_t write_macro_last) {
    if (!L) return fail(FFB_E_INVALID, "ffb_step: null handle");
    if (n_steps < 0) return fail(FFB_E_INVALID, "ffb_step: n_steps < 0");
    if (n_steps == 0) return FFB_OK;
    HIPCHK(hipSetDevice(L->cfg.device));
    if (!L->timing) {
        HIPCHK(hipEventRecord(L->ev_start, L->stream));
        L->timing = true;
    }
    const bool pairs = L->two_step && L->cfg.sim_mode != FFB_POISEUILLE_FLOW;
    const bool triples = pairs && L->steps_per_pass >= 3;
    const bool quads = pairs && L->steps_per_pass >= 4 &&
                       (L->cfg.precision == FFB_PREC_FP32 || L->cfg.precision == FFB_PREC_FP64 ||
                        L->cfg.precision == FFB_PREC_FP32_DEV || L->cfg.precision == FFB_PREC_FP16_DEV);
    const bool pentas = quads && L->steps_per_pass >= 5 &&
                        (L->cfg.precision == FFB_PREC_FP32 || L->cfg.precision == FFB_PREC_FP32_DEV);
    for (int32_t k = 0; k < n_steps;) {
        if (pentas && n_steps - k >= 5) {  // steps k+1 .. k+5 in one pass
            const int wm = (write_macro_last && k + 5 == n_steps) ? 1 : 0;
            int rc = dispatch_step5(L, wm);
            if (rc) return rc;
            L->cur = 1 - L->cur;
            L->steps += 5;
            k += 5;
            continue;
        }
        if (quads && n_steps - k >= 4) {  // steps k+1 .. k+4 in one pass
            const int wm = (write_macro_last && k + 4 == n_steps) ? 1 : 0;
            int rc = dispatch_step4(L, wm);
            if (rc) return rc;
            L->cur = 1 - L->cur;
            L->steps += 4;
            k += 4;
            continue;
        }
        if (triples && n_steps - k >= 3) {  // steps k+1 .. k+3 in one pass
            const int wm = (write_macro_last && k + 3 == n_steps) ? 1 : 0;
            int rc = dispatch_step3(L, wm);
            if (rc) return rc;
            L->cur = 1 - L->cur;
            L->steps += 3;
            k += 3;
            continue;
        }
        if (pairs && n_steps - k >= 2) {  // steps k+1, k+2 in one pass; the result lands in the other set
            const int wm = (write_macro_last && k + 2 == n_steps) ? 1 : 0;
            int rc = dispatch_step2(L, wm);
            if (rc) return rc;
            L->cur = 1 - L->cur;
            L->steps += 2;
            k += 2;
            continue;
        }
        const int wm = (write_macro_last && k == n_steps - 1) ? 1 : 0;
        int rc = dispatch_step(L, wm);
        if (rc) return rc;
        L->cur = 1 - L->cur;
        if (L->cfg.sim_mode == FFB_POISEUILLE_FLOW) L->strip_old = 1 - L->strip_old;
        ++L->steps;
        ++k;
    }
    HIPCHK(hipEventRecord(L->ev_stop, L->stream));
    return FFB_OK;
}

int ffb_set_parameters(ffb_lattice_t *L, double omega, double body_force, double u_wall, double u_lid) {
    if (!L) return fail(FFB_E_INVALID, "ffb_set_parameters: null handle");
    L->cfg.omega = omega;
    L->cfg.body_force = body_force;
    L->cfg.u_wall = u_wall;
    L->cfg.u_lid = u_lid;
    return FFB_OK;
}

int ffb_sync(ffb_lattice_t *L) {
    if (!L) return fail(FFB_E_INVALID, "ffb_sync: null handle");
    HIPCHK(hipStreamSynchronize(L->stream));
    if (L->cfg.density_checks) {
        unsigned st = 0;
        HIPCHK(hipMemcpy(&st, L->status, sizeof(st), hipMemcpyDeviceToHost));
        // the reference's message (simulation.f90:231)
        if (st & 1u) return fail(FFB_E_DENSITY, "error: density is zero in collision/streaming step (rho_val <= 0)");
        if (st & 2u) return fail(FFB_E_DENSITY, "error: density is NaN in collision/streaming step");
    }
    return FFB_OK;
}

int ffb_read_macros(ffb_lattice_t *L, void *rho, void *u_x, void *u_y) {
    if (!L) return fail(FFB_E_INVALID, "ffb_read_macros: null handle");
    HIPCHK(hipSetDevice(L->cfg.device));
    const size_t bytes = (size_t)L->cfg.n_x * (size_t)L->cfg.n_y * L->celem;
    void *dst[3] = {rho, u_x, u_y};
    for (int k = 0; k < 3; ++k)
        if (dst[k]) HIPCHK(hipMemcpyAsync(dst[k], L->macro[k], bytes, hipMemcpyDeviceToHost, L->stream));
    HIPCHK(hipStreamSynchronize(L->stream));
    return FFB_OK;
}

int ffb_read_populations(ffb_lattice_t *L, void *f_halo) {
    if (!L || !f_halo) return fail(FFB_E_INVALID, "ffb_read_populations: null argument");
    HIPCHK(hipSetDevice(L->cfg.device));
    const int nx = L->cfg.n_x, ny = L->cfg.n_y;
    std::vector<unsigned char> stage((size_t)L->set_elems * L->elem);
    HIPCHK(hipMemcpyAsync(stage.data(), L->buf[L->cur], stage.size(), hipMemcpyDeviceToHost, L->stream));
    HIPCHK(hipStreamSynchronize(L->stream));
    const size_t hp = (size_t)(nx + 2), hplane = hp * (size_t)(ny + 2);
    const int s = work_scale_log2(L->cfg);
    for (int d = 0; d < 9; ++d) {
        for (int y = 0; y < ny; ++y) {
            for (int x = 0; x < nx; ++x) {
                const size_t src = (size_t)d * (size_t)L->plane + (size_t)y * (size_t)L->pitch + (size_t)x;
                const size_t dst = (size_t)d * hplane + (size_t)(y + 1) * hp + (size_t)(x + 1);
                switch (L->cfg.precision) {
                case FFB_PREC_FP64: ((double *)f_halo)[dst] = ((const double *)stage.data())[src]; break;
                case FFB_PREC_FP32: ((float *)f_halo)[dst] = ((const float *)stage.data())[src]; break;
                case FFB_PREC_FP32_DEV:
                    ((float *)f_halo)[dst] = weightf(d) + ((const float *)stage.data())[src] * (float)std::ldexp(1.0, -s);
                    break;
                case FFB_PREC_FP16_DEV:
                    ((float *)f_halo)[dst] = weightf(d) + half_to_float(((const uint16_t *)stage.data())[src]) *
                                                                    (float)std::ldexp(1.0, -s);
                    break;
                case FFB_PREC_E3M12_DEV:  // stored value times 2^-off of its direction, as the kernels read it
                    ((float *)f_halo)[dst] =
                        weightf(d) + (e3m12_to_float(((const uint16_t *)stage.data())[src]) *
                                      pow2f(-store_off<e3m12>(d))) * (float)std::ldexp(1.0, -s);
                    break;
                default:
                    ((float *)f_halo)[dst] = weightf(d) + bf16_to_float(((const uint16_t *)stage.data())[src]) *
                                                                    (float)std::ldexp(1.0, -s);
                    break;
                }
            }
        }
    }
    // halo entries: zero
    const size_t ce = L->celem;
    for (int d = 0; d < 9; ++d) {
        unsigned char *base = (unsigned char *)f_halo + (size_t)d * hplane * ce;
        std::memset(base, 0, hp * ce);
        std::memset(base + (size_t)(ny + 1) * hp * ce, 0, hp * ce);
        for (int y = 1; y <= ny; ++y) {
            std::memset(base + (size_t)y * hp * ce, 0, ce);
            std::memset(base + ((size_t)y * hp + (size_t)(nx + 1)) * ce, 0, ce);
        }
    }
    return FFB_OK;
}

int ffb_write_populations(ffb_lattice_t *L, const void *f_halo) {
    if (!L || !f_halo) return fail(FFB_E_INVALID, "ffb_write_populations: null argument");
    HIPCHK(hipSetDevice(L->cfg.device));
    const int nx = L->cfg.n_x, ny = L->cfg.n_y;
    std::vector<unsigned char> stage((size_t)L->set_elems * L->elem, 0);
    const size_t hp = (size_t)(nx + 2), hplane = hp * (size_t)(ny + 2);
    const int s = work_scale_log2(L->cfg);
    for (int d = 0; d < 9; ++d) {
        for (int y = 0; y < ny; ++y) {
            for (int x = 0; x < nx; ++x) {
                const size_t dst = (size_t)d * (size_t)L->plane + (size_t)y * (size_t)L->pitch + (size_t)x;
                const size_t src = (size_t)d * hplane + (size_t)(y + 1) * hp + (size_t)(x + 1);
                switch (L->cfg.precision) {
                case FFB_PREC_FP64: ((double *)stage.data())[dst] = ((const double *)f_halo)[src]; break;
                case FFB_PREC_FP32: ((float *)stage.data())[dst] = ((const float *)f_halo)[src]; break;
                case FFB_PREC_FP32_DEV:
                    ((float *)stage.data())[dst] = (((const float *)f_halo)[src] - weightf(d)) * (float)std::ldexp(1.0, s);
                    break;
                case FFB_PREC_FP16_DEV:
                    ((uint16_t *)stage.data())[dst] =
                        float_to_half((((const float *)f_halo)[src] - weightf(d)) * (float)std::ldexp(1.0, s));
                    break;
                case FFB_PREC_E3M12_DEV:
                    ((uint16_t *)stage.data())[dst] = float_to_e3m12(
                        ((((const float *)f_halo)[src] - weightf(d)) * (float)std::ldexp(1.0, s)) *
                        pow2f(store_off<e3m12>(d)));
                    break;
                default:
                    ((uint16_t *)stage.data())[dst] =
                        float_to_bf16((((const float *)f_halo)[src] - weightf(d)) * (float)std::ldexp(1.0, s));
                    break;
                }
            }
        }
    }
    HIPCHK(hipMemcpyAsync(L->buf[L->cur], stage.data(), stage.size(), hipMemcpyHostToDevice, L->stream));
    HIPCHK(hipStreamSynchronize(L->stream));
    return FFB_OK;
}

int ffb_read_pf_strips(ffb_lattice_t *L, void *macro_left, void *macro_right) {
    if (!L) return fail(FFB_E_INVALID, "ffb_read_pf_strips: null handle");
    HIPCHK(hipSetDevice(L->cfg.device));
    const size_t bytes = (size_t)3 * (size_t)L->cfg.n_y * L->celem;
    unsigned char *base = (unsigned char *)L->strips[L->strip_old];
    if (macro_left) HIPCHK(hipMemcpyAsync(macro_left, base, bytes, hipMemcpyDeviceToHost, L->stream));
    if (macro_right) HIPCHK(hipMemcpyAsync(macro_right, base + bytes, bytes, hipMemcpyDeviceToHost, L->stream));
    HIPCHK(hipStreamSynchronize(L->stream));
    return FFB_OK;
}

int ffb_elapsed_ms(ffb_lattice_t *L, double *ms) {
    if (!L || !ms) return fail(FFB_E_INVALID, "ffb_elapsed_ms: null argument");
    *ms = 0.0;
    if (!L->timing) return FFB_OK;
    HIPCHK(hipEventSynchronize(L->ev_stop));
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, L->ev_start, L->ev_stop));
    *ms = (double)t;
    return FFB_OK;
}

int ffb_reset_timer(ffb_lattice_t *L) {
    if (!L) return fail(FFB_E_INVALID, "ffb_reset_timer: null handle");
    L->timing = false;
    return FFB_OK;
}

int ffb_step_count(ffb_lattice_t *L, int64_t *steps) {
    if (!L || !steps) return fail(FFB_E_INVALID, "ffb_step_count: null argument");
    *steps = L->steps;
    return FFB_OK;
}

int32_t ffb_default_scale_log2(int32_t precision, double speed) {
    if (!(speed >= 0.0 && speed <= 1.0)) return INT32_MIN;
    const double u = speed, rho_dev = 0.05;
    switch (precision) {
    case FFB_PREC_FP16_DEV:
        return (int32_t)std::floor(std::log2(16384.0 / ((4.0 / 9.0) * (rho_dev + 3.0 * u + 6.0 * u * u))));
    case FFB_PREC_E3M12_DEV: {
        // per class w (rho_dev + 3 |c.u| + 4.5 (c.u)^2 + 1.5 u^2) with |c.u| <= |c| u, times 2^off
        const double rest = (4.0 / 9.0) * (rho_dev + 1.5 * u * u) / 2.0;
        const double axis = (1.0 / 9.0) * (rho_dev + 3.0 * u + 6.0 * u * u);
        const double diag = (1.0 / 36.0) * (rho_dev + 3.0 * std::sqrt(2.0) * u + 10.5 * u * u) * 2.0;
        return (int32_t)std::floor(std::log2(1.5 / std::max(rest, std::max(axis, diag))));
    }
    case FFB_PREC_FP32:
    case FFB_PREC_FP64:
    case FFB_PREC_BF16_DEV:
    case FFB_PREC_FP32_DEV: return 0;
    default: return INT32_MIN;
    }
}

int ffb_deviation_extent(ffb_lattice_t *L, double max_dev[3], double limit[3]) {
    if (!L || !max_dev || !limit) return fail(FFB_E_INVALID, "ffb_deviation_extent: null argument");
    HIPCHK(hipSetDevice(L->cfg.device));
    int rc;
    switch (L->cfg.precision) {
    case FFB_PREC_FP64: rc = launch_extent<double, double>(L); break;
    case FFB_PREC_FP16_DEV: rc = launch_extent<__half, float>(L); break;
    case FFB_PREC_BF16_DEV: rc = launch_extent<__hip_bfloat16, float>(L); break;
    case FFB_PREC_FP32_DEV: rc = launch_extent<f32dev, float>(L); break;
    case FFB_PREC_E3M12_DEV: rc = launch_extent<e3m12, float>(L); break;
    default: rc = launch_extent<float, float>(L); break;
    }
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(max_dev, L->mom_out, 3 * sizeof(double), hipMemcpyDeviceToHost, L->stream));
    HIPCHK(hipStreamSynchronize(L->stream));
    // largest |f - w| the stored form represents, per class
    const int s = L->cfg.storage_scale_log2;
    for (int j = 0; j < 3; ++j) {
        const int d = j == 0 ? 0 : (j == 1 ? 1 : 5);
        double lim;
        switch (L->cfg.precision) {
        case FFB_PREC_FP64: lim = DBL_MAX; break;
        case FFB_PREC_FP32: lim = FLT_MAX; break;
        case FFB_PREC_FP32_DEV: lim = std::ldexp((double)FLT_MAX, -s); break;
        case FFB_PREC_FP16_DEV: lim = std::ldexp(65504.0, -s); break;
        case FFB_PREC_BF16_DEV: lim = std::ldexp(3.3895313892515355e38, -s); break;
        default: lim = std::ldexp(2.0 - std::ldexp(1.0, -12), -s - store_off<e3m12>(d)); break;
        }
        limit[j] = lim;
    }
    return FFB_OK;
}

int ffb_moments(ffb_lattice_t *L, double *mass, double *mom_x, double *mom_y) {
    if (!L) return fail(FFB_E_INVALID, "ffb_moments: null handle");
    HIPCHK(hipSetDevice(L->cfg.device));
    int rc;
    switch (L->cfg.precision) {
    case FFB_PREC_FP64: rc = launch_moments<double, double>(L); break;
    case FFB_PREC_FP16_DEV: rc = launch_moments<__half, float>(L); break;
    case FFB_PREC_BF16_DEV: rc = launch_moments<__hip_bfloat16, float>(L); break;
    case FFB_PREC_FP32_DEV: rc = launch_moments<f32dev, float>(L); break;
    case FFB_PREC_E3M12_DEV: rc = launch_moments<e3m12, float>(L); break;
    default: rc = launch_moments<float, float>(L); break;
    }
    if (rc) return rc;
    double h[3];
    HIPCHK(hipMemcpyAsync(h, L->mom_out, sizeof(h), hipMemcpyDeviceToHost, L->stream));
    HIPCHK(hipStreamSynchronize(L->stream));
    if (mass) *mass = h[0];
    if (mom_x) *mom_x = h[1];
    if (mom_y) *mom_y = h[2];
    return FFB_OK;
}

int ffb_moments_noalloc(const ffb_config_t *cfg, const void *populations, int64_t pitch, int64_t plane,
                        void *stream, double *out) {
    if (!cfg || !populations || !out) return fail(FFB_E_INVALID, "ffb_moments_noalloc: null argument");
    if (cfg->n_x < 1 || cfg->n_y < 1 || cfg->n_x > (1 << 24) || (long long)cfg->n_x * cfg->n_y > (1LL << 34))
        return fail(FFB_E_INVALID, "ffb_moments_noalloc: n_x, n_y outside 1 .. 2^24, n_x * n_y above 2^34");
    if (cfg->precision < FFB_PREC_FP32 || cfg->precision > FFB_PREC_E3M12_DEV)
        return fail(FFB_E_INVALID, "ffb_moments_noalloc: unknown precision " + std::to_string(cfg->precision));
    if (pitch < cfg->n_x || pitch > INT_MAX || plane < cfg->n_x)
        return fail(FFB_E_INVALID, "ffb_moments_noalloc: pitch or plane smaller than n_x (or pitch above 2^31-1)");
    HIPCHK(hipSetDevice(cfg->device));
    double *partial = nullptr;
    if (int rc = scratch_on(cfg->device, &partial)) return rc;
    // the workgroup count of ffb_moments (same partial sums, same order)
    const long long cells = (long long)cfg->n_x * cfg->n_y;
    const int blocks = (int)std::min<long long>(kMomentBlocks, (cells + 255) / 256);
    hipStream_t st = static_cast<hipStream_t>(stream);
    switch (cfg->precision) {
    case FFB_PREC_FP64: launch_moments_into<double, double>(*cfg, populations, pitch, plane, blocks, partial, out, st); break;
    case FFB_PREC_FP16_DEV: launch_moments_into<__half, float>(*cfg, populations, pitch, plane, blocks, partial, out, st); break;
    case FFB_PREC_BF16_DEV:
        launch_moments_into<__hip_bfloat16, float>(*cfg, populations, pitch, plane, blocks, partial, out, st);
        break;
    case FFB_PREC_FP32_DEV: launch_moments_into<f32dev, float>(*cfg, populations, pitch, plane, blocks, partial, out, st); break;
    case FFB_PREC_E3M12_DEV: launch_moments_into<e3m12, float>(*cfg, populations, pitch, plane, blocks, partial, out, st); break;
    default: launch_moments_into<float, float>(*cfg, populations, pitch, plane, blocks, partial, out, st); break;
    }
    HIPCHK(hipGetLastError());
    return FFB_OK;
}

int ffb_device_info(ffb_lattice_t *L, void **cur, void **stream, int64_t *pitch, int64_t *plane) {
    if (!L) return fail(FFB_E_INVALID, "ffb_device_info: null handle");
    if (cur) *cur = L->buf[L->cur];
    if (stream) *stream = (void *)L->stream;
    if (pitch) *pitch = L->pitch;
    if (plane) *plane = L->plane;
    return FFB_OK;
}

int ffb_bench_copy(int32_t device, int64_t bytes, int32_t iters, double *gbps) {
    if (!gbps || bytes < 16 || iters < 1) return fail(FFB_E_INVALID, "ffb_bench_copy: bad argument");
    HIPCHK(hipSetDevice(device));
    const long long n = bytes / 16;
    void *a = nullptr, *b = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = FFB_OK;
    float ms = 0.f;
    auto run = [&]() -> hipError_t {
        hipError_t e;
        if ((e = hipMalloc(&a, (size_t)n * 16)) != hipSuccess) return e;
        if ((e = hipMalloc(&b, (size_t)n * 16)) != hipSuccess) return e;
        if ((e = hipMemset(a, 0, (size_t)n * 16)) != hipSuccess) return e;
        if ((e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking)) != hipSuccess) return e;
        if ((e = hipEventCreate(&e0)) != hipSuccess) return e;
        if ((e = hipEventCreate(&e1)) != hipSuccess) return e;
        float best = 0.f;
        for (int variant = 0; variant < 4; ++variant) {
            const unsigned grid = variant == 0 ? 256u * 32u : (unsigned)((n + 255) / 256);
            auto launch = [&]() {
                switch (variant) {
                case 0: hipLaunchKernelGGL(copy_kernel<0>, dim3(grid), dim3(256), 0, st, (const v4f *)a, (v4f *)b, n); break;
                case 1: hipLaunchKernelGGL(copy_kernel<1>, dim3(grid), dim3(256), 0, st, (const v4f *)a, (v4f *)b, n); break;
                case 2: hipLaunchKernelGGL(copy_kernel<2>, dim3(grid), dim3(256), 0, st, (const v4f *)a, (v4f *)b, n); break;
                default: hipLaunchKernelGGL(copy_kernel<3>, dim3(grid), dim3(256), 0, st, (const v4f *)a, (v4f *)b, n); break;
                }
            };
            launch();
            if ((e = hipEventRecord(e0, st)) != hipSuccess) return e;
            for (int k = 0; k < iters; ++k) launch();
            if ((e = hipGetLastError()) != hipSuccess) return e;
            if ((e = hipEventRecord(e1, st)) != hipSuccess) return e;
            if ((e = hipEventSynchronize(e1)) != hipSuccess) return e;
            float t = 0.f;
            if ((e = hipEventElapsedTime(&t, e0, e1)) != hipSuccess) return e;
            if (best == 0.f || t < best) best = t;
        }
        ms = best;
        return hipSuccess;
    };
    hipError_t e = run();
    if (e != hipSuccess) rc = fail(FFB_E_HIP, std::string("ffb_bench_copy: ") + hipGetErrorString(e));
    else *gbps = 2.0 * (double)n * 16.0 * iters / ((double)ms * 1e-3) / 1e9;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return rc;
}

const char *ffb_last_error(void) { return g_last_error.c_str(); }

const char *ffb_version(void) { return "ffb-hip 0.1.0 gfx950"; }

}  // extern "C"
