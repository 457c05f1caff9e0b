// ffb_batch.hip -- batched form: many independent lattices of uneven size,
// each with its own omega and boundary kind, advanced together.  Consecutive
// steps run as step triples (or pairs), three (two) steps per launch of one
// kernel over the tiles of every lattice (one launch per step is the option,
// and runs a remainder); the launches of a ping-pong cycle are captured once
// in a hipGraph and replayed.  A project-written autotuner picks the launch shape
// of the kernel the batch runs by timing the candidates on the batch.  No
// reference counterpart (the reference runs one domain per process).
//
// Layout: every lattice keeps the single-domain row-blocked layout (per row
// the nine direction rows, each n_x rounded up to 64 elements; no halo) inside
// two pools (ping-pong); a descriptor table holds each lattice's offsets and
// parameters.  Multi-step launches: a tile table names (lattice, tile) for
// every output tile (pairs 62 x 14 or 62 x 30, FP64 62 x 6; triples 60 x 12 or
// 60 x 28, FP64 60 x 12), lattice by lattice, grouped by kind.  One launch per step: work is cut into row segments of 64 cells (one
// wave each, so every load of a wave is one aligned stretch of one row);
// segs[t] names the lattice and row of segment t, and a workgroup of TW
// threads takes TW/64 consecutive segments per iteration of its persistent
// loop.  The tables are built once, at creation; a change of tuning rewrites
// the tile table in place (after dropping the graph that refers to it) and the
// graph is captured again before the next step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "d2q9_kernels.h"
#include "ffb.h"
#include "ffb_internal.h"

using namespace ffb;

#define BHIPCHK(expr)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return ffb::fail(FFB_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));    \
    } while (0)

namespace {

template <typename C> struct BatchLat {
    long long off;        // element offset of the lattice's planes in a pool
    long long plane;      // direction stride in elements (row stride: pitch)
    long long macro_off;  // element offset in the macro arrays
    long long prof_off;   // element offset of the shear-wave profile row table
    int nx, ny, pitch, kind;
    int tile_start;       // first row segment
    int segs_per_row;     // ceil(nx / 64)
    C omega, corr8, corr9, force, rho_0;
    C scale, inv_scale;   // 2^s and 2^-s of the deviation forms (s - 120 for the 1-3-12 form)
    // deviation forms: the collide_dev constants of the lattice, computed once on
    // the host (StepArgs' kw, kw_lo, kf, om1, half_force)
    C kw[9], kw_lo[9], kf[9];
    C om1, half_force;
};

struct Seg {
    int lat, y;  // lattice index and row of one 64-cell row segment
};

struct PTile {
    int lat, tx, ty, pad;  // lattice index and tile of one step-pair tile (pair_tile)
};

template <typename C> struct BatchArgs {
    const void *src_pool;
    void *dst_pool;
    const BatchLat<C> *lats;
    const Seg *segs;      // (lattice, row) of each row segment
    int n_tiles;          // row segments
    int write_macro;
    C *rho_o, *ux_o, *uy_o;
    const C *prof;        // shear-wave u_x(y) rows of all SW lattices
    const unsigned *step_ctr;  // steps done before this launch (16-bit forms: stochastic rounding)
    int stochastic;       // 16-bit forms: 1 = stochastic rounding of the stored values
    C w[9];
};

// The segment and descriptor tables are read-only during a launch and every
// wave reads one entry of each: read word by word through the constant
// address space, the compiler loads them with scalar loads (uniform per wave).
template <typename T> __device__ __forceinline__ T load_uniform(const T *p) {
    static_assert(sizeof(T) % 4 == 0, "table entries are whole dwords");
    T v;
    const __attribute__((address_space(4))) int *q = (const __attribute__((address_space(4))) int *)(p);
    int *d = reinterpret_cast<int *>(&v);
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / 4); ++i) d[i] = q[i];
    return v;
}

template <typename S, typename C>
__device__ __forceinline__ StepArgs<C> lattice_args(const BatchArgs<C> &b, const BatchLat<C> &L) {
    StepArgs<C> a;
    using R = typename Store<S>::R;
    a.src = static_cast<const R *>(b.src_pool) + L.off;
    a.dst = static_cast<R *>(b.dst_pool) + L.off;
    a.nx = L.nx;
    a.ny = L.ny;
    a.pitch = L.pitch;
    a.plane = L.plane;
    a.omega = L.omega;
    a.corr8 = L.corr8;
    a.corr9 = L.corr9;
    a.rho_in = 0;
    a.rho_out = 0;
    a.m_old = nullptr;
    a.m_new = nullptr;
    a.rho_o = b.rho_o + L.macro_off;
    a.ux_o = b.ux_o + L.macro_off;
    a.uy_o = b.uy_o + L.macro_off;
    a.write_macro = b.write_macro;
#pragma unroll
    for (int d = 0; d < 9; ++d) a.w[d] = b.w[d];
    a.scale = L.scale;
    a.inv_scale = L.inv_scale;
    a.n_tiles_x = 0;
    a.n_blocks = 0;
    a.xcd_group = 0;
    a.check_rho = 0;
    a.status = nullptr;
    a.force = L.force;
    a.stochastic = 0;
    a.step_seed = 0;
    if constexpr (Store<S>::dev) {  // deviation-space constants of collide_dev, from the host
        a.om1 = L.om1;
#pragma unroll
        for (int d = 0; d < 9; ++d) {
            a.kw[d] = L.kw[d];
            a.kw_lo[d] = L.kw_lo[d];
            a.kf[d] = L.kf[d];
        }
        a.half_force = L.half_force;
        if constexpr (sizeof(R) == 2) {
            a.stochastic = b.stochastic;
            a.step_seed = load_uniform(b.step_ctr) + 1u;  // the step this launch computes
        }
    }
    return a;
}

// Step counter of the stochastic rounding: advanced by the steps of every
// step launch (inside the captured graph too), read by the next launch.
__global__ void batch_count_kernel(unsigned *ctr, int reset, unsigned inc) { ctr[0] = reset ? 0u : ctr[0] + inc; }

// One step of every lattice: a persistent grid walks the row segments, one
// wave per segment.  The 16-bit forms are compiled for 8 waves per SIMD (the
// descriptor table's scalar registers would otherwise hold them to 7; the
// excess is kept in vector-register lanes) and store a cell's nine values
// through one descriptor (the nine direction rows of a lattice row lie within
// 9 * plane elements, checked at creation): together measured 5-9 % faster on
// 512 lattices, while FP32 ran 2 % slower that way and keeps the defaults.
template <typename S, typename C, int TW>
__global__ void __launch_bounds__(TW)
    __attribute__((amdgpu_waves_per_eu(sizeof(typename Store<S>::R) == 2 ? 8 : 1))) batch_step_kernel(BatchArgs<C> b) {
    constexpr bool NARROW = sizeof(typename Store<S>::R) == 2;
    constexpr int WAVES = TW / 64;
    const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6), lane = (int)threadIdx.x & 63;
    for (int t = (int)blockIdx.x * WAVES + wave; t < b.n_tiles; t += (int)gridDim.x * WAVES) {
        const Seg sg = load_uniform(b.segs + t);
        const BatchLat<C> L = load_uniform(b.lats + sg.lat);
        const int y = sg.y;
        const int x = (t - L.tile_start - y * L.segs_per_row) * 64 + lane;
        if (x >= L.nx) continue;
        const StepArgs<C> a = lattice_args<S, C>(b, L);
        using R = typename Store<S>::R;
        const R *f = static_cast<const R *>(a.src);
        R *g = static_cast<R *>(a.dst);
        switch (L.kind) {
        case SW: step_cell<S, C, SW, NARROW>(a, f, g, x, y); break;
        case CF: step_cell<S, C, CF, NARROW>(a, f, g, x, y); break;
        case SL: step_cell<S, C, SL, NARROW>(a, f, g, x, y); break;
        default: step_cell<S, C, BF, NARROW>(a, f, g, x, y); break;
        }
    }
}

// K steps (pairs, K = 2, or triples, K = 3) of every lattice: workgroups walk
// the tile table (pair_tile: step 1 on the tile and its one-cell ring into
// LDS, step 2 of the tile from LDS; multi_tile for three steps, a two-cell
// ring), a persistent grid (a multiple of 8 workgroups) or one workgroup per
// tile, the tiles mapped to workgroups in XCD groups of G (tile_of_block).
template <typename S, typename C, int WAVES, int RPT, bool NT, int K>
__global__ void __launch_bounds__(64 * WAVES) batch_pair_kernel(BatchArgs<C> b, const PTile *tiles, int n, int G) {
    __shared__ typename Store<S>::R lds[9][WAVES * RPT][64];
    RawPop<S> raw[RPT][9];  // multi_tile's pipelined form only
    for (int p = (int)blockIdx.x; p < n; p += (int)gridDim.x) {
        const int t = tile_of_block(p, n, G);
        const PTile pt = load_uniform(tiles + t);
        const BatchLat<C> L = load_uniform(b.lats + pt.lat);
        StepArgs<C> a = lattice_args<S, C>(b, L);
        if constexpr (K == 2) {
            switch (L.kind) {  // uniform per workgroup
            case SW: pair_tile<S, C, SW, WAVES, RPT, NT>(a, pt.tx, pt.ty, lds); break;
            case CF: pair_tile<S, C, CF, WAVES, RPT, NT>(a, pt.tx, pt.ty, lds); break;
            case SL: pair_tile<S, C, SL, WAVES, RPT, NT>(a, pt.tx, pt.ty, lds); break;
            default: pair_tile<S, C, BF, WAVES, RPT, NT>(a, pt.tx, pt.ty, lds); break;
            }
        } else {
            switch (L.kind) {
            case SW: multi_tile<S, C, SW, WAVES, RPT, K>(a, pt.tx, pt.ty, lds, raw); break;
            case CF: multi_tile<S, C, CF, WAVES, RPT, K>(a, pt.tx, pt.ty, lds, raw); break;
            case SL: multi_tile<S, C, SL, WAVES, RPT, K>(a, pt.tx, pt.ty, lds, raw); break;
            default: multi_tile<S, C, BF, WAVES, RPT, K>(a, pt.tx, pt.ty, lds, raw); break;
            }
        }
        if (p + (int)gridDim.x < n) __syncthreads();  // the next tile rewrites the LDS
    }
}

// Equilibrium initial state of every lattice (as init_kernel: shear-wave
// lattices start from their u_x(y) profile, the others at rest).
template <typename S, typename C, int TW>
__global__ void __launch_bounds__(TW) batch_init_kernel(BatchArgs<C> b) {
    constexpr int WAVES = TW / 64;
    const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6), lane = (int)threadIdx.x & 63;
    for (int t = (int)blockIdx.x * WAVES + wave; t < b.n_tiles; t += (int)gridDim.x * WAVES) {
        const Seg sg = load_uniform(b.segs + t);
        const BatchLat<C> L = load_uniform(b.lats + sg.lat);
        const int y = sg.y;
        const int x = (t - L.tile_start - y * L.segs_per_row) * 64 + lane;
        if (x >= L.nx) continue;
        const int q = y * L.nx + x;
        typename Store<S>::R *g = static_cast<typename Store<S>::R *>(b.dst_pool) + L.off;
        const size_t c = (size_t)y * (size_t)L.pitch + (size_t)x;
        const bool sw = L.kind == SW;
        const C ux = sw ? b.prof[L.prof_off + y] : (C)0;
        const C uy = 0;
        const C usq = ux * ux;
#pragma unroll
        for (int d = 0; d < 9; ++d) {
            C feq;
            if (sw) {
                const C cu = (C)cx_of(d) * ux + (C)cy_of(d) * uy;
                feq = b.w[d] * L.rho_0 * ((((C)1 + (C)3 * cu) + (C)4.5 * cu * cu) - (C)1.5 * usq);
            } else {
                feq = b.w[d] * L.rho_0;
            }
            g[(size_t)d * (size_t)L.plane + c] = pop_out<S>(Store<S>::from_f(feq, b.w[d], L.scale), d);
        }
        const size_t m = (size_t)L.macro_off + (size_t)q;
        b.rho_o[m] = L.rho_0;
        b.ux_o[m] = ux;
        b.uy_o[m] = uy;
    }
}

}  // namespace

struct ffb_batch {
    int precision = FFB_PREC_FP32;
    int device = 0;
    size_t elem = 4;          // bytes per stored population
    size_t celem = 4;         // bytes per macro / profile value (the compute type)
    int rounding = FFB_ROUND_STOCHASTIC;
    void *step_ctr = nullptr; // steps done (one unsigned), for the stochastic rounding
    std::vector<ffb_batch_lattice_t> desc;
    std::vector<long long> off, plane, macro_off, prof_off;
    std::vector<int> pitch;
    long long pool_elems = 0, macro_elems = 0, prof_elems = 0, cells = 0;
    void *pool[2] = {nullptr, nullptr};
    int cur = 0;
    void *macro[3] = {nullptr, nullptr, nullptr};
    void *prof = nullptr;
    void *lats = nullptr;     // BatchLat<C>[count]
    void *segs = nullptr;     // Seg per row segment: (lattice, row)
    int n_tiles = 0;
    bool pairs = false;       // two or three steps per launch (batch_pair_kernel)
    int per_pass = 1;         // steps per launch: 1, or 2 .. 5 (multi-step launches)
    void *ptiles = nullptr;   // PTile per multi-step tile (capacity ptile_cap)
    int n_ptiles = 0, ptile_cap = 0;
    int ptile_shape = 0;      // 4 pair_shape + steps per launch of the tile table
    int cus = 0;              // compute units of the device
    ffb_batch_tuning_t tuning{};
    hipStream_t stream = nullptr;
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    bool timing = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    int graph_pass = 0;       // B->per_pass when the graph was captured
    long long steps = 0;
};

namespace {

// workgroups of a single-step launch: one per TW/64 segments, or the persistent grid
int grid_for(const ffb_batch *B) {
    const int waves = B->tuning.tile_w / 64;
    const int all = (B->n_tiles + waves - 1) / waves;
    const int g = B->tuning.grid_blocks;
    return (g <= 0 || g > all) ? all : g;
}

// region rows of a multi-step tile: 16 (8 waves x 2 rows), 32 (16 waves x 2
// rows, shape 1); four steps 24 (12 waves x 2 rows) or 32; five steps 24 (12
// waves x 2 rows) or 32 (8 waves x 4 rows, shape 1); FP64 pairs 8 (8 waves x 1
// row), FP64 triples 16.  Output tiles of K steps are (64 - 2(K-1)) x
// (rows - 2(K-1)): pairs 62 x 14 / 62 x 30 (FP64 62 x 6), triples 60 x 12 /
// 60 x 28 (FP64 60 x 12), four steps 58 x 18 / 58 x 26, five steps 56 x 16 /
// 56 x 24.
int region_rows(const ffb_batch *B, int shape, int k) {
    if (B->elem == 8) return k >= 3 ? 16 : 8;
    if (k >= 4) return shape == 1 ? 32 : 24;
    return shape == 1 ? 32 : 16;
}

// lattices in table order: grouped by kind (stable), so resident waves share a code path
std::vector<int> lattice_order(const ffb_batch *B) {
    std::vector<int> order(B->desc.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return B->desc[(size_t)a].kind < B->desc[(size_t)b].kind; });
    return order;
}

// Step-pair tiles of a region shape, lattice by lattice in table order.
std::vector<PTile> pair_tiles(const ffb_batch *B, int shape, int k) {
    std::vector<PTile> ptiles;
    const int out_rows = region_rows(B, shape, k) - 2 * (k - 1), out_w = 64 - 2 * (k - 1);
    for (int i : lattice_order(B)) {
        const ffb_batch_lattice_t &d = B->desc[(size_t)i];
        for (int ty = 0; ty < (d.n_y + out_rows - 1) / out_rows; ++ty)
            for (int tx = 0; tx < (d.n_x + out_w - 1) / out_w; ++tx) ptiles.push_back(PTile{i, tx, ty, 0});
    }
    return ptiles;
}

// Descriptor, segment and tile tables, built once at creation: the tile
// table's allocation holds the larger of the two shapes' tile lists.
template <typename C> int upload_tables(ffb_batch *B) {
    const int n = (int)B->desc.size();
    std::vector<BatchLat<C>> lats((size_t)n);
    std::vector<Seg> tiles;
    for (int i : lattice_order(B)) {
        const ffb_batch_lattice_t &d = B->desc[(size_t)i];
        BatchLat<C> &L = lats[(size_t)i];
        L.off = B->off[(size_t)i];
        L.plane = B->plane[(size_t)i];
        L.macro_off = B->macro_off[(size_t)i];
        L.prof_off = B->prof_off[(size_t)i];
        L.nx = d.n_x;
        L.ny = d.n_y;
        L.pitch = B->pitch[(size_t)i];
        L.kind = d.kind;
        L.segs_per_row = (d.n_x + 63) / 64;
        L.tile_start = (int)tiles.size();
        for (int y = 0; y < d.n_y; ++y)
            for (int k = 0; k < L.segs_per_row; ++k) tiles.push_back(Seg{i, y});
        L.omega = (C)d.omega;
        const C w6 = (C)1 / (C)36;
        L.force = (C)d.body_force;
        L.rho_0 = (C)d.rho_0;
        // deviation forms: the lattice's 2^s (the 1-3-12 form works on D * 2^-120)
        const bool dev = B->precision != FFB_PREC_FP32 && B->precision != FFB_PREC_FP64;
        const int sw = dev ? d.storage_scale_log2 - (B->precision == FFB_PREC_E3M12_DEV ? 120 : 0) : 0;
        L.scale = (C)std::ldexp(1.0, sw);
        L.inv_scale = (C)std::ldexp(1.0, -sw);
        // as couette_flow.f90:118-119, ((6*w_6)*rho_0)*u, with u*2^s in the deviation forms
        L.corr8 = (C)6 * w6 * (C)d.rho_0 * (dev ? (C)d.u_wall * L.scale : (C)d.u_wall);
        L.corr9 = L.corr8;
        // collide_dev constants, as the single-domain form computes them:
        // omega w_d 2^s with the exact weights as a float pair, the Guo source
        // factors (w_d ((1 - omega/2) F)) 2^s
        L.om1 = (C)1 - L.omega;
        const C pre = ((C)1 - (C)0.5 * L.omega) * L.force;
        for (int dd = 0; dd < 9; ++dd) {
            const double wd = dd == 0 ? 4.0 / 9.0 : (dd <= 4 ? 1.0 / 9.0 : 1.0 / 36.0);
            const double kd = (double)L.omega * wd * (double)L.scale;
            L.kw[dd] = (C)kd;
            L.kw_lo[dd] = (C)(kd - (double)L.kw[dd]);
            const C w = dd == 0 ? (C)4 / (C)9 : (dd <= 4 ? (C)1 / (C)9 : (C)1 / (C)36);
            L.kf[dd] = (w * pre) * L.scale;
        }
        L.half_force = (C)0.5 * L.force;
    }
    size_t cap = 0;
    for (int k = 2; k <= 5; ++k)
        for (int sh = 1; sh <= 2; ++sh) cap = std::max(cap, pair_tiles(B, sh, k).size());
    B->ptile_cap = (int)cap;
    BHIPCHK(hipMalloc(&B->ptiles, (size_t)B->ptile_cap * sizeof(PTile)));
    BHIPCHK(hipMalloc(&B->lats, lats.size() * sizeof(BatchLat<C>)));
    BHIPCHK(hipMalloc(&B->segs, tiles.size() * sizeof(Seg)));
    BHIPCHK(hipMemcpy(B->lats, lats.data(), lats.size() * sizeof(BatchLat<C>), hipMemcpyHostToDevice));
    BHIPCHK(hipMemcpy(B->segs, tiles.data(), tiles.size() * sizeof(Seg), hipMemcpyHostToDevice));
    B->n_tiles = (int)tiles.size();
    B->ptile_shape = 0;
    return FFB_OK;
}

int drop_graph(ffb_batch *B) {
    if (B->graph_exec) (void)hipGraphExecDestroy(B->graph_exec);
    if (B->graph) (void)hipGraphDestroy(B->graph);
    B->graph_exec = nullptr;
    B->graph = nullptr;
    return FFB_OK;
}

// Writes the tile list of B->tuning.pair_shape into the tile table (in place,
// synchronously; the graph, whose launches read the table, is dropped first).
int write_pair_tiles(ffb_batch *B) {
    const int shape = B->tuning.pair_shape, k = std::max(2, B->per_pass);
    if (B->ptile_shape == 4 * shape + k) return FFB_OK;
    drop_graph(B);
    BHIPCHK(hipStreamSynchronize(B->stream));
    const std::vector<PTile> pt = pair_tiles(B, shape, k);
    if ((int)pt.size() > B->ptile_cap) return fail(FFB_E_STATE, "ffb_batch: tile table too small");
    BHIPCHK(hipMemcpy(B->ptiles, pt.data(), pt.size() * sizeof(PTile), hipMemcpyHostToDevice));
    B->n_ptiles = (int)pt.size();
    B->ptile_shape = 4 * shape + k;
    return FFB_OK;
}

template <typename C> BatchArgs<C> batch_args(const ffb_batch *B, int parity, int write_macro) {
    BatchArgs<C> b{};
    b.src_pool = B->pool[parity];
    b.dst_pool = B->pool[1 - parity];
    b.lats = static_cast<const BatchLat<C> *>(B->lats);
    b.segs = static_cast<const Seg *>(B->segs);
    b.n_tiles = B->n_tiles;
    b.write_macro = write_macro;
    b.rho_o = static_cast<C *>(B->macro[0]);
    b.ux_o = static_cast<C *>(B->macro[1]);
    b.uy_o = static_cast<C *>(B->macro[2]);
    b.prof = static_cast<const C *>(B->prof);
    b.step_ctr = static_cast<const unsigned *>(B->step_ctr);
    b.stochastic = B->rounding == FFB_ROUND_STOCHASTIC;
    b.w[0] = (C)4 / (C)9;
    for (int d = 1; d <= 4; ++d) b.w[d] = (C)1 / (C)9;
    for (int d = 5; d <= 8; ++d) b.w[d] = (C)1 / (C)36;
    return b;
}

template <typename S, typename C> void launch_step_c(const ffb_batch *B, int parity, int write_macro) {
    const BatchArgs<C> b = batch_args<C>(B, parity, write_macro);
    const dim3 grid((unsigned)grid_for(B));
    switch (B->tuning.tile_w) {
    case 64: hipLaunchKernelGGL((batch_step_kernel<S, C, 64>), grid, dim3(64), 0, B->stream, b); break;
    case 128: hipLaunchKernelGGL((batch_step_kernel<S, C, 128>), grid, dim3(128), 0, B->stream, b); break;
    case 512: hipLaunchKernelGGL((batch_step_kernel<S, C, 512>), grid, dim3(512), 0, B->stream, b); break;
    case 1024: hipLaunchKernelGGL((batch_step_kernel<S, C, 1024>), grid, dim3(1024), 0, B->stream, b); break;
    default: hipLaunchKernelGGL((batch_step_kernel<S, C, 256>), grid, dim3(256), 0, B->stream, b); break;
    }
}

bool is_16bit(int p) { return p == FFB_PREC_FP16_DEV || p == FFB_PREC_BF16_DEV || p == FFB_PREC_E3M12_DEV; }

void count_steps(const ffb_batch *B, unsigned n) {
    if (is_16bit(B->precision))  // the next launch computes the next step
        hipLaunchKernelGGL(batch_count_kernel, dim3(1), dim3(1), 0, B->stream, static_cast<unsigned *>(B->step_ctr), 0,
                           n);
}

void launch_step(const ffb_batch *B, int parity, int write_macro) {
    switch (B->precision) {
    case FFB_PREC_FP64: launch_step_c<double, double>(B, parity, write_macro); break;
    case FFB_PREC_FP32_DEV: launch_step_c<f32dev, float>(B, parity, write_macro); break;
    case FFB_PREC_FP16_DEV: launch_step_c<__half, float>(B, parity, write_macro); break;
    case FFB_PREC_BF16_DEV: launch_step_c<__hip_bfloat16, float>(B, parity, write_macro); break;
    case FFB_PREC_E3M12_DEV: launch_step_c<e3m12, float>(B, parity, write_macro); break;
    default: launch_step_c<float, float>(B, parity, write_macro); break;
    }
    count_steps(B, 1u);
}

// One launch of K steps (pairs or triples) of every lattice.
template <typename S, typename C, int WAVES, int RPT, bool NT, int K>
void launch_pair_c(const ffb_batch *B, int parity, int write_macro) {
    const BatchArgs<C> b = batch_args<C>(B, parity, write_macro);
    const int n = B->n_ptiles, pg = B->tuning.pair_grid;
    const int grid = pg > 0 ? std::min(n, ((pg + 7) / 8) * 8) : n;
    hipLaunchKernelGGL((batch_pair_kernel<S, C, WAVES, RPT, NT, K>), dim3((unsigned)grid), dim3(64 * WAVES), 0,
                       B->stream, b, static_cast<const PTile *>(B->ptiles), n, B->tuning.pair_group);
}

template <typename S, typename C, int K> void launch_pair_shape(const ffb_batch *B, int parity, int write_macro) {
    if (B->tuning.pair_shape == 1) {  // 64 x 32 regions
        if constexpr (K >= 3) launch_pair_c<S, C, 8, 4, false, K>(B, parity, write_macro);
        else launch_pair_c<S, C, 16, 2, false, K>(B, parity, write_macro);
    } else if constexpr (K == 5) {
        launch_pair_c<S, C, 12, 2, false, K>(B, parity, write_macro);
    } else if constexpr (K == 4) {
        launch_pair_c<S, C, 12, 2, false, K>(B, parity, write_macro);
    } else {
        launch_pair_c<S, C, 8, 2, false, K>(B, parity, write_macro);
    }
}

template <int K> void launch_pair_k(const ffb_batch *B, int parity, int write_macro) {
    if constexpr (K >= 4) {  // 4-byte forms only (per_pass is at most 3 for the others)
        if (B->precision == FFB_PREC_FP32_DEV) launch_pair_shape<f32dev, float, K>(B, parity, write_macro);
        else launch_pair_shape<float, float, K>(B, parity, write_macro);
        return;
    }
    switch (B->precision) {
    case FFB_PREC_FP64:
        if constexpr (K == 2) launch_pair_c<double, double, 8, 1, true, 2>(B, parity, write_macro);
        else launch_pair_c<double, double, 8, 2, false, K>(B, parity, write_macro);
        break;
    case FFB_PREC_FP32_DEV: launch_pair_shape<f32dev, float, K>(B, parity, write_macro); break;
    case FFB_PREC_FP16_DEV: launch_pair_shape<__half, float, K>(B, parity, write_macro); break;
    case FFB_PREC_BF16_DEV: launch_pair_shape<__hip_bfloat16, float, K>(B, parity, write_macro); break;
    case FFB_PREC_E3M12_DEV: launch_pair_shape<e3m12, float, K>(B, parity, write_macro); break;
    default: launch_pair_shape<float, float, K>(B, parity, write_macro); break;
    }
}

// One launch of B->per_pass (2 or 3) steps.  Default-policy stores as the
// single domain's pairs (FP64 pairs: nontemporal); regions 8 waves x 2 rows or
// 16 waves x 2 rows, FP64 pairs 8 waves x 1 row, FP64 triples 8 waves x 2 rows.
void launch_pair(const ffb_batch *B, int parity, int write_macro) {
    if (B->per_pass >= 5) launch_pair_k<5>(B, parity, write_macro);
    else if (B->per_pass == 4) launch_pair_k<4>(B, parity, write_macro);
    else if (B->per_pass == 3) launch_pair_k<3>(B, parity, write_macro);
    else launch_pair_k<2>(B, parity, write_macro);
    count_steps(B, (unsigned)B->per_pass);
}

// Captures the launches of one ping-pong cycle (parity 0 then 1): two
// multi-step launches (four or six steps) or two single steps.
int build_graph(ffb_batch *B) {
    drop_graph(B);
    BHIPCHK(hipStreamBeginCapture(B->stream, hipStreamCaptureModeThreadLocal));
    if (B->pairs) {
        launch_pair(B, 0, 0);
        launch_pair(B, 1, 0);
    } else {
        launch_step(B, 0, 0);
        launch_step(B, 1, 0);
    }
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(B->stream, &g);
    if (e != hipSuccess) return fail(FFB_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    B->graph = g;
    BHIPCHK(hipGraphInstantiate(&B->graph_exec, B->graph, nullptr, nullptr, 0));
    B->graph_pass = B->per_pass;
    return FFB_OK;
}

// The graph for the current tuning and pass mode, captured if missing or stale.
int ensure_graph(ffb_batch *B) {
    if (B->graph_exec && B->graph_pass == B->per_pass) return FFB_OK;
    return build_graph(B);
}

template <typename S, typename C> int init_c(ffb_batch *B) {
    BatchArgs<C> b = batch_args<C>(B, 1, 0);  // dst = pool[0]
    const dim3 grid((unsigned)((B->n_tiles + 3) / 4));
    hipLaunchKernelGGL((batch_init_kernel<S, C, 256>), grid, dim3(256), 0, B->stream, b);
    BHIPCHK(hipGetLastError());
    BHIPCHK(hipStreamSynchronize(B->stream));
    return FFB_OK;
}

template <typename S, typename C> int init_any(ffb_batch *B) {
    if (is_16bit(B->precision)) {
        hipLaunchKernelGGL(batch_count_kernel, dim3(1), dim3(1), 0, B->stream, static_cast<unsigned *>(B->step_ctr), 1,
                           0u);
        BHIPCHK(hipGetLastError());
    }
    return init_c<S, C>(B);
}

int apply_tuning(ffb_batch *B, const ffb_batch_tuning_t &t) {
    if (t.tile_w != 64 && t.tile_w != 128 && t.tile_w != 256 && t.tile_w != 512 && t.tile_w != 1024)
        return fail(FFB_E_INVALID, "ffb_batch: tile_w must be 64, 128, 256, 512 or 1024");
    if (t.grid_blocks < 0 || t.pair_grid < 0 || t.pair_group < 0)
        return fail(FFB_E_INVALID, "ffb_batch: grid_blocks, pair_grid and pair_group must be >= 0");
    if (t.pair_shape != 1 && t.pair_shape != 2) return fail(FFB_E_INVALID, "ffb_batch: pair_shape must be 1 or 2");
    // the graph refers to the tuning (and to the tile table): dropped first,
    // captured again by the next ffb_batch_step
    drop_graph(B);
    BHIPCHK(hipStreamSynchronize(B->stream));
    B->tuning = t;
    if (B->elem == 8) B->tuning.pair_shape = 2;  // FP64 has one region shape (its tiles do not depend on it)
    return write_pair_tiles(B);
}

// Tuning before the autotuner runs: 256-thread workgroups, one per four row
// segments, for single steps; step pairs in 62 x 14 tiles (FP64 62 x 6), XCD
// groups of 32 tiles (FP64 4), one workgroup per tile.  FFB_BATCH_PAIR_GROUP,
// FFB_BATCH_PAIR_SHAPE and FFB_BATCH_PAIR_GRID override the pair settings.
ffb_batch_tuning_t ffb_batch_default_tuning(const ffb_batch *B) {
    auto env = [](const char *name, int dflt) {
        const char *v = std::getenv(name);
        return v && *v ? std::atoi(v) : dflt;
    };
    ffb_batch_tuning_t t{};
    t.tile_w = 256;
    t.grid_blocks = 0;
    // region shape: three or more steps 64 x 32 regions of 8 waves x 4 rows
    // (shape 1), pairs 64 x 16 of 8 waves x 2 rows (shape 2)
    t.pair_shape = env("FFB_BATCH_PAIR_SHAPE", B->per_pass >= 3 ? 1 : 2) == 1 ? 1 : 2;
    t.pair_group = std::max(0, env("FFB_BATCH_PAIR_GROUP", B->elem >= 8 ? 4 : 32));
    t.pair_grid = std::max(0, env("FFB_BATCH_PAIR_GRID", 0));
    return t;
}

template <typename C> int read_macros_c(ffb_batch *B, int i, void *rho, void *ux, void *uy) {
    const size_t n = (size_t)B->desc[(size_t)i].n_x * (size_t)B->desc[(size_t)i].n_y;
    void *dst[3] = {rho, ux, uy};
    for (int k = 0; k < 3; ++k)
        if (dst[k])
            BHIPCHK(hipMemcpyAsync(dst[k], static_cast<C *>(B->macro[k]) + B->macro_off[(size_t)i], n * sizeof(C),
                                   hipMemcpyDeviceToHost, B->stream));
    BHIPCHK(hipStreamSynchronize(B->stream));
    return FFB_OK;
}

}  // namespace

extern "C" {

int ffb_batch_create(const ffb_batch_lattice_t *lattices, int32_t count, int32_t precision, int32_t device,
                     ffb_batch_t **out) {
    if (!lattices || !out || count < 1) return fail(FFB_E_INVALID, "ffb_batch_create: bad argument");
    *out = nullptr;
    if (precision < FFB_PREC_FP32 || precision > FFB_PREC_E3M12_DEV)
        return fail(FFB_E_INVALID, "ffb_batch_create: unknown precision " + std::to_string(precision));
    for (int i = 0; i < count; ++i) {
        const ffb_batch_lattice_t &d = lattices[i];
        if (precision != FFB_PREC_FP32 && precision != FFB_PREC_FP64 &&
            (d.storage_scale_log2 < -30 || d.storage_scale_log2 > 30))
            return fail(FFB_E_INVALID, "ffb_batch_create: lattice " + std::to_string(i) +
                                           " storage_scale_log2 outside [-30, 30]");
        // (n_x <= 2^24 keeps the nine direction rows of a lattice row within
        // 4 GiB, what the one-descriptor stores of the 16-bit kernels address)
        if (d.n_x < 1 || d.n_y < 1 || d.n_x > (1 << 24) || (long long)d.n_x * d.n_y > (1LL << 30))
            return fail(FFB_E_INVALID, "ffb_batch_create: lattice " + std::to_string(i) + " has a bad size");
        if (d.kind != FFB_SHEAR_WAVE && d.kind != FFB_COUETTE_FLOW && d.kind != FFB_SLIDING_LID &&
            d.kind != FFB_BODY_FORCE_CHANNEL)
            return fail(FFB_E_INVALID, "ffb_batch_create: lattice " + std::to_string(i) +
                                           " kind must be shear wave, Couette, sliding lid or body-force channel");
    }
    int ndev = 0;
    BHIPCHK(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(FFB_E_INVALID, "ffb_batch_create: device not present");
    BHIPCHK(hipSetDevice(device));

    ffb_batch *B = new ffb_batch();
    B->precision = precision;
    B->device = device;
    B->elem = precision == FFB_PREC_FP64 ? 8 : (is_16bit(precision) ? 2 : 4);
    B->celem = precision == FFB_PREC_FP64 ? 8 : 4;
    // step pairs in one launch for the 4- and 8-byte forms, one launch per step
    // for the 16-bit forms (512 mixed lattices, autotuned, MLUPS: FP32 121,000
    // in pairs against 83,200; FP32 deviations 127,000 / 83,700; FP64 63,000 /
    // 39,100; FP16 127,000 in pairs against 144,100, BF16 129,000 / 141,900,
    // 1-3-12 120,500 / 138,200); FFB_BATCH_PAIRS=0/1 at creation overrides
    // and multi-step launches of 2..5 steps (FFB_BATCH_STEPS_PER_PASS; default
    // 5 for the 4-byte forms, 3 for FP64 and the 16-bit forms; four and five
    // steps only for the 4-byte forms).  512 mixed lattices, autotuned, one
    // box: FP32 five steps 169,700-178,000 MLUPS against 150,700-168,200 with
    // four; FP16 deviations three steps in 64 x 32 regions (8 waves x 4 rows)
    // 177,500 against 144,100 with one launch per step (129,800 in pairs).
    {
        const char *e = std::getenv("FFB_BATCH_PAIRS");
        B->pairs = e && *e ? std::atoi(e) != 0 : true;
        const char *k = std::getenv("FFB_BATCH_STEPS_PER_PASS");
        int kk = k && *k ? std::atoi(k) : (B->pairs ? (B->elem == 4 ? 5 : 3) : 1);
        if (!(e && *e) && k && *k) B->pairs = kk >= 2;  // the steps per pass alone selects
        kk = std::max(1, std::min(B->elem == 4 ? 5 : 3, kk));
        B->per_pass = B->pairs ? std::max(2, kk) : 1;
    }
    B->desc.assign(lattices, lattices + count);
    long long off = 0, moff = 0, poff = 0;
    // each lattice in the row-blocked layout of the single domain: per row y
    // the nine direction rows (n_x rounded up to 64) one after another
    for (int i = 0; i < count; ++i) {
        const ffb_batch_lattice_t &d = lattices[i];
        const long long plane = ((d.n_x + 63) / 64) * 64;  // direction stride
        const int pitch = (int)(9 * plane);                  // row stride
        B->pitch.push_back(pitch);
        B->plane.push_back(plane);
        B->off.push_back(off);
        B->macro_off.push_back(moff);
        B->prof_off.push_back(poff);
        off += (long long)pitch * d.n_y;
        moff += (long long)d.n_x * d.n_y;
        poff += d.kind == FFB_SHEAR_WAVE ? d.n_y : 0;
        B->cells += (long long)d.n_x * d.n_y;
    }
    B->pool_elems = off;
    B->macro_elems = moff;
    B->prof_elems = std::max(1LL, poff);

    auto bail = [&](hipError_t e, const char *what) {
        ffb_batch_destroy(B);
        return fail(FFB_E_HIP, std::string("ffb_batch_create: ") + what + ": " + hipGetErrorString(e));
    };
    hipError_t e;
    for (int k = 0; k < 2; ++k)
        if ((e = hipMalloc(&B->pool[k], (size_t)B->pool_elems * B->elem)) != hipSuccess) return bail(e, "pools");
    for (int k = 0; k < 3; ++k)
        if ((e = hipMalloc(&B->macro[k], (size_t)B->macro_elems * B->celem)) != hipSuccess) return bail(e, "macros");
    if ((e = hipMalloc(&B->prof, (size_t)B->prof_elems * B->celem)) != hipSuccess) return bail(e, "profiles");
    if ((e = hipMalloc(&B->step_ctr, 4)) != hipSuccess) return bail(e, "step counter");
    if ((e = hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking)) != hipSuccess) return bail(e, "stream");
    if ((e = hipEventCreate(&B->ev_start)) != hipSuccess) return bail(e, "event");
    if ((e = hipEventCreate(&B->ev_stop)) != hipSuccess) return bail(e, "event");

    // shear-wave profiles u_x(y) = u_max sin(2 pi n_sin y / n_y), evaluated on the
    // host with the C library's sin in the storage precision (as the single-domain
    // initial condition, initialization.f90:42)
    {
        std::vector<unsigned char> prof((size_t)B->prof_elems * B->celem, 0);
        for (int i = 0; i < count; ++i) {
            const ffb_batch_lattice_t &d = lattices[i];
            if (d.kind != FFB_SHEAR_WAVE) continue;
            for (int y = 0; y < d.n_y; ++y) {
                const size_t k = (size_t)B->prof_off[(size_t)i] + (size_t)y;
                if (precision == FFB_PREC_FP64) {
                    const double pi = 3.141592653589793238462643383279502884197;
                    ((double *)prof.data())[k] = d.u_max * std::sin(((2.0 * pi) * d.n_sin * (double)y) / (double)d.n_y);
                } else {
                    const float pi = (float)3.141592653589793238462643383279502884197L;
                    ((float *)prof.data())[k] = (float)d.u_max *
                        std::sin((((float)2 * pi) * (float)d.n_sin * (float)y) / (float)d.n_y);
                }
            }
        }
        if ((e = hipMemcpy(B->prof, prof.data(), prof.size(), hipMemcpyHostToDevice)) != hipSuccess)
            return bail(e, "profiles");
    }
    if ((e = hipDeviceGetAttribute(&B->cus, hipDeviceAttributeMultiprocessorCount, device)) != hipSuccess)
        return bail(e, "device attributes");
    int rc = B->precision == FFB_PREC_FP64 ? upload_tables<double>(B) : upload_tables<float>(B);
    if (!rc) rc = apply_tuning(B, ffb_batch_default_tuning(B));
    if (rc) {
        ffb_batch_destroy(B);
        return rc;
    }
    *out = B;
    return FFB_OK;
}

int ffb_batch_destroy(ffb_batch_t *B) {
    if (!B) return FFB_OK;
    (void)hipSetDevice(B->device);
    if (B->stream) (void)hipStreamSynchronize(B->stream);
    drop_graph(B);
    for (void *p : B->pool) if (p) (void)hipFree(p);
    for (void *p : B->macro) if (p) (void)hipFree(p);
    if (B->prof) (void)hipFree(B->prof);
    if (B->step_ctr) (void)hipFree(B->step_ctr);
    if (B->lats) (void)hipFree(B->lats);
    if (B->segs) (void)hipFree(B->segs);
    if (B->ptiles) (void)hipFree(B->ptiles);
    if (B->ev_start) (void)hipEventDestroy(B->ev_start);
    if (B->ev_stop) (void)hipEventDestroy(B->ev_stop);
    if (B->stream) (void)hipStreamDestroy(B->stream);
    delete B;
    return FFB_OK;
}

int ffb_batch_init(ffb_batch_t *B) {
    if (!B) return fail(FFB_E_INVALID, "ffb_batch_init: null handle");
    BHIPCHK(hipSetDevice(B->device));
    B->cur = 0;
    B->steps = 0;
    B->timing = false;
    switch (B->precision) {
    case FFB_PREC_FP64: return init_any<double, double>(B);
    case FFB_PREC_FP32_DEV: return init_any<f32dev, float>(B);
    case FFB_PREC_FP16_DEV: return init_any<__half, float>(B);
    case FFB_PREC_BF16_DEV: return init_any<__hip_bfloat16, float>(B);
    case FFB_PREC_E3M12_DEV: return init_any<e3m12, float>(B);
    default: return init_any<float, float>(B);
    }
}

int ffb_batch_step(ffb_batch_t *B, int32_t n_steps, int32_t write_macro_last) {
    if (!B) return fail(FFB_E_INVALID, "ffb_batch_step: null handle");
    if (n_steps < 0) return fail(FFB_E_INVALID, "ffb_batch_step: n_steps < 0");
    if (n_steps == 0) return FFB_OK;
    BHIPCHK(hipSetDevice(B->device));
    if (!B->timing) {
        BHIPCHK(hipEventRecord(B->ev_start, B->stream));
        B->timing = true;
    }
    int left = n_steps;
    // the last step runs outside the graph when it writes the macro fields
    const int tail = write_macro_last ? 1 : 0;
    const int k = B->per_pass;
    if (left - tail >= 2 * k) {
        int rc = ensure_graph(B);
        if (rc) return rc;
    }
    if (B->pairs) {
        // launches of k steps each; the captured graph holds two of them
        // (parity 0 then 1, 2k steps); a remainder below k runs as single steps
        if (B->cur == 1 && left - tail >= k) {
            launch_pair(B, 1, 0);
            B->cur = 0;
            left -= k;
        }
        while (left - tail >= 2 * k) {
            BHIPCHK(hipGraphLaunch(B->graph_exec, B->stream));
            left -= 2 * k;
        }
        while (left - tail >= k) {
            launch_pair(B, B->cur, 0);
            B->cur = 1 - B->cur;
            left -= k;
        }
    } else {
        // align to an even parity so the captured pair (0 then 1) applies
        if (B->cur == 1 && left - tail > 0) {
            launch_step(B, 1, 0);
            B->cur = 0;
            --left;
        }
        while (left - tail >= 2) {
            BHIPCHK(hipGraphLaunch(B->graph_exec, B->stream));
            left -= 2;
        }
    }
    while (left > 0) {
        launch_step(B, B->cur, (left == 1 && write_macro_last) ? 1 : 0);
        B->cur = 1 - B->cur;
        --left;
    }
    BHIPCHK(hipGetLastError());
    BHIPCHK(hipEventRecord(B->ev_stop, B->stream));
    B->steps += n_steps;
    return FFB_OK;
}

int ffb_batch_sync(ffb_batch_t *B) {
    if (!B) return fail(FFB_E_INVALID, "ffb_batch_sync: null handle");
    BHIPCHK(hipStreamSynchronize(B->stream));
    return FFB_OK;
}

int ffb_batch_elapsed_ms(ffb_batch_t *B, double *ms) {
    if (!B || !ms) return fail(FFB_E_INVALID, "ffb_batch_elapsed_ms: null argument");
    *ms = 0;
    if (!B->timing) return FFB_OK;
    BHIPCHK(hipEventSynchronize(B->ev_stop));
    float t = 0.f;
    BHIPCHK(hipEventElapsedTime(&t, B->ev_start, B->ev_stop));
    *ms = t;
    return FFB_OK;
}

int ffb_batch_reset_timer(ffb_batch_t *B) {
    if (!B) return fail(FFB_E_INVALID, "ffb_batch_reset_timer: null handle");
    B->timing = false;
    return FFB_OK;
}

int ffb_batch_info(ffb_batch_t *B, int64_t *cells, int32_t *tiles, ffb_batch_tuning_t *tuning) {
    if (!B) return fail(FFB_E_INVALID, "ffb_batch_info: null handle");
    if (cells) *cells = B->cells;
    if (tiles) *tiles = B->n_tiles;
    if (tuning) *tuning = B->tuning;
    return FFB_OK;
}

int ffb_batch_set_tuning(ffb_batch_t *B, const ffb_batch_tuning_t *t) {
    if (!B || !t) return fail(FFB_E_INVALID, "ffb_batch_set_tuning: null argument");
    BHIPCHK(hipSetDevice(B->device));
    return apply_tuning(B, *t);
}

int ffb_batch_autotune(ffb_batch_t *B, int32_t trial_steps, ffb_batch_tuning_t *best_out, double *best_ms) {
    if (!B) return fail(FFB_E_INVALID, "ffb_batch_autotune: null handle");
    if (trial_steps < 2) trial_steps = 2;
    trial_steps += trial_steps & 1;  // whole ping-pong cycles of single steps; multi-step: trial_steps / k launches
    BHIPCHK(hipSetDevice(B->device));
    // Candidates of the kernel the batch runs for consecutive steps, each timed
    // by direct launches (no graph is captured until the winner is known):
    //   multi-step launches: region shape (1: 64 x 32, 2: 64 x 16 or 64 x 24;
    //     FP64 one shape) x XCD group (4, 6, 8, 12, 16, 32, 64 tiles) x grid (one
    //     workgroup per tile, or a persistent grid of 1, 2 or 4 rounds of 2 or 4
    //     workgroups per CU)
    //   one launch per step: workgroup size (64 .. 1024 threads) x persistent
    //     grid (0: one workgroup per tile; 2 .. 384 workgroups per CU)
    std::vector<ffb_batch_tuning_t> cands;
    const ffb_batch_tuning_t base = B->tuning;
    if (B->pairs) {
        const int shapes[] = {2, 1};
        const int groups[] = {4, 6, 8, 12, 16, 32, 64};
        for (int sh : shapes) {
            if (B->elem == 8 && sh == 1) continue;
            const int resident = (sh == 1 ? 2 : 4) * B->cus;
            const int grids[] = {0, resident, 2 * resident, 4 * resident};
            for (int g : groups)
                for (int gr : grids) {
                    ffb_batch_tuning_t t = base;
                    t.pair_shape = sh;
                    t.pair_group = g;
                    t.pair_grid = gr;
                    cands.push_back(t);
                }
        }
    } else {
        const int tws[] = {64, 128, 256, 512, 1024};
        const int per_cu[] = {0, 2, 4, 6, 7, 8, 12, 14, 16, 28, 48, 96, 192, 384};
        for (int tw : tws)
            for (int pc : per_cu) {
                ffb_batch_tuning_t t = base;
                t.tile_w = tw;
                t.grid_blocks = pc ? pc * B->cus : 0;
                cands.push_back(t);
            }
    }
    int rc = ffb_batch_init(B);
    if (rc) return rc;
    ffb_batch_tuning_t best = base;
    double best_t = 1e30;
    const int launches = B->pairs ? std::max(1, trial_steps / B->per_pass) : trial_steps;
    for (const ffb_batch_tuning_t &t : cands) {
        if ((rc = apply_tuning(B, t))) return rc;
        // one warm-up launch, then the timed ones, alternating the pools
        auto one = [&](int k) {
            if (B->pairs) launch_pair(B, (B->cur + k) & 1, 0);
            else launch_step(B, (B->cur + k) & 1, 0);
        };
        one(0);
        BHIPCHK(hipEventRecord(B->ev_start, B->stream));
        for (int k = 1; k <= launches; ++k) one(k);
        BHIPCHK(hipGetLastError());
        BHIPCHK(hipEventRecord(B->ev_stop, B->stream));
        BHIPCHK(hipEventSynchronize(B->ev_stop));
        B->cur = (B->cur + launches + 1) & 1;
        float ms = 0.f;
        BHIPCHK(hipEventElapsedTime(&ms, B->ev_start, B->ev_stop));
        if (ms < best_t) {
            best_t = ms;
            best = t;
        }
    }
    if ((rc = apply_tuning(B, best))) return rc;
    if ((rc = ffb_batch_init(B))) return rc;  // trials advanced the state: start again
    B->timing = false;
    if (best_out) *best_out = B->tuning;
    if (best_ms) *best_ms = best_t / trial_steps;
    return FFB_OK;
}

int ffb_batch_read_macros(ffb_batch_t *B, int32_t index, void *rho, void *u_x, void *u_y) {
    if (!B) return fail(FFB_E_INVALID, "ffb_batch_read_macros: null handle");
    if (index < 0 || index >= (int)B->desc.size()) return fail(FFB_E_INVALID, "ffb_batch_read_macros: bad index");
    BHIPCHK(hipSetDevice(B->device));
    return B->precision == FFB_PREC_FP64 ? read_macros_c<double>(B, index, rho, u_x, u_y)
                                         : read_macros_c<float>(B, index, rho, u_x, u_y);
}

int ffb_batch_read_populations(ffb_batch_t *B, int32_t index, void *f_halo) {
    if (!B || !f_halo) return fail(FFB_E_INVALID, "ffb_batch_read_populations: null argument");
    if (index < 0 || index >= (int)B->desc.size())
        return fail(FFB_E_INVALID, "ffb_batch_read_populations: bad index");
    BHIPCHK(hipSetDevice(B->device));
    const ffb_batch_lattice_t &d = B->desc[(size_t)index];
    const size_t plane = (size_t)B->plane[(size_t)index], pitch = (size_t)B->pitch[(size_t)index];
    std::vector<unsigned char> stage(pitch * (size_t)d.n_y * B->elem);
    BHIPCHK(hipMemcpyAsync(stage.data(), (unsigned char *)B->pool[B->cur] + (size_t)B->off[(size_t)index] * B->elem,
                           stage.size(), hipMemcpyDeviceToHost, B->stream));
    BHIPCHK(hipStreamSynchronize(B->stream));
    const size_t hp = (size_t)d.n_x + 2, hplane = hp * ((size_t)d.n_y + 2);
    std::memset(f_halo, 0, 9 * hplane * B->celem);
    if (B->precision == FFB_PREC_FP32 || B->precision == FFB_PREC_FP64) {
        for (int dd = 0; dd < 9; ++dd)
            for (int y = 0; y < d.n_y; ++y)
                std::memcpy((unsigned char *)f_halo + ((size_t)dd * hplane + (size_t)(y + 1) * hp + 1) * B->elem,
                            stage.data() + ((size_t)dd * plane + (size_t)y * pitch) * B->elem,
                            (size_t)d.n_x * B->elem);
        return FFB_OK;
    }
    // deviation forms: f = w + D * 2^-s, D as the step reads it (the 1-3-12
    // form: the stored value times 2^-off of its direction, on the 2^-120 scale)
    const int s = d.storage_scale_log2 - (B->precision == FFB_PREC_E3M12_DEV ? 120 : 0);
    const float inv = (float)std::ldexp(1.0, -s);
    float *f = static_cast<float *>(f_halo);
    for (int dd = 0; dd < 9; ++dd) {
        const float w = dd == 0 ? 4.0f / 9.0f : (dd <= 4 ? 1.0f / 9.0f : 1.0f / 36.0f);
        for (int y = 0; y < d.n_y; ++y) {
            const size_t src = (size_t)dd * plane + (size_t)y * pitch;
            float *row = f + (size_t)dd * hplane + (size_t)(y + 1) * hp + 1;
            for (size_t x = 0; x < (size_t)d.n_x; ++x) {
                float v;
                if (B->precision == FFB_PREC_FP32_DEV) {
                    v = ((const float *)stage.data())[src + x];
                } else {
                    const uint16_t h = ((const uint16_t *)stage.data())[src + x];
                    uint32_t u;
                    if (B->precision == FFB_PREC_FP16_DEV) {
                        __half hv;
                        std::memcpy(&hv, &h, 2);
                        v = __half2float(hv);
                        std::memcpy(&u, &v, 4);
                    } else if (B->precision == FFB_PREC_BF16_DEV) {
                        u = (uint32_t)h << 16;
                    } else {
                        u = ((uint32_t)(int32_t)(int16_t)h << 11) & 0x83ffffffu;
                    }
                    std::memcpy(&v, &u, 4);
                    if (B->precision == FFB_PREC_E3M12_DEV && store_off<e3m12>(dd) != 0)
                        v = v * pow2f(-store_off<e3m12>(dd));
                }
                row[x] = w + v * inv;
            }
        }
    }
    return FFB_OK;
}

}  // extern "C"
