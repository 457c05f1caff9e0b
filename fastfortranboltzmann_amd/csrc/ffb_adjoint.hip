// ffb_adjoint.hip -- differentiation form: a hand-written discrete adjoint
// (reverse-time pull-stream-collide) of the fused step, with checkpoints
// every K steps and recomputation inside each segment.  Gradients of a loss
// of the final velocity field with respect to omega, the body force and the
// moving-wall (lid) speed.  No reference counterpart.
//
// Forward step t:  p^t = Pull(f^{t-1})  (streaming + boundary rules; each
// population is read exactly once, so Pull is a permutation plus the moving-
// wall constants), f^t = C(p^t; omega, F)  (BGK, Guo forcing for the body-
// force channel).  The loss depends on u^N, the velocity of p^N (the macro
// field the last step writes).  Reverse sweep with lambda_p^t = dL/dp^t:
//   seed:     lambda_p^N = (dL/du) . (c_k - u) / rho                (u of p^N)
//   step t:   lambda_f^t = Pull^T lambda_p^{t+1}   (gather from the reader of
//             each population: the cell at x + c_j, or the cell itself in the
//             opposite direction at a wall)
//             lambda_p^t = (dC/dp)^T lambda_f^t     (local 9x9 transpose
//             product, written out in closed form below)
//   gradients accumulate dC/domega, dC/dF along the way and dPull/du_wall
//   (+-6 w rho_0 on the two top-wall populations of the top row).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "d2q9_kernels.h"
#include "ffb.h"
#include "ffb_internal.h"

using namespace ffb;

#define AHIPCHK(expr)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return ffb::fail(FFB_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));    \
    } while (0)

namespace {

constexpr int ATW = 256;  // threads per workgroup of the adjoint kernels

template <typename C> struct AdjArgs {
    StepArgs<C> s;        // geometry and parameters of the forward step (src = f^{t-1})
    const C *lam_in;      // lambda_p^{t+1}
    C *lam_out;           // lambda_p^t
    const C *g_ux;        // seed: dL/du_x, dL/du_y (n_x*n_y each)
    const C *g_uy;
    double *acc;          // per-workgroup gradient sums [block][omega, F, u_wall]
    C dk_du;              // d(top-wall term)/d(u_wall) = 6 w_6 rho_0 (0 without a moving wall)
};

// Reader of population f_j(x, y) in the next step's pull: (tx, ty, tj).
template <int MODE>
__device__ __forceinline__ void reader_of(int x, int y, int j, int nx, int ny, int &tx, int &ty, int &tj) {
    int xt = x + cx_of(j), yt = y + cy_of(j);
    const bool xo = (xt < 0) | (xt >= nx);
    const bool yo = (yt < 0) | (yt >= ny);
    if (MODE == SW) {
        tx = xt < 0 ? xt + nx : (xt >= nx ? xt - nx : xt);
        ty = yt < 0 ? yt + ny : (yt >= ny ? yt - ny : yt);
        tj = j;
        return;
    }
    if ((MODE == SL) ? (xo | yo) : yo) {  // bounce-back: read by the cell itself
        tx = x;
        ty = y;
        tj = opp_of(j);
        return;
    }
    tx = xo ? (xt < 0 ? xt + nx : xt - nx) : xt;  // periodic x (CF, BF)
    ty = yt;
    tj = j;
}

// Pulled values of cell (x, y) from s.src by the forward rule: populations p
// (FP32 / FP64 storage) or scaled deviations D = (p - w) 2^s (deviation form).
template <typename S, typename C, int MODE>
__device__ __forceinline__ void pull_cell(const StepArgs<C> &s, int x, int y, C (&p)[9]) {
    pull_pops<S, C, MODE>(s, static_cast<const typename Store<S>::R *>(s.src), x, y, p);
}

template <typename C> struct Moments {
    C rho, ux, uy, usq;
    C drho;  // rho - 1, from the deviations in the deviation form
};

template <typename S, typename C, int MODE>
__device__ __forceinline__ Moments<C> moments_of(const StepArgs<C> &s, const C (&p)[9]) {
    Moments<C> m;
    const C F = MODE == BF ? s.force : (C)0;
    if constexpr (Store<S>::dev) {
        const C inv = s.inv_scale;
        m.drho = (((p[0] + (p[1] + p[3])) + (p[2] + p[4])) + ((p[5] + p[7]) + (p[6] + p[8]))) * inv;
        m.rho = (C)1 + m.drho;
        const C jx = (((p[1] - p[3]) + (p[5] - p[7])) - (p[6] - p[8])) * inv;
        const C jy = (((p[2] - p[4]) + (p[5] - p[7])) + (p[6] - p[8])) * inv;
        m.ux = (jx + (C)0.5 * F) / m.rho;
        m.uy = jy / m.rho;
    } else {
        m.rho = p[0] + p[1] + p[2] + p[3] + p[4] + p[5] + p[6] + p[7] + p[8];
        m.drho = m.rho - (C)1;
        const C jx = p[1] - p[3] + p[5] - p[6] - p[7] + p[8];
        const C jy = p[2] - p[4] + p[5] + p[6] - p[7] - p[8];
        m.ux = (jx + (C)0.5 * F) / m.rho;
        m.uy = jy / m.rho;
    }
    m.usq = m.ux * m.ux + m.uy * m.uy;
    return m;
}

// Block sum of three doubles into acc[blockIdx.x*3 + k] (+=): butterfly sums
// inside each 64-wide wave, then the ATW/64 wave sums in a fixed order (the
// result does not depend on timing).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
__device__ __forceinline__ void block_accumulate(double *acc, double g0, double g1, double g2) {
    constexpr int W = ATW / 64;
    __shared__ double sh[3][W];
    g0 = wave_sum(g0);
    g1 = wave_sum(g1);
    g2 = wave_sum(g2);
    const int wave = (int)threadIdx.x >> 6;
    if (((int)threadIdx.x & 63) == 0) {
        sh[0][wave] = g0;
        sh[1][wave] = g1;
        sh[2][wave] = g2;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        double t = 0;
#pragma unroll
        for (int k = 0; k < W; ++k) t += sh[threadIdx.x][k];
        acc[3 * blockIdx.x + threadIdx.x] += t;
    }
}

// Loss seed at the last step: lambda_p^N and the direct dL/dF, dL/du_wall terms.
template <typename S, typename C, int MODE>
__global__ void __launch_bounds__(ATW) adjoint_seed_kernel(AdjArgs<C> a) {
    const StepArgs<C> &s = a.s;
    const int y = (int)blockIdx.x / s.n_tiles_x;
    const int x = ((int)blockIdx.x - y * s.n_tiles_x) * ATW + (int)threadIdx.x;
    double gF = 0, gW = 0;
    if (x < s.nx) {
        C p[9];
        pull_cell<S, C, MODE>(s, x, y, p);
        const Moments<C> m = moments_of<S, C, MODE>(s, p);
        const C r = (C)1 / m.rho;
        const size_t mi = (size_t)y * (size_t)s.nx + (size_t)x;
        const C gx = a.g_ux[mi], gy = a.g_uy[mi];
        const size_t c = (size_t)y * (size_t)s.pitch + (size_t)x;
        C lam[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            lam[k] = (gx * ((C)cx_of(k) - m.ux) + gy * ((C)cy_of(k) - m.uy)) * r;
            a.lam_out[(size_t)k * (size_t)s.plane + c] = lam[k];
        }
        if (MODE == BF) gF = (double)gx * 0.5 * (double)r;
        if ((MODE == CF || MODE == SL) && y == s.ny - 1)
            gW = -(double)lam[7] * (double)a.dk_du + (double)lam[8] * (double)a.dk_du;
    }
    block_accumulate(a.acc, 0.0, gF, gW);
}

// One reverse step: lambda_p^{t+1} (lam_in) -> lambda_p^t (lam_out), with
// p^t recomputed from f^{t-1} (s.src).
template <typename S, typename C, int MODE>
__global__ void __launch_bounds__(ATW) adjoint_step_kernel(AdjArgs<C> a) {
    const StepArgs<C> &s = a.s;
    const int y = (int)blockIdx.x / s.n_tiles_x;
    const int x = ((int)blockIdx.x - y * s.n_tiles_x) * ATW + (int)threadIdx.x;
    double gO = 0, gF = 0, gW = 0;
    if (x < s.nx) {
        // lambda_f^t = Pull^T lambda_p^{t+1}
        C lf[9];
#pragma unroll
        for (int j = 0; j < 9; ++j) {
            int tx, ty, tj;
            reader_of<MODE>(x, y, j, s.nx, s.ny, tx, ty, tj);
            lf[j] = a.lam_in[(size_t)tj * (size_t)s.plane + (size_t)ty * (size_t)s.pitch + (size_t)tx];
        }
        C p[9];
        pull_cell<S, C, MODE>(s, x, y, p);
        const Moments<C> m = moments_of<S, C, MODE>(s, p);
        const C r = (C)1 / m.rho;
        const C om = s.omega;
        const C F = MODE == BF ? s.force : (C)0;
        const C half_pre = (C)1 - (C)0.5 * om;  // 1 - omega/2
        C A = 0, Bx = 0, By = 0, Dx = 0, Dy = 0, GO = 0, GF = 0;
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            const C cx = (C)cx_of(i), cy = (C)cy_of(i), w = s.w[i];
            const C cu = cx * m.ux + cy * m.uy;
            const C feq = w * m.rho * (((C)1 + (C)3 * cu) + (C)4.5 * cu * cu - (C)1.5 * m.usq);
            // f_eq - p: in the deviation form from rho - 1 and D, without the
            // cancellation of two values near w
            const C neq = Store<S>::dev
                              ? w * (m.drho + m.rho * (((C)3 * cu + (C)4.5 * cu * cu) - (C)1.5 * m.usq)) -
                                    p[i] * s.inv_scale
                              : feq - p[i];
            A += lf[i] * feq * r;
            Bx += lf[i] * w * m.rho * ((C)3 * cx + (C)9 * cu * cx - (C)3 * m.ux);
            By += lf[i] * w * m.rho * ((C)3 * cy + (C)9 * cu * cy - (C)3 * m.uy);
            if (MODE == BF) {
                const C sfac = (C)3 * (cx - m.ux) + (C)9 * cu * cx;  // [3(c-u) + 9(c.u)c] . e_x
                Dx += lf[i] * half_pre * w * F * ((C)9 * cx * cx - (C)3);
                Dy += lf[i] * half_pre * w * F * ((C)9 * cx * cy);
                GO += lf[i] * (neq - (C)0.5 * w * F * sfac);
                GF += lf[i] * half_pre * w * sfac;
            } else {
                GO += lf[i] * neq;
            }
        }
        const C ax = om * Bx + Dx, ay = om * By + Dy;
        const size_t c = (size_t)y * (size_t)s.pitch + (size_t)x;
        C lp[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            lp[k] = ((C)1 - om) * lf[k] + om * A + (ax * ((C)cx_of(k) - m.ux) + ay * ((C)cy_of(k) - m.uy)) * r;
            a.lam_out[(size_t)k * (size_t)s.plane + c] = lp[k];
        }
        gO = (double)GO;
        if (MODE == BF) gF = (double)GF + (double)ax * 0.5 * (double)r;
        if ((MODE == CF || MODE == SL) && y == s.ny - 1)
            gW = -(double)lp[7] * (double)a.dk_du + (double)lp[8] * (double)a.dk_du;
    }
    block_accumulate(a.acc, gO, gF, gW);
}

}  // namespace

struct ffb_adjoint {
    ffb_lattice *lat = nullptr;
    int n_steps = 0, interval = 0, n_ck = 0;
    size_t state_bytes = 0;
    void *ck = nullptr;     // checkpoints f^{kK}, k = 0 .. n_ck-1
    void *seg = nullptr;    // recomputed states of one segment (K slots)
    void *lam[2] = {nullptr, nullptr};
    void *gseed[2] = {nullptr, nullptr};
    double *acc = nullptr;  // [blocks][3]
    int blocks = 0;
    bool forward_done = false;
    hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr, e3 = nullptr;
    double fwd_ms = 0, bwd_ms = 0;
};

namespace {

template <typename S, typename C> StepArgs<C> adj_step_args(const ffb_adjoint *A, const void *src) {
    const ffb_lattice *L = A->lat;
    const ffb_config_t &c = L->cfg;
    StepArgs<C> s{};
    s.src = src;
    s.dst = nullptr;
    s.nx = c.n_x;
    s.ny = c.n_y;
    s.pitch = L->pitch;
    s.plane = L->plane;
    s.omega = (C)c.omega;
    const C w6 = (C)1 / (C)36;
    const C u = (C)(c.sim_mode == FFB_COUETTE_FLOW ? c.u_wall : c.u_lid);
    s.corr8 = (C)6 * w6 * (C)c.rho_0 * u;
    s.corr9 = s.corr8;
    s.w[0] = (C)4 / (C)9;
    for (int d = 1; d <= 4; ++d) s.w[d] = (C)1 / (C)9;
    for (int d = 5; d <= 8; ++d) s.w[d] = (C)1 / (C)36;
    s.scale = 1;
    s.inv_scale = 1;
    if constexpr (Store<S>::dev) {  // pulled deviations: scaled wall terms (as the forward step)
        s.scale = (C)std::ldexp(1.0, c.storage_scale_log2);
        s.inv_scale = (C)std::ldexp(1.0, -c.storage_scale_log2);
        s.corr8 = s.corr8 * s.scale;
        s.corr9 = s.corr9 * s.scale;
    }
    s.force = (C)c.body_force;
    s.n_tiles_x = (c.n_x + ATW - 1) / ATW;
    s.n_blocks = s.n_tiles_x * c.n_y;
    return s;
}

template <typename S, typename C, int MODE>
int launch_adj(ffb_adjoint *A, bool seed, const void *state, int lam_in, int lam_out) {
    AdjArgs<C> a{};
    a.s = adj_step_args<S, C>(A, state);
    a.lam_in = static_cast<const C *>(A->lam[lam_in]);
    a.lam_out = static_cast<C *>(A->lam[lam_out]);
    a.g_ux = static_cast<const C *>(A->gseed[0]);
    a.g_uy = static_cast<const C *>(A->gseed[1]);
    a.acc = A->acc;
    const ffb_config_t &c = A->lat->cfg;
    a.dk_du = (c.sim_mode == FFB_COUETTE_FLOW || c.sim_mode == FFB_SLIDING_LID)
                  ? (C)6 * ((C)1 / (C)36) * (C)c.rho_0 : (C)0;
    const dim3 grid((unsigned)a.s.n_blocks);
    if (seed) hipLaunchKernelGGL((adjoint_seed_kernel<S, C, MODE>), grid, dim3(ATW), 0, A->lat->stream, a);
    else hipLaunchKernelGGL((adjoint_step_kernel<S, C, MODE>), grid, dim3(ATW), 0, A->lat->stream, a);
    AHIPCHK(hipGetLastError());
    return FFB_OK;
}

template <typename S, typename C>
int launch_adj_mode(ffb_adjoint *A, bool seed, const void *state, int lam_in, int lam_out) {
    switch (A->lat->cfg.sim_mode) {
    case FFB_SHEAR_WAVE: return launch_adj<S, C, SW>(A, seed, state, lam_in, lam_out);
    case FFB_COUETTE_FLOW: return launch_adj<S, C, CF>(A, seed, state, lam_in, lam_out);
    case FFB_SLIDING_LID: return launch_adj<S, C, SL>(A, seed, state, lam_in, lam_out);
    default: return launch_adj<S, C, BF>(A, seed, state, lam_in, lam_out);
    }
}

int adj(ffb_adjoint *A, bool seed, const void *state, int lam_in, int lam_out) {
    switch (A->lat->cfg.precision) {
    case FFB_PREC_FP64: return launch_adj_mode<double, double>(A, seed, state, lam_in, lam_out);
    case FFB_PREC_FP32_DEV: return launch_adj_mode<f32dev, float>(A, seed, state, lam_in, lam_out);
    default: return launch_adj_mode<float, float>(A, seed, state, lam_in, lam_out);
    }
}

unsigned char *slot(void *base, size_t bytes, int k) { return static_cast<unsigned char *>(base) + bytes * (size_t)k; }

}  // namespace

extern "C" {

int ffb_adjoint_create(ffb_lattice_t *L, int32_t n_steps, int32_t checkpoint_interval, ffb_adjoint_t **out) {
    if (!L || !out) return fail(FFB_E_INVALID, "ffb_adjoint_create: null argument");
    *out = nullptr;
    if (n_steps < 1 || checkpoint_interval < 1) return fail(FFB_E_INVALID, "ffb_adjoint_create: bad step counts");
    const int mode = L->cfg.sim_mode;
    if (mode == FFB_POISEUILLE_FLOW)
        return fail(FFB_E_INVALID, "ffb_adjoint_create: the pressure-driven Poiseuille mode has no adjoint; "
                                   "use FFB_BODY_FORCE_CHANNEL");
    if (L->cfg.precision != FFB_PREC_FP32 && L->cfg.precision != FFB_PREC_FP64 &&
        L->cfg.precision != FFB_PREC_FP32_DEV)
        return fail(FFB_E_INVALID, "ffb_adjoint_create: FP32, FP32 deviation or FP64 storage only");
    AHIPCHK(hipSetDevice(L->cfg.device));
    ffb_adjoint *A = new ffb_adjoint();
    A->lat = L;
    A->n_steps = n_steps;
    A->interval = checkpoint_interval;
    A->n_ck = (n_steps - 1) / checkpoint_interval + 1;
    A->state_bytes = (size_t)L->set_elems * L->elem;
    const int seg_slots = std::min(checkpoint_interval, n_steps);
    A->blocks = ((L->cfg.n_x + ATW - 1) / ATW) * L->cfg.n_y;
    auto bail = [&](hipError_t e, const char *what) {
        ffb_adjoint_destroy(A);
        return fail(FFB_E_HIP, std::string("ffb_adjoint_create: ") + what + ": " + hipGetErrorString(e));
    };
    hipError_t e;
    if ((e = hipMalloc(&A->ck, A->state_bytes * (size_t)A->n_ck)) != hipSuccess) return bail(e, "checkpoints");
    if ((e = hipMalloc(&A->seg, A->state_bytes * (size_t)seg_slots)) != hipSuccess) return bail(e, "segment");
    for (int k = 0; k < 2; ++k)
        if ((e = hipMalloc(&A->lam[k], A->state_bytes)) != hipSuccess) return bail(e, "adjoint fields");
    const size_t mb = (size_t)L->cfg.n_x * (size_t)L->cfg.n_y * L->celem;
    for (int k = 0; k < 2; ++k)
        if ((e = hipMalloc(&A->gseed[k], mb)) != hipSuccess) return bail(e, "seed");
    if ((e = hipMalloc(&A->acc, (size_t)A->blocks * 3 * sizeof(double))) != hipSuccess) return bail(e, "sums");
    hipEvent_t *ev[4] = {&A->e0, &A->e1, &A->e2, &A->e3};
    for (auto *p : ev)
        if ((e = hipEventCreate(p)) != hipSuccess) return bail(e, "events");
    *out = A;
    return FFB_OK;
}

int ffb_adjoint_destroy(ffb_adjoint_t *A) {
    if (!A) return FFB_OK;
    if (A->lat && A->lat->stream) (void)hipStreamSynchronize(A->lat->stream);
    void *bufs[] = {A->ck, A->seg, A->lam[0], A->lam[1], A->gseed[0], A->gseed[1], A->acc};
    for (void *p : bufs) if (p) (void)hipFree(p);
    hipEvent_t ev[] = {A->e0, A->e1, A->e2, A->e3};
    for (hipEvent_t x : ev) if (x) (void)hipEventDestroy(x);
    delete A;
    return FFB_OK;
}

int ffb_adjoint_forward(ffb_adjoint_t *A, void *rho, void *u_x, void *u_y) {
    if (!A) return fail(FFB_E_INVALID, "ffb_adjoint_forward: null handle");
    ffb_lattice *L = A->lat;
    AHIPCHK(hipSetDevice(L->cfg.device));
    int rc = ffb_init(L);
    if (rc) return rc;
    AHIPCHK(hipEventRecord(A->e0, L->stream));
    // f^0 is checkpoint 0; then N steps through the handle's ping-pong buffers
    AHIPCHK(hipMemcpyAsync(slot(A->ck, A->state_bytes, 0), L->buf[0], A->state_bytes, hipMemcpyDeviceToDevice,
                           L->stream));
    int cur = 0;
    for (int t = 1; t <= A->n_steps; ++t) {
        if ((rc = step_between(L, L->buf[cur], L->buf[1 - cur], t == A->n_steps ? 1 : 0))) return rc;
        cur = 1 - cur;
        if (t % A->interval == 0 && t < A->n_steps)
            AHIPCHK(hipMemcpyAsync(slot(A->ck, A->state_bytes, t / A->interval), L->buf[cur], A->state_bytes,
                                   hipMemcpyDeviceToDevice, L->stream));
    }
    L->cur = cur;
    L->steps = A->n_steps;
    AHIPCHK(hipEventRecord(A->e1, L->stream));
    const size_t mb = (size_t)L->cfg.n_x * (size_t)L->cfg.n_y * L->celem;
    void *dst[3] = {rho, u_x, u_y};
    for (int k = 0; k < 3; ++k)
        if (dst[k]) AHIPCHK(hipMemcpyAsync(dst[k], L->macro[k], mb, hipMemcpyDefault, L->stream));
    AHIPCHK(hipStreamSynchronize(L->stream));
    float ms = 0.f;
    AHIPCHK(hipEventElapsedTime(&ms, A->e0, A->e1));
    A->fwd_ms = ms;
    A->forward_done = true;
    return FFB_OK;
}

int ffb_adjoint_backward(ffb_adjoint_t *A, const void *grad_u_x, const void *grad_u_y, double *d_omega,
                         double *d_body_force, double *d_u_wall) {
    if (!A || !grad_u_x || !grad_u_y) return fail(FFB_E_INVALID, "ffb_adjoint_backward: null argument");
    if (!A->forward_done) return fail(FFB_E_STATE, "ffb_adjoint_backward: call ffb_adjoint_forward first");
    ffb_lattice *L = A->lat;
    AHIPCHK(hipSetDevice(L->cfg.device));
    const size_t mb = (size_t)L->cfg.n_x * (size_t)L->cfg.n_y * L->celem;
    AHIPCHK(hipMemcpyAsync(A->gseed[0], grad_u_x, mb, hipMemcpyDefault, L->stream));
    AHIPCHK(hipMemcpyAsync(A->gseed[1], grad_u_y, mb, hipMemcpyDefault, L->stream));
    AHIPCHK(hipMemsetAsync(A->acc, 0, (size_t)A->blocks * 3 * sizeof(double), L->stream));
    AHIPCHK(hipEventRecord(A->e2, L->stream));
    const int N = A->n_steps, K = A->interval;
    int lin = 0;  // lam[lin] holds lambda_p^{t+1}
    int rc;
    for (int k = A->n_ck - 1; k >= 0; --k) {
        const int s0 = k * K, s1 = std::min((k + 1) * K, N);  // states f^{s0} .. f^{s1-1}
        // recompute the segment: slot j holds f^{s0+j}
        AHIPCHK(hipMemcpyAsync(slot(A->seg, A->state_bytes, 0), slot(A->ck, A->state_bytes, k), A->state_bytes,
                               hipMemcpyDeviceToDevice, L->stream));
        for (int j = 1; j < s1 - s0; ++j)
            if ((rc = step_between(L, slot(A->seg, A->state_bytes, j - 1), slot(A->seg, A->state_bytes, j), 0)))
                return rc;
        for (int s = s1 - 1; s >= s0; --s) {
            const void *state = slot(A->seg, A->state_bytes, s - s0);  // f^s
            if (s == N - 1) {
                if ((rc = adj(A, true, state, 1 - lin, lin))) return rc;  // lambda_p^N into lam[lin]
            } else {
                if ((rc = adj(A, false, state, lin, 1 - lin))) return rc;  // lambda_p^{s+1} from lambda_p^{s+2}
                lin = 1 - lin;
            }
        }
    }
    AHIPCHK(hipEventRecord(A->e3, L->stream));
    std::vector<double> acc((size_t)A->blocks * 3);
    AHIPCHK(hipMemcpyAsync(acc.data(), A->acc, acc.size() * sizeof(double), hipMemcpyDeviceToHost, L->stream));
    AHIPCHK(hipStreamSynchronize(L->stream));
    double g[3] = {0, 0, 0};
    for (int b = 0; b < A->blocks; ++b)
        for (int q = 0; q < 3; ++q) g[q] += acc[(size_t)b * 3 + q];
    if (d_omega) *d_omega = g[0];
    if (d_body_force) *d_body_force = g[1];
    if (d_u_wall) *d_u_wall = g[2];
    float ms = 0.f;
    AHIPCHK(hipEventElapsedTime(&ms, A->e2, A->e3));
    A->bwd_ms = ms;
    return FFB_OK;
}

int ffb_adjoint_timing(ffb_adjoint_t *A, double *forward_ms, double *backward_ms) {
    if (!A) return fail(FFB_E_INVALID, "ffb_adjoint_timing: null handle");
    if (forward_ms) *forward_ms = A->fwd_ms;
    if (backward_ms) *backward_ms = A->bwd_ms;
    return FFB_OK;
}

}  // extern "C"
