/*
 * ffb_oracle_impl.h -- body of the CPU oracle, included once per precision by
 * ffb_oracle.c with REAL (double | float) and NAME(x) (ffbo_d_x | ffbo_s_x).
 *
 * TEST INFRASTRUCTURE ONLY.  This is a scalar C restatement of the reference
 * D2Q9 BGK path, in the reference's own "halo" form: populations live in
 * f(0:NX+1, 0:NY+1, 1:9) (x fastest, then y, then direction), the boundary
 * conditions are written into the one-cell halo before the pull, and the
 * fused pull-stream-collide kernel then reads the halo like any other cell.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it,
 * as the checker; the product path never calls it.
 *
 * Every arithmetic expression keeps the reference's left-to-right evaluation
 * order and is compiled with -ffp-contract=off, so that no fused multiply-add
 * changes the rounding.  Each routine cites the reference file:line it
 * follows.
 */

/* f(x, y, i) with 0-based halo coordinates x = 0..nx+1, y = 0..ny+1 and the
 * reference's 1-based direction i = 1..9 (source/layouts/xyd/simulation.f90:27). */
#define FI(f, x, y, i) (f)[(size_t)((i) - 1) * plane + (size_t)(y) * (size_t)pitch + (size_t)(x)]

/* D2Q9 weights, folded exactly like the reference's 4.0_FP/9.0_FP etc.
 * (source/settings.f90:114-123). */
REAL NAME(w)(int i)
{
    switch (i) {
    case 1: return (REAL)4.0 / (REAL)9.0;
    case 2: case 3: case 4: case 5: return (REAL)1.0 / (REAL)9.0;
    default: return (REAL)1.0 / (REAL)36.0;
    }
}

/* Equilibrium initial conditions, source/layouts/xyd/initialization.f90:9-198.
 * n_y_global is the reference's N_Y (the sine period), y_global_start the
 * domain's first global row (1 for a single domain). */
void NAME(init)(int mode, int nx, int ny, int y_global_start, int n_y_global,
                REAL rho_0, REAL u_max, REAL n_sin,
                REAL *f, REAL *rho, REAL *u_x, REAL *u_y)
{
    const long pitch = nx + 2;
    const size_t plane = (size_t)(nx + 2) * (size_t)(ny + 2);
    const REAL pi = (REAL)3.141592653589793238462643383279502884197L; /* settings.f90:126 */
    static const int cx[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
    static const int cy[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};

    if (mode == 3) {
        /* Poiseuille: halo also starts at equilibrium (initialization.f90:137-140) */
        for (int i = 1; i <= 9; ++i) {
            REAL feq = NAME(w)(i) * rho_0;
            for (size_t k = 0; k < plane; ++k) f[(size_t)(i - 1) * plane + k] = feq;
        }
    } else {
        memset(f, 0, sizeof(REAL) * plane * 9);
    }

    for (int y = 1; y <= ny; ++y) {
        REAL ux = 0, uy = 0, usq = 0;
        if (mode == 1) {
            /* initialization.f90:41-44 */
            int y_global = y_global_start + y - 1;
            ux = u_max * SIN_FN((((REAL)2.0 * pi) * n_sin * (REAL)(y_global - 1)) / (REAL)n_y_global);
            uy = 0;
            usq = ux * ux;
        }
        for (int x = 1; x <= nx; ++x) {
            size_t c = (size_t)(y - 1) * (size_t)nx + (size_t)(x - 1);
            rho[c] = rho_0;
            u_x[c] = ux;
            u_y[c] = uy;
            for (int i = 1; i <= 9; ++i) {
                REAL feq;
                if (mode == 1) {
                    /* initialization.f90:56-63: c.u with real(C, FP) products */
                    REAL cu = (REAL)cx[i - 1] * ux + (REAL)cy[i - 1] * uy;
                    feq = NAME(w)(i) * rho_0 *
                          ((((REAL)1.0 + (REAL)3.0 * cu) + (REAL)4.5 * cu * cu) - (REAL)1.5 * usq);
                } else {
                    feq = NAME(w)(i) * rho_0; /* initialization.f90:90-110,177-197 */
                }
                FI(f, x, y, i) = feq;
            }
        }
    }
}

/* One-domain halo exchange: exchange_halos() with every neighbour equal to the
 * own image (source/layouts/xyd/exchange.f90:213-291, self-copy branches
 * :234-236,245-247,270-274,282-286), with the plan of build_exchange_plan()
 * (:67-114) for a domain that touches all four boundaries. */
static void NAME(exchange_1domain)(int mode, int nx, int ny, REAL *f)
{
    const long pitch = nx + 2;
    const size_t plane = (size_t)(nx + 2) * (size_t)(ny + 2);
    int plan_x = (mode == 1 || mode == 2 || mode == 3 || mode == 5); /* SW, CF, PF, BF: left/right */
    int plan_y = (mode == 1);                            /* SW only: bottom/top */
    if (plan_x) {
        for (int y = 1; y <= ny; ++y) {
            /* window(SEND_RIGHT) = column nx dirs 2,6,9 -> halo column 0 */
            FI(f, 0, y, 2) = FI(f, nx, y, 2);
            FI(f, 0, y, 6) = FI(f, nx, y, 6);
            FI(f, 0, y, 9) = FI(f, nx, y, 9);
        }
        for (int y = 1; y <= ny; ++y) {
            /* window(SEND_LEFT) = column 1 dirs 4,7,8 -> halo column nx+1 */
            FI(f, nx + 1, y, 4) = FI(f, 1, y, 4);
            FI(f, nx + 1, y, 7) = FI(f, 1, y, 7);
            FI(f, nx + 1, y, 8) = FI(f, 1, y, 8);
        }
    }
    if (plan_y) {
        for (int x = 0; x <= nx + 1; ++x) {
            FI(f, x, 0, 3) = FI(f, x, ny, 3);
            FI(f, x, 0, 6) = FI(f, x, ny, 6);
            FI(f, x, 0, 7) = FI(f, x, ny, 7);
        }
        for (int x = 0; x <= nx + 1; ++x) {
            FI(f, x, ny + 1, 5) = FI(f, x, 1, 5);
            FI(f, x, ny + 1, 8) = FI(f, x, 1, 8);
            FI(f, x, ny + 1, 9) = FI(f, x, 1, 9);
        }
    }
}

/* prepare_shear_wave_halos_SW, modes/shear_wave.f90:9-56 (one domain in x and y). */
static void NAME(halos_sw)(int nx, int ny, REAL *f)
{
    const long pitch = nx + 2;
    const size_t plane = (size_t)(nx + 2) * (size_t)(ny + 2);
    for (int y = 1; y <= ny; ++y) {
        FI(f, 0, y, 2) = FI(f, nx, y, 2);
        FI(f, 0, y, 6) = FI(f, nx, y, 6);
        FI(f, 0, y, 9) = FI(f, nx, y, 9);
        FI(f, nx + 1, y, 4) = FI(f, 1, y, 4);
        FI(f, nx + 1, y, 7) = FI(f, 1, y, 7);
        FI(f, nx + 1, y, 8) = FI(f, 1, y, 8);
    }
    for (int x = 0; x <= nx + 1; ++x) {
        FI(f, x, 0, 3) = FI(f, x, ny, 3);
        FI(f, x, 0, 6) = FI(f, x, ny, 6);
        FI(f, x, 0, 7) = FI(f, x, ny, 7);
        FI(f, x, ny + 1, 5) = FI(f, x, 1, 5);
        FI(f, x, ny + 1, 8) = FI(f, x, 1, 8);
        FI(f, x, ny + 1, 9) = FI(f, x, 1, 9);
    }
}

/* Halfway bounce-back on the bottom row, couette_flow.f90:49-55,
 * sliding_lid.f90:55-61, poiseuille_flow.f90:37-43. */
static void NAME(bb_bottom)(int nx, int ny, REAL *f)
{
    const long pitch = nx + 2;
    const size_t plane = (size_t)(nx + 2) * (size_t)(ny + 2);
    for (int x = 1; x <= nx; ++x) {
        FI(f, x, 0, 3) = FI(f, x, 1, 5);
        FI(f, x - 1, 0, 6) = FI(f, x, 1, 8);
        FI(f, x + 1, 0, 7) = FI(f, x, 1, 9);
    }
}

/* Moving-wall bounce-back on the top row, couette_flow.f90:58-67,
 * sliding_lid.f90:64-73 (poiseuille_flow.f90:46-52 with a resting wall). */
static void NAME(bb_top)(int nx, int ny, REAL rho_0, REAL u_wall, int moving, REAL *f)
{
    const long pitch = nx + 2;
    const size_t plane = (size_t)(nx + 2) * (size_t)(ny + 2);
    REAL corr8 = (REAL)6.0 * NAME(w)(6) * rho_0 * u_wall;
    REAL corr9 = (REAL)6.0 * NAME(w)(7) * rho_0 * u_wall;
    for (int x = 1; x <= nx; ++x) {
        FI(f, x, ny + 1, 5) = FI(f, x, ny, 3);
        if (moving) {
            FI(f, x + 1, ny + 1, 8) = FI(f, x, ny, 6) - corr8;
            FI(f, x - 1, ny + 1, 9) = FI(f, x, ny, 7) + corr9;
        } else {
            FI(f, x + 1, ny + 1, 8) = FI(f, x, ny, 6);
            FI(f, x - 1, ny + 1, 9) = FI(f, x, ny, 7);
        }
    }
}

/* prepare_couette_flow_halos_CF, modes/couette_flow.f90:9-68. */
static void NAME(halos_cf)(int nx, int ny, REAL rho_0, REAL u_wall, REAL *f)
{
    const long pitch = nx + 2;
    const size_t plane = (size_t)(nx + 2) * (size_t)(ny + 2);
    for (int y = 1; y <= ny; ++y) {
        FI(f, 0, y, 2) = FI(f, nx, y, 2);
        FI(f, 0, y, 6) = FI(f, nx, y, 6);
        FI(f, 0, y, 9) = FI(f, nx, y, 9);
        FI(f, nx + 1, y, 4) = FI(f, 1, y, 4);
        FI(f, nx + 1, y, 7) = FI(f, 1, y, 7);
        FI(f, nx + 1, y, 8) = FI(f, 1, y, 8);
    }
    NAME(bb_bottom)(nx, ny, f);
    NAME(bb_top)(nx, ny, rho_0, u_wall, 1, f);
}

/* prepare_sliding_lid_halos_SL, modes/sliding_lid.f90:9-74 (write order
 * left, right, bottom, top: the lid wins at the two top corners). */
static void NAME(halos_sl)(int nx, int ny, REAL rho_0, REAL u_lid, REAL *f)
{
    const long pitch = nx + 2;
    const size_t plane = (size_t)(nx + 2) * (size_t)(ny + 2);
    for (int y = 1; y <= ny; ++y) {
        FI(f, 0, y, 2) = FI(f, 1, y, 4);
        FI(f, 0, y - 1, 6) = FI(f, 1, y, 8);
        FI(f, 0, y + 1, 9) = FI(f, 1, y, 7);
    }
    for (int y = 1; y <= ny; ++y) {
        FI(f, nx + 1, y, 4) = FI(f, nx, y, 2);
        FI(f, nx + 1, y - 1, 7) = FI(f, nx, y, 9);
        FI(f, nx + 1, y + 1, 8) = FI(f, nx, y, 6);
    }
    NAME(bb_bottom)(nx, ny, f);
    NAME(bb_top)(nx, ny, rho_0, u_lid, 1, f);
}

/* pressure_periodic_distribution, modes/poiseuille_flow.f90:170-195. */
static REAL NAME(ppd)(int i, REAL f_src, REAL rho_src, REAL ux_src, REAL uy_src, REAL rho_b)
{
    static const int cx[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
    static const int cy[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
    REAL usq = ux_src * ux_src + uy_src * uy_src;
    REAL cu = (REAL)cx[i - 1] * ux_src + (REAL)cy[i - 1] * uy_src;
    REAL pf = (((REAL)1.0 + (REAL)3.0 * cu) + (REAL)4.5 * cu * cu) - (REAL)1.5 * usq;
    return f_src + NAME(w)(i) * (rho_b - rho_src) * pf;
}

/* prepare_poiseuille_flow_halos_PF, modes/poiseuille_flow.f90:9-89.
 * macro_left / macro_right are (ny, 3) column-major strips: rho, u_x, u_y. */
static void NAME(halos_pf)(int nx, int ny, REAL rho_in, REAL rho_out,
                           const REAL *ml, const REAL *mr, REAL *f)
{
    const long pitch = nx + 2;
    const size_t plane = (size_t)(nx + 2) * (size_t)(ny + 2);
#define ML(y, k) ml[(size_t)((k) - 1) * (size_t)ny + (size_t)((y) - 1)]
#define MR(y, k) mr[(size_t)((k) - 1) * (size_t)ny + (size_t)((y) - 1)]
    NAME(bb_bottom)(nx, ny, f);
    NAME(bb_top)(nx, ny, 0, 0, 0, f);
    for (int y = 1; y <= ny; ++y) {
        FI(f, 0, y, 2) = NAME(ppd)(2, FI(f, 0, y, 2), ML(y, 1), ML(y, 2), ML(y, 3), rho_in);
        FI(f, 0, y, 6) = NAME(ppd)(6, FI(f, 0, y, 6), ML(y, 1), ML(y, 2), ML(y, 3), rho_in);
        FI(f, 0, y, 9) = NAME(ppd)(9, FI(f, 0, y, 9), ML(y, 1), ML(y, 2), ML(y, 3), rho_in);
    }
    FI(f, 0, 0, 6) = NAME(ppd)(6, FI(f, 0, 0, 6), ML(1, 1), ML(1, 2), ML(1, 3), rho_in);
    FI(f, 0, ny + 1, 9) = NAME(ppd)(9, FI(f, 0, ny + 1, 9), ML(ny, 1), ML(ny, 2), ML(ny, 3), rho_in);
    for (int y = 1; y <= ny; ++y) {
        FI(f, nx + 1, y, 4) = NAME(ppd)(4, FI(f, nx + 1, y, 4), MR(y, 1), MR(y, 2), MR(y, 3), rho_out);
        FI(f, nx + 1, y, 7) = NAME(ppd)(7, FI(f, nx + 1, y, 7), MR(y, 1), MR(y, 2), MR(y, 3), rho_out);
        FI(f, nx + 1, y, 8) = NAME(ppd)(8, FI(f, nx + 1, y, 8), MR(y, 1), MR(y, 2), MR(y, 3), rho_out);
    }
    FI(f, nx + 1, 0, 7) = NAME(ppd)(7, FI(f, nx + 1, 0, 7), MR(1, 1), MR(1, 2), MR(1, 3), rho_out);
    FI(f, nx + 1, ny + 1, 8) = NAME(ppd)(8, FI(f, nx + 1, ny + 1, 8), MR(ny, 1), MR(ny, 2), MR(ny, 3), rho_out);
#undef ML
#undef MR
}

/* update_poiseuille_flow_macro_strips_PF, modes/poiseuille_flow.f90:92-167:
 * pulled (rho, u_x, u_y) of column x into a (ny, 3) strip. */
static void NAME(macro_strip)(int nx, int ny, int x, const REAL *f, REAL *strip)
{
    const long pitch = nx + 2;
    const size_t plane = (size_t)(nx + 2) * (size_t)(ny + 2);
    for (int y = 1; y <= ny; ++y) {
        REAL f1 = FI(f, x, y, 1), f2 = FI(f, x - 1, y, 2), f3 = FI(f, x, y - 1, 3);
        REAL f4 = FI(f, x + 1, y, 4), f5 = FI(f, x, y + 1, 5), f6 = FI(f, x - 1, y - 1, 6);
        REAL f7 = FI(f, x + 1, y - 1, 7), f8 = FI(f, x + 1, y + 1, 8), f9 = FI(f, x - 1, y + 1, 9);
        REAL rho = f1 + f2 + f3 + f4 + f5 + f6 + f7 + f8 + f9;
        REAL ux = f2 - f4 + f6 - f7 - f8 + f9;
        REAL uy = f3 - f5 + f6 + f7 - f8 - f9;
        strip[(size_t)(y - 1)] = rho;
        strip[(size_t)ny + (size_t)(y - 1)] = ux / rho;
        strip[2 * (size_t)ny + (size_t)(y - 1)] = uy / rho;
    }
}

/* fuzed_pull_streaming_collision_local_unrolled_universal,
 * source/layouts/xyd/simulation.f90:171-296.  Interior of f_next only; the
 * macro fields are (nx, ny) x-fastest and written when write_macro != 0. */
static void NAME(kernel)(int nx, int ny, int write_macro, REAL omega,
                         const REAL *f, REAL *f_next, REAL *rho_o, REAL *ux_o, REAL *uy_o)
{
    const long pitch = nx + 2;
    const size_t plane = (size_t)(nx + 2) * (size_t)(ny + 2);
    const REAL w0 = (REAL)4.0 / (REAL)9.0, w1 = (REAL)1.0 / (REAL)9.0, w2 = (REAL)1.0 / (REAL)36.0;
    const REAL one = 1, three = 3, c45 = (REAL)4.5, c15 = (REAL)1.5;
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int y = 1; y <= ny; ++y) {
        for (int x = 1; x <= nx; ++x) {
            REAL f1 = FI(f, x, y, 1);
            REAL f2 = FI(f, x - 1, y, 2);
            REAL f3 = FI(f, x, y - 1, 3);
            REAL f4 = FI(f, x + 1, y, 4);
            REAL f5 = FI(f, x, y + 1, 5);
            REAL f6 = FI(f, x - 1, y - 1, 6);
            REAL f7 = FI(f, x + 1, y - 1, 7);
            REAL f8 = FI(f, x + 1, y + 1, 8);
            REAL f9 = FI(f, x - 1, y + 1, 9);
            REAL rho = f1 + f2 + f3 + f4 + f5 + f6 + f7 + f8 + f9;
            REAL ux = f2 - f4 + f6 - f7 - f8 + f9;
            REAL uy = f3 - f5 + f6 + f7 - f8 - f9;
            ux = ux / rho;
            uy = uy / rho;
            REAL usq = ux * ux + uy * uy;
            if (write_macro) {
                size_t c = (size_t)(y - 1) * (size_t)nx + (size_t)(x - 1);
                rho_o[c] = rho;
                ux_o[c] = ux;
                uy_o[c] = uy;
            }
            REAL s;
            FI(f_next, x, y, 1) = f1 + omega * (w0 * rho * (one - c15 * usq) - f1);
            FI(f_next, x, y, 2) = f2 + omega * (w1 * rho * (one + three * ux + c45 * ux * ux - c15 * usq) - f2);
            FI(f_next, x, y, 3) = f3 + omega * (w1 * rho * (one + three * uy + c45 * uy * uy - c15 * usq) - f3);
            FI(f_next, x, y, 4) = f4 + omega * (w1 * rho * (one - three * ux + c45 * ux * ux - c15 * usq) - f4);
            FI(f_next, x, y, 5) = f5 + omega * (w1 * rho * (one - three * uy + c45 * uy * uy - c15 * usq) - f5);
            s = ux + uy;
            FI(f_next, x, y, 6) = f6 + omega * (w2 * rho * (one + three * s + c45 * s * s - c15 * usq) - f6);
            s = -ux + uy;
            FI(f_next, x, y, 7) = f7 + omega * (w2 * rho * (one + three * s + c45 * s * s - c15 * usq) - f7);
            s = ux + uy;
            FI(f_next, x, y, 8) = f8 + omega * (w2 * rho * (one - three * s + c45 * s * s - c15 * usq) - f8);
            s = ux - uy;
            FI(f_next, x, y, 9) = f9 + omega * (w2 * rho * (one + three * s + c45 * s * s - c15 * usq) - f9);
        }
    }
}

/* Body-force channel collision with Guo forcing, F = (force, 0) (no
 * reference counterpart; Guo, Zheng & Shi, Phys. Rev. E 65, 046308 (2002)):
 * u = (sum c f + F/2)/rho, f' = f + omega (f_eq(rho, u) - f) + S,
 * S_i = (1 - omega/2) w_i [3 (c_i - u) + 9 (c_i . u) c_i] . F. */
static void NAME(kernel_guo)(int nx, int ny, int write_macro, REAL omega, REAL force,
                             const REAL *f, REAL *f_next, REAL *rho_o, REAL *ux_o, REAL *uy_o)
{
    const long pitch = nx + 2;
    const size_t plane = (size_t)(nx + 2) * (size_t)(ny + 2);
    static const int cx[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
    static const int cy[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
    const REAL one = 1, three = 3, c45 = (REAL)4.5, c15 = (REAL)1.5, nine = 9, half = (REAL)0.5;
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int y = 1; y <= ny; ++y) {
        for (int x = 1; x <= nx; ++x) {
            REAL v[9];
            v[0] = FI(f, x, y, 1);
            v[1] = FI(f, x - 1, y, 2);
            v[2] = FI(f, x, y - 1, 3);
            v[3] = FI(f, x + 1, y, 4);
            v[4] = FI(f, x, y + 1, 5);
            v[5] = FI(f, x - 1, y - 1, 6);
            v[6] = FI(f, x + 1, y - 1, 7);
            v[7] = FI(f, x + 1, y + 1, 8);
            v[8] = FI(f, x - 1, y + 1, 9);
            REAL rho = v[0] + v[1] + v[2] + v[3] + v[4] + v[5] + v[6] + v[7] + v[8];
            REAL ux = v[1] - v[3] + v[5] - v[6] - v[7] + v[8];
            REAL uy = v[2] - v[4] + v[5] + v[6] - v[7] - v[8];
            ux = (ux + half * force) / rho;
            uy = uy / rho;
            REAL usq = ux * ux + uy * uy;
            if (write_macro) {
                size_t c = (size_t)(y - 1) * (size_t)nx + (size_t)(x - 1);
                rho_o[c] = rho;
                ux_o[c] = ux;
                uy_o[c] = uy;
            }
            REAL pre = (one - half * omega) * force;
            for (int d = 0; d < 9; ++d) {
                REAL ccx = (REAL)cx[d], ccy = (REAL)cy[d];
                REAL cu = ccx * ux + ccy * uy;
                REAL feq = NAME(w)(d + 1) * rho * (((one + three * cu) + c45 * cu * cu) - c15 * usq);
                REAL src = NAME(w)(d + 1) * pre * (three * (ccx - ux) + nine * cu * ccx);
                FI(f_next, x, y, d + 1) = v[d] + omega * (feq - v[d]) + src;
            }
        }
    }
}

/* One time step of the single-domain driver loop, source/layouts/xyd/main.f90:194-262:
 *   exchange_halos (exchange.f90:161-378, one domain),
 *   [PF] prepare_PF(recv_macro) + update strips + macro exchange (main.f90:212-246,
 *        exchange.f90:381-431: recv_macro_left <- strip at x=nx, recv_macro_right <- strip at x=1),
 *   execute_local_sim_step (simulation.f90:18-73): mode halo prep, then the fused kernel.
 * pf_state holds four (ny, 3) strips: [recv_macro_left, recv_macro_right,
 * window(MACRO_LEFT), window(MACRO_RIGHT)]. */
void NAME(step)(int mode, int nx, int ny, REAL omega, REAL rho_0, REAL u_wall, REAL u_lid,
                REAL rho_in, REAL rho_out, REAL *f, REAL *f_next, REAL *pf_state,
                int write_macro, REAL *rho, REAL *u_x, REAL *u_y, REAL force)
{
    NAME(exchange_1domain)(mode, nx, ny, f);
    if (mode == 3) {
        size_t s = (size_t)ny * 3;
        REAL *recv_l = pf_state, *recv_r = pf_state + s;
        REAL *win_l = pf_state + 2 * s, *win_r = pf_state + 3 * s;
        NAME(halos_pf)(nx, ny, rho_in, rho_out, recv_l, recv_r, f);
        NAME(macro_strip)(nx, ny, 1, f, win_l);
        NAME(macro_strip)(nx, ny, nx, f, win_r);
        memcpy(recv_l, win_r, s * sizeof(REAL));
        memcpy(recv_r, win_l, s * sizeof(REAL));
    }
    switch (mode) {
    case 1: NAME(halos_sw)(nx, ny, f); break;
    case 2: NAME(halos_cf)(nx, ny, rho_0, u_wall, f); break;
    case 3: NAME(halos_pf)(nx, ny, rho_in, rho_out, pf_state, pf_state + (size_t)ny * 3, f); break;
    case 5: /* body-force channel: periodic x (exchange above), resting walls in y */
        NAME(bb_bottom)(nx, ny, f);
        NAME(bb_top)(nx, ny, 0, 0, 0, f);
        break;
    default: NAME(halos_sl)(nx, ny, rho_0, u_lid, f); break;
    }
    if (mode == 5)
        NAME(kernel_guo)(nx, ny, write_macro, omega, force, f, f_next, rho, u_x, u_y);
    else
        NAME(kernel)(nx, ny, write_macro, omega, f, f_next, rho, u_x, u_y);
}

/* PF macro-strip seeds, main.f90:139-155. */
void NAME(pf_seed)(int ny, REAL rho_in, REAL rho_out, REAL *pf_state)
{
    size_t s = (size_t)ny * 3;
    for (int k = 0; k < 4; ++k) {
        REAL r = (k == 0 || k == 2) ? rho_in : rho_out;
        REAL *m = pf_state + (size_t)k * s;
        for (int y = 0; y < ny; ++y) {
            m[y] = r;
            m[(size_t)ny + y] = 0;
            m[2 * (size_t)ny + y] = 0;
        }
    }
}

/* n_steps of the loop with ping-pong buffers (main.f90:194-262).  The final
 * state is left in f_a (copied from f_b after an odd count); the macro fields
 * are those of the last step when write_macro_last != 0. */
void NAME(run)(int mode, int nx, int ny, int n_steps, REAL omega, REAL rho_0, REAL u_wall,
               REAL u_lid, REAL rho_in, REAL rho_out, REAL *f_a, REAL *f_b, REAL *pf_state,
               int write_macro_last, REAL *rho, REAL *u_x, REAL *u_y, REAL force)
{
    REAL *cur = f_a, *nxt = f_b;
    for (int step = 1; step <= n_steps; ++step) {
        int wm = write_macro_last && step == n_steps;
        NAME(step)(mode, nx, ny, omega, rho_0, u_wall, u_lid, rho_in, rho_out, cur, nxt,
                   pf_state, wm, rho, u_x, u_y, force);
        REAL *t = cur;
        cur = nxt;
        nxt = t;
    }
    if (cur != f_a)
        memcpy(f_a, cur, sizeof(REAL) * (size_t)(nx + 2) * (size_t)(ny + 2) * 9);
}

#undef FI
