/*
 * ffb.h -- C ABI of the MI355X (gfx950) D2Q9 pull-stream-collide BGK solver
 * (libffb_hip.so).
 *
 * The reference advances one domain per time step with
 *   exchange_halos(...)                                   source/layouts/xyd/exchange.f90:161-378
 *   [Poiseuille] prepare_poiseuille_flow_halos_PF + update_poiseuille_flow_macro_strips_PF
 *                + exchange_poiseuille_macro_halos        source/layouts/xyd/main.f90:212-246
 *   execute_local_sim_step(domain_info, halo_buffers, n_x_local, n_y_local,
 *                          write_macro_fields, f, f_next, rho, u_x, u_y)
 *                                                         source/layouts/xyd/simulation.f90:18-73
 * on host arrays f(0:NX+1, 0:NY+1, 9) it owns.  Here one handle owns a
 * device-resident domain (two ping-pong population sets, no halo: the halo
 * writes of the reference become index rules inside the kernel) and
 * ffb_step() performs that whole sequence, n_steps times, on the GPU.
 *
 * Conventions
 *   - Every function returns 0 on success and a non-zero FFB_E_* code on
 *     failure; ffb_last_error() then holds a message (thread-local).  The
 *     reference stops all images with `error stop "..."` instead
 *     (simulation.f90:59-60, domain.f90:79-81); a Fortran host maps a
 *     non-zero status to `error stop ffb_last_error()`.
 *   - Calls on one handle are not thread-safe; different handles are
 *     independent.  Work is enqueued on the handle's own HIP stream and
 *     ffb_step() returns without waiting; the ffb_read_* calls and
 *     ffb_sync() wait.
 *   - Host arrays use the reference's Fortran layouts: macro fields
 *     rho(NX, NY) with x fastest (simulation.f90:32-34), populations
 *     f(0:NX+1, 0:NY+1, 1:9) with x fastest, then y, then direction
 *     (simulation.f90:27).  Element type is double for FFB_PREC_FP64 and
 *     float for every other precision.
 */
#ifndef FFB_H
#define FFB_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* sim modes, source/settings.f90:56-59 */
#define FFB_SHEAR_WAVE 1
#define FFB_COUETTE_FLOW 2
#define FFB_POISEUILLE_FLOW 3
#define FFB_SLIDING_LID 4
/* body-force channel: periodic in x, halfway bounce-back walls in y, uniform
 * body force along x (Guo forcing).  No reference counterpart: the batched
 * and differentiation forms use it for their Poiseuille lattices. */
#define FFB_BODY_FORCE_CHANNEL 5

/* storage precision of the populations.  FP64 / FP32 are the reference's
 * FFB_FP64 on / off builds (source/settings.f90:7-13).  The deviation forms
 * store D_i = (f_i - w_i) * 2^storage_scale_log2 as FP16, BF16, FP32 or the
 * 1-3-12 layout below and run
 * the step on D in FP32 (boundary rules and BGK collision in deviation space,
 * FP32 arithmetic on values of the size of the deviations); they have no
 * reference counterpart.  FFB_PREC_FP32_DEV has the FP32 form's traffic and
 * an error against FP64 about ten times smaller for small velocities and
 * forces.  FFB_PREC_E3M12_DEV stores D in 16 bits as sign, 3-bit exponent and
 * 12-bit significand with subnormals, with one more binade of range for the
 * rest population and one more bit of precision for the diagonal ones:
 * |f - w| < 2^(1 - s) (axis), 2^(2 - s) (rest), 2^-s (diagonal), s =
 * storage_scale_log2, with 12-bit precision over the top seven binades and a
 * uniform step below (two significand bits more than FP16 over the range a
 * flow with a known largest speed uses); larger deviations are clamped to the
 * largest magnitude, so s must leave room for the flow's largest |f - w|
 * (s = 5 for speeds up to 0.1, including the density peak at a sliding lid's
 * corners). */
#define FFB_PREC_FP32 0
#define FFB_PREC_FP64 1
#define FFB_PREC_FP16_DEV 2
#define FFB_PREC_BF16_DEV 3
#define FFB_PREC_FP32_DEV 4
#define FFB_PREC_E3M12_DEV 5

/* rounding of the 16-bit deviation forms' stored values (FP16, BF16, 1-3-12).
 * Nearest-even loses every update smaller than half a storage step, so slow
 * flows stagnate (a 256^2 shear wave at omega 1 does not decay at all in
 * FP16); stochastic rounding, with random bits from a hash of (cell, step),
 * keeps those updates on average and stays reproducible. */
#define FFB_ROUND_NEAREST 0
#define FFB_ROUND_STOCHASTIC 1

/* status codes */
#define FFB_OK 0
#define FFB_E_INVALID 1   /* bad argument or configuration */
#define FFB_E_HIP 2       /* HIP runtime error (allocation, launch, copy) */
#define FFB_E_STATE 3     /* call not valid in the handle's current state */
#define FFB_E_DENSITY 4   /* density check (ffb_config_t.density_checks) failed: a cell had
                             rho <= 0 or rho NaN in a collision (ffb_sync reports it) */

/* Run configuration: the compile-time parameters of source/settings.f90:21-53
 * as run-time values. */
typedef struct ffb_config {
    int32_t n_x;          /* N_X, cells in x (settings.f90:23) */
    int32_t n_y;          /* N_Y (settings.f90:24) */
    int32_t sim_mode;     /* SIM_MODE, FFB_SHEAR_WAVE .. FFB_SLIDING_LID (settings.f90:61) */
    int32_t precision;    /* FFB_PREC_* */
    double omega;         /* OMEGA, BGK relaxation rate (settings.f90:37) */
    double rho_0;         /* RHO_0 (settings.f90:36) */
    double u_max;         /* U_MAX, shear-wave amplitude (settings.f90:38) */
    double n_sin;         /* N_SIN, shear-wave periods over N_Y (settings.f90:39) */
    double u_wall;        /* U_WALL, Couette top wall (settings.f90:40) */
    double u_lid;         /* U_LID, sliding lid (settings.f90:41) */
    double rho_in;        /* RHO_IN, Poiseuille inlet density (settings.f90:42) */
    double rho_out;       /* RHO_OUT (settings.f90:43) */
    int32_t storage_scale_log2; /* s of the deviation forms; ignored otherwise */
    int32_t device;       /* HIP device ordinal */
    double body_force;    /* FFB_BODY_FORCE_CHANNEL: force density along x; ignored otherwise */
    int32_t storage_rounding; /* 16-bit forms (FP16, BF16, 1-3-12): FFB_ROUND_STOCHASTIC or FFB_ROUND_NEAREST */
    int32_t density_checks;   /* FFB_DENSITY_CHECKS (simulation.f90:229-233, poiseuille_flow.f90:156-160):
                                 non-zero: every collision and Poiseuille strip flags a density rho <= 0
                                 (or NaN) and ffb_sync returns FFB_E_DENSITY */
} ffb_config_t;

typedef struct ffb_lattice ffb_lattice_t;

/* Fills *cfg with the reference defaults of source/settings.f90:27-63
 * (2520 x 2520, sliding lid, omega 1.1, ...), FP32 storage, device 0. */
void ffb_config_default(ffb_config_t *cfg);

/* Allocates the device domain.  Replaces the reference's allocate(f_a, f_b,
 * rho, u_x, u_y) and halo buffers (main.f90:96-105, exchange.f90:117-135).
 * Bounds: n_x <= 2^24 (row and byte offsets fit 32 bits), n_x * n_y <= 2^34. */
int ffb_create(const ffb_config_t *cfg, ffb_lattice_t **out);
int ffb_destroy(ffb_lattice_t *lat);

/* Initial condition of cfg->sim_mode on the device, plus the Poiseuille strip
 * seeds.  Replaces apply_condition_*_local (initialization.f90:9-198) and
 * main.f90:139-155.  Resets the step counter. */
int ffb_init(ffb_lattice_t *lat);

/* Enqueues n_steps time steps (main.f90:194-262 loop body, incl.
 * exchange_halos and the Poiseuille pre-pass, then execute_local_sim_step,
 * simulation.f90:18-73).  When write_macro_last != 0 the last step also
 * writes rho, u_x, u_y (the reference's write_macro_fields).  In every form
 * (not in Poiseuille flow) consecutive steps run several per kernel launch
 * (five for FP32 and FP32 deviations, four for FP64, three for FP16 and BF16
 * deviations, two for the 1-3-12 form), with fewer for a remainder; the
 * results are bitwise those of one launch per step (FFB_STEPS_PER_PASS=1..5 at
 * ffb_create selects the steps per launch; FFB_TWO_STEP=0 one launch per
 * step). */
int ffb_step(ffb_lattice_t *lat, int32_t n_steps, int32_t write_macro_last);

/* Changes the physical parameters of an existing handle (the next steps use
 * them): omega, body-force density, Couette wall speed, lid speed. */
int ffb_set_parameters(ffb_lattice_t *lat, double omega, double body_force, double u_wall, double u_lid);

/* Waits for all work of the handle.  With cfg.density_checks set, returns
 * FFB_E_DENSITY once a collision has met rho <= 0 (message: the reference's
 * "error: density is zero in collision/streaming step (rho_val <= 0)") or a
 * NaN density; the flag stays set until ffb_init.  Replaces the reference's
 * `error stop` under FFB_DENSITY_CHECKS (simulation.f90:229-233, :133-137,
 * modes/poiseuille_flow.f90:156-160). */
int ffb_sync(ffb_lattice_t *lat);

/* Copies the macro fields written by the last write_macro step to host
 * arrays of n_x*n_y elements each, x fastest (any pointer may be NULL). */
int ffb_read_macros(ffb_lattice_t *lat, void *rho, void *u_x, void *u_y);

/* Current populations as the reference array f(0:NX+1, 0:NY+1, 1:9) of the
 * buffer the next step reads.  Halo entries are returned as 0 (on the device
 * they do not exist). */
int ffb_read_populations(ffb_lattice_t *lat, void *f_halo);

/* Sets the current populations from f(0:NX+1, 0:NY+1, 1:9); halo entries are
 * ignored.  16-bit forms convert to the stored deviation with rounding. */
int ffb_write_populations(ffb_lattice_t *lat, const void *f_halo);

/* Poiseuille strips that the next step will use as "old" (recv_macro_left,
 * recv_macro_right of exchange.f90:33-40), each (NY, 3) column-major: rho,
 * u_x, u_y.  Either pointer may be NULL. */
int ffb_read_pf_strips(ffb_lattice_t *lat, void *macro_left, void *macro_right);

/* Device time of the steps enqueued since ffb_init() or the last
 * ffb_reset_timer(), in milliseconds (HIP events around each ffb_step()). */
int ffb_elapsed_ms(ffb_lattice_t *lat, double *ms);
int ffb_reset_timer(ffb_lattice_t *lat);

/* Number of steps taken since ffb_init(). */
int ffb_step_count(ffb_lattice_t *lat, int64_t *steps);

/* FP64 sums over the owned cells of the current state: mass = sum rho,
 * mom_x = sum rho u_x, mom_y = sum rho u_y (conservation drift). */
int ffb_moments(ffb_lattice_t *lat, double *mass, double *mom_x, double *mom_y);

/* Largest |f_i - w_i| of the current state over the owned cells, per
 * direction class (max_dev[0] rest, [1] axis, [2] diagonal), in FP64, and
 * the largest |f - w| the storage form holds for that class (limit[]): for
 * the 1-3-12 form (2 - 2^-12) 2^(-s - off) (off = -1, 0, +1), values beyond it
 * are clamped, so max_dev == limit shows clamping and max_dev near limit a
 * scale s without headroom; FP16 65504 2^-s (overflow beyond); the FP32/BF16
 * forms their type's largest finite value times 2^-s; plain FP32/FP64 the
 * type's largest value.  Synchronises.  No reference counterpart (a check of
 * the storage_scale_log2 choice). */
int ffb_deviation_extent(ffb_lattice_t *lat, double max_dev[3], double limit[3]);

/* Default storage_scale_log2 of a storage form for flows with largest speed
 * `speed` (the lid, wall or wave speed) and density deviations up to 0.05:
 * FP16 puts the largest expected |f - w|, (4/9)(0.05 + 3u + 6u^2), near 2^14
 * (a factor 4 below its maximum) so small deviations stay normal; the 1-3-12
 * form puts the largest per-class bound w (0.05 + 3|c|u + ...) 2^off at 3/4 of
 * its largest magnitude (s = 5 for u = 0.1); BF16, FP32 deviations and the
 * plain forms 0.  Host arithmetic only (no GPU).  Returns the scale, or
 * INT32_MIN for an unknown precision or a speed outside [0, 1].  No reference
 * counterpart. */
int32_t ffb_default_scale_log2(int32_t precision, double speed);

/* ffb_moments without allocating: the FP64 sums mass, mom_x, mom_y of the
 * populations at `populations` (device memory, the layout ffb_device_info
 * reports: element d * plane + y * pitch + x, stored form of cfg->precision and
 * cfg->storage_scale_log2, cfg->n_x by cfg->n_y cells), written to out[0..2]
 * (device memory) by kernels enqueued on `stream` (a hipStream_t; NULL: the
 * null stream) of device cfg->device; returns without waiting.  The same
 * kernels and partial-sum order as ffb_moments, so the sums are bitwise
 * equal.  No counterpart in the reference; for callers that cannot allocate
 * device memory (code under hipGraph capture, serving loops, harnesses that
 * forbid hipMalloc): the partial sums live in a 48 KiB array of the library
 * on each device, so the call makes no allocation.  Limit: n_x <= 2^24,
 * n_x * n_y <= 2^34 (at most 2048 partial sums, any domain size within that).
 * Calls that overlap on one device (other streams, other threads) share that
 * array and race; each call overwrites every partial sum it reads. */
int ffb_moments_noalloc(const ffb_config_t *cfg, const void *populations, int64_t pitch, int64_t plane,
                        void *stream, double *out);

/* Device pointers for zero-copy interop (torch, hipGraph capture):
 * *cur = populations the next step reads, *stream = the handle's hipStream_t;
 * population d of cell (x, y) is element d * plane + y * pitch + x of *cur
 * (*pitch = row stride, *plane = direction stride, in elements). */
int ffb_device_info(ffb_lattice_t *lat, void **cur, void **stream, int64_t *pitch, int64_t *plane);

/* Measurement helper with no reference counterpart: device-to-device copy of
 * `bytes` bytes by plain HIP copy kernels (16-byte loads and stores; four
 * variants -- grid-stride, one element per thread with plain, nontemporal-store
 * and nontemporal load+store -- the fastest counts), repeated `iters` times on `device`;
 * *gbps = (read + written bytes) / time.  Gives the measured HBM ceiling the
 * step kernel is compared with. */
int ffb_bench_copy(int32_t device, int64_t bytes, int32_t iters, double *gbps);

/* ------------------------------------------------------------------------
 * Batched form (no reference counterpart: the reference runs one domain per
 * process).  Many independent lattices of uneven size, each with its own
 * omega and boundary kind, advance together: consecutive steps run several
 * per launch of one kernel over the tiles of all lattices (five for the 4-byte
 * forms, three for FP64 and the 16-bit forms; FFB_BATCH_STEPS_PER_PASS=2..5 at
 * creation selects another count), or with FFB_BATCH_PAIRS=0 one step per
 * launch of one kernel over the row segments of all lattices (all bitwise the
 * same; either kernel runs one workgroup per tile or a persistent grid), and
 * the launches of a
 * ping-pong cycle are captured once in a hipGraph and replayed.  ffb_batch_autotune()
 * picks the launch shape of the kernel the batch runs by timing the
 * candidates on the batch.
 * ------------------------------------------------------------------------ */
typedef struct ffb_batch_lattice {
    int32_t n_x, n_y;     /* n_x <= 2^24, n_x * n_y <= 2^30 */
    int32_t kind;         /* FFB_SHEAR_WAVE, FFB_COUETTE_FLOW, FFB_SLIDING_LID or
                             FFB_BODY_FORCE_CHANNEL (Poiseuille flow needs the
                             single-domain API) */
    int32_t storage_scale_log2; /* s of the deviation forms (D = (f - w) 2^s); ignored otherwise */
    double omega;
    double rho_0;
    double u_wall;        /* Couette top wall / sliding lid speed */
    double body_force;    /* body-force channel: force density along x */
    double u_max, n_sin;  /* shear wave: initial u_x = u_max sin(2 pi n_sin y / n_y) */
} ffb_batch_lattice_t;

typedef struct ffb_batch_tuning {
    /* one launch per step (also a remainder of a multi-step batch) */
    int32_t tile_w;       /* threads per workgroup (64-cell row segments x 64): 64 .. 1024 */
    int32_t grid_blocks;  /* persistent grid size; 0 = one workgroup per tile */
    /* multi-step launches (K = 2 .. 5 steps each) */
    int32_t pair_shape;   /* 1: 64 x 32 regions (output tiles (64 - 2(K-1)) x (32 - 2(K-1)); 8 waves x 4
                             rows for three or more steps, 16 x 2 for pairs), 2: 64 x 16 regions (8 waves x
                             2 rows; four and five steps 64 x 24, 12 x 2); FP64 one shape (62 x 6, 60 x 12) */
    int32_t pair_group;   /* consecutive tiles per XCD (0: dispatch order) */
    int32_t pair_grid;    /* persistent grid size (rounded up to a multiple of 8); 0 = one workgroup per tile */
    int32_t reserved;
} ffb_batch_tuning_t;

typedef struct ffb_batch ffb_batch_t;

/* precision: any FFB_PREC_* form.  The deviation forms use each lattice's
 * storage_scale_log2 (ffb_default_scale_log2 at the lattice's speed is a good
 * choice; 0 for FP32 deviations) and round the 16-bit stores stochastically,
 * with random bits from (cell of the lattice, step), as the single-domain form;
 * populations read back as f = w + D 2^-s in FP32. */
int ffb_batch_create(const ffb_batch_lattice_t *lattices, int32_t count, int32_t precision, int32_t device,
                     ffb_batch_t **out);
int ffb_batch_destroy(ffb_batch_t *batch);
/* Equilibrium initial state of every lattice (shear wave: its sine profile;
 * the others at rest with rho_0). */
int ffb_batch_init(ffb_batch_t *batch);
/* Enqueues n_steps steps of every lattice (graph replays; the graph is
 * captured again after a change of tuning). */
int ffb_batch_step(ffb_batch_t *batch, int32_t n_steps, int32_t write_macro_last);
int ffb_batch_sync(ffb_batch_t *batch);
int ffb_batch_elapsed_ms(ffb_batch_t *batch, double *ms);
int ffb_batch_reset_timer(ffb_batch_t *batch);
/* Total cells, number of tiles, and the tuning in use. */
int ffb_batch_info(ffb_batch_t *batch, int64_t *cells, int32_t *tiles, ffb_batch_tuning_t *tuning);
int ffb_batch_set_tuning(ffb_batch_t *batch, const ffb_batch_tuning_t *tuning);
/* Times every candidate launch shape of the kernel the batch runs (step
 * multi-step launches: pair_shape x pair_group x pair_grid; one launch per step: tile_w x
 * grid_blocks) for trial_steps steps, keeps the fastest, and re-initialises
 * the batch (the trials advance it).  *best_ms_per_step is per step. */
int ffb_batch_autotune(ffb_batch_t *batch, int32_t trial_steps, ffb_batch_tuning_t *best, double *best_ms_per_step);
/* Macro fields of lattice `index` after a write_macro step (n_x*n_y each). */
int ffb_batch_read_macros(ffb_batch_t *batch, int32_t index, void *rho, void *u_x, void *u_y);
/* Populations of lattice `index` as f(0:NX+1, 0:NY+1, 1:9) (halo zero). */
int ffb_batch_read_populations(ffb_batch_t *batch, int32_t index, void *f_halo);

/* ------------------------------------------------------------------------
 * Differentiation form (no reference counterpart): discrete adjoint of n_steps
 * steps of a handle (shear wave, Couette, sliding lid or body-force channel;
 * FP32, FP32 deviations or FP64), with a checkpoint every checkpoint_interval steps and
 * recomputation of each segment during the reverse sweep.  The loss is any
 * function of the velocity field u written by the last step; the caller
 * passes dL/du and gets dL/domega, dL/d(body force) and dL/d(wall or lid
 * speed).  Parameters are those of the handle (ffb_set_parameters).
 * ------------------------------------------------------------------------ */
typedef struct ffb_adjoint ffb_adjoint_t;

int ffb_adjoint_create(ffb_lattice_t *lat, int32_t n_steps, int32_t checkpoint_interval, ffb_adjoint_t **out);
int ffb_adjoint_destroy(ffb_adjoint_t *adj);
/* ffb_init + n_steps steps, storing the checkpoints; copies the final macro
 * fields (n_x*n_y each, element type of the handle; host or device pointers,
 * any may be NULL). */
int ffb_adjoint_forward(ffb_adjoint_t *adj, void *rho, void *u_x, void *u_y);
/* Reverse sweep for the seed dL/du_x, dL/du_y (host or device arrays of
 * n_x*n_y); the three gradients are returned as FP64 sums. */
int ffb_adjoint_backward(ffb_adjoint_t *adj, const void *grad_u_x, const void *grad_u_y, double *d_omega,
                         double *d_body_force, double *d_u_wall);
/* Device time of the last forward and backward calls, in milliseconds. */
int ffb_adjoint_timing(ffb_adjoint_t *adj, double *forward_ms, double *backward_ms);

/* Name ("<marketing name> (<gfx arch>)", truncated to name_len - 1 characters),
 * compute units and memory of HIP device `device`, for the hardware report the
 * reference prints from hardware_info.f90:132-145 (any pointer may be NULL). */
int ffb_device_properties(int32_t device, char *name, int32_t name_len, int32_t *compute_units,
                          int64_t *memory_bytes);

/* Device bytes the handle holds: both population sets (row padding
 * included), the macro fields rho, u_x, u_y, and the rest (Poiseuille strips,
 * profile row, partial sums, status word).  Counterpart of the reference's
 * memory_usage report (main.f90:109-119, export.f90:237-247). */
int ffb_memory_usage(ffb_lattice_t *lat, int64_t *population_bytes, int64_t *macro_bytes, int64_t *other_bytes);

/* Message of the last failing call on this thread ("" if none). */
const char *ffb_last_error(void);

/* Version string of the library. */
const char *ffb_version(void);

#ifdef __cplusplus
}
#endif

#endif /* FFB_H */
