// ffb_internal.h -- state and helpers shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <string>

#include "ffb.h"

namespace ffb {
// Records msg as the thread's last error (returned by ffb_last_error) and
// returns code.
int fail(int code, const std::string &msg);
}  // namespace ffb

// One device-resident domain (the handle behind ffb_lattice_t).
struct ffb_lattice {
    ffb_config_t cfg;
    // element (d, x, y) of a population set is at d * plane + y * pitch + x:
    //   layout 0 (planes): 9 planes of n_y rows, pitch = n_x rounded up to 64,
    //     plane = pitch * n_y (+ plane_pad);
    //   layout 1 (rows): per row y the 9 direction rows one after another,
    //     plane = n_x rounded up to 64 (+ plane_pad), pitch = 9 * plane.
    int pitch = 0;           // row stride in elements
    long long plane = 0;     // direction stride in elements
    long long set_elems = 0; // elements of one population set
    int layout = 0;
    size_t elem = 0;         // bytes per stored population
    size_t celem = 0;        // bytes per compute value (macro fields, strips)
    void *pop_alloc = nullptr;  // one allocation holding both population sets
    void *buf[2] = {nullptr, nullptr};
    int cur = 0;             // buffer the next step reads
    void *macro[3] = {nullptr, nullptr, nullptr};
    void *strips[2] = {nullptr, nullptr};  // Poiseuille strips, [side][3][ny] each
    int strip_old = 0;
    void *ux_row = nullptr;  // shear-wave initial profile
    double *mom_partial = nullptr;
    double *mom_out = nullptr;
    unsigned *status = nullptr;  // density-check flags (StepArgs::status)
    int mom_blocks = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    bool timing = false;
    long long steps = 0;
    int tile_w = 512;        // threads per workgroup
    int vec = 1;             // cells per thread of the step kernel (1, 2, 4, 8)
    int xcd_group = 0;       // G consecutive tiles per XCD, 0: plain order (tile_of_block)
    int two_step = 0;        // 1: step pairs in one pass (step2_kernel) where the mode allows
    int steps_per_pass = 2;  // 3: step triples (stepk_kernel) where pairs are allowed
    int triple_shape = 2;    // triple region: 2 = 8 waves x 2 rows, 1 = 16 waves x 2 rows, 3 = 8 x 4 (<= 4-byte elements)
    int quad_shape = 2;      // four-step region: 2 = 12 waves x 2 rows (default), 1 = 16 x 2, 3 = 8 x 2 (FP64 default),
                             // 4 = 8 x 4 (<= 4-byte elements)
    int penta_shape = 1;     // five-step region (4-byte elements): 1 = 8 waves x 4 rows, 2 = 12 x 2
    unsigned lds_pad = 0;    // dynamic LDS bytes added to multi-step launches (fewer workgroups per CU)
    int multi_pipe = 0;      // 1: three/four-step launches by stepk_pipe_kernel (persistent, loads one tile ahead)
    int multi_grid = 0;      // stepk_pipe_kernel workgroups, 0: occupancy x compute units
    int n_cu = 1;            // compute units of the device
    int pair_xcd_group = 0;  // G consecutive pair tiles per XCD, 0: plain order
    int pair_xcd_rows = 1;   // H > 1: blocks of G x H pair tiles per XCD
    int pair_shape = 1;      // pair region (4-byte and narrower): 1 = 16 waves x 2 rows, 2 = 8 waves x 2 rows,
                             // 3 = 8 waves x 4 rows (stepk_kernel)
    int pair_nt = 1;         // pair stores: 1 nontemporal, 0 default cache policy
    int pair_grid = 0;       // pair workgroups: 0 = one per tile, else a persistent grid
    int pair_kernel = 0;     // step pairs: 0 = tiles (step2_kernel), 1 = column march (march2_kernel)
    int march_shape = 1;     // march band: 1 = 8 waves x 1 row, 2 = 8 x 2, 3 = 16 x 1, 4 = 4 x 2
    int march_rows = 126;    // march: output rows per workgroup
    int march_inner = 1;     // march: strips away from the x edges pull without edge rules
    long long plane_pad = 0; // elements added to each plane stride (multiple of 64)
    long long buf_pad = 0;   // elements between the two population sets (multiple of 64)
};

namespace ffb {
// Enqueues one step of L's configuration reading populations at src and
// writing dst (both in L's layout), on L's stream; Poiseuille flow and the
// 16-bit forms are not supported (FP32 deviations are).
int step_between(ffb_lattice *L, const void *src, void *dst, int write_macro);
}  // namespace ffb
