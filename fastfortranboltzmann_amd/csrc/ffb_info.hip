// ffb_info.hip -- host-side queries of the C ABI (include/ffb.h) for the
// reports a host prints: the device it runs on and the device memory a handle
// holds.  No kernels.
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>

#include "ffb.h"
#include "ffb_internal.h"

using ffb::fail;

#define IHIPCHK(expr)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return ffb::fail(FFB_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));    \
    } while (0)

extern "C" {

int ffb_device_properties(int32_t device, char *name, int32_t name_len, int32_t *compute_units,
                          int64_t *memory_bytes) {
    int ndev = 0;
    IHIPCHK(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(FFB_E_INVALID, "ffb_device_properties: device not present");
    hipDeviceProp_t p{};
    IHIPCHK(hipGetDeviceProperties(&p, device));
    if (name && name_len > 0) {
        const std::string n = std::string(p.name) + " (" + p.gcnArchName + ")";
        std::strncpy(name, n.c_str(), (size_t)name_len - 1);
        name[name_len - 1] = '\0';
    }
    if (compute_units) *compute_units = p.multiProcessorCount;
    if (memory_bytes) *memory_bytes = (int64_t)p.totalGlobalMem;
    return FFB_OK;
}

// Bytes of the allocations ffb_create makes: both population sets with their
// row padding and the gap between them; rho, u_x, u_y; two Poiseuille strip
// sets of [side][rho, u_x, u_y][n_y]; the shear-wave profile row; the moment
// partial sums and result; the status word.
int ffb_memory_usage(ffb_lattice_t *L, int64_t *population_bytes, int64_t *macro_bytes, int64_t *other_bytes) {
    if (!L) return fail(FFB_E_INVALID, "ffb_memory_usage: null handle");
    const long long cells = (long long)L->cfg.n_x * L->cfg.n_y;
    const long long ce = (long long)L->celem, ny = L->cfg.n_y;
    if (population_bytes) *population_bytes = 2 * (L->set_elems + L->buf_pad) * (long long)L->elem;
    if (macro_bytes) *macro_bytes = 3 * cells * ce;
    if (other_bytes)
        *other_bytes = 2 * 6 * ny * ce + ny * ce + 3LL * L->mom_blocks * (long long)sizeof(double) +
                       3 * (long long)sizeof(double) + (long long)sizeof(unsigned);
    return FFB_OK;
}

}  // extern "C"
