// fsgh_noalloc.hip -- energy tallies of caller-held device arrays with no
// allocation, for callers that cannot allocate device memory (code under HIP
// graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// The tally is the one fsgh_energy computes (the reference's IE/KE sums,
// K/sgh_solve.cpp:66-104; 2D RZ: KE = 1/2 sum m R |v|^2), with the same two
// kernels (fixed-order block partial sums, then one block over the partials),
// so both give bitwise-identical results for the same arrays in the same
// order.  The partial sums live in a file-scope device array instead of the
// context's scratch buffer.  Its device address is looked up once per device
// with hipGetSymbolAddress (host code never passes the array's name itself to
// a kernel or copy: on the host that name is only a stand-in).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>

#include "fsgh_internal.h"
#include "../../include/fsgh.h"

namespace {

// Partial-sum slots: one (IE, KE) pair per block of the first kernel.
constexpr int32_t kPartials = 1024;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[2 * kPartials];

std::mutex g_mutex;
double* g_scratch_ptr[kMaxDevices] = {};

// The device address of this device's copy of `scratch`, looked up on the
// first call with that device current.
int scratch_for_device(int device, double** out)
{
    std::lock_guard<std::mutex> lock(g_mutex);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
            (void)hipGetLastError();
            return FSGH_E_HIP;
        }
        g_scratch_ptr[device] = static_cast<double*>(p);
    }
    *out = g_scratch_ptr[device];
    return FSGH_OK;
}

}  // namespace

extern "C" {

int fsgh_energy_noalloc(int32_t hip_device, int32_t num_dims, int64_t num_elems,
                        const double* elem_mass, const double* elem_sie, int64_t num_nodes,
                        const double* node_mass, const double* node_u, const double* node_v,
                        const double* node_w, const double* node_radius, double* ie_ke,
                        void* hip_stream)
{
    if (num_dims != 2 && num_dims != 3) return FSGH_E_INVALID;
    if (num_elems < 0 || num_nodes < 0 || num_elems > FSGH_NOALLOC_MAX_ENTRIES ||
        num_nodes > FSGH_NOALLOC_MAX_ENTRIES)
        return FSGH_E_INVALID;
    if (!ie_ke || (num_elems > 0 && (!elem_mass || !elem_sie)) ||
        (num_nodes > 0 && (!node_mass || !node_u || !node_v)) ||
        (num_nodes > 0 && num_dims == 3 && !node_w) ||
        (num_nodes > 0 && num_dims == 2 && !node_radius))
        return FSGH_E_INVALID;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess) return FSGH_E_HIP;
    if (hip_device < 0 || hip_device >= ndev || hip_device >= kMaxDevices) return FSGH_E_INVALID;

    // make the arrays' device current for the look-up and the launches, then
    // give the caller back its own current device
    int prev = 0;
    if (hipGetDevice(&prev) != hipSuccess) return FSGH_E_HIP;
    if (prev != hip_device && hipSetDevice(hip_device) != hipSuccess) return FSGH_E_HIP;
    double* part = nullptr;
    int rc = scratch_for_device(hip_device, &part);
    if (rc == FSGH_OK) {
        // only the fields the tally reads; the kernels index [0, E) and [0, N)
        fsgh::DState s{};
        s.emass = const_cast<double*>(elem_mass);
        s.sie1 = const_cast<double*>(elem_sie);
        s.nmass = const_cast<double*>(node_mass);
        s.u1 = const_cast<double*>(node_u);
        s.v1 = const_cast<double*>(node_v);
        s.w1 = const_cast<double*>(node_w);
        s.y1 = const_cast<double*>(node_radius);
        const int32_t eoff[2] = {0, static_cast<int32_t>(num_elems)};
        const int32_t noff[2] = {0, static_cast<int32_t>(num_nodes)};
        fsgh::DMesh m{};
        // every partial slot is written by the first kernel (a block with no
        // entries writes zeros), so the array needs no clearing between calls
        fsgh::launch_energy_sums(m, s, eoff, noff, 0, ie_ke, part, kPartials,
                                 num_dims == 2 ? 1 : 0, static_cast<hipStream_t>(hip_stream));
        if (hipGetLastError() != hipSuccess) rc = FSGH_E_HIP;
    }
    if (prev != hip_device) (void)hipSetDevice(prev);
    return rc;
}

}  // extern "C"
