// fsgh_lds.h -- device helpers shared by the LDS-staged element kernels of
// the 3D and the 2D RZ cycles: streaming 16-byte accesses and the block's
// node-list range (lists built in fsgh_stage.hip).
#pragma once

#include "fsgh_internal.h"

namespace fsgh {
namespace {

// Streaming 16-byte accesses (non-temporal): the corner forces are written
// once by the force kernel and read once per element by the energy kernel,
// gigabytes apart at large sizes, so they gain nothing from the caches.
typedef double v2d_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void stream_store(double2* p, double2 v)
{
    __builtin_nontemporal_store(v2d_t{v.x, v.y}, reinterpret_cast<v2d_t*>(p));
}
__device__ __forceinline__ double2 stream_load(const double2* p)
{
    const v2d_t q = __builtin_nontemporal_load(reinterpret_cast<const v2d_t*>(p));
    return make_double2(q.x, q.y);
}

// the block's node range in its list, or nn = cap + 1 when the mesh view has
// no lists (block-uniform)
__device__ __forceinline__ void stage_range(const DMesh& m, const BlockItem& bi, int cap, int& q0,
                                            int& nn)
{
    q0 = 0;
    nn = cap + 1;
    if (m.sn_start != nullptr) {
        q0 = m.sn_start[bi.lid];
        nn = m.sn_start[bi.lid + 1] - q0;
    }
}

// true when this wave's entries r0 + (threadIdx.x & ~63) ... start inside a
// list of nn entries (wave-uniform, a scalar branch); lanes past the end
// then load clamped entries whose values are not stored
__device__ __forceinline__ bool wave_in_list(int r0, int nn)
{
    const int first = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) & ~63);
    return r0 + first < nn;
}

}  // namespace
}  // namespace fsgh
