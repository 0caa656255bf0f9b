// fsgh_stage.hip -- per-block node lists for the LDS-staged element kernels.
//
// The element-centric fused kernels (corner forces; energy + volume + EOS +
// divergence) read eight nodes' coordinates and velocities per element.  With
// one thread per element those reads are 8-byte gathers whose 64 lanes touch
// many cache lines each, and they -- not HBM -- bound the kernels: the vector
// memory address and data units are 90-99 % busy while DRAM moves 3.6 TB/s.
// A workgroup of BS Morton-ordered elements touches far fewer distinct nodes
// than 8 BS (a 256-element Morton brick 8x8x4 has 9x9x5 = 405), so the
// kernels load each distinct node of the block once, with the block's lanes
// reading consecutive entries of a sorted node list (coalesced), into LDS,
// and then read the element's eight nodes from LDS.
//
// For every logical element block of a partition (BS elements per block,
// blocks numbered as BlockItem::lid), this file builds
//   start[b] .. start[b+1]  the block's distinct internal node ids, ascending
//                           (nodes[]),
//   lconn[npe e + l]        the position of node l of element e in its
//                           block's list (16 bits; npe = 8 nodes per hex,
//                           4 per RZ quad),
// on the GPU: one workgroup per block sorts the block's npe BS node ids in LDS
// (bitonic), marks the first of each run, ranks them by a block scan, and
// finds each corner's rank by binary search.  The kernels that read them fall
// back to global gathers for a block whose list exceeds their LDS capacity,
// so any element order stays correct (only slower).
//
// The node-centric velocity kernel sums each node's corner forces in its
// corner-list order.  The corners of a block of consecutive nodes occupy one
// contiguous range [cin_start[n0], cin_start[n1]) of the corner list; for
// every logical node block this file also builds
//   scor[q]   the block's corner ids sorted ascending (q in the block's range),
//   spos[q]   for list entry q, the position of its corner in scor (16 bits),
// so the kernel loads the block's corner forces in address order (coalesced)
// into LDS and each node reads its corners from there in list order.
#include <algorithm>
#include <type_traits>
#include <vector>

#include "fsgh_ctx.h"

namespace fsgh {

namespace {

constexpr int kStageThreads = 256;

// KMAX = power of two >= NPE x block size (NPE nodes per element: 8 for
// hexes, 4 for RZ quads); one workgroup of 256 threads per block
template <int KMAX, int NPE>
__global__ __launch_bounds__(kStageThreads) void k_stage_lists(const BlockItem* blocks,
                                                               const int32_t* conn,
                                                               int32_t* counts,
                                                               const int32_t* start,
                                                               int32_t* nodes, uint16_t* lconn)
{
    __shared__ int32_t key[KMAX];
    __shared__ int32_t part[kStageThreads + 1];
    const BlockItem b = blocks[blockIdx.x];
    const int nk = NPE * (b.end - b.start);
    const int32_t* src = conn + NPE * static_cast<size_t>(b.start);
    for (int i = threadIdx.x; i < KMAX; i += kStageThreads) key[i] = i < nk ? src[i] : INT32_MAX;
    __syncthreads();
    // bitonic sort, ascending
    for (int k = 2; k <= KMAX; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < KMAX; i += kStageThreads) {
                const int p = i ^ j;
                if (p > i) {
                    const int32_t a = key[i], c = key[p];
                    const bool up = (i & k) == 0;
                    if ((a > c) == up) {
                        key[i] = c;
                        key[p] = a;
                    }
                }
            }
            __syncthreads();
        }
    // each thread owns a contiguous chunk of the sorted keys
    constexpr int kChunk = KMAX / kStageThreads;
    const int c0 = threadIdx.x * kChunk;
    int cnt = 0;
    for (int i = c0; i < c0 + kChunk; ++i)
        cnt += (i < nk && (i == 0 || key[i] != key[i - 1])) ? 1 : 0;
    part[threadIdx.x] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
        int run = 0;
        for (int t = 0; t < kStageThreads; ++t) {
            const int v = part[t];
            part[t] = run;
            run += v;
        }
        part[kStageThreads] = run;
    }
    __syncthreads();
    const int nu = part[kStageThreads];
    if (counts != nullptr) {
        if (threadIdx.x == 0) counts[b.lid] = nu;
        return;
    }
    // the distinct ids, in order, to the global list (and kept in registers
    // until the barrier: the first nu slots of key[] are then overwritten
    // with the compacted list for the searches)
    int32_t mine[kChunk];
    int r = part[threadIdx.x];
    int32_t* out = nodes + start[b.lid];
    for (int i = c0, k = 0; i < c0 + kChunk; ++i, ++k) {
        const bool first = i < nk && (i == 0 || key[i] != key[i - 1]);
        mine[k] = first ? r : -1;
        if (first) out[r++] = key[i];
    }
    int32_t val[kChunk];
    for (int k = 0; k < kChunk; ++k) val[k] = key[c0 + k];
    __syncthreads();
    for (int k = 0; k < kChunk; ++k)
        if (mine[k] >= 0) key[mine[k]] = val[k];
    __syncthreads();
    uint16_t* lc = lconn + NPE * static_cast<size_t>(b.start);
    for (int i = threadIdx.x; i < nk; i += kStageThreads) {
        const int32_t v = src[i];
        int lo = 0, hi = nu - 1;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (key[mid] < v) lo = mid + 1;
            else hi = mid;
        }
        lc[i] = static_cast<uint16_t>(lo);
    }
}

// one workgroup per logical node block: sort the block's corner-list entries
// by corner id (key = corner << 32 | entry), KMAX = power of two >= the
// capacity; blocks with more entries are skipped (their kernels gather)
template <int KMAX>
__global__ __launch_bounds__(kStageThreads) void k_corner_stage(const BlockItem* blocks,
                                                                const int32_t* cin_start,
                                                                const int32_t* cin, int32_t* scor,
                                                                uint16_t* spos)
{
    __shared__ uint64_t key[KMAX];
    const BlockItem b = blocks[blockIdx.x];
    const int q0 = cin_start[b.start];
    const int nc = cin_start[b.end] - q0;
    if (nc > KMAX) return;
    for (int i = threadIdx.x; i < KMAX; i += kStageThreads)
        key[i] = i < nc ? (static_cast<uint64_t>(static_cast<uint32_t>(cin[q0 + i])) << 32) |
                              static_cast<uint32_t>(i)
                        : ~0ull;
    __syncthreads();
    for (int k = 2; k <= KMAX; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < KMAX; i += kStageThreads) {
                const int p = i ^ j;
                if (p > i) {
                    const uint64_t a = key[i], c = key[p];
                    const bool up = (i & k) == 0;
                    if ((a > c) == up) {
                        key[i] = c;
                        key[p] = a;
                    }
                }
            }
            __syncthreads();
        }
    for (int k = threadIdx.x; k < nc; k += kStageThreads) {
        const uint64_t v = key[k];
        scor[q0 + k] = static_cast<int32_t>(v >> 32);
        spos[q0 + static_cast<int>(v & 0xffffffffu)] = static_cast<uint16_t>(k);
    }
}

template <int KMAX, int NPE>
void launch_stage(int32_t nb, const BlockItem* blocks, const int32_t* conn, int32_t* counts,
                  const int32_t* start, int32_t* nodes, uint16_t* lconn, hipStream_t st)
{
    hipLaunchKernelGGL((k_stage_lists<KMAX, NPE>), dim3(nb), dim3(kStageThreads), 0, st, blocks,
                       conn, counts, start, nodes, lconn);
}

}  // namespace

int build_node_stage(fsgh_ctx* c, int32_t bs)
{
    if (c->host_only || c->stages.count(bs)) return FSGH_OK;
    if (bs != 64 && bs != 128 && bs != 256 && bs != 512) return FSGH_OK;
    std::vector<BlockItem> logical;
    std::vector<int32_t> ib;
    block_table(c->L.elem_off, logical, ib, bs, 0);
    const int32_t nb = static_cast<int32_t>(logical.size());
    if (nb == 0) return FSGH_OK;
    fsgh_ctx::NodeStage ns;
    BlockItem* d_blocks = nullptr;
    int32_t* d_counts = nullptr;
    auto fail = [&](const char* what) {
        set_error(c, std::string("hipMalloc: ") + what);
        for (void* p : {static_cast<void*>(d_blocks), static_cast<void*>(d_counts),
                        static_cast<void*>(ns.start), static_cast<void*>(ns.nodes),
                        static_cast<void*>(ns.lconn)})
            if (p) hipFree(p);
        return FSGH_E_NOMEM;
    };
    if (hipMalloc(reinterpret_cast<void**>(&d_blocks), nb * sizeof(BlockItem)) != hipSuccess)
        return fail("stage blocks");
    if (hipMalloc(reinterpret_cast<void**>(&d_counts), nb * sizeof(int32_t)) != hipSuccess)
        return fail("stage counts");
    if (hipMalloc(reinterpret_cast<void**>(&ns.start), (nb + 1) * sizeof(int32_t)) != hipSuccess)
        return fail("stage start");
    const int npe = c->dims == 2 ? 4 : 8;   // nodes per element
    if (hipMalloc(reinterpret_cast<void**>(&ns.lconn), npe * static_cast<size_t>(c->E) * sizeof(uint16_t)) !=
        hipSuccess)
        return fail("stage lconn");
    HIPCHK(c, hipMemcpyAsync(d_blocks, logical.data(), nb * sizeof(BlockItem), hipMemcpyHostToDevice,
                             c->stream));
    auto run = [&](int32_t* counts, const int32_t* start, int32_t* nodes) {
        const int32_t* cn = c->d_conn;
        if (npe == 4) {
            switch (bs) {
            case 64: launch_stage<256, 4>(nb, d_blocks, cn, counts, start, nodes, ns.lconn, c->stream); break;
            case 128: launch_stage<512, 4>(nb, d_blocks, cn, counts, start, nodes, ns.lconn, c->stream); break;
            case 256: launch_stage<1024, 4>(nb, d_blocks, cn, counts, start, nodes, ns.lconn, c->stream); break;
            default: launch_stage<2048, 4>(nb, d_blocks, cn, counts, start, nodes, ns.lconn, c->stream); break;
            }
            return;
        }
        switch (bs) {
        case 64: launch_stage<512, 8>(nb, d_blocks, cn, counts, start, nodes, ns.lconn, c->stream); break;
        case 128: launch_stage<1024, 8>(nb, d_blocks, cn, counts, start, nodes, ns.lconn, c->stream); break;
        case 256: launch_stage<2048, 8>(nb, d_blocks, cn, counts, start, nodes, ns.lconn, c->stream); break;
        default: launch_stage<4096, 8>(nb, d_blocks, cn, counts, start, nodes, ns.lconn, c->stream); break;
        }
    };
    run(d_counts, nullptr, nullptr);
    std::vector<int32_t> cnt(static_cast<size_t>(nb)), st(static_cast<size_t>(nb) + 1, 0);
    HIPCHK(c, hipMemcpyAsync(cnt.data(), d_counts, nb * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    int64_t total = 0;
    int32_t mx = 0;
    for (int32_t b = 0; b < nb; ++b) {
        st[static_cast<size_t>(b)] = static_cast<int32_t>(total);
        total += cnt[static_cast<size_t>(b)];
        mx = std::max(mx, cnt[static_cast<size_t>(b)]);
    }
    if (total > INT32_MAX) {
        // list offsets are 32-bit: such a mesh keeps the global gathers
        hipFree(d_blocks);
        hipFree(d_counts);
        hipFree(ns.start);
        hipFree(ns.lconn);
        return FSGH_OK;
    }
    st[static_cast<size_t>(nb)] = static_cast<int32_t>(total);
    if (hipMalloc(reinterpret_cast<void**>(&ns.nodes), std::max<int64_t>(total, 1) * sizeof(int32_t)) !=
        hipSuccess)
        return fail("stage nodes");
    HIPCHK(c, hipMemcpyAsync(ns.start, st.data(), st.size() * sizeof(int32_t), hipMemcpyHostToDevice,
                             c->stream));
    run(nullptr, ns.start, ns.nodes);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    hipFree(d_blocks);
    hipFree(d_counts);
    ns.total = total;
    ns.max_nodes = mx;
    c->stages[bs] = ns;
    return FSGH_OK;
}

int build_corner_stage(fsgh_ctx* c, int32_t bs)
{
    if (c->host_only || c->cstages.count(bs)) return FSGH_OK;
    if (bs != 64 && bs != 128 && bs != 256 && bs != 512) return FSGH_OK;
    std::vector<BlockItem> logical;
    std::vector<int32_t> ib;
    block_table(c->L.node_off, logical, ib, bs, 0);
    const int32_t nb = static_cast<int32_t>(logical.size());
    if (nb == 0) return FSGH_OK;
    fsgh_ctx::CornerStage cs;
    BlockItem* d_blocks = nullptr;
    const size_t C = 8 * static_cast<size_t>(c->E);
    if (hipMalloc(reinterpret_cast<void**>(&d_blocks), nb * sizeof(BlockItem)) != hipSuccess ||
        hipMalloc(reinterpret_cast<void**>(&cs.scor), C * sizeof(int32_t)) != hipSuccess ||
        hipMalloc(reinterpret_cast<void**>(&cs.spos), C * sizeof(uint16_t)) != hipSuccess) {
        for (void* p : {static_cast<void*>(d_blocks), static_cast<void*>(cs.scor),
                        static_cast<void*>(cs.spos)})
            if (p) hipFree(p);
        set_error(c, "hipMalloc: corner stage");
        return FSGH_E_NOMEM;
    }
    HIPCHK(c, hipMemcpyAsync(d_blocks, logical.data(), nb * sizeof(BlockItem), hipMemcpyHostToDevice,
                             c->stream));
    const int cap = corner_stage_cap(bs);
    auto go = [&](auto kmax) {
        constexpr int K = decltype(kmax)::value;
        hipLaunchKernelGGL(k_corner_stage<K>, dim3(nb), dim3(kStageThreads), 0, c->stream, d_blocks,
                           c->d_cin_start, c->d_cin, cs.scor, cs.spos);
    };
    if (cap <= 1024) go(std::integral_constant<int, 1024>());
    else if (cap <= 2048) go(std::integral_constant<int, 2048>());
    else if (cap <= 4096) go(std::integral_constant<int, 4096>());
    else go(std::integral_constant<int, 8192>());
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    hipFree(d_blocks);
    c->cstages[bs] = cs;
    return FSGH_OK;
}

void free_node_stages(fsgh_ctx* c)
{
    for (auto& kv : c->cstages)
        for (void* p : {static_cast<void*>(kv.second.scor), static_cast<void*>(kv.second.spos)})
            if (p) hipFree(p);
    c->cstages.clear();
    for (auto& kv : c->stages)
        for (void* p : {static_cast<void*>(kv.second.start), static_cast<void*>(kv.second.nodes),
                        static_cast<void*>(kv.second.lconn)})
            if (p) hipFree(p);
    c->stages.clear();
}

void attach_corner_stage(fsgh_ctx* c, DMesh& m, int32_t nbs)
{
    auto it = c->cstages.find(nbs);
    m.cs_cor = it == c->cstages.end() ? nullptr : it->second.scor;
    m.cs_pos = it == c->cstages.end() ? nullptr : it->second.spos;
}

void attach_node_stage(fsgh_ctx* c, DMesh& m, int32_t ebs)
{
    auto it = c->stages.find(ebs);
    if (it == c->stages.end()) {
        m.sn_start = nullptr;
        m.sn_nodes = nullptr;
        m.sn_lconn = nullptr;
        return;
    }
    m.sn_start = it->second.start;
    m.sn_nodes = it->second.nodes;
    m.sn_lconn = it->second.lconn;
}

}  // namespace fsgh
