// fsgh_mesh.hip -- host-side mesh preparation (renumbering, CSR, work tables, BC sets).
#include "fsgh_mesh.h"

#include <algorithm>
#include <cmath>
#include <numeric>

namespace fsgh {

namespace {

// node lids of the six faces of a hex (K/mesh.h:353-358) and of the four
// edges of a quad (K/mesh.h:366-377)
const int kFace[6][4] = {{0, 4, 7, 3}, {1, 2, 6, 5}, {0, 1, 5, 4},
                         {2, 3, 7, 6}, {0, 3, 2, 1}, {4, 5, 6, 7}};
const int kQuadEdge[4][4] = {{0, 3, 0, 0}, {1, 2, 0, 0}, {0, 1, 0, 0}, {3, 2, 0, 0}};

}  // namespace

void morton_renumber(const Layout& L, const int64_t* conn, const double* xyz,
                     std::vector<int32_t>& new_of_old_elem, std::vector<int32_t>& new_of_old_node,
                     int npe)
{
    const int64_t E = L.elem_off.back(), N = L.node_off.back();
    new_of_old_elem.assign(static_cast<size_t>(E), -1);
    new_of_old_node.assign(static_cast<size_t>(N), -1);
    std::vector<std::pair<uint64_t, int32_t>> key;
    for (int32_t i = 0; i < L.ninst; ++i) {
        const int64_t e0 = L.elem_off[i], e1 = L.elem_off[i + 1];
        const int64_t n0 = L.node_off[i], n1 = L.node_off[i + 1];
        double lo[3] = {HUGE_VAL, HUGE_VAL, HUGE_VAL}, hi[3] = {-HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
        for (int64_t n = n0; n < n1; ++n)
            for (int d = 0; d < 3; ++d) {
                lo[d] = std::min(lo[d], xyz[3 * n + d]);
                hi[d] = std::max(hi[d], xyz[3 * n + d]);
            }
        key.resize(static_cast<size_t>(e1 - e0));
        for (int64_t e = e0; e < e1; ++e) {
            double c[3] = {0.0, 0.0, 0.0};
            for (int a = 0; a < npe; ++a)
                for (int d = 0; d < 3; ++d) c[d] += xyz[3 * conn[npe * e + a] + d];
            uint64_t code = 0;
            for (int d = 0; d < 3; ++d)
                code |= morton_spread3(morton_quantize(c[d] / static_cast<double>(npe), lo[d], hi[d])) << d;
            key[static_cast<size_t>(e - e0)] = {code, static_cast<int32_t>(e)};
        }
        std::sort(key.begin(), key.end());
        int32_t next_node = static_cast<int32_t>(n0);
        for (size_t k = 0; k < key.size(); ++k) {
            const int32_t e = key[k].second;
            new_of_old_elem[static_cast<size_t>(e)] = static_cast<int32_t>(e0 + static_cast<int64_t>(k));
            for (int a = 0; a < npe; ++a) {
                const int64_t n = conn[npe * static_cast<int64_t>(e) + a];
                if (new_of_old_node[static_cast<size_t>(n)] < 0)
                    new_of_old_node[static_cast<size_t>(n)] = next_node++;
            }
        }
        // nodes referenced by no element keep their relative order at the end
        for (int64_t n = n0; n < n1; ++n)
            if (new_of_old_node[static_cast<size_t>(n)] < 0)
                new_of_old_node[static_cast<size_t>(n)] = next_node++;
    }
}

void corner_csr(int64_t N, int64_t E, const int64_t* conn, std::vector<int64_t>& start,
                std::vector<int64_t>& corners, int npe)
{
    const int64_t C = npe * E;
    start.assign(static_cast<size_t>(N + 1), 0);
    for (int64_t c = 0; c < C; ++c) start[static_cast<size_t>(conn[c] + 1)]++;
    for (int64_t n = 0; n < N; ++n) start[static_cast<size_t>(n + 1)] += start[static_cast<size_t>(n)];
    std::vector<int64_t> fill(start.begin(), start.end() - 1);
    corners.resize(static_cast<size_t>(C));
    // corners are visited in ascending id = ascending element, then node_lid
    for (int64_t c = 0; c < C; ++c) corners[static_cast<size_t>(fill[static_cast<size_t>(conn[c])]++)] = c;
}

void block_table(const std::vector<int64_t>& off, std::vector<BlockItem>& table,
                 std::vector<int32_t>& inst_begin, int32_t bs, int32_t deal)
{
    std::vector<BlockItem> logical;
    const int32_t ninst = static_cast<int32_t>(off.size()) - 1;
    inst_begin.assign(static_cast<size_t>(ninst + 1), 0);
    for (int32_t i = 0; i < ninst; ++i) {
        inst_begin[static_cast<size_t>(i)] = static_cast<int32_t>(logical.size());
        for (int64_t s = off[i]; s < off[i + 1]; s += bs) {
            BlockItem b;
            b.inst = i;
            b.start = static_cast<int32_t>(s);
            b.end = static_cast<int32_t>(std::min<int64_t>(s + bs, off[i + 1]));
            b.lid = static_cast<int32_t>(logical.size());
            logical.push_back(b);
        }
    }
    inst_begin[static_cast<size_t>(ninst)] = static_cast<int32_t>(logical.size());
    const int64_t nb = static_cast<int64_t>(logical.size());
    if (deal == 0) {
        table = logical;
        return;
    }
    table.resize(logical.size());
    // Chunks pay when every XCD works through many rounds of resident
    // workgroups (its tail then ends with the others'); a table of fewer
    // than 512 work items per XCD runs in a round or two, where one
    // contiguous range per XCD keeps more of each block's neighbours in the
    // same L2 (the 24,576-element anvil: 23.5 against 25.7 ms for 400 cycles).
    if (deal >= 2 && nb < kMinChunkedBlocks) deal = 1;
    if (deal == 1) {
        // one contiguous range of logical blocks per XCD
        int64_t base[kNumXcd + 1];
        base[0] = 0;
        for (int x = 0; x < kNumXcd; ++x)
            base[x + 1] = base[x] + nb / kNumXcd + (x < nb % kNumXcd ? 1 : 0);
        for (int64_t b = 0; b < nb; ++b) {
            const int x = static_cast<int>(b % kNumXcd);
            table[static_cast<size_t>(b)] = logical[static_cast<size_t>(base[x] + b / kNumXcd)];
        }
        return;
    }
    // chunks of `deal` consecutive logical blocks dealt round-robin to the
    // XCDs (chunk c to XCD c mod 8): neighbouring blocks share an L2 and
    // every XCD sweeps the whole mesh, so their tails end together; at
    // least 4 chunks per XCD, so no XCD idles.  The physical order
    // interleaves the XCDs' queues (workgroup p is dispatched to XCD p mod
    // 8); queues that run out are skipped.
    const int64_t k = std::max<int64_t>(1, std::min<int64_t>(deal, nb / (4 * kNumXcd)));
    std::vector<std::vector<int64_t>> q(kNumXcd);
    for (int64_t l = 0; l < nb; ++l) q[static_cast<size_t>((l / k) % kNumXcd)].push_back(l);
    std::vector<size_t> pos(kNumXcd, 0);
    int64_t p = 0;
    while (p < nb)
        for (int x = 0; x < kNumXcd && p < nb; ++x)
            if (pos[x] < q[x].size()) table[static_cast<size_t>(p++)] = logical[static_cast<size_t>(q[x][pos[x]++])];
}

int boundary_sets(const Layout& L, int32_t inst, const int64_t* conn, const double* xyz,
                  const std::vector<int64_t>& cstart, const std::vector<int64_t>& corners,
                  int32_t nsets, const fsgh_boundary* bc, std::vector<std::vector<int64_t>>& out,
                  int npe)
{
    // patches: the 6 four-node faces of a hex, or the 4 two-node edges of a quad
    const int nface = npe == 4 ? 4 : 6, fnodes = npe == 4 ? 2 : 4;
    const int(*face)[4] = npe == 4 ? kQuadEdge : kFace;
    const int64_t e0 = L.elem_off[inst], e1 = L.elem_off[inst + 1];
    const int64_t n0 = L.node_off[inst], n1 = L.node_off[inst + 1];
    out.assign(static_cast<size_t>(nsets), {});
    std::vector<char> on(static_cast<size_t>(n1 - n0)), mark(static_cast<size_t>(n1 - n0));
    for (int32_t s = 0; s < nsets; ++s) {
        const int tag = bc[s].surface;
        const double val = bc[s].value;
        // check_bdy's per-node surface test (K/setup.cpp:514-556)
        for (int64_t n = n0; n < n1; ++n) {
            const double* c = xyz + 3 * n;
            bool ok = false;
            if (tag >= 0 && tag <= 2) ok = std::fabs(c[tag] - val) <= 1.0e-7;
            else if (tag == 3) ok = std::fabs(std::sqrt(c[0] * c[0] + c[1] * c[1]) - val) <= 1.0e-7;
            else if (tag == 4)
                ok = std::fabs(std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]) - val) <= 1.0e-7;
            on[static_cast<size_t>(n - n0)] = ok;
        }
        std::fill(mark.begin(), mark.end(), 0);
        for (int64_t e = e0; e < e1; ++e) {
            const int64_t* en = conn + npe * e;
            for (int f = 0; f < nface; ++f) {
                int64_t fn[4];
                bool all = true;
                for (int q = 0; q < fnodes; ++q) {
                    fn[q] = en[face[f][q]];
                    all = all && on[static_cast<size_t>(fn[q] - n0)];
                }
                if (!all) continue;
                // a face is on the mesh boundary when no other element holds all its nodes
                bool shared = false;
                for (int64_t q = cstart[static_cast<size_t>(fn[0])];
                     q < cstart[static_cast<size_t>(fn[0] + 1)] && !shared; ++q) {
                    const int64_t e2 = corners[static_cast<size_t>(q)] / npe;
                    if (e2 == e) continue;
                    int hit = 0;
                    for (int k = 1; k < fnodes; ++k)
                        for (int a = 0; a < npe; ++a)
                            if (conn[npe * e2 + a] == fn[k]) { ++hit; break; }
                    shared = hit == fnodes - 1;
                }
                if (shared) continue;
                // nodes in first-seen order over (element, face, face node), the
                // order build_boundry_node_sets stores them in (K/setup.cpp:591-642)
                for (int q = 0; q < fnodes; ++q)
                    if (!mark[static_cast<size_t>(fn[q] - n0)]) {
                        mark[static_cast<size_t>(fn[q] - n0)] = 1;
                        out[static_cast<size_t>(s)].push_back(fn[q]);
                    }
            }
        }
    }
    return 0;
}

}  // namespace fsgh
