// fsgh_mesh.h -- host-side mesh preparation for the SGH library.
#pragma once

#include <cstdint>
#include <vector>

#include "fsgh_internal.h"

namespace fsgh {

// Ragged instance layout: instance i owns [off[i], off[i+1]).
struct Layout {
    int32_t ninst = 1;
    std::vector<int64_t> node_off, elem_off;
};

#ifndef FSGH_HD
#define FSGH_HD __host__ __device__
#endif

// interleave the low 21 bits of v with two zero bits between each
FSGH_HD inline uint64_t morton_spread3(uint32_t v)
{
    uint64_t x = v & 0x1fffffu;
    x = (x | x << 32) & 0x1f00000000ffffULL;
    x = (x | x << 16) & 0x1f0000ff0000ffULL;
    x = (x | x << 8) & 0x100f00f00f00f00fULL;
    x = (x | x << 4) & 0x10c30c30c30c30c3ULL;
    x = (x | x << 2) & 0x1249249249249249ULL;
    return x;
}

// a centroid coordinate in [lo, hi] quantised to 21 bits
FSGH_HD inline uint32_t morton_quantize(double c, double lo, double hi)
{
    if (!(hi > lo)) return 0;
    double t = (c - lo) / (hi - lo);
    t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
    return static_cast<uint32_t>(t * 2097151.0);
}

// Morton renumbering inside each instance.  new_of_old_elem / new_of_old_node
// map caller ids to internal ids; old_of_new_* are their inverses.
// npe = nodes per element (8 hexes, 4 RZ quads); coords are N x 3 (z = 0 in 2D)
void morton_renumber(const Layout& L, const int64_t* nodes_in_elem, const double* coords,
                     std::vector<int32_t>& new_of_old_elem, std::vector<int32_t>& new_of_old_node,
                     int npe = 8);

// corners_in_node of the caller's mesh, each node's list in ascending element
// order (K/mesh.h:218-241), as CSR over caller node ids.
void corner_csr(int64_t N, int64_t E, const int64_t* nodes_in_elem, std::vector<int64_t>& start,
                std::vector<int64_t>& corners, int npe = 8);

// Work items of `bs` entries per instance, in a physical order for the
// hardware's round-robin dealing of workgroups to XCDs (physical block p runs
// on XCD p % 8): deal = 0 keeps the logical order; 1 gives each XCD one
// contiguous range of logical blocks; k >= 2 deals chunks of k consecutive
// logical blocks round-robin to the XCDs when the table has at least
// kMinChunkedBlocks blocks (fewer per chunk when it has fewer than 32 k
// blocks, so that every XCD gets at least 4 chunks), and is 1 otherwise.
constexpr int64_t kMinChunkedBlocks = 4096;
void block_table(const std::vector<int64_t>& off, std::vector<BlockItem>& table,
                 std::vector<int32_t>& inst_begin, int32_t bs = kBlock, int32_t deal = 1);

// Boundary node sets of one instance (K/setup.cpp:431-665): nodes of the
// boundary faces whose 4 nodes all satisfy the surface test.  Each set holds
// its nodes in the reference's first-seen order: boundary faces by ascending
// caller element id, then face, then the face's node order (K/mesh.h:353-376,
// K/setup.cpp:591-642).
int boundary_sets(const Layout& L, int32_t inst, const int64_t* nodes_in_elem,
                  const double* coords, const std::vector<int64_t>& cstart,
                  const std::vector<int64_t>& corners, int32_t nsets, const fsgh_boundary* bc,
                  std::vector<std::vector<int64_t>>& out, int npe = 8);

}  // namespace fsgh
