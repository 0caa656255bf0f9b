"""Algorithmic HBM bytes of the SGH kernels (unique bytes a launch must move).

Counting rule: every array element a kernel must read or write is counted
once (FP64 = 8 B, int32 index = 4 B, 16-bit index = 2 B); values recomputed
on chip or re-read from cache count 0; index arrays count as the kernel
reads them (the LDS-staged kernels read local indices and per-block lists in
place of global connectivity).  The ideal-gas, strength-free path reads no stress; with
one material the material id is not read.  E elements, N nodes, C = 8E.
"""
from __future__ import annotations


def div_extremes(E, N):
    # corner CSR (4C + 4(N+1)) + div 8E + two (min, 2nd min)/(max, 2nd max) pairs 32N (w)
    return (32 * E + 4 * (N + 1)) + 8 * E + 32 * N


def force(E, N, list_entries=None):
    # connectivity + x1,v1 48N + vol,div,den,sspd,pres 40E + the divergence-extreme
    # pairs 16N + F 24C (w).  Connectivity: 16-bit local connectivity 16E + the
    # blocks' node lists 4 x list_entries when the nodes are staged in LDS (both
    # extreme pairs are then staged: 32N), else conn 32E.
    if list_entries is None:
        return 32 * E + 48 * N + 40 * E + 16 * N + 192 * E
    return 16 * E + 4 * list_entries + 48 * N + 40 * E + 32 * N + 192 * E


def vel_bc_pos(E, N, corner_staged=True):
    # F 24C + corner CSR (starts 4(N+1) + the blocks' sorted corner lists 4C)
    # + the entries' positions in them 2C when the block's corner forces are
    # staged in LDS + m 8N + v0 24N + v1 24N (w) + x0 24N + x1 24N (w) + BC mask 1N
    return (192 * E + (32 * E + 4 * (N + 1)) + (16 * E if corner_staged else 0) + 8 * N
            + 96 * N + N)


def energy_state_div(E, N, list_entries=None):
    # F 24C + connectivity + v1,v0,x1 72N + M, e0 16E + e1,V,rho,p,c,div 48E (w)
    # + gamma,csmin 16E.  Connectivity: with the blocks' nodes staged in LDS,
    # 16-bit local connectivity 16E + the blocks' node lists 4 x list_entries;
    # otherwise conn 32E.
    conn = 32 * E if list_entries is None else 16 * E + 4 * list_entries
    return 192 * E + conn + 72 * N + 16 * E + 48 * E + 16 * E


KERNELS = {"div_extremes": div_extremes, "force": force, "vel_bc_pos": vel_bc_pos,
           "energy_state_div": energy_state_div}


# ---- reverse pass of one RK stage (fsgh_adjoint.hip), accumulating form ----
# (the first stage reversed in a cycle writes some accumulators instead of
# adding to them; counted here as read + write throughout).  Full work: every
# element and node as if its adjoint were nonzero.  The kernels skip what the
# reverse of a zero adjoint leaves at zero (k_adj_energy per element, the
# element passes per wave, k_adj_node's corner forces per block,
# k_adj_gather's absent rows and row-less blocks), so where much of the mesh
# is not yet reached by the adjoint they read and write less than these
# figures.
def adj_energy(E, N):
    # e-bar 8E + e0-bar 16E (rw) + M 8E + g 8E (w) + conn 32E + F 24C + M-bar 16E (rw)
    # + v0, v1 48N
    return 8 * E + 16 * E + 8 * E + 8 * E + 32 * E + 192 * E + 16 * E + 48 * N


def adj_node(E, N):
    # corner CSR starts 4(N+1) + sorted corner list 4C + positions 2C + F 24C + g 8E
    # + m 8N + BC mask 1N + v0, v1 48N + x-bar, v-bar of the stage output 48N
    # + x0-bar, v0-bar 96N (rw) + m-bar 16N (rw) + the (h, kv) records 48N (w)
    # + kv flags 1N (w)
    return (4 * (N + 1) + 32 * E + 16 * E + 192 * E + 8 * E + 8 * N + N + 48 * N + 48 * N
            + 96 * N + 16 * N + 48 * N + N)


def arg_extremes(E, N):
    # corner CSR (4C + 4(N+1)) + div 8E + the extreme pairs 32N and their
    # elements 16N (w)
    return 32 * E + 4 * (N + 1) + 8 * E + 48 * N


def adj_ext(E, N):
    # conn 32E + x, v 48N + vol, div, den, sspd, pres 40E + extreme pairs and their
    # elements 48N + g 8E + (h, kv) records 48N + kv flags 1N + ext-bar 8E (w)
    # + chosen extreme 4E (w)
    return 32 * E + 48 * N + 40 * E + 48 * N + 8 * E + 48 * N + N + 8 * E + 4 * E


def adj_ext_node(E, N):
    # corner CSR (4C + 4(N+1)) + chosen extreme 4E + ext-bar 8E + per-slot sums 32N (w)
    # + slot-sum flags 1N (w)
    return 32 * E + 4 * (N + 1) + 4 * E + 8 * E + 32 * N + N


def adj_force_derived(E, N):
    # conn 32E + x, v 48N + vol, div, den, sspd, pres 40E + extreme pairs and their
    # elements 48N + g 8E + (h, kv) records 48N + per-slot sums 32N + kv and
    # slot-sum flags 2N + gamma, csmin 16E + e 8E + e-bar 16E (rw) + gamma-bar 16E (rw)
    # + M-bar (or rho0-bar) 16E (rw) + corner velocity rows 24C (w) + corner position
    # rows 24C (w) + row flag 1E (w)
    return (32 * E + 48 * N + 40 * E + 48 * N + 8 * E + 48 * N + 32 * N + 2 * N + 16 * E
            + 8 * E + 16 * E + 16 * E + 16 * E + 192 * E + 192 * E + E)


def adj_gather(E, N):
    # corner CSR starts 4(N+1) + sorted corner list 4C + positions 2C + row flags
    # 1C + corner position and velocity rows 48C + x-bar, v-bar 96N (rw)
    return 4 * (N + 1) + 32 * E + 16 * E + 8 * E + 384 * E + 96 * N


ADJ_KERNELS = {"adj_energy": adj_energy, "adj_node": adj_node, "arg_extremes": arg_extremes,
               "adj_ext": adj_ext, "adj_ext_node": adj_ext_node,
               "adj_force_derived": adj_force_derived, "adj_gather": adj_gather}


def adj_stage(E, N):
    """Bytes of the reverse of one RK stage (every reverse kernel once)."""
    return sum(f(E, N) for f in ADJ_KERNELS.values())


# ---- 2D axisymmetric (RZ) path: quads, 4 corners per element, 2 components ----
def rz_div_extremes(E, N):
    # corner CSR (16E + 4(N+1)) + div 8E + the two extreme pairs 32N (w)
    return (16 * E + 4 * (N + 1)) + 8 * E + 32 * N


def rz_force(E, N, list_entries=None):
    # connectivity + x,y,u,v 32N + den,pres,sspd,div 32E + the divergence-extreme
    # pairs + F 64E (w).  With the blocks' nodes staged in LDS: 16-bit local
    # connectivity 8E + the node lists 4 x list_entries, and both extreme pairs
    # are staged (32N); otherwise conn 16E and one pair per node (16N).
    if list_entries is None:
        return 16 * E + 32 * N + 32 * E + 16 * N + 64 * E
    return 8 * E + 4 * list_entries + 32 * N + 32 * E + 32 * N + 64 * E


def rz_node(E, N):
    # F 64E + corner CSR + m 8N (r) + m 8N (w) + BC mask 1N + u0,v0,x0,y0 32N
    # + stage-input radius 8N + u1,v1,x1,y1 32N (w) + radius copy 8N (w) + extensive mass 8N
    return 64 * E + (16 * E + 4 * (N + 1)) + 16 * N + N + 32 * N + 8 * N + 32 * N + 8 * N + 8 * N


def rz_energy_state_div(E, N, list_entries=None):
    # F 64E + connectivity + u1,v1,u0,v0 32N + radius 8N + x1,y1 16N
    # + M,e0,gamma,csmin 32E + e1,V,rho,p,c,div 48E (w).  Connectivity: conn 16E,
    # or with the blocks' nodes staged in LDS the 16-bit local connectivity 8E
    # + the node lists 4 x list_entries.
    conn = 16 * E if list_entries is None else 8 * E + 4 * list_entries
    return 64 * E + conn + 32 * N + 8 * N + 16 * N + 32 * E + 48 * E


RZ_KERNELS = {"rz_div_extremes": rz_div_extremes, "rz_force": rz_force, "rz_node": rz_node,
              "rz_energy_state_div": rz_energy_state_div}


def cycle(E, N, lists=None, corner_staged=True):
    """Bytes of one fused RK2 cycle (2 stages; dt and control kernels are O(1)
    per block).  lists: node-list entries per element kernel
    ({"force": n, "energy_state_div": n}; absent = global gathers)."""
    lists = lists or {}
    return 2 * (div_extremes(E, N) + force(E, N, lists.get("force"))
                + vel_bc_pos(E, N, corner_staged)
                + energy_state_div(E, N, lists.get("energy_state_div")))


def kernel_bytes(name, E, N, list_entries=None, corner_staged=True):
    """Algorithmic bytes of one launch of a fused 3D kernel (list_entries:
    the node-list entries of that kernel's element partition, if staged)."""
    if name == "vel_bc_pos":
        return vel_bc_pos(E, N, corner_staged)
    if name == "energy_state_div":
        return energy_state_div(E, N, list_entries)
    if name == "force":
        return force(E, N, list_entries)
    return KERNELS[name](E, N)
