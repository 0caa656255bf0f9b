"""Calibration of point blasts by the discrete adjoint.

Objective of one instance: the L2 misfit of its final radial density
profile against a target profile,

    J = sum_k (rho_k - rho*_k)^2,  rho_k = mean density of the elements whose
    initial centroid radius lies in bin k (nbins bins over [0, rmax]).

Bin membership is fixed by the initial (Lagrangian) centroid, so J is a
smooth function of the final state (density M / V of the final mesh).
Controls per instance: the deposited energy E0 (origin element energy
e = E0 / (rho0 V0)), gamma (every element) and the uniform initial density
rho0 (element masses, node masses, pressure, sound speed and the origin
element's energy).  All instances run in one batched context: one
checkpointed forward and one reverse pass give every instance's gradient.
"""
from __future__ import annotations

import numpy as np

NBINS = 12
RMAX = 0.3


def centroids(coords, conn):
    """Element centroids (mean of the nodes), summed node by node so that no
    E x nodes x 3 temporary is formed."""
    c = np.zeros((conn.shape[0], coords.shape[1]))
    for a in range(conn.shape[1]):
        c += coords[conn[:, a]]
    return c / conn.shape[1]


def bin_index(coords, conn, nbins=NBINS, rmax=RMAX):
    """Radial bin of each element by initial centroid radius (-1 beyond rmax)."""
    r = np.linalg.norm(centroids(coords, conn), axis=1)
    k = np.floor(r / rmax * nbins).astype(np.int64)
    k[k >= nbins] = -1
    return k


def profile(den, bins, nbins=NBINS):
    """Mean density per bin (0 for an empty bin)."""
    ok = bins >= 0
    s = np.bincount(bins[ok], weights=den[ok], minlength=nbins)
    c = np.bincount(bins[ok], minlength=nbins)
    return np.where(c > 0, s / np.maximum(c, 1), 0.0), c


def misfit(den, bins, target, nbins=NBINS):
    """J and dJ/d(den) for one instance."""
    prof, cnt = profile(den, bins, nbins)
    r = prof - target
    seed = np.zeros_like(den)
    ok = bins >= 0
    seed[ok] = 2.0 * r[bins[ok]] / cnt[bins[ok]]
    return float(np.sum(r * r)), seed


def deposit_sie(spec):
    """Specific internal energy of the energy deposit: the blast's last fill,
    a sphere around the origin holding only the origin element."""
    from fierro_amd import problems as P

    f = spec.fills[-1]
    if f["volume"] != P.SPHERE:
        raise ValueError(f"{spec.name}: the last fill is not the energy deposit")
    return float(f["sie"])


def origin_element(coords, conn):
    """The element holding the deposited energy (centroid nearest the origin)."""
    return int(np.argmin(np.linalg.norm(centroids(coords, conn), axis=1)))


def target_profile(Solver, spec, cycles, nbins=NBINS, rmax=RMAX):
    """The target's final radial density profile (fused loop, same cycles)."""
    g = Solver(spec)
    g.sgh_solve(cycles=cycles)
    prof, _ = profile(g.download("elem_den"), bin_index(g.coords0, g.conn, nbins, rmax), nbins)
    g.close()
    return prof


class Layout:
    """What the objective needs of a solver's initial mesh, computed once:
    per instance its element range, radial bins and origin element."""

    def __init__(self, g, nbins=NBINS, rmax=RMAX):
        self.inst = []
        for i in range(g.ninst):
            e0, e1 = int(g.elem_off[i]), int(g.elem_off[i + 1])
            n0, n1 = int(g.node_off[i]), int(g.node_off[i + 1])
            x0 = g.coords0[n0:n1]
            conn = g.conn[e0:e1] - n0
            c = centroids(x0, conn)
            r = np.linalg.norm(c, axis=1)
            k = np.floor(r / rmax * nbins).astype(np.int64)
            k[k >= nbins] = -1
            self.inst.append((e0, e1, k, int(np.argmin(r))))


def objective(g, target, cycles, dt_replay=None):
    """J of every instance of solver g after ``cycles`` cycles from the set-up
    state (``dt_replay`` [cycles][instances] freezes the steps)."""
    g.forward(cycles, checkpoint_every=cycles, dt_replay=dt_replay)
    den = g.download("elem_den")
    J = np.zeros(g.ninst)
    for i in range(g.ninst):
        e0, e1 = g.elem_off[i], g.elem_off[i + 1]
        n0 = g.node_off[i]
        bins = bin_index(g.coords0[n0:g.node_off[i + 1]], g.conn[e0:e1] - n0)
        J[i], _ = misfit(den[e0:e1], bins, target)
    return J


def gradient(g, params, target, cycles, checkpoint_every=10, dt_replay=None, layout=None):
    """One batched forward + reverse pass over every instance of solver g,
    which must hold its set-up state (a new solver, or one set up again).
    ``layout`` (a ``Layout`` of g, built once) holds the bins and origin
    elements; without it they are computed here.

    Returns (J [inst], dJ/dE0 [inst], dJ/dgamma [inst], dJ/drho0 [inst]) and
    the recorded time steps."""
    lay = layout if layout is not None else Layout(g)
    g.forward(cycles, checkpoint_every=checkpoint_every, dt_replay=dt_replay)
    den = g.download("elem_den")
    seed = np.zeros(g.E)
    J = np.zeros(g.ninst)
    for i, (e0, e1, bins, _) in enumerate(lay.inst):
        J[i], seed[e0:e1] = misfit(den[e0:e1], bins, target)
    gr = g.backward(seed_den=seed)
    dE0, dgam, drho = np.zeros(g.ninst), np.zeros(g.ninst), np.zeros(g.ninst)
    for i, (e0, e1, _, o) in enumerate(lay.inst):
        n, E0, gamma, rho0 = params[i]
        # the deposit is e = E0 / (rho0 V) for whatever volume V the fill used:
        # de/dE0 = e / E0 and de/drho0 = -e / rho0
        sie_o = deposit_sie(g.specs[i])
        g_sie_o = gr["sie0"][e0 + o]
        dE0[i] = g_sie_o * sie_o / E0
        dgam[i] = float(np.sum(gr["gamma"][e0:e1]))
        drho[i] = float(np.sum(gr["den0"][e0:e1])) - g_sie_o * sie_o / rho0
    return J, dE0, dgam, drho, g.dt_history()
