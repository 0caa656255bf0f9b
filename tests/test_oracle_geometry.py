"""Hex geometry pins: the oracle's get_vol_hex / get_bmatrix restatement and
the library's generated hex_geometry.inc, checked against exact rational
arithmetic (independent of either formula) and against each other bitwise.

Exact volume: det J of a trilinear hex has degree <= 2 in each reference
coordinate, so the 3x3x3 tensor Simpson rule integrates it exactly; with
Fraction arithmetic on the (exactly representable) double inputs this gives
V exactly.  V is linear in every single coordinate, so B_pd = dV/dx_pd =
V(x + e_pd) - V(x) exactly.
"""
import ctypes as C
import os
import subprocess
from fractions import Fraction
from itertools import product

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NODES = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
SIMPSON = [(Fraction(0), Fraction(1, 6)), (Fraction(1, 2), Fraction(4, 6)),
           (Fraction(1), Fraction(1, 6))]


def _grad(p, xi):
    s = NODES[p]
    f = [xi[c] if s[c] else 1 - xi[c] for c in range(3)]
    d = [1 if s[c] else -1 for c in range(3)]
    return [d[0] * f[1] * f[2], f[0] * d[1] * f[2], f[0] * f[1] * d[2]]


def exact_volume(X):
    """X: 8 x 3 list of Fractions."""
    V = Fraction(0)
    for (a, wa), (b, wb), (c, wc) in product(SIMPSON, repeat=3):
        xi = (a, b, c)
        J = [[Fraction(0)] * 3 for _ in range(3)]
        for p in range(8):
            g = _grad(p, xi)
            for i in range(3):
                for k in range(3):
                    J[i][k] += X[p][i] * g[k]
        det = (J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1])
               - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0])
               + J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]))
        V += wa * wb * wc * det
    return V


def exact_B(X):
    V0 = exact_volume(X)
    B = [[Fraction(0)] * 3 for _ in range(8)]
    for p in range(8):
        for d in range(3):
            Y = [row[:] for row in X]
            Y[p][d] += 1
            B[p][d] = exact_volume(Y) - V0
    return V0, B


def random_hex(rng, h=1.0, jitter=0.2, offset=(0.0, 0.0, 0.0)):
    base = np.array(NODES, float) * h + np.array(offset)
    return base + rng.uniform(-jitter * h, jitter * h, size=(8, 3))


@pytest.fixture(scope="module")
def hexes():
    rng = np.random.default_rng(7)
    hs = [random_hex(rng) for _ in range(3)]
    hs.append(random_hex(rng, h=1.2 / 256, jitter=0.05, offset=(0.9, 0.3, 1.1)))
    hs.append(np.array(NODES, float) * 0.5 + 0.25)
    return hs


def test_bmatrix_and_volume_exact(oracle, hexes):
    for X in hexes:
        XF = [[Fraction(float(v)) for v in row] for row in X]
        V, B = exact_B(XF)
        Bx = np.array([[float(b) for b in row] for row in B])
        Bo = oracle.bmatrix(X[:, 0], X[:, 1], X[:, 2])
        Vo = oracle.vol_hex(X[:, 0], X[:, 1], X[:, 2])
        # cancellation grows as (|x|/h)^2 for an offset small cell
        cond = (np.abs(X).max() / np.ptp(X, axis=0).max()) ** 2
        assert abs(Vo - float(V)) <= 1e-14 * cond * abs(float(V))
        assert np.max(np.abs(Bo - Bx)) <= 1e-14 * cond * np.max(np.abs(Bx))


def test_bmatrix_identities(oracle):
    rng = np.random.default_rng(3)
    for _ in range(20):
        X = random_hex(rng)
        B = oracle.bmatrix(X[:, 0], X[:, 1], X[:, 2])
        V = oracle.vol_hex(X[:, 0], X[:, 1], X[:, 2])
        assert np.allclose(B.sum(0), 0.0, atol=1e-14)
        # Jacobi / Euler: sum_p x_p (x) B_p = V I
        assert np.allclose(X.T @ B, V * np.eye(3), atol=1e-13)
        # affine map: V(Ax+b) = det(A) V(x)
        A = np.eye(3) + 0.1 * rng.standard_normal((3, 3))
        Y = X @ A.T + rng.standard_normal(3)
        assert np.isclose(oracle.vol_hex(Y[:, 0], Y[:, 1], Y[:, 2]), np.linalg.det(A) * V,
                          rtol=1e-13)
    # a cube of side h: B_p = +-h^2/4 per component, outward
    h = 0.3
    X = np.array(NODES, float) * h
    B = oracle.bmatrix(X[:, 0], X[:, 1], X[:, 2])
    assert np.allclose(B, (2 * np.array(NODES) - 1) * h * h / 4, rtol=1e-15, atol=0)
    assert oracle.vol_hex(X[:, 0], X[:, 1], X[:, 2]) == pytest.approx(h ** 3, rel=1e-15)


_SRC = r"""
#define FSGH_HD
#include "hex_geometry.inc"
extern "C" void geo(const double* x, const double* y, const double* z, double* B, double* V) {
    hex_bmatrix(x, y, z, B);
    V[0] = hex_volume(x, y, z);
    V[1] = hex_bmatrix_volume(x, y, z, B + 24);
}
extern "C" void geo_vjp(const double* x, const double* y, const double* z, const double* Bb,
                        double* xb, double* yb, double* zb) {
    hex_bmatrix_vjp(x, y, z, Bb, xb, yb, zb);
}
"""


@pytest.fixture(scope="module")
def generated(tmp_path_factory):
    d = tmp_path_factory.mktemp("geo")
    src = d / "geo.cpp"
    src.write_text(_SRC)
    so = d / "libgeo.so"
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-shared", "-fPIC",
                    "-I", os.path.join(ROOT, "fierro_amd", "csrc"), str(src), "-o", str(so)],
                   check=True)
    L = C.CDLL(str(so))
    return L


def test_generated_geometry_bitwise_equals_oracle(oracle, generated):
    """hex_geometry.inc (derived from the shape functions, used by the HIP
    kernels) evaluates in the same order as the oracle: identical bits."""
    rng = np.random.default_rng(11)
    f64 = np.ctypeslib.ndpointer(np.float64)
    generated.geo.argtypes = [f64, f64, f64, f64, f64]
    for k in range(200):
        X = random_hex(rng, h=10.0 ** rng.uniform(-3, 0), offset=rng.uniform(0, 2, 3))
        x, y, z = (np.ascontiguousarray(X[:, d]) for d in range(3))
        B = np.zeros(48)
        V = np.zeros(2)
        generated.geo(x, y, z, B, V)
        Bo = oracle.bmatrix(x, y, z)
        assert np.array_equal(B[:24].reshape(8, 3), Bo), k
        assert np.array_equal(B[24:].reshape(8, 3), Bo), k
        Vo = oracle.vol_hex(x, y, z)
        assert V[0] == Vo and V[1] == Vo, k


def test_bmatrix_vjp_dot_product(generated):
    """<Bb, dB(x)[dx]> == <vjp(Bb), dx>.  B is a quadratic polynomial of the
    coordinates, so the central difference gives dB[dx] exactly up to rounding."""
    rng = np.random.default_rng(13)
    f64 = np.ctypeslib.ndpointer(np.float64)
    generated.geo.argtypes = [f64, f64, f64, f64, f64]
    generated.geo_vjp.argtypes = [f64] * 7

    def B_of(X):
        B = np.zeros(48)
        generated.geo(*(np.ascontiguousarray(X[:, d]) for d in range(3)), B, np.zeros(2))
        return B[:24]

    for _ in range(10):
        X = random_hex(rng)
        dX = rng.standard_normal((8, 3))
        Bb = rng.standard_normal(24)
        t = 1e-3
        dB = (B_of(X + t * dX) - B_of(X - t * dX)) / (2 * t)
        xb, yb, zb = np.zeros(8), np.zeros(8), np.zeros(8)
        generated.geo_vjp(*(np.ascontiguousarray(X[:, d]) for d in range(3)), Bb, xb, yb, zb)
        lhs = float(Bb @ dB)
        rhs = float(np.stack([xb, yb, zb], 1).ravel() @ dX.ravel())
        assert abs(lhs - rhs) <= 1e-12 * (abs(lhs) + 1)
