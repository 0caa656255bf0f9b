"""Generate hex_geometry.inc: exact trilinear-hex volume and corner normals.

The volume of a trilinear hexahedron is V = sum_{i,j,k} C_ijk x_i y_j z_k with
C_ijk = int_[0,1]^3 eps_abc dN_i/dxi_a dN_j/dxi_b dN_k/dxi_c, and the outward
corner area normals are B_p = dV/dx_p.  This script derives C exactly (rational
arithmetic over the trilinear shape functions, local node order
(0,0,0),(1,0,0),(1,1,0),(0,1,0),(0,0,1),(1,0,1),(1,1,1),(0,1,1)), checks that
12*C has entries in {-1,0,1}, and emits straight-line code.

Floating-point order.  The reference evaluates the same polynomials in a fixed
order (K/geometry.cpp:66-232 for B, :298-306 for V): for B(p,.) the outer sum
runs over j ascending and each inner sum over k ascending; the volume runs its
outer sum over x-nodes in the order VOL_OUTER with the y-node orders VOL_JORDER.
Reproducing that order (and compiling without FMA contraction) makes the
device results bit-identical to the reference arithmetic, which matters: at
h = 1.2/256 the volume polynomial cancels by ~1e5 and any other order differs
by ~1e-11 relative.

Run:  python gen_hex_geometry.py > hex_geometry.inc
"""
from fractions import Fraction
from itertools import product

NODES = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
# summation order of the reference's volume expression (K/geometry.cpp:299-306)
VOL_OUTER = [1, 7, 3, 5, 6, 0, 2, 4]
VOL_JORDER = {1: [3, 4, 0, 6, 5, 2], 7: [0, 6, 2, 3, 5, 4], 3: [1, 7, 6, 2, 4, 0],
              5: [0, 7, 2, 1, 4, 6], 6: [1, 7, 5, 4, 3, 2], 0: [2, 7, 5, 1, 3, 4],
              2: [0, 5, 1, 7, 6, 3], 4: [1, 7, 3, 0, 6, 5]}


def lin(s):
    """1-D factor of a shape function as polynomial coefficients [c0, c1]."""
    return [Fraction(0), Fraction(1)] if s == 1 else [Fraction(1), Fraction(-1)]


def dlin(s):
    return [Fraction(1)] if s == 1 else [Fraction(-1)]


def pmul(a, b):
    out = [Fraction(0)] * (len(a) + len(b) - 1)
    for i, x in enumerate(a):
        for j, y in enumerate(b):
            out[i + j] += x * y
    return out


def pint(a):
    """integral over [0,1]"""
    return sum(c / (i + 1) for i, c in enumerate(a))


def grad_factors(p):
    """dN_p/dxi_a as a tuple of three 1-D polynomials (one per coordinate)."""
    s = NODES[p]
    out = []
    for a in range(3):
        out.append(tuple(dlin(s[c]) if c == a else lin(s[c]) for c in range(3)))
    return out


def perm_sign(a, b, c):
    if len({a, b, c}) < 3:
        return 0
    return 1 if (a, b, c) in ((0, 1, 2), (1, 2, 0), (2, 0, 1)) else -1


def coeff_tensor():
    G = [grad_factors(p) for p in range(8)]
    Cc = {}
    for i, j, k in product(range(8), repeat=3):
        tot = Fraction(0)
        for a, b, c in product(range(3), repeat=3):
            sg = perm_sign(a, b, c)
            if sg == 0:
                continue
            val = Fraction(1)
            for d in range(3):
                val *= pint(pmul(pmul(G[i][a][d], G[j][b][d]), G[k][c][d]))
            tot += sg * val
        if tot:
            Cc[(i, j, k)] = tot
    return Cc


def main():
    Cc = coeff_tensor()
    S = {}
    for key, v in Cc.items():
        s = v * 12
        assert s.denominator == 1 and abs(s.numerator) == 1, (key, v)
        S[key] = int(s)
    # antisymmetry checks that a mistaken derivation would fail
    for (i, j, k), s in S.items():
        assert S.get((j, i, k), 0) == -s and S.get((i, k, j), 0) == -s

    def ks(p, j):
        return sorted(k for (a, b, k) in S if a == p and b == j)

    def inner(p, j, var):
        terms = None
        for q, k in enumerate(ks(p, j)):
            sg = S[(p, j, k)]
            if q == 0:
                terms = f"{var}[{k}]" if sg > 0 else f"-{var}[{k}]"
            else:
                terms = f"({terms} {'+' if sg > 0 else '-'} {var}[{k}])"
        return terms

    def js(p):
        return sorted({b for (a, b, k) in S if a == p})

    # Inner sums are antisymmetric: inner(j, p) evaluates the same terms with
    # every sign flipped, and IEEE rounding commutes with negation, so
    # inner(j, p) == -inner(p, j) bit for bit.  Each unordered pair is summed
    # once per component.
    COMP = [("y", "z"), ("z", "x"), ("x", "y")]

    def shared_inners(w, d):
        A, Bv = COMP[d]
        for p in range(8):
            for j in js(p):
                if p < j:
                    w(f"    const T i{d}_{p}{j} = {inner(p, j, Bv)};")

    def iref(d, p, j):
        return f"i{d}_{p}{j}" if p < j else f"(-i{d}_{j}{p})"

    def brow_shared(d, p):
        A, _ = COMP[d]
        acc = None
        for j in js(p):
            t = f"{A}[{j}] * {iref(d, p, j)}"
            acc = t if acc is None else f"({acc}) + {t}"
        return acc

    out = []
    w = out.append
    w("// Generated by gen_hex_geometry.py -- do not edit.")
    w("// Exact trilinear-hex volume and outward corner area normals B_p = dV/dx_p,")
    w("// evaluated in the reference's summation order (K/geometry.cpp:40-310).")
    w("// Requires -ffp-contract=off for bit-identical results.")
    w("#pragma once")
    w("#ifndef FSGH_HD")
    w("#define FSGH_HD")
    w("#endif")
    w("")
    w("template <typename T>")
    w("FSGH_HD inline void hex_bmatrix(const T* x, const T* y, const T* z, T* B) {")
    w("    const double twelfth = 1. / 12.;")
    for d in range(3):
        shared_inners(w, d)
        for p in range(8):
            w(f"    B[{3 * p + d}] = ({brow_shared(d, p)}) * twelfth;")
    w("}")
    w("")
    # volume: x_i * (sum_j y_j * inner(i, j, z)) in the reference's order; the
    # products y_j * inner(i, j, z) are the ones B(i, 0) sums.
    def vol_expr(prod):
        outer = None
        for i in VOL_OUTER:
            acc = None
            assert sorted(VOL_JORDER[i]) == js(i)
            for j in VOL_JORDER[i]:
                t = prod(i, j)
                acc = t if acc is None else f"({acc}) + {t}"
            t = f"x[{i}] * ({acc})"
            outer = t if outer is None else f"({outer})\n        + {t}"
        return outer

    w("template <typename T>")
    w("FSGH_HD inline T hex_volume(const T* x, const T* y, const T* z) {")
    w("    const double twelfth = 1. / 12.;")
    shared_inners(w, 0)
    w(f"    return ({vol_expr(lambda i, j: f'y[{j}] * {iref(0, i, j)}')}) * twelfth;")
    w("}")
    w("")
    w("// B and V together: V reuses the products of B's x-component rows.")
    w("template <typename T>")
    w("FSGH_HD inline T hex_bmatrix_volume(const T* x, const T* y, const T* z, T* B) {")
    w("    const double twelfth = 1. / 12.;")
    shared_inners(w, 0)
    for p in range(8):
        for j in js(p):
            w(f"    const T t_{p}{j} = y[{j}] * {iref(0, p, j)};")
    for p in range(8):
        acc = None
        for j in js(p):
            acc = f"t_{p}{j}" if acc is None else f"({acc}) + t_{p}{j}"
        w(f"    B[{3 * p}] = ({acc}) * twelfth;")
    for d in (1, 2):
        shared_inners(w, d)
        for p in range(8):
            w(f"    B[{3 * p + d}] = ({brow_shared(d, p)}) * twelfth;")
    w(f"    return ({vol_expr(lambda i, j: f't_{i}{j}')}) * twelfth;")
    w("}")
    w("")
    # reverse mode: B is bilinear, B(p,d) = 1/12 sum_j A_j sum_k s_pjk C_k with
    # (A, C) = (y, z), (z, x), (x, y) for d = 0, 1, 2.
    w("// Reverse mode of hex_bmatrix: accumulates (dB . Bb)/dx into xb, yb, zb.")
    w("FSGH_HD inline void hex_bmatrix_vjp(const double* x, const double* y, const double* z,")
    w("                                    const double* Bb, double* xb, double* yb, double* zb) {")
    w("    const double twelfth = 1. / 12.;")
    for d in range(3):
        A, Bv = COMP[d]
        for p in range(8):
            w(f"    {{ const double g = Bb[{3 * p + d}] * twelfth;")
            for j in js(p):
                w(f"      {A}b[{j}] += g * {inner(p, j, Bv)};")
                for k in ks(p, j):
                    sg = S[(p, j, k)]
                    w(f"      {Bv}b[{k}] {'+' if sg > 0 else '-'}= g * {A}[{j}];")
            w("    }")
    w("}")
    print("\n".join(out))


if __name__ == "__main__":
    main()
