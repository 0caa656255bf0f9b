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

    def inner(p, j, var):
        ks = sorted(k for (a, b, k) in S if a == p and b == j)
        terms = []
        for q, k in enumerate(ks):
            s = S[(p, j, k)]
            if q == 0:
                terms.append(f"{var}[{k}]" if s > 0 else f"-{var}[{k}]")
            else:
                terms = [f"({terms[0]} {'+' if s > 0 else '-'} {var}[{k}])"]
        return terms[0]

    def js(p):
        return sorted({b for (a, b, k) in S if a == p})

    def brow(p, A, Bv):
        acc = None
        for j in js(p):
            t = f"{A}[{j}] * {inner(p, j, Bv)}"
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
    for p in range(8):
        w(f"FSGH_HD inline void hex_brow_{p}(const double* x, const double* y, const double* z,")
        w("                                 double& b0, double& b1, double& b2) {")
        w("    const double twelfth = 1. / 12.;")
        w(f"    b0 = ({brow(p, 'y', 'z')}) * twelfth;")
        w(f"    b1 = ({brow(p, 'z', 'x')}) * twelfth;")
        w(f"    b2 = ({brow(p, 'x', 'y')}) * twelfth;")
        w("}")
        w("")
    w("FSGH_HD inline void hex_bmatrix(const double* x, const double* y, const double* z,")
    w("                                double* B) {")
    for p in range(8):
        w(f"    hex_brow_{p}(x, y, z, B[{3 * p}], B[{3 * p + 1}], B[{3 * p + 2}]);")
    w("}")
    w("")
    w("FSGH_HD inline double hex_volume(const double* x, const double* y, const double* z) {")
    w("    const double twelfth = 1. / 12.;")
    outer = None
    for i in VOL_OUTER:
        acc = None
        assert sorted(VOL_JORDER[i]) == js(i)
        for j in VOL_JORDER[i]:
            t = f"y[{j}] * {inner(i, j, 'z')}"
            acc = t if acc is None else f"({acc}) + {t}"
        t = f"x[{i}] * ({acc})"
        outer = t if outer is None else f"({outer})\n        + {t}"
    w(f"    return ({outer}) * twelfth;")
    w("}")
    print("\n".join(out))


if __name__ == "__main__":
    main()
