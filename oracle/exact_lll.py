"""Exact lowest-Landau-level matrix elements for tiny droplets (test infrastructure only).

For N <= 4 the edge states p_lambda * Psi_L, Psi_L = prod_{i<j} (z_i - z_j)^m, are expanded
into monomials z^a = prod_i z_i^{a_i}.  With the Gaussian weight exp(-sum|z|^2/2),
monomials are orthogonal and

    ||z^a||^2 / (2 pi)^N = prod_i 2^{a_i} a_i!,

and a radial one-body potential is diagonal in them:

    <r^2>_a = 2 (a + 1),   <r^4>_a = 4 (a + 1)(a + 2),   <r^{2q}>_a = 2^q (a+q)!/a!.

So S = (p_l Psi | p_m Psi)/(Psi|Psi) and H = (p_l Psi | sum_i V(r_i) | p_m Psi)/(Psi|Psi) are
exact here (up to the final float conversion).  This is the closed-form anchor for the
plain estimator; nothing here follows reference code.  Power sums are taken of z / R,
so a sector-dL block is the raw block divided by R^{2 dL}.
"""
from fractions import Fraction
from math import factorial
from itertools import combinations


def _mul(p, q):
    out = {}
    for ea, ca in p.items():
        for eb, cb in q.items():
            e = tuple(x + y for x, y in zip(ea, eb))
            out[e] = out.get(e, 0) + ca * cb
    return {e: c for e, c in out.items() if c != 0}


def _pow(p, r, n):
    out = {tuple([0] * n): 1}
    for _ in range(r):
        out = _mul(out, p)
    return out


def laughlin(n, m):
    poly = {tuple([0] * n): 1}
    for i, j in combinations(range(n), 2):
        ei = [0] * n
        ej = [0] * n
        ei[i] = 1
        ej[j] = 1
        diff = {tuple(ei): 1, tuple(ej): -1}
        poly = _mul(poly, _pow(diff, m, n))
    return poly


def power_sum(n, k):
    if k == 0:
        return {tuple([0] * n): n}
    out = {}
    for i in range(n):
        e = [0] * n
        e[i] = k
        out[tuple(e)] = 1
    return out


def monomial_poly(n, pairs):
    poly = {tuple([0] * n): 1}
    for k, r in pairs:
        if k == 0:
            continue
        poly = _mul(poly, _pow(power_sum(n, k), r, n))
    return poly


def _norm(a):
    v = 1
    for x in a:
        v *= (2 ** x) * factorial(x)
    return v


def _radial(a, w, eps, strength=0.0, expo=0, R=1.0):
    """sum_i <V(r_i)>_a with V = w r^2/2 + eps r^4 + strength (r/R)^expo (expo even)."""
    v = Fraction(0)
    w = Fraction(w)
    eps = Fraction(eps)
    for x in a:
        v += w * (x + 1) + 4 * eps * (x + 1) * (x + 2)
    if strength:
        assert expo % 2 == 0, "exact radial moments need an even exponent"
        q = expo // 2
        s = 0.0
        for x in a:
            s += (2 ** q) * factorial(x + q) / factorial(x)
        return float(v) + strength * s / R ** expo
    return float(v)


def exact_blocks(n, m, states, w, eps, strength=0.0, expo=0):
    """S and H for a list of (dL, pairs) states of ONE sector, with z/R power sums.

    Returns (S, H) as lists of lists of floats (real: the exact matrices are real)."""
    R = (2.0 * m * n) ** 0.5
    psi = laughlin(n, m)
    norm0 = sum(c * c * _norm(a) for a, c in psi.items())
    polys = [_mul(monomial_poly(n, p), psi) for _, p in states]
    dls = [dl for dl, _ in states]
    d = len(states)
    S = [[0.0] * d for _ in range(d)]
    H = [[0.0] * d for _ in range(d)]
    for i in range(d):
        for j in range(d):
            if dls[i] != dls[j]:
                continue
            si = 0
            hi = 0.0
            pi, pj = polys[i], polys[j]
            small, big = (pi, pj) if len(pi) < len(pj) else (pj, pi)
            for a, c in small.items():
                c2 = big.get(a)
                if c2 is None:
                    continue
                nrm = c * c2 * _norm(a)
                si += nrm
                hi += nrm * _radial(a, w, eps, strength, expo, R)
            scale = R ** (dls[i] + dls[j])
            S[i][j] = si / norm0 / scale
            H[i][j] = hi / norm0 / scale
    return S, H


def exact_occupation(n, m, states, lcut):
    """<a| sum_i |l><l|_i |b> / (Psi|Psi) for one sector (z / R power sums), l < lcut.

    The one-body projector on orbital l is diagonal in the monomial basis with eigenvalue
    n_l(alpha) = #{i : alpha_i = l}.  Returns O[a][b][l] (real)."""
    R = (2.0 * m * n) ** 0.5
    psi = laughlin(n, m)
    norm0 = sum(c * c * _norm(a) for a, c in psi.items())
    polys = [_mul(monomial_poly(n, p), psi) for _, p in states]
    dls = [dl for dl, _ in states]
    d = len(states)
    O = [[[0.0] * lcut for _ in range(d)] for _ in range(d)]
    for i in range(d):
        for j in range(d):
            acc = [0] * lcut
            pi, pj = polys[i], polys[j]
            for a, c in pi.items():
                c2 = pj.get(a)
                if c2 is None:
                    continue
                w = c * c2 * _norm(a)
                for x in a:
                    if x < lcut:
                        acc[x] += w
            scale = R ** (dls[i] + dls[j]) * norm0
            for l in range(lcut):
                O[i][j][l] = acc[l] / scale
    return O


def exact_removal(n, m, l, states):
    """Particle removal from the n-particle Laughlin state, exactly, for tiny n.

    c_l Psi_L^n, as a polynomial of the remaining n - 1 coordinates, is the z_n^l coefficient
    of Psi_L^n (the orbital z^l exp(-|z|^2/4) is orthogonal to every other power of z_n).
    states: (dL, pairs) of ONE (n-1)-particle sector, dL = m (n - 1) - l.  Returns (G, Hq, b, qq)
    with G = (B_a | B_b), b = (B_a | Q), qq = (Q | Q) for B_a = p_a Psi_L^{n-1} and Q = c_l Psi
    (raw power sums, one common normalization; only ratios are meaningful), and Hq the matrix
    of sum_i r_i^2 / 2 + r_i^4 (w = 1, eps = 1) in the same basis."""
    nm = n - 1
    psi = laughlin(n, m)
    q = {}
    for a, c in psi.items():
        if a[n - 1] == l:
            q[a[:nm]] = q.get(a[:nm], 0) + c
    q = {a: c for a, c in q.items() if c != 0}
    psi1 = laughlin(nm, m)
    basis = [_mul(monomial_poly(nm, p), psi1) for _, p in states]

    def dot(p1, p2, v=False):
        small, big = (p1, p2) if len(p1) < len(p2) else (p2, p1)
        acc = Fraction(0)
        for a, c in small.items():
            c2 = big.get(a)
            if c2 is not None:
                acc += c * c2 * _norm(a) * (Fraction(_radial(a, 1, 1)) if v else 1)
        return acc

    d = len(states)
    G = [[float(dot(basis[i], basis[j])) for j in range(d)] for i in range(d)]
    Hq = [[float(dot(basis[i], basis[j], True)) for j in range(d)] for i in range(d)]
    b = [float(dot(basis[i], q)) for i in range(d)]
    return G, Hq, b, float(dot(q, q))
