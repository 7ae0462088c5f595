"""CPU tests that pin the FP64 oracle (no GPU).

The reference ships no tests, golden vectors or fixtures, so each oracle function is
pinned by something that does not restate it: published known-answer vectors
(Philox), closed-form counts (p(n)), brute force, exact LLL algebra (small N) and
textbook library routines (scipy.linalg.eigh).
"""
import itertools
import math
import os

import numpy as np
import pytest
import scipy.linalg

from oracle.exact_lll import exact_blocks
from oracle.partitions import Basis, accel_asc, compress, sector_basis

P_COUNTS = [1, 1, 2, 3, 5, 7, 11, 15, 22, 30, 42, 56, 77, 101, 135, 176, 231, 297, 385, 490, 627]


def test_philox_known_answers(oracle_lib):
    # Random123 kat_vectors, philox4x32_10
    kat = [
        ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
        ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
        ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
         [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
    ]
    for ctr, key, want in kat:
        assert oracle_lib.philox(ctr, key) == want


def test_partition_counts():
    assert [len(accel_asc(n)) for n in range(21)] == P_COUNTS


def _brute_partitions(n):
    out = set()

    def rec(rem, mx, cur):
        if rem == 0:
            out.add(tuple(sorted(cur)))
            return
        for k in range(min(rem, mx), 0, -1):
            rec(rem - k, k, cur + [k])
    rec(n, n, [])
    return out if n else {(0,)}


@pytest.mark.parametrize("n", list(range(0, 13)))
def test_accel_asc_is_every_partition_once(n):
    got = [tuple(p) for p in accel_asc(n)]
    assert len(got) == len(set(got))
    assert set(got) == _brute_partitions(n)
    assert all(list(p) == sorted(p) for p in got)


def test_compression_and_order_dl4():
    # accelAsc order + (part, multiplicity) pairs sorted by increasing multiplicity
    assert sector_basis(10, 4) == [[(1, 4)], [(2, 1), (1, 2)], [(1, 1), (3, 1)], [(2, 2)], [(4, 1)]]
    assert sector_basis(10, 0) == [[(0, 1)]]


def test_length_cutoff():
    # at most N parts: N = 3 drops only (1,1,1,1) from dL = 4
    b = sector_basis(3, 4)
    assert all(sum(r for _, r in p) <= 3 for p in b)
    assert len(b) == 4 and [(1, 4)] not in b
    for n in range(1, 12):
        for N in range(1, 8):
            assert len(sector_basis(N, n)) == sum(1 for p in _brute_partitions(n) if len(p) <= N)


def test_compress_stability():
    # equal multiplicities keep increasing part order
    assert compress([1, 2, 2, 3, 3, 5]) == [(1, 1), (5, 1), (2, 2), (3, 2)]


def test_exact_quadratic_closed_form():
    """eps = 0: H = w (L0 + dL + N) S exactly (sum_i r_i^2 = 2 (L + N) in the LLL)."""
    for N, m, dl in [(3, 3, 3), (3, 2, 4), (4, 2, 3)]:
        L0 = m * N * (N - 1) // 2
        b = Basis(N, [dl])
        S, H = exact_blocks(N, m, b.states, 0.05, 0.0)
        S, H = np.array(S), np.array(H)
        assert np.allclose(H, 0.05 * (L0 + dl + N) * S, rtol=1e-12)


def test_exact_radial_moments_by_quadrature():
    """<r^2>_a = 2(a+1), <r^4>_a = 4(a+1)(a+2) for |z^a|^2 e^{-|z|^2/2}."""
    from scipy.integrate import quad
    for a in range(6):
        w = lambda r, p: r ** (2 * a + p + 1) * math.exp(-r * r / 2)
        n0 = quad(lambda r: w(r, 0), 0, 60)[0]
        assert abs(quad(lambda r: w(r, 2), 0, 60)[0] / n0 - 2 * (a + 1)) < 1e-8
        assert abs(quad(lambda r: w(r, 4), 0, 60)[0] / n0 - 4 * (a + 1) * (a + 2)) < 1e-7


def _numpy_features(pos, N, m, states):
    """independent FP64 evaluation of p_lambda (z / R power sums) by direct products"""
    R = math.sqrt(2 * m * N)
    z = (pos[..., 0] + 1j * pos[..., 1]) / R
    out = []
    for _, pairs in states:
        v = np.ones(z.shape[0], dtype=complex)
        for k, r in pairs:
            if k == 0:
                continue
            v = v * np.sum(z ** k, axis=1) ** r
        out.append(v)
    return np.stack(out, 1)


def test_fixed_accumulation_vs_numpy(oracle_lib):
    N, m, sectors = 7, 3, [0, 2, 3, 5]
    w, eps = 0.05, 1e-3
    o = oracle_lib.Oracle(N, m, sectors, w=w, eps=eps)
    rng = np.random.default_rng(1)
    pos = rng.normal(size=(500, N, 2)) * 3
    So, Ho, vs = o.accumulate_fixed(pos)
    P = _numpy_features(pos, N, m, o.basis.states)
    r2 = (pos ** 2).sum(-1)
    V = (0.5 * w * r2 + eps * r2 ** 2).sum(1)
    assert abs(vs - V.sum()) < 1e-9 * abs(V.sum())
    off = 0
    for s, d in enumerate(o.dims):
        blk = P[:, off:off + d]
        S = blk.conj().T @ blk
        H = blk.conj().T @ (V[:, None] * blk)
        assert np.allclose(So[s], S, rtol=1e-12, atol=1e-12 * np.abs(S).max())
        assert np.allclose(Ho[s], H, rtol=1e-12, atol=1e-12 * np.abs(H).max())
        off += d


def test_weighted_accumulation_vs_numpy(oracle_lib):
    """sum_s u_s conj(p_a) p_b between two sectors (harmonic and user-weight blocks)."""
    N, m, sectors = 7, 3, [0, 2, 3]
    o = oracle_lib.Oracle(N, m, sectors, w=0.05, eps=1e-3)
    rng = np.random.default_rng(4)
    pos = rng.normal(size=(400, N, 2)) * 3
    u = rng.normal(size=400)
    P = _numpy_features(pos, N, m, o.basis.states)
    offs = np.concatenate([[0], np.cumsum(o.dims)])
    for a, b in ((0, 1), (1, 2), (2, 1), (2, 2)):
        Pa, Pb = P[:, offs[a]:offs[a + 1]], P[:, offs[b]:offs[b + 1]]
        ref = Pa.conj().T @ (u[:, None] * Pb)
        got = o.accumulate_weighted(pos, u, a, b)
        assert np.allclose(got, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max()), (a, b)


def test_geneig_vs_scipy(oracle_lib):
    rng = np.random.default_rng(7)
    for d in (1, 2, 3, 8, 25):
        A = rng.normal(size=(d, d + 3)) + 1j * rng.normal(size=(d, d + 3))
        S = A @ A.conj().T + 0.1 * np.eye(d)
        B = rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))
        H = B + B.conj().T
        ev, vec = oracle_lib.geneig(S, H, vectors=True)
        ref = scipy.linalg.eigh(H, S, eigvals_only=True)
        assert np.allclose(ev, ref, rtol=1e-10, atol=1e-10)
        assert np.allclose(H @ vec, S @ vec * ev[None, :], atol=1e-9 * np.abs(H).max())
        assert np.allclose(vec.conj().T @ S @ vec, np.eye(d), atol=1e-9)
    # a non-positive-definite metric is reported, not turned into NaN
    with pytest.raises(np.linalg.LinAlgError):
        oracle_lib.geneig(np.diag([1.0, -1.0]).astype(complex), np.eye(2, dtype=complex))


def test_dsf_weights_vs_scipy(oracle_lib):
    rng = np.random.default_rng(11)
    d = 6
    A = rng.normal(size=(d + 1, d + 4)) + 1j * rng.normal(size=(d + 1, d + 4))
    M = A @ A.conj().T
    B = rng.normal(size=(d + 1, d + 1)) + 1j * rng.normal(size=(d + 1, d + 1))
    H = B + B.conj().T
    VC = rng.normal(size=(d + 1, d + 1)) + 1j * rng.normal(size=(d + 1, d + 1))
    VS = rng.normal(size=(d + 1, d + 1)) + 1j * rng.normal(size=(d + 1, d + 1))
    e, wts = oracle_lib.dsf_weights(H, M, VC, VS)
    ev, vec = scipy.linalg.eigh(H[1:, 1:], M[1:, 1:])
    v = VC[0, 1:] - 1j * VS[0, 1:]
    assert np.allclose(e, ev)
    assert np.allclose(wts, np.abs(v @ vec) ** 2, rtol=1e-9)


def test_plain_chain_r2_sum_rule(oracle_lib):
    """FP64 Metropolis chain: <sum_i r_i^2> = 2 (L0 + N) exactly."""
    N, m = 6, 3
    o = oracle_lib.Oracle(N, m, [0])
    L0 = m * N * (N - 1) // 2
    vals = []
    for seed in range(4):
        _, _, acc, r2, _ = o.plain_run(64, 200, 400, 0.8, 100 + seed)
        vals.append(r2)
    mean, err = np.mean(vals), np.std(vals, ddof=1) / 2
    assert abs(mean - 2 * (L0 + N)) < 4 * err + 2e-3 * 2 * (L0 + N)


def test_plain_chain_vs_exact_small_n(oracle_lib):
    """N = 3, m = 3, dL = 2: MC eigenvalues and normalised metric vs the exact expansion."""
    N, m, dl = 3, 3, 2
    w, eps = 0.05, 2e-3
    o = oracle_lib.Oracle(N, m, [dl], w=w, eps=eps)
    Se, He = exact_blocks(N, m, o.basis.states, w, eps)
    Se, He = np.array(Se, dtype=complex), np.array(He, dtype=complex)
    exact = scipy.linalg.eigh(He, Se, eigvals_only=True)
    ests = []
    for seed in range(6):
        Sb, Hb, _, _, _ = o.plain_run(128, 200, 1500, 0.9, 2000 + seed)
        ests.append(oracle_lib.geneig(Sb[0], Hb[0]))
    ests = np.array(ests)
    mean = ests.mean(0)
    err = ests.std(0, ddof=1) / np.sqrt(len(ests))
    assert np.all(np.abs(mean - exact) < 5 * err + 1e-9), (mean, exact, err)


def test_rowwise_estimator_small_n(oracle_lib):
    """Reference row-wise estimator (|p_f Psi|^2 chains, g/f ratios, /sqrt<|g/f|^2>,
    symmetrised) reproduces the exact normalised metric and the spectrum."""
    N, m, dl = 3, 3, 2
    w, eps = 0.05, 2e-3
    o = oracle_lib.Oracle(N, m, [dl], w=w, eps=eps)
    Se, He = exact_blocks(N, m, o.basis.states, w, eps)
    Se, He = np.array(Se), np.array(He)
    dn = np.sqrt(np.diag(Se))
    Mexact = Se / np.outer(dn, dn)
    exact = scipy.linalg.eigh(He, Se, eigvals_only=True)
    H, M, _, _, acc = o.rowwise_sector(0, 64, 200, 1500, 0.9, 31)
    assert np.allclose(np.diag(M).real, 1.0, atol=1e-12)      # exact unit diagonal
    assert np.allclose(M, M.conj().T) and np.allclose(H, H.conj().T)
    assert np.abs(M - Mexact).max() < 0.03
    ev = oracle_lib.geneig(M, H)
    assert np.allclose(ev, exact, rtol=0.01)
    assert 0.2 < acc < 0.9


def test_mixture_chain_without_tail_terms_is_the_plain_chain(oracle_lib):
    """w = c0 (every tail term underflows to 0): the mixture chain makes the plain chain's moves
    and its 1/w-weighted, sum(1/w)-normalised S and H are the plain means."""
    N, m = 8, 3
    o = oracle_lib.Oracle(N, m, [1, 2, 3], w=0.05, eps=1e-4)
    mix = (3, 0.5, [3.0, 1.5, 1.0], [-1e300] * 3)
    Sp, Hp, acc_p, _, pos_p = o.plain_run(16, 20, 40, 0.8, 77, want_pos=True)
    Sm, Hm, ws, acc_m, pos_m = o.mixture_run(mix, 16, 20, 40, 0.8, 77, want_pos=True)
    assert acc_m == acc_p and np.array_equal(pos_m, pos_p)
    assert ws == pytest.approx(16 * 40 / 0.5, rel=1e-15)
    for a, b in zip(Sm + Hm, Sp + Hp):
        assert np.abs(a - b).max() <= 1e-13 * np.abs(b).max()


def test_mixture_density_bounds_every_monomial(oracle_lib):
    """Weighted AM-GM: with t_k = c_k (|p_k|^2 / n_k)^(L/k) and w = rho + sum_k t_k, every
    partition lambda of every sector up to L has |p_lambda|^2 / w <= prod_k n_k^(r_k)
    max(1 / rho, max_k (k r_k / L) / c_k).  Checked on configurations far into the tails (the
    plain estimator's integrand |p_1|^(2 L) is unbounded there)."""
    N, m, L = 12, 3, 6
    o = oracle_lib.Oracle(N, m, list(range(1, L + 1)))
    rng = np.random.default_rng(5)
    nk = rng.uniform(0.2, 2.0, size=L)
    rho = 0.3
    kmix, c0, a, b = oracle_lib.mixture_coefficients(L, nk, rho)
    ck = np.exp(np.array(b) + np.array(a) * np.log(nk))
    worst = 0.0
    for trial in range(400):
        pos = rng.normal(size=(N, 2)) * o.R * rng.uniform(0.05, 3.0)
        if trial % 4 == 0:
            pos[:, 0] += o.R * rng.uniform(0.5, 4.0)          # centre of mass far out: |p_1| large
        w = o.mixture_weights((kmix, c0, a, b), pos[None])[0]
        pl, _ = o.features(np.ascontiguousarray(pos.reshape(-1)))
        for t, (lam, pairs) in enumerate(o.basis.states):
            r = dict(pairs)
            bound = np.prod([nk[k - 1] ** q for k, q in r.items()]) * max(
                [1.0 / rho] + [(k * q / L) / ck[k - 1] for k, q in r.items()])
            ratio = abs(pl[t]) ** 2 / w / bound
            assert lam <= L and ratio <= 1.0 + 1e-9, (trial, st, ratio)
            worst = max(worst, ratio)
    assert worst > 1e-3          # the configurations reach the bound's scale


@pytest.mark.parametrize("com", [False, True])
def test_mixture_chain_vs_exact_small_n(oracle_lib, com):
    """N = 3, m = 3, dL = 1, 2, 3: S and H of the mixture chain (sums of p_a* p_b / w over sums
    of 1 / w) and the spectra agree with the exact expansion, also with the exact centre-of-mass
    random walk and rotation after every kept sample."""
    N, m, sectors = 3, 3, [1, 2, 3]
    w, eps = 0.05, 2e-3
    o = oracle_lib.Oracle(N, m, sectors, w=w, eps=eps)
    Se, He = exact_blocks(N, m, o.basis.states, w, eps)
    Se, He = np.array(Se, dtype=complex), np.array(He, dtype=complex)
    pos, _ = o.plain_positions(256, 200, 20, 0.9, 3)
    u = (pos[..., 0] + 1j * pos[..., 1]) / o.R
    nk = [float(np.mean(np.abs((u ** k).sum(-1)) ** 2)) for k in range(1, 4)]
    mix = oracle_lib.mixture_coefficients(3, nk, 0.5)
    runs = [o.mixture_run(mix, 128, 300, 1000, 0.9, 500 + seed, com_moves=com) for seed in range(6)]
    off = 0
    for s, d in enumerate(o.dims):
        Sx, Hx = Se[off:off + d, off:off + d], He[off:off + d, off:off + d]
        off += d
        Ss = np.array([r[0][s] for r in runs])
        Hs = np.array([r[1][s] for r in runs])
        for est, ex in ((Ss, Sx), (Hs, Hx)):
            err = est.std(0, ddof=1) / np.sqrt(len(runs))
            dev = np.abs(est.mean(0) - ex)
            assert np.all(dev <= 5 * np.abs(err) + 1e-9 * np.abs(ex).max()), (s, dev.max(), np.abs(err).max())
        ev = np.array([oracle_lib.geneig(Ss[i], Hs[i]) for i in range(len(runs))])
        exact = scipy.linalg.eigh(Hx, Sx, eigvals_only=True)
        err = ev.std(0, ddof=1) / np.sqrt(len(runs))
        assert np.all(np.abs(ev.mean(0) - exact) < 5 * err + 1e-9), (s, ev.mean(0), exact, err)


def _block_system(rng, dims, E0=7.0, spread=0.3):
    D = sum(dims)
    M = np.zeros((D, D), complex)
    H = np.zeros((D, D), complex)
    o = 0
    for d in dims:
        A = rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))
        Mb = A @ A.conj().T / d + np.eye(d)
        B = rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))
        Hb = E0 * Mb + spread * (B + B.conj().T) / 2
        M[o:o + d, o:o + d] = Mb
        H[o:o + d, o:o + d] = Hb
        o += d
    B = rng.normal(size=(D, D)) + 1j * rng.normal(size=(D, D))
    U = (B + B.conj().T) / 2
    rhos = [rng.normal(size=(D, D)) + 1j * rng.normal(size=(D, D)) for _ in range(2)]
    return M, H, U, rhos


def test_evolve_oracle_free_closed_form(oracle_lib):
    """F = 0: C(t) = sum_j e^{-i E_j t} v_j (v_j^H M e_0); norm conserved (dynamics.cpp:73-83)."""
    rng = np.random.default_rng(2)
    M, H, U, rhos = _block_system(rng, [1, 2, 3])
    rho, nrm, _ = oracle_lib.evolve(M, H, U, rhos, lambda t: 0.0, 3.0, 0.25)
    E, V = scipy.linalg.eigh(H, M)
    a0 = V.conj().T @ M[:, 0]
    for k, t in enumerate(np.arange(12) * 0.25):
        C = V @ (np.exp(-1j * E * t) * a0)
        for l, R in enumerate(rhos):
            assert abs(rho[k, l] - C.conj() @ R @ C) < 1e-6 * max(1, abs(C.conj() @ R @ C))
        assert abs(nrm[k] - M[0, 0].real) < 1e-6   # C^H M C is conserved (= M_00 at t = 0)


def test_evolve_oracle_driven_vs_dense_expm(oracle_lib):
    """Driven case against a fine-grained exact propagator (product of matrix exponentials
    at midpoints, step 1e-3) - an independent integrator."""
    rng = np.random.default_rng(4)
    M, H, U, rhos = _block_system(rng, [1, 2, 2])
    F = lambda t: 0.3 * np.exp(-(t - 1.0) ** 2 / (2 * 0.3 ** 2))
    rho, nrm, _ = oracle_lib.evolve(M, H, U, rhos, F, 2.0, 0.5)
    Mi = np.linalg.inv(M)
    C = np.zeros(M.shape[0], complex)
    C[0] = 1
    t, h = 0.0, 1e-3
    for k in range(4):
        while t < 0.5 * k - 1e-12:
            C = scipy.linalg.expm(-1j * h * Mi @ (H + F(t + h / 2) * U)) @ C
            t += h
        for l, R in enumerate(rhos):
            ref = C.conj() @ R @ C
            assert abs(rho[k, l] - ref) < 2e-5 * max(1.0, abs(ref))


def test_exact_occupation_sum_rules():
    """sum_l O_ab(l) = N S_ab and sum_l V_l O_ab(l) = H_ab with V_l = w(l+1) + 4 eps (l+1)(l+2)."""
    from oracle.exact_lll import exact_occupation
    N, m, dl = 3, 3, 3
    b = Basis(N, [dl])
    lcut = m * (N - 1) + dl + 1
    O = np.array(exact_occupation(N, m, b.states, lcut))
    S, H = exact_blocks(N, m, b.states, 0.05, 1e-3)
    S, H = np.array(S), np.array(H)
    l = np.arange(lcut)
    V = 0.05 * (l + 1) + 4e-3 * (l + 1) * (l + 2)
    assert np.allclose(O.sum(-1), N * S, rtol=1e-12)
    assert np.allclose((O * V).sum(-1), H, rtol=1e-12)


def test_learned_oracle_gradient_by_finite_differences():
    """the oracle's analytic loss gradient (through both MLP evaluations, the reverse
    density and the features at z') against central finite differences"""
    from oracle import learned
    rng = np.random.default_rng(3)
    N, m = 12, 3
    R = np.sqrt(2 * m * N)
    p = learned.init_params(rng)
    p += rng.normal(size=p.shape) * 0.05
    configs, idx, noise = [], [], []
    for b in range(6):
        r = R * np.sqrt(rng.random(N))
        t = 2 * np.pi * rng.random(N)
        configs.append(r * np.exp(1j * t))
        idx.append(int(rng.integers(N)))
        noise.append(rng.normal(size=2))
    L, G = learned.batch_loss_grad(p, configs, idx, noise, m, R, 0.7)
    for k in rng.choice(learned.NPARAM, 25, replace=False).tolist() + [0, 100, 4400, 4545]:
        e = np.zeros_like(p)
        e[k] = 1e-6
        Lp, _ = learned.batch_loss_grad(p + e, configs, idx, noise, m, R, 0.7)
        Lm, _ = learned.batch_loss_grad(p - e, configs, idx, noise, m, R, 0.7)
        fd = (Lp - Lm) / 2e-6
        assert abs(fd - G[k]) < 1e-5 + 1e-4 * abs(fd), (k, fd, G[k])


def test_learned_features_are_nearest_neighbour_distances():
    """f2, f3 = the nearest and the mean of the 8 nearest distances (capped at 8), against
    scipy's k-d tree queries; the feature gradient against central differences"""
    from scipy.spatial import cKDTree
    from oracle import learned
    rng = np.random.default_rng(11)
    for N in (5, 9, 40, 200):
        R = np.sqrt(6.0 * N)
        z = R * np.sqrt(rng.random(N)) * np.exp(2j * np.pi * rng.random(N))
        pts = np.stack([z.real, z.imag], 1)
        tree = cKDTree(pts)
        for i in rng.choice(N, min(N, 6), replace=False):
            k = min(9, N)
            d, _ = tree.query(pts[i], k=k)
            d = np.concatenate([np.asarray(d)[1:], np.full(9 - k, np.inf)])   # drop the point itself
            near = np.minimum(d, learned.D_CAP)
            f, df = learned.features(z, int(i), z[i], R)
            assert abs(f[0] - abs(z[i]) / R) < 1e-14
            assert abs(f[1] - near[0]) < 1e-12 and abs(f[2] - near.mean()) < 1e-12, (N, i)
            for c, dz in ((0, 1e-6), (1, 1e-6j)):
                fp, _ = learned.features(z, int(i), z[i] + dz, R)
                fm, _ = learned.features(z, int(i), z[i] - dz, R)
                assert np.allclose((fp - fm) / 2e-6, df[:, c], atol=1e-6), (N, i, c)


GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_golden_geneig(oracle_lib):
    """oracle eigensolver against the committed scipy.linalg.eigh vectors"""
    g = np.load(os.path.join(GOLDEN, "geneig_scipy.npz"))
    for k in range(3):
        ev = oracle_lib.geneig(g[f"S{k}"], g[f"H{k}"])
        assert np.allclose(ev, g[f"E{k}"], rtol=1e-10, atol=1e-10)


def test_golden_exact_blocks():
    """the exact N = 4 fixture is what oracle/exact_lll.py computes, and it obeys
    H = w (L0 + dL + N) S at eps = 0 up to the eps r^4 part"""
    from oracle.exact_lll import exact_blocks
    from oracle.partitions import Basis
    g = np.load(os.path.join(GOLDEN, "exact_n4_m3.npz"))
    N, m, w, eps = int(g["N"]), int(g["m"]), float(g["w"]), float(g["eps"])
    for dl in range(5):
        b = Basis(N, [dl])
        S, H = exact_blocks(N, m, b.states, w, eps)
        assert np.allclose(S, g[f"S{dl}"], rtol=1e-13)
        assert np.allclose(H, g[f"H{dl}"], rtol=1e-13)
        S0, H0 = exact_blocks(N, m, b.states, w, 0.0)
        L0 = m * N * (N - 1) // 2
        assert np.allclose(H0, w * (L0 + dl + N) * np.array(S0), rtol=1e-12)


def test_removal_newton_vs_coefficient_extraction():
    """Removing orbital l from the 4-particle Laughlin state (m = 3) leaves exactly
    e_k(M) Psi_L^3, k = 9 - l, M = {z_j, each 3 times}: the z_0^l coefficient of the expanded
    polynomial equals the Newton-identity combination of power-sum states (b = +-G a and
    |Q|^2 = a^T G a), so the Newton-route removal weights are the exact ones and sum to 1.
    A wrong z_lambda, sign or power of m in the coefficients breaks b = +-G a."""
    from scipy.linalg import eigh
    from oracle.exact_lll import exact_removal
    from oracle.partitions import sector_basis
    from oracle.spectral import newton_coefficients, removal_weights
    n, m = 4, 3
    for k in range(0, 4):
        l = m * (n - 1) - k
        states = [(k, p) for p in sector_basis(n - 1, k)]
        G, Hq, b, qq = exact_removal(n, m, l, states)
        G, Hq, b = np.array(G), np.array(Hq), np.array(b)
        a = newton_coefficients(k, m, states)
        ga = G @ a
        sgn = np.sign(b @ ga)
        assert np.allclose(b, sgn * ga, rtol=1e-12, atol=1e-12 * np.abs(ga).max()), (k, b, ga)
        assert abs(qq - a @ ga) <= 1e-12 * qq, k
        ev, w = removal_weights(G, Hq, a)
        assert abs(w.sum() - 1) < 1e-12
        e2, V = eigh(Hq, G)
        assert np.allclose(w, (V.T @ b) ** 2 / qq, atol=1e-12)
        if k == 0:
            assert len(w) == 1 and abs(w[0] - 1) < 1e-15
