"""GPU parity tests: HIP path (through the C ABI) against the FP64 CPU oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel_fro(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _disc_positions(rng, n, N, R):
    r = R * np.sqrt(rng.random((n, N)))
    t = 2 * np.pi * rng.random((n, N))
    return np.stack([r * np.cos(t), r * np.sin(t)], -1)


@pytest.mark.parametrize("prec,tol", [(0, 1e-6), (1, 1e-12), (2, 2e-5), (3, 1e-6)])
def test_fixed_sample_moments(oracle_lib, prec, tol):
    """Same FP64 configurations through the GPU features+Gram path and the oracle."""
    from laughlinedge_amd import Engine
    N, m, sectors = 12, 3, [0, 1, 2, 3, 4, 5, 6]
    w, eps = 0.05, 1e-4
    print(f"case prec={prec} N={N} sectors={sectors}", flush=True)
    o = oracle_lib.Oracle(N, m, sectors, w=w, eps=eps)
    pos, _ = o.plain_positions(16, 50, 40, 0.6, 99)
    pos = pos.reshape(-1, N, 2)
    So, Ho, _ = o.accumulate_fixed(pos)
    n = pos.shape[0]
    with Engine(N, m, sectors, w=w, eps=eps, walkers=64, precision=prec, samples=1) as e:
        # D = 30 states = 60 real columns, one 64-column tile.  The sector squares meet the
        # 16x16 blocks (0,0) (0,1) | (1,1) (1,2) | (2,2) (2,3) | (3,3) above the diagonal; the odd
        # block of row block 3 is computed twice: 8 blocks x 16^2 x 2 weights x 2 FLOP per row
        assert e.info.gram_mfma_flop_per_row == 8 * 256 * 2 * 2
        e.accumulate_positions(pos)
        mom = e.moments()
    errs = [(_rel_fro(Sg, So[s] / n), _rel_fro(Hg, Ho[s] / n)) for s, (Sg, Hg) in enumerate(mom)]
    print(f"prec={prec} max rel Frobenius error S {max(e[0] for e in errs):.3e} H {max(e[1] for e in errs):.3e}")
    for s, (es, eh) in enumerate(errs):
        assert es < tol and eh < tol, (s, es, eh)
    # entry (p, q), p > q, of the accumulated tiles is read as (q, p): the raw records
    # (no host symmetrization) are exactly Hermitian
    for Sg, Hg in mom:
        assert np.array_equal(Sg, Sg.conj().T) and np.array_equal(Hg, Hg.conj().T)


@pytest.mark.parametrize("prec,tol", [(0, 1e-6), (2, 2e-5), (3, 1e-6)])
def test_fixed_sample_large_tiles(oracle_lib, prec, tol):
    """N=40, dL 10..12 spans several 64-column Gram tiles, 3000 samples (not a multiple of
    the 32-sample staging step); FP32 and both BF16-split paths vs the FP64 oracle."""
    from laughlinedge_amd import Engine
    N, m, sectors = 40, 3, [10, 11, 12]
    print(f"case N={N} sectors={sectors}", flush=True)
    o = oracle_lib.Oracle(N, m, sectors, w=0.05, eps=5e-5)
    rng = np.random.default_rng(5)
    pos = _disc_positions(rng, 3000, N, o.R)
    So, Ho, _ = o.accumulate_fixed(pos)
    with Engine(N, m, sectors, w=0.05, eps=5e-5, walkers=256, samples=1, precision=prec) as e:
        assert e.info.n_tiles > 3
        e.accumulate_positions(pos)
        mom = e.moments()
    errs = [(_rel_fro(Sg, So[s] / 3000), _rel_fro(Hg, Ho[s] / 3000)) for s, (Sg, Hg) in enumerate(mom)]
    print(f"prec={prec} max rel Frobenius error S {max(e[0] for e in errs):.3e} H {max(e[1] for e in errs):.3e}")
    for s, (es, eh) in enumerate(errs):
        assert es < tol and eh < tol, (s, es, eh)


@pytest.mark.parametrize("prec,tol", [(0, 1e-6), (1, 1e-12)])
def test_fixed_sample_mixed_layout(oracle_lib, prec, tol):
    """The sample-row layout is chosen per sector set; N = 200, dL 0..8 (dims 1, 1, 2, 3, 5, 7,
    11, 15, 22) takes the mixed one: the sectors of 16 real columns or more (dL 6, 7, 8) start
    on multiples of 16, the small ones stay packed.  Its tiles compute 24 16x16 blocks against
    28 packed and 30 with every sector aligned (all in 5 tiles).  S and H of 2048 configurations
    match the FP64 oracle, and the raw records stay exactly Hermitian."""
    from laughlinedge_amd import Engine
    N, m, sectors = 200, 3, list(range(0, 9))
    print(f"case mixed layout prec={prec} N={N} sectors=0..8", flush=True)
    o = oracle_lib.Oracle(N, m, sectors, w=0.05, eps=5e-5)
    rng = np.random.default_rng(17)
    pos = _disc_positions(rng, 2048, N, o.R)
    So, Ho, _ = o.accumulate_fixed(pos)
    with Engine(N, m, sectors, w=0.05, eps=5e-5, walkers=256, samples=1, precision=prec) as e:
        assert e.info.gram_mfma_flop_per_row == 24 * 16 * 16 * 2 * 2
        e.accumulate_positions(pos)
        mom = e.moments()
    errs = [(_rel_fro(Sg, So[s] / 2048), _rel_fro(Hg, Ho[s] / 2048)) for s, (Sg, Hg) in enumerate(mom)]
    print(f"prec={prec} max rel Frobenius error S {max(e[0] for e in errs):.3e} H {max(e[1] for e in errs):.3e}")
    for s, (es, eh) in enumerate(errs):
        assert es < tol and eh < tol, (s, es, eh)
    for Sg, Hg in mom:
        assert np.array_equal(Sg, Sg.conj().T) and np.array_equal(Hg, Hg.conj().T)


def test_fp64_chain_matches_oracle(oracle_lib):
    """FP64 sampler: same Philox streams as the oracle chain -> same positions."""
    from laughlinedge_amd import Engine
    N, m, W = 10, 3, 64
    print("case fp64 chain N=10", flush=True)
    o = oracle_lib.Oracle(N, m, [0])
    ref, _ = o.plain_positions(W, 0, 1, 0.5, 1234, thin=6)
    with Engine(N, m, [0], walkers=W, precision=1, step=0.5, seed=1234, lanes_per_walker=2) as e:
        e.init_walkers()
        e.burnin(6)
        got = e.positions()
    err = np.abs(got - ref[:, 0]).max(axis=(1, 2))
    assert np.mean(err < 1e-9) >= 0.95, err


def test_fp32_chain_tracks_oracle(oracle_lib):
    """FP32 sampler follows the FP64 chain for a couple of sweeps (same random streams)."""
    from laughlinedge_amd import Engine
    N, m, W = 16, 3, 64
    print("case fp32 chain N=16", flush=True)
    o = oracle_lib.Oracle(N, m, [0])
    ref, _ = o.plain_positions(W, 0, 1, 0.5, 77, thin=2)
    for G in (2, 4, 8, 16, 32, 64):
        with Engine(N, m, [0], walkers=W, precision=0, step=0.5, seed=77, lanes_per_walker=G) as e:
            e.init_walkers()
            e.burnin(2)
            got = e.positions()
        err = np.abs(got - ref[:, 0]).max(axis=(1, 2))
        assert np.mean(err < 1e-3) >= 0.9, (G, err)


def test_geneig_gpu_vs_oracle(oracle_lib):
    rng = np.random.default_rng(3)
    from laughlinedge_amd import geneig
    for d in (1, 2, 5, 17, 64, 131):
        print(f"case geneig d={d}", flush=True)
        A = rng.standard_normal((d, 2 * d)) + 1j * rng.standard_normal((d, 2 * d))
        S = A @ A.conj().T / (2 * d) + 0.1 * np.eye(d)
        B = rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))
        H = (B + B.conj().T) / 2
        ev, vec = geneig(S, H, vectors=True)
        ref = oracle_lib.geneig(S, H)
        assert np.allclose(ev, ref, rtol=1e-9, atol=1e-9)
        # residual and S-orthonormality
        r = H @ vec - S @ vec * ev[None, :]
        assert np.abs(r).max() < 1e-8 * max(1.0, np.abs(H).max())
        g = vec.conj().T @ S @ vec
        assert np.allclose(g, np.eye(d), atol=1e-8)


def test_geneig_values_only_path():
    """eigenvalue-only solves (Householder tridiagonalization + bisection) vs scipy.linalg.eigh
    up to the 640-state limit, a fully degenerate pencil H = c S, and a block-diagonal pencil
    whose Householder columns vanish"""
    from scipy.linalg import eigh
    from laughlinedge_amd import geneig
    rng = np.random.default_rng(8)
    for d in (1, 2, 3, 17, 131, 300, 640):
        print(f"case geneig values d={d}", flush=True)
        A = rng.standard_normal((d, 2 * d)) + 1j * rng.standard_normal((d, 2 * d))
        S = A @ A.conj().T / (2 * d) + 0.1 * np.eye(d)
        B = rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))
        H = (B + B.conj().T) / 2
        ev = geneig(S, H)
        ref = eigh(H, S, eigvals_only=True)
        assert np.allclose(ev, ref, rtol=0, atol=1e-9 * np.abs(ref).max()), d
    d = 40
    A = rng.standard_normal((d, 2 * d)) + 1j * rng.standard_normal((d, 2 * d))
    S = A @ A.conj().T / (2 * d) + 0.1 * np.eye(d)
    assert np.allclose(geneig(S, 3.25 * S), 3.25, rtol=1e-12)
    Hb = np.diag(rng.standard_normal(d)).astype(complex)
    Hb[:10, :10] += (B := rng.standard_normal((10, 10)) + 1j * rng.standard_normal((10, 10))) + B.conj().T
    ev = geneig(np.eye(d, dtype=complex), Hb)
    assert np.allclose(ev, np.linalg.eigvalsh(Hb), atol=1e-11)


def test_geneig_vectors_large_and_clustered():
    """eigenvectors (inverse iteration on the tridiagonal form + reflectors + L^-H): residual
    and S-orthonormality up to 640 states, for a fully degenerate pencil (one cluster of 60)
    and for clusters split by 1e-10 relative"""
    from laughlinedge_amd import geneig
    rng = np.random.default_rng(21)

    def check(S, H, tag):
        ev, V = geneig(S, H, vectors=True)
        scale = max(1.0, np.abs(H).max())
        r = np.abs(H @ V - S @ V * ev[None, :]).max()
        g = np.abs(V.conj().T @ S @ V - np.eye(len(ev))).max()
        print(f"{tag}: residual {r / scale:.2e}, S-orthonormality {g:.2e}", flush=True)
        assert r < 1e-8 * scale and g < 1e-8, tag
        return ev

    for d in (300, 640):
        A = rng.standard_normal((d, 2 * d)) + 1j * rng.standard_normal((d, 2 * d))
        S = A @ A.conj().T / (2 * d) + 0.1 * np.eye(d)
        B = rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))
        check(S, (B + B.conj().T) / 2, f"random d={d}")
    d = 60
    A = rng.standard_normal((d, 2 * d)) + 1j * rng.standard_normal((d, 2 * d))
    S = A @ A.conj().T / (2 * d) + 0.1 * np.eye(d)
    ev = check(S, 2.5 * S, "degenerate d=60")
    assert np.allclose(ev, 2.5, rtol=1e-12)
    Q, _ = np.linalg.qr(rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d)))
    lam = np.repeat(np.linspace(1.0, 4.0, 12), 5) * (1 + 1e-10 * rng.standard_normal(d))
    Hc = Q @ np.diag(lam) @ Q.conj().T
    ev = check(np.eye(d, dtype=complex), (Hc + Hc.conj().T) / 2, "clusters d=60")
    assert np.allclose(ev, np.sort(lam), atol=1e-12)


@pytest.mark.parametrize("sym", [0, 1, 2])
def test_sample_graphs_match_eager(sym):
    """Sampling rounds replayed from hipGraphs give exactly the eager rounds' moments, traces
    and walkers: several replays per buffer, a partial (eager) round, then a thinning change
    that forces a new capture, with Gram accumulation and block closing in between."""
    from laughlinedge_amd import Engine
    N, m, sectors, W = 24, 3, [1, 2, 3], 2048
    print(f"case graphs vs eager N=24 symmetry={sym}", flush=True)
    out = []
    for graphs in (0, 1):
        with Engine(N, m, sectors, w=0.05, eps=1e-4, walkers=W, samples=400, seed=9, tune_sweeps=20, step=0.5,
                    trace_walkers=32, n_blocks=3, graphs=graphs, harmonics=(1, 2), symmetry_moves=sym) as e:
            e.init_walkers()
            e.tune_step()
            e.burnin(20)
            spl = e.info.samples_per_launch
            e.sample(2 * spl + 3)
            e.set_thin(2)
            e.sample(2 * spl)
            e.sample(spl)
            st = e.stats()
            out.append(([e.moments(b) for b in range(e.n_blocks())], e.trace(), e.positions(),
                        e.block(0x100 | 1, 0, 1), st))
    (m0, t0, p0, h0, s0), (m1, t1, p1, h1, s1) = out
    assert s1["ms_sampler_graph"] > 0 and s0["ms_sampler_graph"] == 0
    assert np.array_equal(p0, p1)
    assert np.array_equal(t0, t1)
    assert np.array_equal(h0, h1)
    for b0, b1 in zip(m0, m1):
        for (S0, H0), (S1, H1) in zip(b0, b1):
            assert np.array_equal(S0, S1) and np.array_equal(H0, H1)


@pytest.mark.parametrize("sym", [0, 1, 2])
def test_sampler_r2_sum_rule(sym):
    """<sum_i r_i^2> = 2 (L0 + N) for |Psi_L|^2 (exact LLL identity), FP32 sampler; the
    centre of mass Z is Gaussian with <N |Z|^2> = 2 (with or without its heat bath).  Since
    the Jastrow factor is homogeneous of degree 2 L0, sum_i r_i^2 / 2 ~ Gamma(L0 + N) exactly:
    the walkers' sum r^2 has mean 2 (L0 + N) and variance 4 (L0 + N) (a dilation move with the
    wrong Gamma shape moves the mean by 2 per unit of shape, > 3 sigma here)."""
    from laughlinedge_amd import Engine
    N, m = 20, 3
    L0 = m * N * (N - 1) // 2
    print(f"case r2 sum rule N=20 symmetry={sym}", flush=True)
    # H with w = 2, eps = 0 is sum_i r_i^2; sector 0 has S = 1
    with Engine(N, m, [0], w=2.0, walkers=8192, precision=0, step=0.6, samples=200,
                burnin_sweeps=300, tune_sweeps=0, n_blocks=4, symmetry_moves=sym) as e:
        e.run()
        blocks = [e.moments(b)[0][1][0, 0].real for b in range(4)]
        st = e.stats()
        pos = e.positions()
    mean = np.mean(blocks)
    err = np.std(blocks, ddof=1) / 2
    exact = 2 * (L0 + N)
    assert abs(mean - exact) < 4 * err + 1e-4 * exact, (mean, exact, err)
    assert 0.2 < st["acceptance"] < 0.9
    z = pos[..., 0] + 1j * pos[..., 1]
    nz2 = N * np.abs(z.mean(axis=1)) ** 2
    assert abs(nz2.mean() - 2.0) < 5 * nz2.std() / np.sqrt(len(nz2)), nz2.mean()
    r2 = np.sum(np.abs(z) ** 2, axis=1)
    n = len(r2)
    mu, var = 2.0 * (L0 + N), 4.0 * (L0 + N)
    assert abs(r2.mean() - mu) < 5 * np.sqrt(var / n), (r2.mean(), mu)
    assert abs(r2.var(ddof=1) / var - 1.0) < 5 * np.sqrt(2.0 / n), (r2.var(ddof=1), var)


def test_harmonic_cross_blocks(oracle_lib):
    """sum_i cos/sin(L theta_i)-weighted blocks between sectors |dL2 - dL1| = L."""
    from laughlinedge_amd import Engine
    from laughlinedge_amd.engine import W_COS, W_SIN
    N, m, sectors = 12, 3, [0, 2, 4]
    print("case harmonics N=12 sectors=[0,2,4] L=[2,4]", flush=True)
    o = oracle_lib.Oracle(N, m, sectors, w=0.05)
    pos, _ = o.plain_positions(16, 50, 40, 0.6, 5)
    pos = pos.reshape(-1, N, 2)
    n = pos.shape[0]
    with Engine(N, m, sectors, w=0.05, walkers=64, samples=1, harmonics=[2, 4]) as e:
        e.accumulate_positions(pos)
        for L, pairs in ((2, [(0, 1), (1, 2)]), (4, [(0, 2)])):
            c, s_ = oracle_lib.harmonic_sums(pos, L)
            for (a, b) in pairs:
                for wid, wv in ((W_COS(L), c), (W_SIN(L), s_)):
                    ref = o.accumulate_weighted(pos, wv, a, b) / n
                    got = e.block(wid, a, b)
                    scale = np.sqrt(np.linalg.norm(o.accumulate_weighted(pos, np.ones(n), a, a)) *
                                    np.linalg.norm(o.accumulate_weighted(pos, np.ones(n), b, b))) / n * N
                    assert np.linalg.norm(got - ref) < 1e-6 * scale, (L, a, b)
                    back = e.block(wid, b, a)
                    assert np.allclose(back, got.conj().T)


def test_harmonic_blocks_five_harmonics(oracle_lib):
    """The feature kernel's widest harmonic instantiation (more than two harmonic pairs in one
    call, as excitationResponse_compute uses for three or more sectors): harmonics L = 1..5
    between sectors 0..5, every |dL2 - dL1| = L block's cos and sin weights against the oracle."""
    from laughlinedge_amd import Engine
    from laughlinedge_amd.engine import W_COS, W_SIN
    N, m, sectors = 12, 3, [0, 1, 2, 3, 4, 5]
    harm = [1, 2, 3, 4, 5]
    print("case harmonics N=12 sectors=0..5 L=1..5", flush=True)
    o = oracle_lib.Oracle(N, m, sectors, w=0.05)
    pos, _ = o.plain_positions(16, 50, 40, 0.6, 8)
    pos = pos.reshape(-1, N, 2)
    n = pos.shape[0]
    with Engine(N, m, sectors, w=0.05, walkers=64, samples=1, harmonics=harm) as e:
        e.accumulate_positions(pos)
        checked = 0
        for L in harm:
            c, s_ = oracle_lib.harmonic_sums(pos, L)
            for a in range(len(sectors)):
                for b in range(a + 1, len(sectors)):
                    if sectors[b] - sectors[a] != L:
                        continue
                    for wid, wv in ((W_COS(L), c), (W_SIN(L), s_)):
                        ref = o.accumulate_weighted(pos, wv, a, b) / n
                        got = e.block(wid, a, b)
                        scale = np.sqrt(np.linalg.norm(o.accumulate_weighted(pos, np.ones(n), a, a)) *
                                        np.linalg.norm(o.accumulate_weighted(pos, np.ones(n), b, b))) / n * N
                        assert np.linalg.norm(got - ref) < 1e-6 * scale, (L, a, b)
                        checked += 1
        assert checked == 2 * 15


def test_user_weight_blocks(oracle_lib):
    """user weights via the C callback, accumulated between every sector pair"""
    import ctypes
    import os
    from laughlinedge_amd import Engine
    from laughlinedge_amd.engine import W_USER
    N, m, sectors = 10, 3, [0, 1, 3]
    print("case user weights N=10", flush=True)
    helper = ctypes.CDLL(os.path.join(os.path.dirname(__file__), "cpp", "libuserweights.so"))
    R = np.sqrt(2.0 * m * N)
    helper.le_test_set_radius(ctypes.c_double(R))
    fn = ctypes.cast(helper.le_test_weight_fn, ctypes.c_void_p).value
    o = oracle_lib.Oracle(N, m, sectors)
    pos, _ = o.plain_positions(8, 50, 50, 0.6, 9)
    pos = pos.reshape(-1, N, 2)
    n = pos.shape[0]
    u0 = (pos[..., 0] ** 2).sum(-1) / R ** 2
    u1 = (pos[..., 0] ** 2 - pos[..., 1] ** 2).sum(-1) / R ** 2
    for prec in (0, 1):
        with Engine(N, m, sectors, walkers=64, samples=1, n_user_weights=2, precision=prec) as e:
            e.set_weight_callback(fn)
            e.accumulate_positions(pos)
            for k, u in ((0, u0), (1, u1)):
                for a in range(3):
                    for b in range(a, 3):
                        ref = o.accumulate_weighted(pos, u, a, b) / n
                        got = e.block(W_USER(k), a, b)
                        tol = 1e-6 if prec == 0 else 1e-12
                        assert np.linalg.norm(got - ref) <= tol * max(np.linalg.norm(ref), 1e-300) * 10, (prec, k, a, b)


def _read_tsv(path, cols=4):
    rows = [l.rstrip("\n").split("\t") for l in open(path)]
    for r in rows:
        assert len(r) == cols
        for v in r[2:]:
            assert len(v.split(".")[-1]) == 6, v     # "{:.6f}"
    n = int(np.sqrt(len(rows)))
    assert [(int(r[0]), int(r[1])) for r in rows] == [(i, j) for i in range(n) for j in range(n)]
    vals = np.array([[float(v) for v in r[2:]] for r in rows])
    return [(vals[:, 2 * k] + 1j * vals[:, 2 * k + 1]).reshape(n, n) for k in range((cols - 2) // 2)]


def test_cpp_shell_driver(oracle_lib, tmp_path):
    """The drop-in C++ interface: spectrum_compute, dsf_compute, excitationResponse_compute
    (device lambda) with the reference's TSV outputs, against the Python path and the oracle."""
    import os
    import subprocess
    from laughlinedge_amd import Engine
    from laughlinedge_amd.engine import W_COS, W_SIN
    print("case cpp shell driver N=10", flush=True)
    drv = os.path.join(os.path.dirname(__file__), "cpp", "shell_driver")
    build = tmp_path / "build"
    build.mkdir()
    out = subprocess.run([drv], cwd=build, capture_output=True, text=True, timeout=100)
    assert out.returncode == 0, out.stdout + out.stderr
    lines = {l.split()[0]: l.split()[1:] for l in out.stdout.splitlines() if l and l.split()[0].isupper()}
    ev = np.array([float(x) for x in lines["SPECTRUM"]])
    root = tmp_path / "output" / "N=10_m=3"
    H, = _read_tsv(root / "dL=4" / "Statistics" / "Hamiltonian" / "7.tsv")
    M, = _read_tsv(root / "dL=4" / "Statistics" / "Metric" / "7.tsv")
    assert np.allclose(np.diag(M).real, 1.0) and np.allclose(M, M.conj().T, atol=2e-6)
    assert np.allclose(H, H.conj().T, atol=2e-6)
    # the same engine run through the Python binding (same seed and parameters)
    step = float(lines["STEP"][0])
    # monteCarloParameters keeps randomStep as float (0.8f), as the reference does
    with Engine(10, 3, [4], w=0.05, eps=1e-4, walkers=512, samples=400, burnin_sweeps=200, seed=1234,
                tune_sweeps=100, step=float(np.float32(0.8))) as e:
        e.run()
        ref = e.spectra()[0]
        st = e.stats()
    assert abs(st["step"] - step) < 1e-6 * step
    assert np.allclose(ev, ref, rtol=1e-9)
    # eigenvalues from the 6-decimal TSVs agree to their rounding
    assert np.allclose(oracle_lib.geneig(M, H), ev, rtol=1e-4)
    L0 = 3 * 10 * 9 // 2
    assert np.all(np.abs(ev / (0.05 * (L0 + 4 + 10)) - 1) < 0.5)
    # the row-wise estimator through the shell: the Python path with the same parameters
    evr = np.array([float(x) for x in lines["ROWWISE"]])
    with Engine(10, 3, [4], w=0.05, eps=1e-4, walkers=64, samples=400, burnin_sweeps=200, seed=4321,
                tune_sweeps=100, step=float(np.float32(0.8)), estimator=1) as e:
        e.run()
        refr = e.spectra()[0]
        M_r, _ = e.moments()[0]
        assert abs(e.stats()["step"] - float(lines["ROWSTEP"][0])) < 1e-6 * e.stats()["step"]
    assert np.allclose(np.diag(M_r).real, 1.0, atol=1e-12)
    assert np.allclose(evr, refr, rtol=1e-9)
    assert np.allclose(evr, ev, rtol=0.02)     # same sector, other estimator and chains
    # the mixture estimator through the shell: the Python path with the same parameters
    evm = np.array([float(x) for x in lines["MIXTURE"]])
    with Engine(10, 3, [4], w=0.05, eps=1e-4, walkers=512, samples=400, burnin_sweeps=200, seed=2468,
                tune_sweeps=100, step=float(np.float32(0.8)), estimator=2) as e:
        e.run()
        refm = e.spectra()[0]
        assert e.get_mixture()[0] == 4
        assert abs(e.stats()["step"] - float(lines["MIXSTEP"][0])) < 1e-6 * e.stats()["step"]
    assert np.allclose(evm, refm, rtol=1e-9)
    assert np.allclose(evm, ev, rtol=0.02)
    # DSF: the shell's weights vs the oracle's rotation of the same blocks
    dsf = np.array([float(x) for x in lines["DSF"]]).reshape(-1, 2)
    with Engine(10, 3, [0, 3], w=0.05, eps=1e-4, walkers=512, samples=400, burnin_sweeps=200, seed=1234,
                tune_sweeps=100, step=step, harmonics=[3]) as e:
        e.run()
        mom = e.moments()
        nrm = np.concatenate([np.sqrt(np.diag(S).real) for S, _ in mom])
        D = len(nrm)
        Hf = np.zeros((D, D), complex); Mf = np.zeros((D, D), complex)
        VC = np.zeros((D, D), complex); VS = np.zeros((D, D), complex)
        Hf[0, 0], Mf[0, 0] = mom[0][1][0, 0], mom[0][0][0, 0]
        Hf[1:, 1:], Mf[1:, 1:] = mom[1][1], mom[1][0]
        VC[0:1, 1:] = e.block(W_COS(3), 0, 1); VC[1:, 0:1] = e.block(W_COS(3), 1, 0)
        VS[0:1, 1:] = e.block(W_SIN(3), 0, 1); VS[1:, 0:1] = e.block(W_SIN(3), 1, 0)
        nn = np.outer(nrm, nrm)
        en, w = oracle_lib.dsf_weights(Hf / nn, Mf / nn, VC / nn, VS / nn)
    assert dsf.shape == (3, 2)
    # the shell re-tunes its step from the stored one, so compare statistically
    assert np.allclose(dsf[:, 0], en, rtol=0.05)
    assert np.all(dsf[:, 1] >= 0) and np.all(np.isfinite(dsf[:, 1]))
    # response files: rho_{L=0} = N * metric on same-sector blocks (closed form)
    ex = tmp_path / "output" / "N=10_m=3" / "ExternalExcitation" / "Statistics"
    Mr, = _read_tsv(ex / "Metric" / "0.tsv")
    Ur, = _read_tsv(ex / "ExcitationPotential" / "0.tsv")
    RC0, RS0 = _read_tsv(ex / "EdgeDenstiy_FourierTransform" / "dL=0" / "0.tsv", cols=6)
    RC2, RS2 = _read_tsv(ex / "EdgeDenstiy_FourierTransform" / "dL=2" / "0.tsv", cols=6)
    assert Mr.shape == (1 + 2 + 5, 8)
    same = np.zeros((8, 8), bool)
    for a, b in ((0, 1), (1, 3), (3, 8)):
        same[a:b, a:b] = True
    assert np.allclose(RC0[same], 10 * Mr[same], atol=2e-5)
    assert np.allclose(RS0, 0, atol=1e-6)
    assert np.all(RC2[same] == 0)                    # selection rule |dL2 - dL1| = 2
    assert np.abs(RC2[0, 1:3]).max() > 0
    assert np.allclose(Ur, Ur.conj().T, atol=2e-6)
    # time evolution from those files vs the CPU RK-CK54 oracle on the same (rounded) matrices
    Hr, = _read_tsv(ex / "Hamiltonian" / "0.tsv")
    rhos = []
    for L in (0, 2, 4):
        c, s_ = _read_tsv(ex / "EdgeDenstiy_FourierTransform" / f"dL={L}" / "0.tsv", cols=6)
        rhos.append(c + 1j * s_)
    F = lambda t: 0.1 * np.exp(-((t - 2.5) / 1.0) ** 2)
    ref, nrm, _ = oracle_lib.evolve(Mr, Hr, Ur, rhos, F, 5.0, 0.5)
    for k, L in enumerate((0, 2, 4)):
        rows = [l.split("\t") for l in open(ex / "edgeDynamics" / f"dL={L}" / "0.tsv")]
        assert len(rows) == 10
        t = np.array([float(r[0]) for r in rows])
        z = np.array([float(r[1]) + 1j * float(r[2]) for r in rows])
        assert np.allclose(t, 0.5 * np.arange(10))
        assert np.allclose(z, ref[:, k], atol=2e-6), (L, np.abs(z - ref[:, k]).max())
    # occupation numbers: format, sum rule, and sum_l V_l O_ab(l) = H_ab (cross-module)
    occ_dir = root / "dL=2" / "Statistics"
    text = open(occ_dir / "OccupationNumbers_normalized" / "5.tsv").read()
    blocks = [b for b in text.split("\n\n") if b.strip()]
    assert len(blocks) == 4                          # d = 2: (a, b) pairs, blank line after each
    lcut = 3 * 9 + 15
    O = np.zeros((2, 2, lcut), complex)
    for blk in blocks:
        rows = [r.split("\t") for r in blk.strip().split("\n")]
        assert len(rows) == lcut and all(len(r) == 5 for r in rows)
        for a, b, l, re_, im_ in rows:
            O[int(a), int(b), int(l)] = float(re_) + 1j * float(im_)
    M2, = _read_tsv(occ_dir / "Metric" / "5.tsv")
    H2, = _read_tsv(occ_dir / "Hamiltonian" / "5.tsv")
    assert np.allclose(O.sum(-1), 10 * M2, atol=lcut * 1e-6 + 1e-5)
    l = np.arange(lcut)
    Vl = 0.05 * (l + 1) + 4e-4 * (l + 1) * (l + 2)
    assert np.allclose((O * Vl).sum(-1), H2, rtol=0.02, atol=0.02 * np.abs(H2).max())
    assert (occ_dir / "OccupationNumbers_not_normalized" / "5.tsv").exists()
    # spectral function (removal, dL' = 0..3 of 9 particles): orbitals 27..24, p(dL') rows each,
    # per-orbital weight sum = occupation n_l in (0, 1] (fermions), omega at dL' = 0 = E_0^10 - E_0^9
    rows = [r.split("\t") for r in open(root / "SpectralFunction" / "6.tsv").read().strip().split("\n")]
    assert all(len(r) == 4 for r in rows) and len(rows) == 1 + 1 + 2 + 3
    orb = np.array([int(r[0]) for r in rows])
    om = np.array([float(r[2]) for r in rows])
    wt = np.array([float(r[3]) for r in rows])
    assert list(orb) == [27, 26, 25, 25, 24, 24, 24]
    for l in (27, 26, 25, 24):
        assert 0 < wt[orb == l].sum() <= 1 + 1e-6, (l, wt[orb == l].sum())
    assert np.all(wt >= 0) and np.all(np.isfinite(om))
    e0 = []
    for n_ in (10, 9):
        with Engine(n_, 3, [0], w=0.05, eps=1e-4, walkers=2048, samples=400, burnin_sweeps=200, seed=5,
                    tune_sweeps=100, step=0.8) as e:
            e.run()
            e0.append(e.spectra()[0][0])
    assert abs(om[0] - (e0[0] - e0[1])) < 0.02 * abs(e0[0] - e0[1]), (om[0], e0)
    assert "DONE" in lines


@pytest.mark.parametrize("dims", [[1, 2, 5], [1, 40, 100]])
def test_evolve_gpu_vs_oracle(oracle_lib, dims):
    """FP64 GPU evolution (interaction picture, batched perturbations) vs the CPU RK-CK54
    restatement of dynamics.cpp, driven and free.  D = 141 spans three 64-row tiles of the
    stage products (the last partial), so both products of an off-diagonal tile are used."""
    from laughlinedge_amd.engine import evolve
    from tests.test_oracle import _block_system
    print(f"case evolve gpu dims={dims}", flush=True)
    rng = np.random.default_rng(8)
    M, H, U, rhos = _block_system(rng, dims)
    B = rng.normal(size=U.shape) + 1j * rng.normal(size=U.shape)
    U2 = (B + B.conj().T) / 2
    F = lambda t: 0.5 * np.exp(-(t - 2.0) ** 2 / (2 * 0.5 ** 2))
    rho, nrm, steps = evolve(M, H, [U, U2], rhos, F, 5.0, 0.25, dims)
    for u, Uu in enumerate((U, U2)):
        ref, rn, _ = oracle_lib.evolve(M, H, Uu, rhos, F, 5.0, 0.25)
        assert np.allclose(rho[u], ref, rtol=1e-6, atol=1e-6), (u, np.abs(rho[u] - ref).max())
        # the oracle integrates the fast phases and drifts by ~1e-8 per unit time; the GPU
        # removes them exactly, so its norm is conserved much more tightly
        assert np.allclose(nrm[u], rn, rtol=1e-6)
        assert np.allclose(nrm[u], M[0, 0].real, rtol=1e-8)      # unitarity in the M metric
    # free evolution is exact in the eigenbasis: compare with the closed form
    # C(t) = sum_j e^{-i E_j t} v_j (v_j^H M C0), here from a superposition start via M, H mixing
    import scipy.linalg
    rho0, nrm0, _ = evolve(M, H, [U], rhos, lambda t: 0.0, 5.0, 0.25, dims)
    E, V = scipy.linalg.eigh(H, M)
    a0 = V.conj().T @ M[:, 0]
    for k in range(20):
        C = V @ (np.exp(-1j * E * 0.25 * k) * a0)
        for l, R in enumerate(rhos):
            assert abs(rho0[0, k, l] - C.conj() @ R @ C) < 1e-10 * max(1.0, abs(C.conj() @ R @ C))
    ref0, _, _ = oracle_lib.evolve(M, H, U, rhos, lambda t: 0.0, 5.0, 0.25)
    assert np.allclose(rho0[0], ref0, rtol=1e-6, atol=1e-6)


def _occ_numpy(pos, N, m, states, lcut):
    """independent FP64 evaluation of the two-sided estimator on given (N+1)-particle samples"""
    from math import lgamma, pi
    R = np.sqrt(2.0 * m * N)
    z = pos[..., 0] + 1j * pos[..., 1]
    zl, zr = z[:, :N] / R, z[:, 1:] / R

    def mono(zz, pairs):
        v = np.ones(zz.shape[0], complex)
        for k, r in pairs:
            if k:
                v = v * np.sum(zz ** k, axis=1) ** r
        return v
    PL = np.stack([mono(zl, p) for _, p in states], 1)
    PR = np.stack([mono(zr, p) for _, p in states], 1)
    z0, zN, zb = z[:, 0], z[:, N], z[:, 1:N]
    q = np.prod((zb - zN[:, None]) * np.conj(z0[:, None] - zb) /
                np.abs((zb - zN[:, None]) * np.conj(z0[:, None] - zb)), axis=1)
    ph = (q / np.abs(q)) ** m
    x = 0.5 * z0 * np.conj(zN)
    g = np.exp(-0.25 * (np.abs(z0) ** 2 + np.abs(zN) ** 2))
    W = np.stack([ph * x ** l * g / (2 * pi * np.exp(lgamma(l + 1))) for l in range(lcut)], 1)
    return np.einsum("sa,sb,sl->abl", PL.conj(), PR, W) / pos.shape[0]


@pytest.mark.parametrize("N,dl,lcut", [(6, 3, 20), (6, 5, 70), (12, 10, 40)])
@pytest.mark.parametrize("prec,tol", [(0, 2e-5), (1, 1e-11), (2, 1e-4), (3, 2e-5)])
def test_occupation_fixed_samples(prec, tol, N, dl, lcut):
    """two-sided features + Gram on given configurations vs an independent numpy estimator.
    dL = 3 (6 real columns per side, 36 pairs) and dL = 5 (14 columns, 196 pairs in four
    64-pair tiles, 140 weights in three 64-weight tiles) take the many-weights kernel in FP32
    and FP64; dL = 10 at N = 12 (84 columns per side) and the BF16 split take one Gram launch
    over all weight pairs."""
    from laughlinedge_amd.engine import OccupationEngine
    from oracle.partitions import Basis
    m = 3
    print(f"case occupation fixed prec={prec} N={N} dl={dl} lcut={lcut}", flush=True)
    rng = np.random.default_rng(12)
    R = np.sqrt(2.0 * m * N)
    r = R * np.sqrt(rng.random((3000, N + 1)))
    t = 2 * np.pi * rng.random((3000, N + 1))
    pos = np.stack([r * np.cos(t), r * np.sin(t)], -1)
    ref = _occ_numpy(pos, N, m, Basis(N, [dl]).states, lcut)
    with OccupationEngine(N, m, dl, lcut, walkers=256, samples=1, precision=prec) as e:
        e.accumulate_positions(pos)
        got = e.occupation()
    assert np.linalg.norm(got - ref) < tol * np.linalg.norm(ref), np.linalg.norm(got - ref) / np.linalg.norm(ref)


def test_occupation_vs_exact_small_n():
    """N = 3, m = 3, dL = 3: two-sided chain -> normalised <a|n_l|b>/(|a||b|) vs the exact
    expansion, and the sum rule sum_l = N M_ab."""
    from laughlinedge_amd.engine import OccupationEngine
    from oracle.exact_lll import exact_occupation, exact_blocks
    from oracle.partitions import Basis
    N, m, dl = 3, 3, 3
    lcut = m * (N - 1) + dl + 1
    print("case occupation exact N=3", flush=True)
    b = Basis(N, [dl])
    Oe = np.array(exact_occupation(N, m, b.states, lcut))
    Se, _ = exact_blocks(N, m, b.states, 0.0, 0.0)
    Se = np.array(Se)
    nrm = np.sqrt(np.diag(Se))
    ref = Oe / np.outer(nrm, nrm)[:, :, None]
    Mexact = Se / np.outer(nrm, nrm)
    with OccupationEngine(N, m, dl, lcut, walkers=8192, samples=400, burnin_sweeps=200, step=0.9,
                          tune_sweeps=50, n_blocks=4) as e:
        e.run()
        blocks = [e.occupation(k) for k in range(4)]
        st = e.stats()
    def normalise(O):
        s = O.sum(-1)
        return O / s[:, :, None] * N * Mexact[:, :, None]
    vals = np.array([normalise(O) for O in blocks])
    mean = vals.mean(0)
    err = vals.std(0, ddof=1) / 2
    assert 0.2 < st["acceptance"] < 0.95
    assert np.allclose(mean.sum(-1), N * Mexact)
    assert np.all(np.abs(mean - ref) < 5 * err + 2e-3), np.abs(mean - ref).max()



GOLDEN = __import__("os").path.join(__import__("os").path.dirname(__import__("os").path.abspath(__file__)), "golden")


def test_geneig_golden():
    """GPU eigensolver against the committed scipy.linalg.eigh vectors"""
    from laughlinedge_amd import geneig
    g = np.load(GOLDEN + "/geneig_scipy.npz")
    print("case geneig golden", flush=True)
    for k in range(3):
        ev = geneig(g[f"S{k}"], g[f"H{k}"])
        assert np.allclose(ev, g[f"E{k}"], rtol=1e-10, atol=1e-10), k


@pytest.mark.parametrize("prec,sym", [(0, 0), (2, 0), (0, 1), (0, 2)])
def test_sampled_blocks_vs_exact_n4(prec, sym):
    """N = 4, m = 3, dL = 0..4: GPU chain + features + Gram against the exact matrices of
    the committed fixture, within 5 batch sigma per entry, and the spectra within 5 sigma;
    also with the centre-of-mass heat bath and random rotations (and the dilation heat bath)
    after every kept sample."""
    from laughlinedge_amd import Engine
    from laughlinedge_amd.stats import batch_sigma
    g = np.load(GOLDEN + "/exact_n4_m3.npz")
    N, m, w, eps = int(g["N"]), int(g["m"]), float(g["w"]), float(g["eps"])
    print(f"case sampled vs exact N=4 prec={prec} symmetry={sym}", flush=True)
    nb = 16     # batch errors of the heavy-tailed dL = 4 entries need more than 8 batches
    with Engine(N, m, [0, 1, 2, 3, 4], w=w, eps=eps, walkers=8192, samples=400, n_blocks=nb, seed=2024,
                tune_sweeps=40, step=0.5, precision=prec, symmetry_moves=sym) as e:
        e.init_walkers()
        e.tune_step()
        e.burnin(300)
        for b in range(nb):
            e.sample(25)
        blocks = [e.moments(b) for b in range(nb)]
        specs = [e.spectra(b) for b in range(nb)]
    for dl in range(5):
        for q, name in ((0, "S"), (1, "H")):
            arr = np.array([bl[dl][q] for bl in blocks])
            mean, sig = arr.mean(0), batch_sigma(arr.real) + 1j * batch_sigma(arr.imag)
            ex = g[f"{name}{dl}"]
            assert np.all(np.abs(mean.real - ex) <= 5 * sig.real + 1e-12 * np.abs(ex).max()), (dl, name)
            assert np.all(np.abs(mean.imag) <= 5 * sig.imag + 1e-12 * np.abs(ex).max()), (dl, name)
        from scipy.linalg import eigh
        ex_e = eigh(g[f"H{dl}"], g[f"S{dl}"], eigvals_only=True)
        sp = np.array([s[dl] for s in specs])
        assert np.all(np.abs(sp.mean(0) - ex_e) <= 5 * batch_sigma(sp) + 1e-12), (dl, sp.mean(0), ex_e)


def test_config0_full_run_matches_oracle(oracle_lib):
    """The small plain configuration in full: N = 10, m = 3, V = 0.025 r^2 + 1e-4 r^4,
    dL = 0..6, 64 walkers, seed 1234, uniform start in the disc of radius R, step 0.5,
    2000 burn-in sweeps, 2e4 samples per walker, FP64.  The GPU chain follows the oracle
    chain move for move (same Philox streams), so S, H, the acceptance and the spectra agree
    to rounding, not only statistically."""
    from laughlinedge_amd import Engine
    from oracle.oracle import geneig as or_geneig
    N, m, sectors = 10, 3, list(range(7))
    print("case config 0 full run (64 walkers x 22000 sweeps, FP64)", flush=True)
    o = oracle_lib.Oracle(N, m, sectors, w=0.05, eps=1e-4)
    So, Ho, acc_o, _, _ = o.plain_run(64, 2000, 20000, 0.5, 1234, threads=8)
    with Engine(N, m, sectors, w=0.05, eps=1e-4, walkers=64, precision=1, samples=20000, seed=1234,
                tune_sweeps=0, step=0.5, lanes_per_walker=2) as e:
        e.init_walkers()
        e.burnin(2000)
        e.sample(20000)
        mom = e.moments()
        acc_g = e.stats()["acceptance"]
        spec = e.spectra(-1)
    assert abs(acc_g - acc_o) < 1e-9, (acc_g, acc_o)
    for s, (Sg, Hg) in enumerate(mom):
        es = np.linalg.norm(Sg - So[s]) / np.linalg.norm(So[s])
        eh = np.linalg.norm(Hg - Ho[s]) / np.linalg.norm(Ho[s])
        assert es < 1e-10 and eh < 1e-10, (s, es, eh)
        assert np.allclose(spec[s], or_geneig(So[s], Ho[s]), rtol=1e-9)


@pytest.mark.parametrize("prec,N,G", [(0, 200, 8), (0, 200, 16), (0, 100, 8), (1, 200, 16),
                                      (0, 150, 8), (0, 60, 8), (1, 150, 16)])
def test_register_sweep_matches_lds(prec, N, G):
    """k_sweep_reg (each lane's particles in registers) reproduces k_sweep (particles in LDS)
    at the same lanes per walker: bit-identical positions and acceptance after burn-in and
    sampling, and the same moments (to 1e-12: the FP64 atomics of the Gram accumulation add
    workgroup results in a run-dependent order); 1003 walkers leave the last workgroup partly
    empty.  N = 150 at 8 lanes (19 slots) and N = 60 (8 slots) run the 25- and 13-slot kernels
    with their unused slots skipped, N = 150 at 16 lanes (10 slots) the 13-slot FP64 kernel."""
    from laughlinedge_amd import Engine
    print(f"case register sweep vs LDS prec={prec} N={N} G={G}", flush=True)
    out = []
    for reg in (0, 1):
        with Engine(N, 3, [1, 2, 3], w=0.05, eps=5e-5, walkers=1003, samples=12, seed=17, tune_sweeps=0,
                    step=0.9, lanes_per_walker=G, sweep_registers=reg, precision=prec) as e:
            assert e.info.sweep_in_registers == reg and e.info.lanes_per_walker == G
            e.init_walkers()
            e.burnin(20)
            e.sample(12)
            out.append((e.positions(), e.stats()["acceptance"], e.moments()))
    (p0, a0, m0), (p1, a1, m1) = out
    assert np.array_equal(p0, p1)
    assert a0 == a1 and 0.2 < a0 < 0.9
    for (S0, H0), (S1, H1) in zip(m0, m1):
        assert np.allclose(S0, S1, rtol=1e-12, atol=1e-12 * np.abs(S0).max())
        assert np.allclose(H0, H1, rtol=1e-12, atol=1e-12 * np.abs(H0).max())


def test_autotuned_sweep_sum_rule():
    """lanes_per_walker = -1 (the bench's setting): whichever sweep the autotuner keeps (LDS or
    registers) samples |Psi_L|^2 -- the walkers' sum r^2 has the exact Gamma(L0 + N) law."""
    from laughlinedge_amd import Engine
    N, m = 200, 3
    L0 = m * N * (N - 1) // 2
    print("case autotuned sweep N=200 sum rule", flush=True)
    with Engine(N, m, [1], w=0.05, eps=5e-5, walkers=8192, samples=4, seed=21, tune_sweeps=30, step=0.5,
                lanes_per_walker=-1) as e:
        e.init_walkers()
        e.tune_step()
        e.burnin(200)
        print(f"lanes {e.info.lanes_per_walker} registers {e.info.sweep_in_registers}")
        assert e.info.lanes_per_walker in (4, 8, 16, 32, 64)
        z = e.positions()
    r2 = np.sum(z[..., 0] ** 2 + z[..., 1] ** 2, axis=1)
    mu, var, n = 2.0 * (L0 + N), 4.0 * (L0 + N), len(r2)
    assert abs(r2.mean() - mu) < 5 * np.sqrt(var / n), (r2.mean(), mu)
    assert abs(r2.var(ddof=1) / var - 1.0) < 5 * np.sqrt(2.0 / n), (r2.var(ddof=1), var)


def test_register_sweep_graphs_match_eager():
    """Sampling rounds with the register-resident sweep replayed from hipGraphs give exactly
    the eager rounds' positions, traces and moments (N = 100, 8 lanes per walker, thin 1 and 2)."""
    from laughlinedge_amd import Engine
    print("case register sweep graphs vs eager N=100", flush=True)
    out = []
    for graphs in (0, 1):
        with Engine(100, 3, [1, 2, 3, 4], w=0.05, eps=5e-5, walkers=4096, samples=40, seed=31, tune_sweeps=0,
                    step=0.8, trace_walkers=32, graphs=graphs, lanes_per_walker=8, sweep_registers=1) as e:
            assert e.info.sweep_in_registers == 1
            e.init_walkers()
            e.burnin(10)
            spl = e.info.samples_per_launch
            e.sample(2 * spl)
            e.set_thin(2)
            e.sample(spl)
            out.append((e.positions(), e.trace(), e.moments(), e.stats()["ms_sampler_graph"]))
    (p0, t0, m0, g0), (p1, t1, m1, g1) = out
    assert g0 == 0 and g1 > 0
    assert np.array_equal(p0, p1) and np.array_equal(t0, t1)
    for (S0, H0), (S1, H1) in zip(m0, m1):
        assert np.allclose(S0, S1, rtol=1e-12, atol=0) and np.allclose(H0, H1, rtol=1e-12, atol=0)


def test_autotuned_lanes_per_chain_kind():
    """The lane autotuner caches its choice per (N, walkers, precision, chain kind): an edge
    engine's mapping (registers need ceil(N / lanes) = 13 or 25) is not reused by the two-sided
    chain, whose register sweep needs ceil((N - 1) / lanes) = 13 or 25."""
    from laughlinedge_amd import Engine
    from laughlinedge_amd.engine import OccupationEngine
    N, m, W = 100, 3, 2048
    print("case autotune edge then occupation N=100", flush=True)
    with Engine(N, m, [1, 2], walkers=W, samples=2, lanes_per_walker=-1, tune_sweeps=10, step=0.5) as e:
        e.init_walkers()
        e.tune_step()
    with OccupationEngine(N, m, 0, 20, walkers=W, samples=2, lanes_per_walker=-1, tune_sweeps=10, step=0.5) as o:
        o.init_walkers()
        o.tune_step()
        o.burnin(2)
        o.sample(1)
        G = o.info.lanes_per_walker
        assert G in (4, 8, 16, 32, 64)
        assert not o.info.sweep_in_registers or (G in (8, 16) and -(-(N - 1) // G) in (13, 25))


@pytest.mark.parametrize("G,prec", [(8, 0), (16, 0), (8, 1), (16, 1)])
def test_twosided_register_sweep_matches_lds(G, prec):
    """The two-sided chain with its bulk particles in registers (N = 200: 199 bulk particles in
    25 slots of 8 lanes or 13 slots of 16) gives the LDS kernel's positions bit for bit."""
    from laughlinedge_amd.engine import OccupationEngine
    N, m = 200, 3
    print(f"case two-sided register sweep G={G} prec={prec}", flush=True)
    out = []
    for reg in (0, 1):
        with OccupationEngine(N, m, 0, 10, walkers=256, samples=1, precision=prec, lanes_per_walker=G,
                              sweep_registers=reg, step=0.9, seed=11) as e:
            assert e.info.sweep_in_registers == reg
            e.init_walkers()
            p0 = e.positions()
            e.burnin(3)
            out.append((p0, e.positions()))
    assert np.array_equal(out[0][0], out[1][0])          # same start
    assert np.array_equal(out[0][1], out[1][1])          # same chain
    moved = np.mean(np.any(out[0][1] != out[0][0], axis=-1))
    assert 0.5 < moved < 1.0, moved                      # three sweeps moved most particles
