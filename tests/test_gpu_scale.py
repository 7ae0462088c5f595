"""GPU parity at the largest configuration: nu = 1/5 (m = 5), N = 400, dL = 0..20 (D = 2714,
sectors up to 627 states, K = 20), 65536 walkers, eight edge-kick weights
U_k = sum_i cos(k theta_i) exp(-(r_i - R)^2 / 2), k = 1..8, accumulated between every sector pair.

Checked at full size: fixed-sample S, H and kick blocks against the FP64 oracle; the GPU
eigensolver on the sampled 627-state sector against scipy; the batched FP64 time evolution
of the eight kicks (A = 1e-3, t0 = 20, tau = 5, 0 <= t <= 100, saved every 0.1) conserves
C^H M C and responds linearly in A."""
import ctypes
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N, M_INV, SECTORS = 400, 5, list(range(21))
W_CONF, EPS_CONF = 0.05, 2e-5
WALKERS = 65536
R = float(np.sqrt(2.0 * M_INV * N))


def _kicks():
    helper = ctypes.CDLL(os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "libuserweights.so"))
    helper.le_test_set_kicks(8, ctypes.c_double(R))
    return helper, ctypes.cast(helper.le_test_kick_fn, ctypes.c_void_p).value


def _kick_numpy(pos):
    z = pos[..., 0] + 1j * pos[..., 1]
    g = np.exp(-0.5 * (np.abs(z) - R) ** 2)
    th = np.angle(z)
    return [np.sum(np.cos(k * th) * g, axis=1) for k in range(1, 9)]


@pytest.fixture(scope="module")
def sampled():
    from laughlinedge_amd import Engine
    print("case scale m=5 N=400 dL=0..20 65536 walkers: sampling", flush=True)
    helper, fn = _kicks()
    e = Engine(N, M_INV, SECTORS, w=W_CONF, eps=EPS_CONF, walkers=WALKERS, samples=8, seed=2024,
               tune_sweeps=20, step=0.5, n_user_weights=8)
    e.set_weight_callback(fn)
    assert e.info.D == 2714 and max(e.dims) == 627
    e.init_walkers()
    e.tune_step()
    e.burnin(150)
    e.sample(8)
    yield e, helper
    e.close()


def test_scale_fixed_samples_vs_oracle(sampled, oracle_lib):
    from laughlinedge_amd import Engine
    from laughlinedge_amd.engine import W_USER
    e, helper = sampled
    pos = e.positions()[:256]
    print("case scale fixed samples vs oracle (256 configurations)", flush=True)
    o = oracle_lib.Oracle(N, M_INV, SECTORS, w=W_CONF, eps=EPS_CONF)
    So, Ho, _ = o.accumulate_fixed(pos)
    n = pos.shape[0]
    _, fn = _kicks()
    with Engine(N, M_INV, SECTORS, w=W_CONF, eps=EPS_CONF, walkers=256, samples=1, n_user_weights=8) as f:
        f.set_weight_callback(fn)
        f.accumulate_positions(pos)
        mom = f.moments()
        worst = 0.0
        for s, (Sg, Hg) in enumerate(mom):
            rs = np.linalg.norm(Sg - So[s] / n) / np.linalg.norm(So[s] / n)
            rh = np.linalg.norm(Hg - Ho[s] / n) / np.linalg.norm(Ho[s] / n)
            worst = max(worst, rs, rh)
            assert rs < 1e-6 and rh < 1e-6, (s, rs, rh)
        print(f"worst S/H relative Frobenius error {worst:.3e}")
        us = _kick_numpy(pos)
        for k in (0, 3, 7):
            for a, b in ((20, 19), (3, 11), (0, 20), (17, 17)):
                ref = o.accumulate_weighted(pos, us[k], a, b) / n
                got = f.block(W_USER(k), a, b)
                err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
                assert err < 1e-5, (k, a, b, err)


def test_scale_spectra_vs_scipy(sampled):
    from scipy.linalg import eigh
    e, _ = sampled
    print("case scale spectra (627-state sector) vs scipy", flush=True)
    spec = e.spectra(-1)
    assert all(sp is not None for sp in spec)
    mom = e.moments()
    for s in (12, 20):
        S, H = mom[s]
        S = (S + S.conj().T) / 2
        H = (H + H.conj().T) / 2
        ref = eigh(H, S, eigvals_only=True)
        assert np.allclose(spec[s], ref, rtol=0, atol=1e-7 * np.abs(ref).max()), (s, np.abs(spec[s] - ref).max())
    # the quadratic part alone gives w (L0 + dL + N) exactly; eps r^4 only raises it
    L0 = M_INV * N * (N - 1) // 2
    for s in range(len(SECTORS)):
        assert spec[s][0] > W_CONF * (L0 + SECTORS[s] + N) * (1 - 1e-4), s


def _matrices(e):
    """dense M, H and the eight Hermitian kick matrices U_k of an engine's mean moments"""
    from laughlinedge_amd.engine import W_USER
    dims, offs = e.dims, e.offsets
    D = e.info.D
    M = np.zeros((D, D), complex)
    H = np.zeros((D, D), complex)
    for s, (S, Hs) in enumerate(e.moments()):
        o, d = offs[s], dims[s]
        M[o:o + d, o:o + d] = (S + S.conj().T) / 2
        H[o:o + d, o:o + d] = (Hs + Hs.conj().T) / 2
    Us = []
    for k in range(8):
        U = np.zeros((D, D), complex)
        for a in range(len(SECTORS)):
            for b in range(a, len(SECTORS)):
                blk = e.block(W_USER(k), a, b)
                U[offs[a]:offs[a] + dims[a], offs[b]:offs[b] + dims[b]] = blk
                U[offs[b]:offs[b] + dims[b], offs[a]:offs[a] + dims[a]] = blk.conj().T
        Us.append((U + U.conj().T) / 2)
    return M, H, Us


def test_scale_fp32_response_vs_fp64(sampled):
    """The FP32-MFMA response matrices and the edge density response they drive, against the
    plain FP64 form on the same 65536 configurations (dynamics.cpp:286-290 integrates whatever
    matrices the sampling wrote): M, H and the eight kick matrices U_k within 1e-6 relative
    (Frobenius); rho_1(t) = C^H U_1 C of every kick over 0 <= t <= 100 within a tenth of its
    Monte-Carlo error, estimated from the FP64 response of the two halves of the configurations
    (the pencils are ill-conditioned, so 1e-7 in the matrices moves rho by ~1e-2 of its response,
    while the sampling error of 65536 configurations is far larger)."""
    from laughlinedge_amd import Engine
    from laughlinedge_amd.engine import evolve
    e, _ = sampled
    pos = e.positions()
    print(f"case scale FP32 vs FP64 response on {pos.shape[0]} configurations", flush=True)
    _, fn = _kicks()
    mats = {}
    half = pos.shape[0] // 2
    for key, prec, p in ((0, 0, pos), (1, 1, pos), ("a", 1, pos[:half]), ("b", 1, pos[half:])):
        with Engine(N, M_INV, SECTORS, w=W_CONF, eps=EPS_CONF, walkers=4096, samples=1, n_user_weights=8,
                    precision=prec) as f:
            f.set_weight_callback(fn)
            f.accumulate_positions(p)
            mats[key] = _matrices(f)
            dims = f.dims
    (M32, H32, U32), (M64, H64, U64) = mats[0], mats[1]
    rel = lambda a, b: np.linalg.norm(a - b) / np.linalg.norm(b)
    errs = [rel(M32, M64), rel(H32, H64)] + [rel(a, b) for a, b in zip(U32, U64)]
    print(f"relative Frobenius FP32 vs FP64: M {errs[0]:.2e} H {errs[1]:.2e} U_k max {max(errs[2:]):.2e}")
    assert max(errs) < 1e-6, errs
    prof = lambda t: 1e-3 * np.exp(-(t - 20.0) ** 2 / (2 * 25.0))
    r32, _, _ = evolve(M32, H32, U32, [U32[0]], prof, 100.0, 0.1, dims)
    r64, _, _ = evolve(M64, H64, U64, [U64[0]], prof, 100.0, 0.1, dims)
    ra, _, _ = evolve(*mats["a"], [mats["a"][2][0]], prof, 100.0, 0.1, dims)
    rb, _, _ = evolve(*mats["b"], [mats["b"][2][0]], prof, 100.0, 0.1, dims)
    resp = np.abs(r64[:, :, 0] - r64[:, :1, 0]).max(axis=1)
    dev = np.abs(r32[:, :, 0] - r64[:, :, 0]).max(axis=1)
    mc = 0.5 * np.abs(ra[:, :, 0] - rb[:, :, 0]).max(axis=1)
    print(f"rho_1(t) response per kick {np.array2string(resp, precision=3)}; max |FP32 - FP64| "
          f"{np.array2string(dev, precision=3)}; Monte-Carlo error (halves) {np.array2string(mc, precision=3)}")
    assert np.all(dev <= 0.1 * mc), (dev, mc)


def test_scale_evolution_of_kicks(sampled):
    from laughlinedge_amd.engine import W_USER, evolve
    e, _ = sampled
    print("case scale FP64 evolution of 8 kicks, D=2714", flush=True)
    dims, offs = e.dims, e.offsets
    D = e.info.D
    M = np.zeros((D, D), complex)
    H = np.zeros((D, D), complex)
    for s, (S, Hs) in enumerate(e.moments()):
        o, d = offs[s], dims[s]
        M[o:o + d, o:o + d] = (S + S.conj().T) / 2
        H[o:o + d, o:o + d] = (Hs + Hs.conj().T) / 2
    Us = []
    for k in range(8):
        U = np.zeros((D, D), complex)
        for a in range(len(SECTORS)):
            for b in range(a, len(SECTORS)):
                blk = e.block(W_USER(k), a, b)
                U[offs[a]:offs[a] + dims[a], offs[b]:offs[b] + dims[b]] = blk
                U[offs[b]:offs[b] + dims[b], offs[a]:offs[a] + dims[a]] = blk.conj().T
        Us.append((U + U.conj().T) / 2)
    # observable: the k = 1 kick operator itself (density response at angular momentum 1)
    rhos = [Us[0]]
    prof = lambda A: (lambda t: A * np.exp(-(t - 20.0) ** 2 / (2 * 25.0)))
    rho1, nrm1, steps1 = evolve(M, H, Us, rhos, prof(1e-3), 100.0, 0.1, dims)
    rho2, nrm2, _ = evolve(M, H, Us, rhos, prof(2e-3), 100.0, 0.1, dims)
    print(f"controller steps {steps1}")
    assert rho1.shape == (8, 1000, 1) and np.all(np.isfinite(rho1))
    # C^H M C is conserved (C(0) = e_0, M_00 = 1 for the constant dL = 0 state)
    assert abs(M[0, 0] - 1.0) < 1e-12
    assert np.max(np.abs(nrm1 - 1.0)) < 1e-7 and np.max(np.abs(nrm2 - 1.0)) < 1e-7
    # linear response of the k = 1 kick (it couples dL = 0 to dL = 1, which rho_1 reads at
    # first order): the change of <rho_1> doubles with A
    d1 = rho1[0, :, 0] - rho1[0, 0, 0]
    d2 = rho2[0, :, 0] - rho2[0, 0, 0]
    assert np.abs(d2).max() > 1e-9
    ratio = np.abs(d2 - 2 * d1).max() / np.abs(d2).max()
    print(f"linear-response deviation {ratio:.2e}")
    assert ratio < 1e-2, ratio


@pytest.fixture(scope="module")
def config1_samples(oracle_lib):
    """nu = 1/3, N = 100, V = 0.025 r^2 + 5e-5 r^4, dL = 0..12 (D = 272), 16384 walkers, seed 7:
    1,048,576 configurations of the GPU chain (64 snapshots of the walkers, 2 sweeps apart, after
    tuning and 200 burn-in sweeps) and their FP64 oracle sums (or_accumulate_fixed)."""
    from laughlinedge_amd import Engine
    n100, secs = 100, list(range(13))
    print("case config 1: 1,048,576 configurations and their FP64 oracle sums", flush=True)
    with Engine(n100, 3, secs, w=0.05, eps=5e-5, walkers=16384, samples=4, seed=7, tune_sweeps=30,
                step=0.5) as e:
        assert e.info.D == 272
        e.init_walkers()
        e.tune_step()
        e.burnin(200)
        pos = []
        for _ in range(64):
            e.burnin(2)
            pos.append(e.positions())
    pos = np.concatenate(pos)
    assert pos.shape == (1 << 20, n100, 2)
    o = oracle_lib.Oracle(n100, 3, secs, w=0.05, eps=5e-5)
    So, Ho, _ = o.accumulate_fixed(pos)
    return pos, So, Ho


@pytest.mark.parametrize("prec,tol", [(0, 1e-6), (2, 2e-5), (3, 1e-6)])
def test_config1_fixed_samples(config1_samples, prec, tol):
    """The configuration-1 fixed-sample comparison at its configured size: the 1,048,576
    configurations through the FP32 (or a BF16-split) accumulation and through the FP64 oracle;
    every sector's S and H within tol relative Frobenius."""
    from laughlinedge_amd import Engine
    pos, So, Ho = config1_samples
    n100, secs = 100, list(range(13))
    print(f"case config 1 fixed samples prec={prec}", flush=True)
    n = pos.shape[0]
    with Engine(n100, 3, secs, w=0.05, eps=5e-5, walkers=16384, samples=1, precision=prec) as f:
        f.accumulate_positions(pos)
        mom = f.moments()
    worst = 0.0
    for s, (Sg, Hg) in enumerate(mom):
        es = np.linalg.norm(Sg - So[s] / n) / np.linalg.norm(So[s] / n)
        eh = np.linalg.norm(Hg - Ho[s] / n) / np.linalg.norm(Ho[s] / n)
        worst = max(worst, es, eh)
        assert es < tol and eh < tol, (s, es, eh)
    print(f"worst relative Frobenius error {worst:.3e} over {n} configurations")


def test_dsf_harmonic_blocks_n200(oracle_lib):
    """DSF blocks at the batched-sector scale: N = 200, m = 3, sectors dL = 0 and 16 (1 + 231
    states), harmonics L = 8 and 16; 2048 configurations of the GPU chain.  The cos/sin(L theta)
    weighted blocks between dL = 0 and dL = L match the FP64 oracle (1e-5 of the block scale),
    and the reverse block is the conjugate transpose."""
    from laughlinedge_amd import Engine
    from laughlinedge_amd.engine import W_COS, W_SIN
    N, m, secs = 200, 3, [0, 8, 16]
    print("case DSF harmonic blocks N=200 dL=0,8,16 L=8,16", flush=True)
    with Engine(N, m, secs, w=0.05, eps=5e-5, walkers=2048, samples=1, seed=13, tune_sweeps=30, step=0.5) as e:
        e.init_walkers()
        e.tune_step()
        e.burnin(100)
        pos = e.positions()
    o = oracle_lib.Oracle(N, m, secs, w=0.05, eps=5e-5)
    n = pos.shape[0]
    with Engine(N, m, secs, w=0.05, eps=5e-5, walkers=n, samples=1, harmonics=[8, 16]) as f:
        f.accumulate_positions(pos)
        for L, (a, b) in ((8, (0, 1)), (16, (0, 2)), (8, (1, 2))):
            c, s_ = oracle_lib.harmonic_sums(pos, L)
            scale = np.sqrt(np.linalg.norm(o.accumulate_weighted(pos, np.ones(n), a, a)) *
                            np.linalg.norm(o.accumulate_weighted(pos, np.ones(n), b, b))) / n * N
            for wid, wv in ((W_COS(L), c), (W_SIN(L), s_)):
                ref = o.accumulate_weighted(pos, wv, a, b) / n
                got = f.block(wid, a, b)
                err = np.linalg.norm(got - ref) / scale
                assert err < 1e-5, (L, a, b, err)
                assert np.allclose(f.block(wid, b, a), got.conj().T)
