"""The batched-sector configuration at full size: nu = 1/3, N = 200, dL = 1..16 accumulated
together (D = 914, sector sizes 1..231, one grouped Gram over 97 tiles), 32768 walkers,
seed 42, V = 0.025 r^2 + 5e-5 r^4.

- Fixed samples: all 32768 walkers' configurations after burn-in go through the FP32, BF16
  hi/mid/lo, BF16 hi/lo and FP64 accumulations and the FP64 CPU oracle (or_accumulate_fixed,
  restating em_spectrum.cu:98-115 for the plain estimator): every sector's S and H within 1e-6
  relative Frobenius (FP32, BF16 hi/mid/lo), 2e-5 (hi/lo), 1e-12 (FP64); every sector's
  generalized eigenvalues (GPU FP64 eigensolver on the GPU moments) within 2e-5 relative of the
  oracle's eigensolve of the oracle's moments (1e-10 for FP64).
- Sampler law at full size: sum_i r_i^2 / 2 ~ Gamma(L0 + N) exactly for |Psi_L|^2, so over the
  walkers the mean is 2 (L0 + N) and the variance 4 (L0 + N).
- Sampling rounds replayed from hipGraphs leave the same walkers and traces, bit for bit, as
  eager rounds, and the same moments up to the order of the FP64 atomic adds.
- Every placement of the Gram tiles over the XCDs (gram_xcd_spread 1, 2, 4, 8) and both tile
  widths (64x64 and 64x32 tiles) accumulate the same FP64 moments up to the order of the FP64
  atomic adds; the 64x32 tiles also meet the oracle at FP32 and FP64; the autotuned engine
  reports the placement and width it chose."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N, M_INV, SECTORS = 200, 3, list(range(1, 17))
W_CONF, EPS_CONF = 0.05, 5e-5
WALKERS, SEED = 32768, 42
L0 = M_INV * N * (N - 1) // 2


def _engine(**kw):
    from laughlinedge_amd import Engine
    base = dict(w=W_CONF, eps=EPS_CONF, walkers=WALKERS, seed=SEED, samples=16, tune_sweeps=50, step=0.5)
    base.update(kw)
    return Engine(N, M_INV, SECTORS, **base)


@pytest.fixture(scope="module")
def chain():
    print("case c2 N=200 dL=1..16 32768 walkers seed 42: tune, burn-in, sample", flush=True)
    e = _engine(lanes_per_walker=-1)
    assert e.info.D == 914 and max(e.dims) == 231 and e.info.n_tiles == 97
    e.init_walkers()
    e.tune_step()
    e.burnin(300)
    e.sample(8)
    pos = e.positions()
    step = e.stats()["step"]
    e.close()
    return pos, step


def test_c2_sampler_gamma_law(chain):
    pos, _ = chain
    r2 = np.sum(pos[..., 0] ** 2 + pos[..., 1] ** 2, axis=1)
    mu, var, n = 2.0 * (L0 + N), 4.0 * (L0 + N), len(r2)
    print(f"sum r^2: mean {r2.mean():.2f} (exact {mu}), var {r2.var(ddof=1):.0f} (exact {var})")
    assert abs(r2.mean() - mu) < 5 * np.sqrt(var / n)
    assert abs(r2.var(ddof=1) / var - 1.0) < 5 * np.sqrt(2.0 / n)


@pytest.mark.parametrize("prec,tol,etol,cols", [(0, 1e-6, 2e-5, 64), (3, 1e-6, 2e-5, 64), (2, 2e-5, 2e-4, 64),
                                                (1, 1e-12, 1e-10, 64), (0, 1e-6, 2e-5, 32), (1, 1e-12, 1e-10, 32)])
def test_c2_fixed_samples_vs_oracle(chain, oracle_lib, prec, tol, etol, cols):
    from laughlinedge_amd import Engine
    pos, _ = chain
    n = pos.shape[0]
    print(f"case c2 fixed samples precision={prec} tile columns={cols} ({n} configurations)", flush=True)
    o = oracle_lib.Oracle(N, M_INV, SECTORS, w=W_CONF, eps=EPS_CONF)
    So, Ho, _ = o.accumulate_fixed(pos)
    with Engine(N, M_INV, SECTORS, w=W_CONF, eps=EPS_CONF, walkers=1024, samples=1, precision=prec,
                gram_tile_cols=cols) as f:
        assert f.info.gram_tile_cols == cols
        f.accumulate_positions(pos)
        mom = f.moments()
        spec = f.spectra()
    worst, eworst = 0.0, 0.0
    for s, (Sg, Hg) in enumerate(mom):
        S64, H64 = So[s] / n, Ho[s] / n
        rs = np.linalg.norm(Sg - S64) / np.linalg.norm(S64)
        rh = np.linalg.norm(Hg - H64) / np.linalg.norm(H64)
        worst = max(worst, rs, rh)
        assert rs < tol and rh < tol, (SECTORS[s], rs, rh)
        e64 = oracle_lib.geneig(S64, H64)
        assert spec[s] is not None, SECTORS[s]
        er = np.max(np.abs(spec[s] - e64) / np.abs(e64))
        eworst = max(eworst, er)
        assert er < etol, (SECTORS[s], er)
    print(f"worst relative Frobenius error {worst:.3e}; worst eigenvalue relative error {eworst:.3e}")


def test_c2_graph_replay_matches_eager(chain):
    _, step = chain
    print("case c2 graphs vs eager at full size", flush=True)
    out = []
    for graphs in (0, 1):
        with _engine(tune_sweeps=0, step=step, graphs=graphs, trace_walkers=256, lanes_per_walker=8,
                     sweep_registers=1) as e:
            e.init_walkers()
            e.burnin(100)
            spl = e.info.samples_per_launch
            e.sample(2 * spl)
            st = e.stats()
            out.append((e.positions(), e.trace(), e.moments(), st))
    (p0, t0, m0, s0), (p1, t1, m1, s1) = out
    assert s1["ms_sampler_graph"] > 0 and s0["ms_sampler_graph"] == 0
    assert np.array_equal(p0, p1)
    assert np.array_equal(t0, t1)
    for (S0, H0), (S1, H1) in zip(m0, m1):
        assert np.linalg.norm(S0 - S1) <= 1e-12 * np.linalg.norm(S0)
        assert np.linalg.norm(H0 - H1) <= 1e-12 * np.linalg.norm(H0)


def test_c2_gram_placements_agree(chain):
    from laughlinedge_amd import Engine
    pos, step = chain
    print("case c2 Gram tile placements over the XCDs", flush=True)
    # FP32 tiles over the four placements (same row splits: only the FP64 atomic order differs);
    # FP64 tiles over both widths (the 64x32 tiles split the rows elsewhere)
    for prec, cases in ((0, ((8, 64), (4, 64), (2, 64), (1, 64))), (1, ((8, 64), (8, 32), (1, 32)))):
        ref = None
        for K, cols in cases:
            with Engine(N, M_INV, SECTORS, w=W_CONF, eps=EPS_CONF, walkers=WALKERS, samples=1, gram_xcd_spread=K,
                        gram_tile_cols=cols, precision=prec) as f:
                f.accumulate_positions(pos)
                mom = f.moments()
            if ref is None:
                ref = mom
                continue
            for (S0, H0), (S1, H1) in zip(ref, mom):
                assert np.linalg.norm(S1 - S0) <= 1e-12 * np.linalg.norm(S0), (prec, K, cols)
                assert np.linalg.norm(H1 - H0) <= 1e-12 * np.linalg.norm(H0), (prec, K, cols)
    with _engine(tune_sweeps=0, step=step, lanes_per_walker=8, sweep_registers=1) as e:
        e.init_walkers()
        e.burnin(20)
        e.sample(e.info.samples_per_launch)
        assert e.info.gram_xcd_spread in (1, 2, 4, 8) and e.info.gram_tile_cols in (32, 64)
        print(f"autotuned placement: tiles of a row split on {e.info.gram_xcd_spread} XCDs, "
              f"{e.info.gram_tile_cols}-column tiles")
