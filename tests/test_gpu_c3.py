"""The learned proposal at its configured size: nu = 1/3, N = 200, dL = 1..16 (D = 914),
32768 walkers, seed 42; the 3 -> 64 -> 64 -> 2 tanh MLP (features |z|/R and the minimum and
mean distance to the 8 nearest neighbours) initialised from seed 99 and trained on the GPU for
2000 Adam steps (lr 1e-3) on 1e5 configurations of the plain chain.

- Exactness at full size: the walkers' sum_i r_i^2 / 2 follows Gamma(L0 + N) (mean and
  variance), and the spectra of dL = 1..4 agree with the plain chain's within 4.5 combined
  batch sigma.
- What the network buys: tau_int of Re/Im p_1..p_4 (Madras-Sokal, per kept sample at thin 1)
  is shorter than the plain chain's at the same 50% target the widths are scaled from; both
  tau and ESS/s at equal wall time are printed.  Both chains run their register-resident
  sweeps (8 lanes per walker)."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N, M_INV, SECTORS = 200, 3, list(range(1, 17))
WALKERS, SEED = 32768, 42
L0 = M_INV * N * (N - 1) // 2
BLOCKS, PER_BLOCK = 8, 16


def _run(kind):
    from laughlinedge_amd import Engine
    from laughlinedge_amd.stats import tau_int
    with Engine(N, M_INV, SECTORS, w=0.05, eps=5e-5, walkers=WALKERS, seed=SEED, samples=BLOCKS * PER_BLOCK + 8,
                n_blocks=BLOCKS, tune_sweeps=50, step=0.5, trace_walkers=512, lanes_per_walker=8,
                sweep_registers=1) as e:
        e.init_walkers()
        e.tune_step()
        e.burnin(300)
        rep = None
        if kind == "learned":
            rep = e.train_proposal(steps=2000, lr=1e-3, batch=4096, n_configs=100000, snapshot_sweeps=2, seed=99)
            assert rep["loss_last"] < rep["loss_first"], rep
            e.burnin(50)
        e.sample(8)
        e.reset_moments()
        t0 = time.perf_counter()
        for _ in range(BLOCKS):
            e.sample(PER_BLOCK)
        wall = time.perf_counter() - t0
        tr = e.trace()
        T = BLOCKS * PER_BLOCK
        tr = tr[tr.shape[0] - T * 512:].reshape(T, 512, 9)
        taus = [tau_int(tr[:, :, q].astype(np.float64)) for q in range(8)]
        spec = [e.spectra(b)[:4] for b in range(BLOCKS)]
        pos = e.positions()
        return {"taus": taus, "wall": wall, "spec": spec, "pos": pos, "acc": e.stats()["acceptance"], "rep": rep}


def test_c3_learned_proposal_full_size():
    from laughlinedge_amd.stats import batch_sigma
    print("case c3 learned proposal N=200 dL=1..16: plain chain", flush=True)
    plain = _run("plain")
    print("case c3 learned proposal N=200 dL=1..16: training + learned chain", flush=True)
    learned = _run("learned")
    # exact sampler law
    z = learned["pos"]
    r2 = np.sum(z[..., 0] ** 2 + z[..., 1] ** 2, axis=1)
    mu, var, n = 2.0 * (L0 + N), 4.0 * (L0 + N), len(r2)
    assert abs(r2.mean() - mu) < 5 * np.sqrt(var / n), (r2.mean(), mu)
    assert abs(r2.var(ddof=1) / var - 1.0) < 5 * np.sqrt(2.0 / n), (r2.var(ddof=1), var)
    # dL = 1..4 spectra of the two chains
    worst = 0.0
    for s in range(4):
        a = np.array([b[s] for b in plain["spec"]])
        c = np.array([b[s] for b in learned["spec"]])
        sig = np.sqrt(batch_sigma(a) ** 2 + batch_sigma(c) ** 2)
        dev = np.abs(a.mean(0) - c.mean(0)) / sig
        worst = max(worst, float(dev.max()))
        assert np.all(dev < 4.5), (SECTORS[s], dev)
    tp, tl = max(plain["taus"]), max(learned["taus"])
    T = BLOCKS * PER_BLOCK * WALKERS
    ess_p, ess_l = T / (2 * tp) / plain["wall"], T / (2 * tl) / learned["wall"]
    print(f"acceptance plain {plain['acc']:.3f} learned {learned['acc']:.3f}; training {learned['rep']}")
    print(f"tau_int Re/Im p1..p4 per kept sample: plain {np.round(plain['taus'], 2)} learned {np.round(learned['taus'], 2)}")
    print(f"ms per kept sample: plain {1e3 * plain['wall'] / (BLOCKS * PER_BLOCK):.3f} "
          f"learned {1e3 * learned['wall'] / (BLOCKS * PER_BLOCK):.3f}; ESS/s at equal wall time (slowest p_k): "
          f"plain {ess_p:.3e} learned {ess_l:.3e} ratio {ess_l / ess_p:.2f}; dL 1..4 spectra worst {worst:.2f} sigma")
    assert tl < tp, (tl, tp)
