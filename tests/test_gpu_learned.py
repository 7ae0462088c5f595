"""GPU tests of the learned proposal: gradients vs the FP64 oracle, exactness of the
Metropolis-Hastings chain, and what training buys."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_mlp_gradient_vs_oracle():
    from laughlinedge_amd import Engine
    from oracle import learned
    print("case learned gradient N=12", flush=True)
    rng = np.random.default_rng(5)
    N, m = 12, 3
    R = np.sqrt(2 * m * N)
    p = (learned.init_params(rng) + rng.normal(size=learned.NPARAM) * 0.05).astype(np.float32)
    B = 64
    cfgs = []
    for b in range(16):
        r = R * np.sqrt(rng.random(N))
        t = 2 * np.pi * rng.random(N)
        cfgs.append(np.stack([r * np.cos(t), r * np.sin(t)], -1))
    cfgs = np.array(cfgs, dtype=np.float32)
    cfg = rng.integers(0, 16, B)
    part = rng.integers(0, N, B)
    noise = rng.normal(size=(B, 2)).astype(np.float32)
    with Engine(N, m, [0], walkers=64, samples=1) as e:
        e.set_proposal_params(p, 0.7)
        assert np.array_equal(e.proposal_params(), p)
        loss, g = e.proposal_loss_grad(cfgs, cfg, part, noise, 0.7)
    zs = [c[:, 0].astype(np.float64) + 1j * c[:, 1].astype(np.float64) for c in cfgs]
    L, G = learned.batch_loss_grad(p.astype(np.float64), [zs[c] for c in cfg], part, noise.astype(np.float64), m, R, 0.7)
    assert abs(loss - L) < 1e-4 * abs(L) + 1e-6, (loss, L)
    assert np.linalg.norm(g - G) < 2e-3 * np.linalg.norm(G), np.linalg.norm(g - G) / np.linalg.norm(G)


@pytest.mark.parametrize("sym", [0, 2])
def test_learned_chain_is_exact(sym):
    """MH with the trained, state-dependent widths keeps |Psi_L|^2: the sum rule
    <sum r^2> = 2 (L0 + N) and the dL = 2..3 spectra agree with the plain chain; with the
    symmetry moves (centre of mass, dilation, rotation) after every kept sample as well, and
    then the walkers' sum r^2 also follows its exact law (sum r^2 / 2 ~ Gamma(L0 + N))."""
    from laughlinedge_amd import Engine
    from laughlinedge_amd.stats import batch_sigma
    N, m = 16, 3
    L0 = m * N * (N - 1) // 2
    print(f"case learned exactness N=16 symmetry={sym}", flush=True)
    res = {}
    for kind in ("plain", "learned"):
        with Engine(N, m, [0, 2, 3], w=2.0, eps=1e-3, walkers=4096, samples=400, burnin_sweeps=200, seed=31,
                    tune_sweeps=40, step=0.6, n_blocks=8, lanes_per_walker=8,
                    symmetry_moves=sym if kind == "learned" else 0) as e:
            e.init_walkers()
            e.tune_step()
            e.burnin(200)
            if kind == "learned":
                rep = e.train_proposal(steps=300, batch=2048, n_configs=16384, snapshot_sweeps=2)
                assert rep["loss_last"] < rep["loss_first"]
                e.burnin(50)
            for b in range(8):
                e.sample(50)
            r2 = [e.moments(b)[0][1][0, 0].real / 2.0 for b in range(8)]   # H with w = 2 on dL=0: sum r^2 (+eps r^4)
            spec = np.array([np.concatenate(e.spectra(b)[1:]) for b in range(8)])
            res[kind] = (np.array(r2), spec, e.stats()["acceptance"])
            if kind == "learned":
                z = e.positions()
                rw = np.sum(z[..., 0] ** 2 + z[..., 1] ** 2, axis=1)
                mu, var, n = 2.0 * (L0 + N), 4.0 * (L0 + N), len(rw)
                assert abs(rw.mean() - mu) < 5 * np.sqrt(var / n), (rw.mean(), mu)
                assert abs(rw.var(ddof=1) / var - 1.0) < 5 * np.sqrt(2.0 / n), (rw.var(ddof=1), var)
    for kind in ("plain", "learned"):
        r2, spec, acc = res[kind]
        assert 0.1 < acc < 0.95
    # spectra: difference of the two chains' means within 4 combined sigma
    sp, sl = res["plain"][1], res["learned"][1]
    dm = np.abs(sp.mean(0) - sl.mean(0))
    sig = np.sqrt(batch_sigma(sp) ** 2 + batch_sigma(sl) ** 2)
    assert np.all(dm < 4.5 * sig + 1e-9), (dm / sig)
    rp, rl = res["plain"][0], res["learned"][0]
    assert abs(rp.mean() - rl.mean()) < 4.5 * np.sqrt(batch_sigma(rp) ** 2 + batch_sigma(rl) ** 2)


def test_training_gain_esjd():
    """What the trained network buys over the proposal without it: the expected squared jump
    distance per particle and sweep of the learned chain against the plain chain whose step
    is tuned to 50% acceptance (the step the network's widths are scaled from).  At N = 200
    the same comparison gives 1.42 vs 0.99 and tau_int(p_1) 2.89 vs 3.66 sweeps."""
    from laughlinedge_amd import Engine
    N, m, W = 48, 3, 4096
    print("case learned ESJD gain N=48", flush=True)
    esjd = {}
    for kind in ("plain", "learned"):
        with Engine(N, m, [0], w=0.05, eps=5e-5, walkers=W, samples=1, seed=17, tune_sweeps=50, step=0.5,
                    lanes_per_walker=8) as e:
            e.init_walkers()
            e.tune_step()
            e.burnin(200)
            if kind == "learned":
                rep = e.train_proposal(steps=600, batch=4096, n_configs=32768, snapshot_sweeps=2)
                assert rep["loss_last"] < rep["loss_first"]
                e.burnin(50)
            jumps = []
            for _ in range(4):
                p0 = e.positions()
                e.burnin(1)
                p1 = e.positions()
                jumps.append(((p1 - p0) ** 2).sum(-1).mean())
            esjd[kind] = float(np.mean(jumps))
    print(f"ESJD per particle-sweep: plain {esjd['plain']:.3f} learned {esjd['learned']:.3f}")
    assert esjd["learned"] > 1.15 * esjd["plain"], esjd


@pytest.mark.parametrize("N,G", [(200, 8), (100, 8), (200, 16)])
def test_learned_register_sweep_matches_lds(N, G):
    """The register-resident learned sweep follows the LDS one: the same trained widths table,
    the same walkers and random streams give the same positions after several sweeps (the
    products, logs and nearest-distance lists are taken in the same order)."""
    from laughlinedge_amd import Engine
    from laughlinedge_amd.engine import PROPOSAL_LEARNED
    print(f"case learned register vs LDS sweep N={N} G={G}", flush=True)
    with Engine(N, 3, [0], w=0.05, eps=5e-5, walkers=1024, samples=1, seed=11, tune_sweeps=20, step=0.5,
                lanes_per_walker=G) as e:
        e.init_walkers()
        step = e.tune_step()
        e.burnin(50)
        e.train_proposal(steps=100, batch=2048, n_configs=8192, snapshot_sweeps=2, seed=99)
        params = e.proposal_params()
        start = e.positions()
    out = []
    for reg in (0, 1):
        with Engine(N, 3, [0], w=0.05, eps=5e-5, walkers=1024, samples=1, seed=23, tune_sweeps=0, step=step,
                    lanes_per_walker=G, sweep_registers=reg) as e:
            assert e.info.sweep_in_registers == reg
            e.set_proposal_params(params, step)
            e.set_proposal(PROPOSAL_LEARNED)
            e.set_positions(start)
            e.burnin(4)
            out.append(e.positions())
    same = np.abs(out[0] - out[1]).max(axis=(1, 2)) == 0
    print(f"walkers with identical positions after 4 sweeps: {same.mean():.4f}")
    assert same.mean() >= 0.99, same.mean()
