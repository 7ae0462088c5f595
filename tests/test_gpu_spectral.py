"""GPU tests of the particle-removal spectral weights (le_removal_weights): the C ABI
against the oracle's Newton-identity route on identical moments, and a sampled tiny droplet
against the exact expansion."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_removal_weights_match_oracle_on_same_moments():
    """A 100-particle droplet (m = 3) losing one particle: the 99-particle sectors dL = 0..12.
    Energies and weights from the C ABI equal the oracle's (Newton coefficients, scipy eigh)
    on the engine's own S and H; the weights sum to 1; the orbital is l = 3 * 99 - dL."""
    from laughlinedge_amd import Engine
    from laughlinedge_amd.engine import basis_describe
    from oracle.spectral import newton_coefficients, removal_weights
    N, m, secs = 99, 3, list(range(13))
    print("case removal weights N=99 dL=0..12 vs oracle on the same moments", flush=True)
    with Engine(N, m, secs, w=0.05, eps=5e-5, walkers=16384, samples=4, seed=5, tune_sweeps=30,
                step=0.5) as e:
        e.init_walkers()
        e.tune_step()
        e.burnin(200)
        e.sample(4)
        mom = e.moments()
        states = basis_describe(N, secs)
        for s, dl in enumerate(secs):
            l, ev, w = e.removal_weights(s)
            assert l == m * N - dl
            S, H = mom[s]
            o, d = e.offsets[s], e.dims[s]
            a = newton_coefficients(dl, m, states[o:o + d])
            ev_o, w_o = removal_weights(S, H, a)
            assert abs(w.sum() - 1.0) < 1e-9, (dl, w.sum())
            assert np.allclose(ev, ev_o, rtol=1e-9, atol=1e-9 * np.abs(ev_o).max()), dl
            assert np.allclose(w, w_o, atol=1e-8), (dl, np.abs(w - w_o).max())


def test_removal_weights_vs_exact_n4():
    """A 4-particle droplet (m = 3, V = r^2 / 2 + r^4) losing one particle: the sampled
    3-particle sectors dL = 0..3 give energies and removal weights within 5 batch sigma of the
    exact expansion (coefficient extraction from the Laughlin polynomial)."""
    from scipy.linalg import eigh
    from laughlinedge_amd import Engine
    from laughlinedge_amd.stats import batch_sigma
    from oracle.exact_lll import exact_removal
    from oracle.partitions import sector_basis
    n, m, secs, nb = 3, 3, [0, 1, 2, 3], 16
    print("case removal weights 4 -> 3 particles vs exact", flush=True)
    with Engine(n, m, secs, w=1.0, eps=1.0, walkers=8192, samples=400, n_blocks=nb, seed=77,
                tune_sweeps=40, step=0.5) as e:
        e.init_walkers()
        e.tune_step()
        e.burnin(300)
        for _ in range(nb):
            e.sample(25)
        for s, dl in enumerate(secs):
            states = [(dl, p) for p in sector_basis(n, dl)]
            G, Hq, b, qq = exact_removal(n + 1, m, m * n - dl, states)
            ev_x, V = eigh(np.array(Hq), np.array(G))
            w_x = (V.T @ np.array(b)) ** 2 / qq
            per = [e.removal_weights(s, block=k) for k in range(nb)]
            ev = np.array([p[1] for p in per])
            w = np.array([p[2] for p in per])
            assert all(p[0] == m * n - dl for p in per)
            assert np.all(np.abs(ev.mean(0) - ev_x) <= 5 * batch_sigma(ev) + 1e-9 * np.abs(ev_x).max()), (dl, ev.mean(0), ev_x)
            assert np.all(np.abs(w.mean(0) - w_x) <= 5 * batch_sigma(w) + 1e-6), (dl, w.mean(0), w_x)
            print(f"dL={dl}: exact weights {np.round(w_x, 5)}, sampled {np.round(w.mean(0), 5)}")
