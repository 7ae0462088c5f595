"""Step tuning on the GPU (edgeMC::optimal_step, em_bisection.cu:49-153) and the eigensolver's
error path.

- le_tune_step measures the acceptance of tune_sweeps sweeps from freshly initialised walkers
  per trial step and stops within target +- tol.  A fresh run of the same length at the tuned
  step (other random numbers, same distribution) measures the same acceptance to within the
  tolerance plus its statistical error.
- A tolerance no measurement can meet gives LE_ERR_BISECTION after the capped refinement (the
  reference loops on), and the engine keeps its step and stays usable.
- An indefinite metric gives LE_ERR_NOT_PD from le_geneig (the reference's Eigen solver returns
  garbage there) and a per-sector status from le_engine_spectra."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("target", [0.5, 0.3])
def test_tune_step_hits_target(target):
    from laughlinedge_amd import Engine
    N, m, walkers, sweeps, tol = 20, 3, 4096, 100, 0.01
    print(f"case tune_step N=20 target {target}", flush=True)
    with Engine(N, m, [0, 1, 2], w=0.05, eps=1e-4, walkers=walkers, samples=sweeps, seed=5, tune_sweeps=sweeps,
                step=0.8, target_acceptance=target, acceptance_tol=tol, graphs=0) as e:
        e.init_walkers()
        step = e.tune_step()
        assert step > 0
        # same protocol as one tuning measurement: fresh walkers, tune_sweeps sweeps
        e.init_walkers()
        e.sample(sweeps)
        acc = e.stats()["acceptance"]
    print(f"tuned step {step:.4f}: acceptance of a fresh {sweeps}-sweep run {acc:.4f}")
    assert abs(acc - target) < tol + 0.005, (acc, target)


def test_tune_step_unreachable_target():
    from laughlinedge_amd import Engine, LEError
    from laughlinedge_amd.engine import LE_ERR_BISECTION
    print("case tune_step unreachable tolerance", flush=True)
    # a measured acceptance is k / (walkers x sweeps x N) = k / 10240: no k lies within 1e-9 of
    # 0.33333, so the search brackets the target, refines 60 times and gives up
    with Engine(8, 3, [0, 1], w=0.05, walkers=256, samples=4, seed=3, tune_sweeps=5, step=0.5,
                target_acceptance=0.33333, acceptance_tol=1e-9, graphs=0) as e:
        e.init_walkers()
        with pytest.raises(LEError) as ei:
            e.tune_step()
        assert ei.value.code == LE_ERR_BISECTION
        e.burnin(5)
        e.sample(4)
        st = e.stats()
        assert abs(st["step"] - 0.5) < 1e-12, st["step"]
        assert 0.0 < st["acceptance"] < 1.0


def test_geneig_not_positive_definite():
    from laughlinedge_amd import LEError, geneig
    from laughlinedge_amd.engine import LE_ERR_NOT_PD
    print("case geneig indefinite metric", flush=True)
    S = np.array([[1.0, 2.0], [2.0, 1.0]], dtype=complex)      # eigenvalues 3, -1
    H = np.eye(2, dtype=complex)
    with pytest.raises(LEError) as ei:
        geneig(S, H)
    assert ei.value.code == LE_ERR_NOT_PD
    # a good pencil still solves afterwards
    ev = geneig(np.eye(2, dtype=complex), np.diag([2.0, 1.0]).astype(complex))
    assert np.allclose(ev, [1.0, 2.0])
