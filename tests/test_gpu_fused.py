"""The fused sampling step (mc.fused_step = 1, launch_step): one launch per kept sample holds
the register sweeps, FP64 power sums and monomials of that sample and the S/H Gram tiles of the
previous one, as work items of one kernel.

- The chains are those of the separate kernels: after the same burn-in and kept samples, the
  walkers, the acceptance and the traced p_1..p_4, V equal a separate-kernel engine's bit for bit
  (same seed, step and 8-lane register mapping).
- The moments equal the separate kernels' up to FP32 accumulation rounding (the Gram row splits
  fall elsewhere): relative Frobenius below 1e-6 in every sector.
- Each sample's S and H (one kept sample per le_sample call, so one block per sample) equal the
  FP64 oracle's accumulation of that sample's configurations (or_accumulate_fixed, restating
  em_spectrum.cu:98-115 for the plain estimator) within 1e-6 relative Frobenius: N = 200,
  dL 1..16 (D = 914, ceil(N / 8) = 25) and N = 100, dL 0..12 (ceil(N / 8) = 13).
- Row splits that do not divide the walkers, a walker count that leaves a partial sweep unit, and
  thinning; engines the fused kernel does not cover are refused with LE_ERR_ARG."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

M_INV = 3
CONF = dict(w=0.05, eps=5e-5)


def _run(N, sectors, walkers, samples, step, fused, rows=0, thin=1, trace=64):
    from laughlinedge_amd import Engine
    kw = dict(walkers=walkers, seed=42, samples=samples, tune_sweeps=0, step=step, trace_walkers=trace, thin=thin)
    if fused:
        kw.update(fused_step=1, fused_rows=rows)
    else:
        kw.update(lanes_per_walker=8, sweep_registers=1)
    with Engine(N, M_INV, sectors, **CONF, **kw) as e:
        assert e.info.lanes_per_walker == 8 and e.info.sweep_in_registers == 1
        e.init_walkers()
        e.burnin(30)
        e.sample(samples)
        st = e.stats()
        if fused:
            assert st["step_launches"] == samples and st["ms_step"] > 0
        return e.positions(), e.trace(), e.moments(), st


@pytest.mark.parametrize("N,sectors,walkers,rows,thin", [
    (200, list(range(1, 17)), 4096, 0, 1),
    (200, list(range(1, 17)), 4000, 992, 2),
    (100, list(range(0, 13)), 3000, 1024, 1),
])
def test_fused_matches_separate_kernels(N, sectors, walkers, rows, thin):
    print(f"case fused vs separate N={N} walkers={walkers} rows={rows} thin={thin}", flush=True)
    pf, tf, mf, sf = _run(N, sectors, walkers, 6, 1.27, True, rows, thin)
    ps, ts, ms, ss = _run(N, sectors, walkers, 6, 1.27, False, thin=thin)
    assert np.array_equal(pf, ps)
    assert np.array_equal(tf, ts)
    assert sf["acceptance"] == ss["acceptance"]
    worst = 0.0
    for (S0, H0), (S1, H1) in zip(ms, mf):
        rs = np.linalg.norm(S1 - S0) / np.linalg.norm(S0)
        rh = np.linalg.norm(H1 - H0) / np.linalg.norm(H0)
        worst = max(worst, rs, rh)
        assert rs < 1e-6 and rh < 1e-6, (rs, rh)
    print(f"walkers and traces identical; worst relative Frobenius moment difference {worst:.2e}")


@pytest.mark.parametrize("N,sectors", [(200, list(range(1, 17))), (100, list(range(0, 13)))])
def test_fused_samples_vs_oracle(N, sectors, oracle_lib):
    from laughlinedge_amd import Engine
    walkers = 2048
    print(f"case fused samples vs FP64 oracle N={N} sectors={sectors[0]}..{sectors[-1]}", flush=True)
    o = oracle_lib.Oracle(N, M_INV, sectors, **CONF)
    with Engine(N, M_INV, sectors, **CONF, walkers=walkers, seed=7, samples=3, tune_sweeps=0, step=1.2,
                fused_step=1) as e:
        e.init_walkers()
        e.burnin(20)
        worst = 0.0
        for b in range(3):
            e.sample(1)
            pos = e.positions()
            So, Ho, _ = o.accumulate_fixed(pos)
            for s, (Sg, Hg) in enumerate(e.moments(block=b)):
                S64, H64 = So[s] / walkers, Ho[s] / walkers
                rs = np.linalg.norm(Sg - S64) / np.linalg.norm(S64)
                rh = np.linalg.norm(Hg - H64) / np.linalg.norm(H64)
                worst = max(worst, rs, rh)
                assert rs < 1e-6 and rh < 1e-6, (b, sectors[s], rs, rh)
    print(f"worst relative Frobenius error against the FP64 oracle {worst:.2e}")


def test_fused_refuses_what_it_does_not_cover():
    from laughlinedge_amd import Engine, LEError
    print("case fused refusals", flush=True)
    base = dict(walkers=256, seed=1, samples=2, tune_sweeps=0, step=1.0, fused_step=1)
    for N, sectors, extra in [
        (200, [1, 2], dict(precision=1)),
        (200, [1, 2], dict(symmetry_moves=1)),
        (200, [1, 2], dict(lanes_per_walker=16)),
        (150, [1, 2], {}),                      # ceil(150 / 8) = 19: no register kernel
        (200, [1, 17], {}),                     # K = 17 > 16
        (200, [1, 2], dict(fused_rows=100)),    # not a multiple of 32
    ]:
        with pytest.raises(LEError):
            Engine(N, M_INV, sectors, **CONF, **{**base, **extra})
    with pytest.raises(LEError):
        Engine(200, M_INV, [1, 2], **CONF, harmonics=[1], **base)
    with Engine(200, M_INV, [1, 2], **CONF, **base) as e:
        with pytest.raises(LEError, match="plain proposal"):
            e.set_proposal(1)
