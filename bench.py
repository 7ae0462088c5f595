"""Benchmark of the edge-state Monte-Carlo hot path on one MI355X.

Workload (N=200 Laughlin droplet, nu=1/3, all sectors dL=1..16 at once):
  32768 walkers, seed 42, V(r) = 0.5*0.05*r^2 + 5e-5*r^4, FP32 sampler + FP32 MFMA tiles.
A step is one Metropolis sweep of every walker followed by one kept sample per walker
accumulated into all sector S and H matrices (one "sample-update" per walker).

Prints ONE JSON line: metric/value = S,H sample-updates/s, plus ESS/s, the roofline
of the dominant kernel (Gram accumulation on MFMA, FP32), the sampler's VALU rate, the
FP32-vs-FP64 precision on a fixed sample set, and a CPU baseline (the FP64 oracle port,
OpenMP on the host cores) on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N, M_INV = 200, 3
SECTORS = list(range(1, 17))
W_CONF, EPS_CONF = 0.05, 5e-5
WALKERS = 32768
SEED = 42
PEAK_FP32 = 157.3e12       # MI355X FP32 (MFMA and VALU), dense
PEAK_HBM = 8.0e12


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=40)
    ap.add_argument("--burnin", type=int, default=300)
    ap.add_argument("--walkers", type=int, default=WALKERS)
    ap.add_argument("--trace-walkers", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-precision", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported "
              "(walker sharding across GPUs is described in DESIGN.md, not built)", file=sys.stderr)
        sys.exit(2)

    from laughlinedge_amd import Engine
    from laughlinedge_amd.stats import tau_int, batch_sigma

    steps, warm = args.steps, args.warmup
    nblk = 4 if steps >= 4 else 1
    eng = Engine(N, M_INV, SECTORS, w=W_CONF, eps=EPS_CONF, walkers=args.walkers, seed=SEED,
                 precision=0, samples=steps, n_blocks=nblk, tune_sweeps=50, step=0.5,
                 trace_walkers=args.trace_walkers)
    info = eng.info
    dims = eng.dims
    sum_d2 = sum(d * d for d in dims)
    print(f"[bench] N={N} sectors=1..16 D={info.D} sum_d2={sum_d2} tiles={info.n_tiles} "
          f"walkers={args.walkers}", file=sys.stderr, flush=True)
    eng.init_walkers()
    step = eng.tune_step()
    eng.burnin(args.burnin)
    print(f"[bench] step={step:.4f} lanes_per_walker={eng.info.lanes_per_walker} burn-in done",
          file=sys.stderr, flush=True)
    if warm > 0:
        eng.sample(warm)
    eng.reset_moments()
    st0 = eng.stats()
    # timed region: inputs (walkers) resident in HBM, steps x (sweep + features + Gram)
    per = steps // nblk
    t0 = time.perf_counter()
    for b in range(nblk):
        eng.sample(per if b < nblk - 1 else steps - per * (nblk - 1))
    t1 = time.perf_counter()
    st1 = eng.stats()
    wall = t1 - t0
    updates = args.walkers * steps
    value = updates / wall
    ms_step = 1e3 * wall / steps

    # kernel-level rates from HIP events recorded on the engine's stream
    gram_ms = st1["ms_gram"] - st0["ms_gram"]
    gram_launches = st1["gram_launches"] - st0["gram_launches"]
    sweep_ms = st1["ms_sweep"] - st0["ms_sweep"]
    feat_ms = st1["ms_features"] - st0["ms_features"]
    gram_flop_per_sample = 16.0 * sum_d2
    gram_rate = gram_flop_per_sample * updates / (gram_ms * 1e-3)
    sweep_flop = 12.0 * N * (N - 1) * updates
    sweep_rate = sweep_flop / (sweep_ms * 1e-3)
    row_bytes = info.x_ld * 4 + 4
    stream_gbs = updates * (8 * info.D + 4) / (feat_ms * 1e-3) / 1e9

    # ESS from the traced walkers: Re/Im p1..p4 and V
    tr = eng.trace()
    tw = min(args.trace_walkers, args.walkers)
    ess_per_s = None
    taus = {}
    if tr.shape[0] >= tw * 8:
        T = tr.shape[0] // tw
        tr = tr[-T * tw:].reshape(T, tw, 9)
        # the warmup samples precede the timed ones in the trace; use the last `steps`
        tr = tr[-min(T, steps):]
        names = ["Re p1", "Im p1", "Re p2", "Im p2", "Re p3", "Im p3", "Re p4", "Im p4", "V"]
        for q, nm in enumerate(names):
            taus[nm] = float(tau_int(tr[:, :, q]))
        tmax = max(taus.values())
        ess_per_s = updates / (2.0 * tmax) / wall

    # spectra and their batch errors (FP64 eigensolve on the GPU)
    blocks = [eng.spectra(b) for b in range(eng.n_blocks())]
    mean_spec = eng.spectra(-1)
    sig = []
    for s in range(len(dims)):
        vals = [bl[s] for bl in blocks if bl[s] is not None]
        sig.append(batch_sigma(np.array(vals)) if len(vals) >= 2 else None)

    out = {
        "metric": "S,H sample-updates/s at N=200, DeltaL<=16 (ESS/s reported alongside)",
        "value": value,
        "unit": "sample-updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warm,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded Metropolis walkers, no dataset)",
        "config": {"workload": "laughlin-edge N=200 nu=1/3 dL=1..16 grouped (D=914), "
                               "V=0.025 r^2+5e-5 r^4",
                   "walkers": args.walkers, "seed": SEED, "burnin_sweeps": args.burnin,
                   "thin": 1, "step": step, "lanes_per_walker": eng.info.lanes_per_walker,
                   "samples_per_launch": info.samples_per_launch, "gram_tiles": info.n_tiles},
        "roofline": {"bound": "mfma", "kernel": "k_gram<float>",
                     "achieved": gram_rate / 1e12, "peak": PEAK_FP32 / 1e12, "unit": "TFLOP/s",
                     "frac": gram_rate / PEAK_FP32, "traffic": None,
                     "algorithmic_flop_per_launch": gram_flop_per_sample * updates / max(1, gram_launches),
                     "avg_launch_ms": gram_ms / max(1, gram_launches)},
        "sampler": {"kernel": "k_sweep<G,float>", "achieved_tflops": sweep_rate / 1e12,
                    "peak_tflops": PEAK_FP32 / 1e12, "frac": sweep_rate / PEAK_FP32,
                    "flop_per_sweep_per_walker": 12 * N * (N - 1),
                    "ms_per_sweep": sweep_ms / steps},
        "features": {"ms_per_step": feat_ms / steps, "stream_gb_per_s": stream_gbs,
                     "bytes_per_sample": 8 * info.D + 4},
        "breakdown_ms_per_step": {"sweep": sweep_ms / steps, "features": feat_ms / steps,
                                  "gram": gram_ms / steps},
        "acceptance": st1["acceptance"],
        "ess_per_s": ess_per_s,
        "tau_int": taus,
        "spectrum_lowest": {str(SECTORS[s]): (float(mean_spec[s][0]) if mean_spec[s] is not None else None)
                            for s in range(len(dims))},
    }

    if not args.no_precision:
        out["precision"] = precision_check(eng, sig)
    eng.close()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


def precision_check(eng, sig, n_fixed=8192):
    """FP32 GPU moments vs the FP64 CPU oracle on identical configurations."""
    from oracle.oracle import Oracle, geneig
    pos = eng.positions()[: n_fixed]
    # decorrelate a little: take configurations from two sweeps apart
    o = Oracle(N, M_INV, SECTORS, w=W_CONF, eps=EPS_CONF)
    t0 = time.perf_counter()
    So, Ho, _ = o.accumulate_fixed(pos)
    t_cpu = time.perf_counter() - t0
    from laughlinedge_amd import Engine
    with Engine(N, M_INV, SECTORS, w=W_CONF, eps=EPS_CONF, walkers=1024, samples=1,
                precision=0) as e32:
        e32.accumulate_positions(pos)
        mom = e32.moments()
        spec32 = e32.spectra()
    n = pos.shape[0]
    rs, rh, eig_rel, eig_sig = [], [], [], []
    for s, (Sg, Hg) in enumerate(mom):
        S64, H64 = So[s] / n, Ho[s] / n
        rs.append(float(np.linalg.norm(Sg - S64) / np.linalg.norm(S64)))
        rh.append(float(np.linalg.norm(Hg - H64) / np.linalg.norm(H64)))
        try:
            e64 = geneig(S64, H64)
        except np.linalg.LinAlgError:
            continue
        if spec32[s] is None:
            continue
        eig_rel.append(float(np.max(np.abs(spec32[s] - e64) / np.abs(e64))))
        if sig[s] is not None:
            eig_sig.append(float(np.max(np.abs(spec32[s] - e64) / np.maximum(sig[s], 1e-300))))
    return {"fixed_samples": int(n), "rel_fro_S_max": max(rs), "rel_fro_H_max": max(rh),
            "eig_rel_err_max": max(eig_rel) if eig_rel else None,
            "eig_err_in_mc_sigma_max": max(eig_sig) if eig_sig else None,
            "oracle_seconds": t_cpu}


def cpu_baseline():
    """FP64 oracle port (plain |Psi_L|^2 chain + S/H accumulation), OpenMP, host cores."""
    from oracle.oracle import Oracle
    threads = min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", "16")))
    o = Oracle(N, M_INV, SECTORS, w=W_CONF, eps=EPS_CONF)
    walkers, burn, samples = 4 * threads, 20, 200
    t0 = time.perf_counter()
    o.plain_run(walkers, burn, samples, 0.5, SEED, threads=threads)
    dt = time.perf_counter() - t0
    # burn-in sweeps cost the same as sampling sweeps minus the accumulation
    rate = walkers * samples / dt
    return {"value": rate, "unit": "sample-updates/s", "cores": threads, "kind": "port",
            "sample": f"{walkers} walkers x ({burn} burn-in + {samples} kept sweeps) at N=200, "
                      f"dL=1..16, FP64 (burn-in time included), {dt:.1f} s"}


if __name__ == "__main__":
    main()
