"""Benchmark of the edge-state Monte-Carlo hot path on one MI355X.

Workload (N=200 Laughlin droplet, nu=1/3, all sectors dL=1..16 at once, D=914):
  32768 walkers, seed 42, V(r) = 0.5*0.05*r^2 + 5e-5*r^4, initial positions uniform in the
  disc of radius R = sqrt(2 m N), step tuned to 50% acceptance, then burn-in.
A step is one Metropolis sweep of every walker followed by one kept sample per walker,
accumulated into every sector's S and H (one "sample-update" per walker and step).

The last stdout line is ONE compact JSON object (a few KB); the full result goes to the
JSON file that --full-out names.  value = S,H sample-updates/s of the
production form (FP32 sampler, FP32 MFMA accumulation, one sweep per kept sample), timed as
one le_sample call of --steps kept samples per walker.  Alongside:
  roofline      the Gram accumulation kernel (MFMA bound), average launch time from HIP
                events on the Gram stream over the timed call
  kernels       sampler (VALU FLOP/s), features (sample-matrix bytes/s), Gram (MFMA FLOP/s)
  ess           tau_int (Madras-Sokal) and ESS/s: the slowest of Re/Im p_1..p_4, V and traced
                S and H entries (|p_1|^2, |p_2|^2, Re/Im conj(p_2) p_1^2, V |p_1|^2, ...), and
                (*_sh) the slowest S or H entry alone: every S and H entry pairs states of
                one dL, so all are invariant under z -> e^{i phi} z, while Re/Im p_k average
                to zero and enter no S or H entry
  forms         plain FP64 / plain FP32 / wide-step FP32 and BF16-split / with exact
                symmetry moves / learned FP32 at their best thinning: ESS/s, ratios to the
                plain FP64 form, spectra vs the FP64 form in MC sigma
  precision     FP32 and BF16-split GPU moments vs the FP64 oracle on one fixed sample set
  cpu_baseline  the reference's row-wise estimator and the plain estimator restated in C
                (FP64, OpenMP, host cores) on bounded samples
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
MIX_RHO = 0.5              # mixture estimator: weight of the constant term of w
WALKERS = 32768
SEED = 42
PEAK_FP32 = 157.3e12       # MI355X FP32 dense (MFMA; VALU counting packed FMA)
PEAK_FP64 = 78.6e12
PEAK_HBM = 8.0e12
METRIC = "ESS/s and S,H sample-updates/s at N=200, DeltaL<=16 on 1 GPU; eigenvalue err"
TRACE_NAMES = ["Re p1", "Im p1", "Re p2", "Im p2", "Re p3", "Im p3", "Re p4", "Im p4", "V"]


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def observables(tr):
    """(T, walkers, 9) trace -> dict name -> (T, walkers) series: Re/Im p_1..p_4 and V, and
    S and H entries of sectors 0..4 (conj(p_lambda) p_mu and conj(p_lambda) V p_mu with
    |lambda| = |mu|, unchanged by a global rotation z -> e^{i phi} z)."""
    p = [tr[:, :, 2 * k].astype(np.float64) + 1j * tr[:, :, 2 * k + 1] for k in range(4)]
    obs = {nm: tr[:, :, q].astype(np.float64) for q, nm in enumerate(TRACE_NAMES)}
    obs["|p1|^2"] = np.abs(p[0]) ** 2
    obs["|p2|^2"] = np.abs(p[1]) ** 2
    obs["|p1|^4"] = np.abs(p[0]) ** 4
    obs["Re conj(p2) p1^2"] = (np.conj(p[1]) * p[0] ** 2).real
    obs["|p3|^2"] = np.abs(p[2]) ** 2
    obs["Re conj(p3) p1 p2"] = (np.conj(p[2]) * p[0] * p[1]).real
    obs["|p4|^2"] = np.abs(p[3]) ** 2
    obs["Im conj(p2) p1^2"] = (np.conj(p[1]) * p[0] ** 2).imag
    obs["V |p1|^2"] = obs["V"] * obs["|p1|^2"]
    obs["V |p2|^2"] = obs["V"] * obs["|p2|^2"]
    return obs


PK_NAMES = TRACE_NAMES[:8]


def tau_sh(taus):
    """largest tau_int over the traced S and H entries (the quantities the spectra use)"""
    return max(v for k, v in taus.items() if k not in PK_NAMES)


def tau_pk(taus):
    """largest tau_int of Re/Im p_1..p_4 (diagnostic: no S or H entry)"""
    return max(taus[k] for k in PK_NAMES)


def mixture_kish(eng):
    """Kish factor (sum 1/w)^2 / (n sum 1/w^2) of the mixture weights 1/w over the walkers' current
    configurations: w = c0 + sum_k exp(a_k ln|p_k|^2 + b_k), p_k = sum_i (z_i / R)^k."""
    L, c0, a, b = eng.get_mixture()
    if L == 0:
        return 1.0
    pos = eng.positions()
    z = (pos[..., 0] + 1j * pos[..., 1]) / np.sqrt(2.0 * M_INV * N)
    w = np.full(z.shape[0], c0)
    zk = np.ones_like(z)
    for k in range(L):
        zk = zk * z
        q = np.abs(zk.sum(axis=1)) ** 2
        with np.errstate(divide="ignore"):
            w += np.where(q > 0, np.exp(a[k] * np.log(np.maximum(q, 1e-300)) + b[k]), 0.0)
    iw = 1.0 / w
    return float(iw.sum() ** 2 / (len(iw) * np.sum(iw * iw)))


def taus_of(eng, n_samples, tw):
    from laughlinedge_amd.stats import tau_int
    tr = eng.trace()
    T = min(tr.shape[0] // tw, n_samples)
    if T < 8:
        return None
    tr = tr[tr.shape[0] - T * tw:].reshape(T, tw, 9)
    return {k: float(tau_int(v)) for k, v in observables(tr).items()}


def spectra_blocks(eng, dims):
    from laughlinedge_amd.stats import batch_sigma
    blocks = [eng.spectra(b) for b in range(eng.n_blocks())]
    mean = eng.spectra(-1)
    sig = []
    for s in range(len(dims)):
        vals = [bl[s] for bl in blocks if bl[s] is not None]
        sig.append(batch_sigma(np.array(vals)) if len(vals) >= 2 else None)
    return mean, sig


def spectra_jackknife(eng, dims):
    """per sector: (bias-corrected eigenvalues, jackknife sigma) from the per-block moments,
    solved by the GPU eigensolver; None where a leave-one-out metric is not positive definite"""
    from laughlinedge_amd import geneig
    from laughlinedge_amd.stats import jackknife_eigenvalues
    per_block = [eng.moments(b) for b in range(eng.n_blocks())]
    out = []
    for s in range(len(dims)):
        try:
            c, sg, _ = jackknife_eigenvalues([bl[s] for bl in per_block], geneig)
            out.append((c, sg))
        except Exception:
            out.append(None)
    return out


def compare_spectra(spec, sig, ref_spec, ref_sig):
    """deviations |E - E_ref| / sqrt(sigma^2 + sigma_ref^2) over every eigenvalue of every sector"""
    dev, lowest, where = [], {}, []
    for s in range(len(SECTORS)):
        a, b, sa, sb = spec[s], ref_spec[s], sig[s], ref_sig[s]
        if a is None or b is None or sa is None or sb is None:
            continue
        ds = np.abs(a - b) / np.sqrt(sa ** 2 + sb ** 2 + 1e-300)
        dev.append(ds)
        lowest[str(SECTORS[s])] = round(float(ds[0]), 3)
        where += [(float(x), SECTORS[s], k, len(ds)) for k, x in enumerate(ds)]
    dev = np.concatenate(dev) if dev else np.zeros(0)
    top = max(where) if where else None
    return {"eigenvalues": int(dev.size),
            "max_dev_in_combined_sigma": float(dev.max()) if dev.size else None,
            "max_at": ({"dL": top[1], "index": top[2], "of": top[3]} if top else None),
            "median_dev_in_combined_sigma": float(np.median(dev)) if dev.size else None,
            "frac_within_1_sigma": float(np.mean(dev <= 1)) if dev.size else None,
            "frac_within_2_sigma": float(np.mean(dev <= 2)) if dev.size else None,
            "lowest_per_sector_dev": lowest}


def make_engine(walkers, precision, samples, n_blocks, tw, target=0.5, graphs=1, sym=0, estimator=0, fused=0):
    from laughlinedge_amd import Engine
    return Engine(N, M_INV, SECTORS, w=W_CONF, eps=EPS_CONF, walkers=walkers, seed=SEED,
                  precision=precision, samples=samples, n_blocks=n_blocks, tune_sweeps=50, step=0.5,
                  trace_walkers=tw, target_acceptance=target, acceptance_tol=0.01, graphs=graphs,
                  lanes_per_walker=-1, symmetry_moves=sym, estimator=estimator, mixture_rho=MIX_RHO,
                  fused_step=fused)


def timed_sampling(eng, steps, nblk):
    eng.reset_moments()
    st0 = eng.stats()
    per = steps // nblk
    t0 = time.perf_counter()
    for b in range(nblk):
        eng.sample(per if b < nblk - 1 else steps - per * (nblk - 1))
    wall = time.perf_counter() - t0
    st1 = eng.stats()
    return wall, {k: st1[k] - st0[k] for k in ("ms_sweep", "ms_features", "ms_gram", "gram_launches",
                                               "sweep_launches", "ms_sampler_graph", "ms_step", "step_launches")}


ERR_BLOCKS, ERR_BLOCK_SAMPLES = 8, 8


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=40)
    ap.add_argument("--burnin", type=int, default=300)
    ap.add_argument("--walkers", type=int, default=WALKERS)
    ap.add_argument("--trace-walkers", type=int, default=512)
    ap.add_argument("--form-steps", type=int, default=96, help="kept samples per (form, thinning)")
    ap.add_argument("--thins", default="1,2,4,8")
    ap.add_argument("--estimator", choices=("plain", "mixture"), default="plain",
                    help="headline estimator: plain |Psi_L|^2 walkers (default) or the mixture w |Psi_L|^2")
    ap.add_argument("--fused", action="store_true",
                    help="headline with the fused step kernel (one launch per kept sample, fused_step = 1): "
                         "the command a kernel-trace profile of k_step runs; its roofline block reports the "
                         "fused launches, not a Gram kernel")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-precision", action="store_true")
    ap.add_argument("--no-forms", action="store_true")
    ap.add_argument("--no-rows", action="store_true", help="skip the timings of the other entry points")
    ap.add_argument("--rows", default="all",
                    help="comma list of rows to time: kicks (excitation response and its evolution), dsf, rowwise, "
                         "mixture, fused, occupation; default all")
    ap.add_argument("--no-eager", action="store_true", help="skip the eager per-kernel split run")
    ap.add_argument("--headline-only", action="store_true",
                    help="only the timed headline run (implies every --no-* flag): the command whose "
                         "Gram launches a kernel-trace profile averages")
    ap.add_argument("--full-out", default=os.path.join(ROOT, "gpurun_out", "bench_full.json"),
                    help="file for the full result (forms, rows, spectra); the last stdout line is a "
                         "compact summary of it")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="copy this many MiB device-to-device first (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported: walkers are "
              "independent, so N GPUs would each run walkers/N chains and sum their FP64 moments "
              "in one all-reduce, which is not built here", file=sys.stderr)
        sys.exit(2)
    if args.headline_only:
        args.no_cpu = args.no_precision = args.no_forms = args.no_rows = args.no_eager = True
    if args.copy_probe_mb > 0:
        import torch
        n = args.copy_probe_mb * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        log(f"copy probe: {args.copy_probe_mb} MiB read + {args.copy_probe_mb} MiB written")
        del a, b

    steps, warm, tw = args.steps, args.warmup, min(args.trace_walkers, args.walkers)
    est = 2 if args.estimator == "mixture" else 0
    eng = make_engine(args.walkers, 0, steps + warm + ERR_BLOCKS * ERR_BLOCK_SAMPLES, 1, tw, estimator=est,
                      fused=int(args.fused))
    info, dims = eng.info, eng.dims
    sum_d2 = sum(d * d for d in dims)
    log(f"N={N} sectors=1..16 D={info.D} sum_d2={sum_d2} tiles={info.n_tiles} walkers={args.walkers} "
        f"samples_per_launch={info.samples_per_launch}")
    eng.init_walkers()
    step = eng.tune_step()
    eng.burnin(args.burnin)
    log(f"step={step:.4f} lanes_per_walker={eng.info.lanes_per_walker} registers={eng.info.sweep_in_registers} burn-in done")
    if warm > 0:
        eng.sample(warm)
    # Timed region: walkers resident in HBM; ONE le_sample call of `steps` kept samples per
    # walker = ceil(steps / samples_per_launch) rounds of (sweeps + features, replayed from a
    # hipGraph) on the sampler stream overlapped with the Gram of the previous round on the
    # Gram stream, then the FP64 extraction of every S and H block and its copy to the host.
    # le_sample returns after both streams are synchronised; the previous call ended the same way.
    wall, d = timed_sampling(eng, steps, 1)
    updates = args.walkers * steps
    value = updates / wall
    st = eng.stats()
    log(f"timed: {value:.4e} sample-updates/s, {1e3 * wall / steps:.3f} ms/step")
    taus = taus_of(eng, steps, tw)
    # error bars: equal untimed blocks after the timed region
    eng.reset_moments()
    for _ in range(ERR_BLOCKS):
        eng.sample(ERR_BLOCK_SAMPLES)
    t_eig = time.perf_counter()
    eng.spectra(-1)
    eig_ms = 1e3 * (time.perf_counter() - t_eig)
    mean_spec, sig = spectra_blocks(eng, dims)

    gram_flop = 16.0 * sum_d2 * updates
    herm_flop = 8.0 * sum(x * (x + 1) for x in dims) * updates
    launches = max(1, d["gram_launches"])
    # fused steps: the Gram tiles run inside the step launches (timed as a whole, ms_step)
    ms_gram = d["ms_gram"] if not args.fused else d["ms_step"]
    avg_launch_ms = ms_gram / launches
    gram_rate = gram_flop / (ms_gram * 1e-3)
    rows_per_launch = info.samples_per_launch * args.walkers
    traffic, traffic_note = gram_traffic(rows_per_launch)
    tau_max = tau_sh(taus) if taus else None
    tau_max_all = max(taus.values()) if taus else None

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "sample-updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warm,
        "ms_per_step": 1e3 * wall / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded Metropolis walkers on |Psi_L|^2, no dataset)",
        "config": {"workload": "laughlin-edge N=200 nu=1/3 dL=1..16 grouped (D=914), V=0.025 r^2+5e-5 r^4",
                   "walkers": args.walkers, "seed": SEED, "burnin_sweeps": args.burnin, "thin": 1,
                   "estimator": args.estimator,
                   "step": step, "lanes_per_walker": eng.info.lanes_per_walker,
                   "sweep_in_registers": bool(eng.info.sweep_in_registers),
                   "samples_per_launch": info.samples_per_launch, "gram_tiles": info.n_tiles,
                   "gram_xcd_spread": eng.info.gram_xcd_spread, "gram_tile_cols": eng.info.gram_tile_cols,
                   "gram_launches_timed": int(d["gram_launches"]),
                   "sampler_rounds": "hipGraph replay"},
        "roofline": {"bound": "mfma",
                     "kernel": "k_gram<float, 2>" if eng.info.gram_tile_cols == 64 else "k_gram<float, 2, 32>",
                     "achieved": gram_rate / 1e12, "peak": PEAK_FP32 / 1e12, "unit": "TFLOP/s",
                     "frac": gram_rate / PEAK_FP32, "traffic": traffic,
                     "traffic_note": traffic_note,
                     "avg_launch_ms": avg_launch_ms,
                     "launch_time_source": "HIP events recorded on the Gram stream around every Gram launch "
                                           "of the timed le_sample call (mean over its launches)",
                     "algorithmic_flop_per_launch": gram_flop / launches,
                     "flop_per_sample": 16.0 * sum_d2,
                     "algorithmic_bytes_per_launch": rows_per_launch * (8 * info.D + 8),
                     "rows_per_launch": rows_per_launch,
                     "hermitian_flop_per_launch": herm_flop / launches,
                     "hermitian_frac": herm_flop / (ms_gram * 1e-3) / PEAK_FP32,
                     "issued_mfma_flop_per_launch": info.gram_mfma_flop_per_row * updates / launches,
                     "issued_frac": info.gram_mfma_flop_per_row * updates / (ms_gram * 1e-3) / PEAK_FP32,
                     "flop_note": "algorithmic = 16 sum d^2 per sample (complex GEMM count over whole "
                                  "Hermitian blocks, S and H); hermitian = 8 sum d(d+1) (upper triangles "
                                  "only); the kernel issues MFMAs only for the 16x16 blocks holding "
                                  "upper-triangle entries (issued_mfma_flop_per_launch), so issued_frac is "
                                  "the MFMA pipe's actual use"},
        "acceptance": st["acceptance"],
        "ess": {"thin": 1, "tau_int": taus, "tau_max": tau_max_all,
                "ess_per_s": updates / (2.0 * tau_max_all) / wall if tau_max_all else None,
                "tau_max_sh": tau_max,
                "ess_per_s_sh": updates / (2.0 * tau_max) / wall if tau_max else None,
                "traced_walkers": tw, "kept_samples": steps},
        "spectrum_lowest": {str(SECTORS[s]): (float(mean_spec[s][0]) if mean_spec[s] is not None else None)
                            for s in range(len(dims))},
        "eigensolve_ms": {"sectors": len(dims), "largest": max(dims), "ms": eig_ms,
                          "note": "FP64 Cholesky + tridiagonalization + bisection on the GPU, all sectors "
                                  "in one launch, host copies included"},
    }
    if args.fused:
        out["config"]["sampler_rounds"] = "fused launches, one per kept sample (fused_step = 1)"
        out["roofline"].update({
            "kernel": "k_step<8, 25, 16> (fused: sweeps, features and Gram tiles)", "bound": "valu+mfma",
            "launch_time_source": "HIP events around all fused launches of the timed le_sample call "
                                  "(the closing Gram launch included); achieved counts the Gram FLOPs only"})
    iso_ms = eng.stats()["ms_gram_launch_min"]
    if iso_ms > 0:
        iso_rate = 16.0 * sum_d2 * rows_per_launch / (iso_ms * 1e-3)
        out["roofline"]["alone"] = {
            "achieved": iso_rate / 1e12, "frac": iso_rate / PEAK_FP32, "launch_ms": iso_ms,
            "issued_frac": info.gram_mfma_flop_per_row * rows_per_launch / (iso_ms * 1e-3) / PEAK_FP32,
            "note": "shortest full-size launch of the run (HIP events): the last launch of each sampling "
                    "call runs with no sweeps beside it"}
    if not args.no_precision:
        log("precision check (fixed samples vs FP64 oracle)")
        out["precision"] = precision_check(eng, sig)
    eng.close()
    if not args.no_eager:
        # per-kernel split of the sampler stream: the same workload without graphs
        ksteps = min(steps, 80)
        eg = make_engine(args.walkers, 0, ksteps + 8, 1, 0, graphs=0)
        eg.init_walkers()
        eg.tune_step()
        eg.burnin(min(args.burnin, 100))
        eg.sample(8)
        wall_eager, de = timed_sampling(eg, ksteps, 1)
        # the sweep alone: 20 sweeps of every walker in one launch, nothing beside it (le_sample
        # returned with both streams synchronised)
        t_sw = time.perf_counter()
        eg.burnin(20)
        ms_sweep_alone = 1e3 * (time.perf_counter() - t_sw) / 20
        eg.close()
        eager_updates = args.walkers * ksteps
        log(f"eager: {eager_updates / wall_eager:.4e} sample-updates/s")
        sweep_rate = 12.0 * N * (N - 1) * eager_updates / (de["ms_sweep"] * 1e-3)
        stream_bytes = (8 * info.D + 4) * eager_updates
        out["kernels"] = {
            "sampler": {"kernel": f"k_sweep{'_reg' if eng.info.sweep_in_registers else ''}<{eng.info.lanes_per_walker},float>",
                        "bound": "valu",
                        "achieved_tflops": sweep_rate / 1e12, "peak_tflops": PEAK_FP32 / 1e12,
                        "frac": sweep_rate / PEAK_FP32, "flop_per_walker_sweep": 12 * N * (N - 1),
                        "ms_per_sweep": de["ms_sweep"] / ksteps,
                        "ms_per_sweep_alone": ms_sweep_alone,
                        "frac_alone": 12.0 * N * (N - 1) * args.walkers / (ms_sweep_alone * 1e-3) / PEAK_FP32,
                        "note": "ms_per_sweep and frac: eager sampling, beside the Gram stream's launches; "
                                "alone: one launch of 20 sweeps (le_burnin), wall time"},
            "features": {"kernel": "k_psums<float,16> + k_monomials<float>", "bound": "hbm",
                         "ms_per_step": de["ms_features"] / ksteps,
                         "sample_matrix_gb_per_s": stream_bytes / (de["ms_features"] * 1e-3) / 1e9,
                         "bytes_per_sample": 8 * info.D + 4},
            "gram": {"ms_per_step": d["ms_gram"] / steps, "launches": d["gram_launches"]},
            "sampler_graph_ms_per_step": d["ms_sampler_graph"] / steps,
            "eager_sample_updates_per_s": eager_updates / wall_eager,
            "note": "sampler and features share one stream (replayed from hipGraphs in the timed run; "
                    "their split is from an eager run of the same workload); Gram runs on a second "
                    "stream on the other sample buffer",
        }
    if not args.no_forms:
        out["forms"] = forms(args, [int(x) for x in args.thins.split(",")])
    if not args.no_rows:
        out["rows"] = rows(set(args.rows.split(",")) if args.rows != "all" else None)
    if not args.no_cpu:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(value)
        log("small configuration in full on CPU and GPU")
        out["config0_full"] = config0_full()
    if args.full_out:
        os.makedirs(os.path.dirname(os.path.abspath(args.full_out)), exist_ok=True)
        with open(args.full_out, "w") as f:
            json.dump(out, f, indent=1)
        log(f"full result written to {args.full_out}")
    print(json.dumps(compact(out, args.full_out)), flush=True)


def _r(x, n=4):
    """round for the compact line"""
    if x is None:
        return None
    if isinstance(x, float):
        return float(f"{x:.{n}g}")
    return x


def compact(out, full_path):
    """The driver-facing summary line (a few KB): headline, roofline, CPU baseline, ESS ratios,
    precision maxima and spectra agreement; the full result is in `full_path`."""
    keep = ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better",
            "scaling", "vs_baseline", "dtype", "data", "config", "acceptance")
    c = {k: out[k] for k in keep}
    c["value"] = _r(out["value"], 6)
    c["ms_per_step"] = _r(out["ms_per_step"], 5)
    rf = out["roofline"]
    c["roofline"] = {k: _r(rf.get(k), 5) for k in ("bound", "kernel", "achieved", "peak", "unit", "frac", "traffic",
                                                  "avg_launch_ms", "launch_time_source", "algorithmic_flop_per_launch",
                                                  "flop_per_sample", "algorithmic_bytes_per_launch", "rows_per_launch",
                                                  "hermitian_frac", "issued_frac")}
    c["roofline"]["traffic_source"] = "rocprofv3 FETCH_SIZE + WRITE_SIZE passes, copy-calibrated, committed profile"
    if "alone" in rf:
        c["roofline"]["alone"] = {k: _r(rf["alone"][k], 4) for k in ("achieved", "frac", "launch_ms", "issued_frac")}
    e = out["ess"]
    c["ess_headline"] = {"tau_max_sh": _r(e["tau_max_sh"]), "ess_per_s_sh": _r(e["ess_per_s_sh"]),
                         "tau_max": _r(e["tau_max"]), "ess_per_s": _r(e["ess_per_s"]),
                         "kept_samples": e["kept_samples"], "traced_walkers": e["traced_walkers"]}
    if "kernels" in out:
        k = out["kernels"]
        c["sampler"] = {"kernel": k["sampler"]["kernel"], "ms_per_sweep": _r(k["sampler"]["ms_per_sweep"]),
                        "frac_fp32_valu": _r(k["sampler"]["frac"]),
                        "ms_per_sweep_alone": _r(k["sampler"]["ms_per_sweep_alone"]),
                        "frac_fp32_valu_alone": _r(k["sampler"]["frac_alone"])}
        c["features_ms_per_step"] = _r(k["features"]["ms_per_step"])
    if "precision" in out:
        p = out["precision"]
        c["precision"] = {"fixed_samples": p["fixed_samples"]}
        for lab in ("fp32", "bf16x3", "bf16x6"):
            c["precision"][lab] = {k: _r(v, 3) for k, v in p[lab].items()}
    if "forms" in out:
        f = out["forms"]
        cf = {}
        for name, r in f.items():
            if not isinstance(r, dict):
                continue
            row = {"ess_sh": _r(r["ess_per_s_sh"]), "thin_sh": r["best_thin_sh"], "ess_all": _r(r["ess_per_s"])}
            if "ess_sh_ratio_vs_plain_fp64" in r:
                row["ratio_sh"] = _r(r["ess_sh_ratio_vs_plain_fp64"], 3)
                row["ratio_all"] = _r(r["ess_ratio_vs_plain_fp64"], 3)
                jk = r["spectra_vs_plain_fp64_jackknife"]
                bt = r["spectra_vs_plain_fp64"]
                row["lowest_max_sigma_jk"] = _r(max(jk["lowest_per_sector_dev"].values()) if jk["lowest_per_sector_dev"] else None, 3)
                row["lowest_max_sigma_batch"] = _r(max(bt["lowest_per_sector_dev"].values()) if bt["lowest_per_sector_dev"] else None, 3)
                row["within_1sigma_jk"] = _r(jk["frac_within_1_sigma"], 3)
            if "spectra_vs_mixture_fp64" in r:
                jk = r["spectra_vs_mixture_fp64_jackknife"]
                bt = r["spectra_vs_mixture_fp64"]
                row["vs_mixture_fp64"] = {
                    "lowest_max_sigma_batch": _r(max(bt["lowest_per_sector_dev"].values()) if bt["lowest_per_sector_dev"] else None, 3),
                    "lowest_max_sigma_jk": _r(max(jk["lowest_per_sector_dev"].values()) if jk["lowest_per_sector_dev"] else None, 3),
                    "within_1sigma_batch": _r(bt["frac_within_1_sigma"], 3),
                    "within_2sigma_batch": _r(bt["frac_within_2_sigma"], 3),
                    "max_sigma_batch": _r(bt["max_dev_in_combined_sigma"] if "max_dev_in_combined_sigma" in bt else None, 3)}
            cf[name] = row
        c["forms"] = cf
    if "rows" in out:
        c["rows"] = {k: {kk: _r(vv) for kk, vv in v.items() if isinstance(vv, (int, float))} for k, v in out["rows"].items()}
        for name in ("rowwise_N200_dL1-16", "mixture_N200_dL1-16"):
            rw = out["rows"].get(name, {}).get("lowest_eigenvalues")
            if rw:
                c["rows"][name]["lowest_dL16"] = rw["lowest"]["16"]
                c["rows"][name]["max_seed_deviation_sigma"] = rw["max_seed_deviation_sigma"]
    if "cpu_baseline" in out:
        cb = out["cpu_baseline"]
        c["cpu_baseline"] = {k: _r(v, 5) for k, v in cb.items() if k != "legs"}
        c["cpu_baseline"]["legs"] = {n: {k: _r(v, 5) for k, v in leg.items()} for n, leg in cb["legs"].items()}
    if "config0_full" in out:
        c["config0_full"] = {k: _r(v) for k, v in out["config0_full"].items()}
    c["full_result"] = os.path.relpath(full_path, ROOT) if full_path else None
    return c


def gram_traffic(rows_per_launch):
    """HBM bytes per Gram launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE passes
    (copy-calibrated, per sample row), scaled to this launch's rows."""
    path = os.path.join(ROOT, "profiles", "k_gram_traffic_N200_dL1-16.json")
    if not os.path.exists(path):
        return None, "no counter summary"
    with open(path) as f:
        t = json.load(f)
    b = t["fetch_bytes_per_row"] * rows_per_launch + t["write_bytes_per_launch"]
    return b, (f"rocprofv3 FETCH_SIZE x {t['calibration']['fetch_factor']} + WRITE_SIZE x "
               f"{t['calibration']['write_factor']} (copy-calibrated) at {t['rows_per_launch_measured']} "
               f"rows/launch, scaled to {rows_per_launch}; {t['algorithmic_bytes_per_row']} algorithmic B/row")


def _timed(eng, warm, steps):
    """(wall s, Gram ms, kept samples) of `steps` sampling steps after `warm` untimed ones"""
    eng.sample(warm)
    eng.reset_moments()
    g0 = eng.stats()["ms_gram"]
    t0 = time.perf_counter()
    eng.sample(steps)
    return time.perf_counter() - t0, eng.stats()["ms_gram"] - g0


def _rows_kicks(helper_path):
    """excitation response with eight edge kicks and the FP64 evolution of the kicks (see rows)"""
    import ctypes
    from laughlinedge_amd import Engine
    from laughlinedge_amd.engine import W_USER, evolve
    out = {}
    n4, m4, sec4, w4 = 400, 5, list(range(21)), 65536
    r4 = float(np.sqrt(2.0 * m4 * n4))
    helper = ctypes.CDLL(helper_path)
    helper.le_test_set_kicks(8, ctypes.c_double(r4))
    fn = ctypes.cast(helper.le_test_kick_fn, ctypes.c_void_p).value
    log("rows: excitation response m=5 N=400 dL=0..20, 65536 walkers, 8 kicks")
    steps = 4
    with Engine(n4, m4, sec4, w=0.05, eps=2e-5, walkers=w4, samples=steps + 2, seed=2024, tune_sweeps=20,
                step=0.5, n_user_weights=8) as e:
        e.set_weight_callback(fn)
        e.init_walkers()
        e.tune_step()
        e.burnin(50)
        wall, ms_gram = _timed(e, 2, steps)
        d = e.dims
        upd = w4 * steps
        alg = 16.0 * sum(x * x for x in d) + 8 * 8.0 * sum(d[a] * d[b] for a in range(len(d)) for b in range(a, len(d)))
        out["excitation_response_m5_N400_dL0-20"] = {
            "sample_updates_per_s": upd / wall, "ms_per_step": 1e3 * wall / steps, "walkers": w4, "D": e.info.D,
            "gram_algorithmic_flop_per_sample": alg,
            "gram_tflops_algorithmic": alg * upd / (ms_gram * 1e-3) / 1e12,
            "gram_issued_frac": e.info.gram_mfma_flop_per_row * upd / (ms_gram * 1e-3) / PEAK_FP32,
            "gram_ms_per_step": ms_gram / steps, "step": e.stats()["step"]}
        log("rows: FP64 evolution of the 8 kicks, D=2714")
        offs, D = e.offsets, e.info.D
        M = np.zeros((D, D), complex)
        H = np.zeros((D, D), complex)
        for s, (S, Hs) in enumerate(e.moments()):
            o, n = offs[s], d[s]
            M[o:o + n, o:o + n] = (S + S.conj().T) / 2
            H[o:o + n, o:o + n] = (Hs + Hs.conj().T) / 2
        Us = []
        for k in range(8):
            U = np.zeros((D, D), complex)
            for a in range(len(d)):
                for b in range(a, len(d)):
                    blk = e.block(W_USER(k), a, b)
                    U[offs[a]:offs[a] + d[a], offs[b]:offs[b] + d[b]] = blk
                    U[offs[b]:offs[b] + d[b], offs[a]:offs[a] + d[a]] = blk.conj().T
            Us.append((U + U.conj().T) / 2)
    t0 = time.perf_counter()
    rho, nrm, nsteps = evolve(M, H, Us, [Us[0]], lambda t: 1e-3 * np.exp(-(t - 20.0) ** 2 / 50.0), 100.0, 0.1, d)
    wall = time.perf_counter() - t0
    out["evolution_m5_N400_dL0-20_8kicks"] = {
        "wall_s": wall, "controller_steps": int(nsteps), "saved_points": int(rho.shape[1]),
        "max_norm_drift": float(np.max(np.abs(nrm - 1.0))),
        "note": "per angular-momentum block eigenbasis on the GPU, then Cash-Karp 5(4) stages batched "
                "over the 8 kicks; host-side matrix assembly excluded"}
    del M, H, Us
    return out


def rows(only=None):
    """Throughput of the other entry points at their working sizes, each a few seconds:
    - excitation response (em_edgeResponse.tpp:129-164): m = 5, N = 400, dL 0..20 (D = 2714,
      sectors up to 627 states), 65536 walkers; S, H and eight edge kicks
      U_k = sum_i cos(k theta_i) exp(-(r_i - R)^2 / 2), k = 1..8, between every pair of
      sectors.  The kick functor is the one the tests build (tests/cpp/user_weights.hip),
      registered as a user's functor would be;
    - the FP64 evolution of those eight kicks (dynamics.cpp:286-290): M dC/dt = -i (H + F(t) U_k) C,
      F = 1e-3 exp(-(t - 20)^2 / 50), 0 <= t <= 100, saved every 0.1, one observable;
    - the DSF at dL = 16 (em_dsf.cu:113-134): N = 200, sectors {0, 16}, harmonic 16, 32768 walkers;
    - the fused step kernel (fused_step = 1) on the headline workload;
    - the occupation numbers of the dL = 0 state (gf_occupation_numbers.cu:16-239): N = 200,
      two-sided chain of 201 coordinates, 598 orbitals, 32768 walkers.
    Algorithmic FLOP per sample-update follow the Gram count of the headline (8 m n per
    complex d_a x d_b block and weight; 16 sum d^2 for S and H)."""
    from laughlinedge_amd import Engine
    from laughlinedge_amd.engine import OccupationEngine
    out = {}
    want = lambda name: only is None or name in only   # noqa: E731
    helper_path = os.path.join(ROOT, "tests", "cpp", "libuserweights.so")
    if not want("kicks"):
        pass
    elif os.path.exists(helper_path):
        out.update(_rows_kicks(helper_path))
    else:
        out["excitation_response_m5_N400_dL0-20"] = {"skipped": f"{helper_path} is not built"}
    if want("dsf"):
        _row_dsf(out)
    if want("rowwise"):
        _row_rowwise(out)
    if want("mixture"):
        _row_mixture(out)
    if want("fused"):
        _row_fused(out)
    if want("occupation"):
        _row_occupation(out)
    return out


def _row_dsf(out):
    from laughlinedge_amd import Engine
    # DSF at dL = 16
    log("rows: DSF dL=16 N=200")
    steps = 40
    with Engine(N, M_INV, [0, 16], w=W_CONF, eps=EPS_CONF, walkers=WALKERS, samples=steps + 4, seed=SEED,
                tune_sweeps=50, step=0.5, harmonics=(16,), lanes_per_walker=-1) as e:
        e.init_walkers()
        e.tune_step()
        e.burnin(100)
        wall, ms_gram = _timed(e, 4, steps)
        out["dsf_N200_dL16"] = {"sample_updates_per_s": WALKERS * steps / wall, "ms_per_step": 1e3 * wall / steps,
                                "D": e.info.D, "gram_ms_per_step": ms_gram / steps}


def _row_rowwise(out):
    from laughlinedge_amd import Engine
    # the reference's row-wise estimator on the same workload
    log("rows: row-wise estimator N=200 dL=1..16")
    steps, W = 16, 36
    with Engine(N, M_INV, SECTORS, w=W_CONF, eps=EPS_CONF, walkers=W, samples=steps + 4, seed=SEED, tune_sweeps=30,
                step=0.5, estimator=1) as e:
        e.init_walkers()
        e.tune_step()
        e.burnin(100)
        e.sample(4)
        e.reset_moments()
        t0 = time.perf_counter()
        e.sample(steps)
        wall = time.perf_counter() - t0
        chains = e.info.chains
        out["rowwise_N200_dL1-16"] = {
            "chains": chains, "walkers_per_row": W, "ms_per_step": 1e3 * wall / steps,
            "row_samples_per_s": chains * steps / wall,
            "sample_updates_per_s": chains * steps / wall / e.info.D,
            "acceptance": e.stats()["acceptance"],
            "note": "one step = one sweep of every chain + its monomials + the per-row g/f sums; a full "
                    "S,H sample-update needs one sample of every one of the D rows"}
    out["rowwise_N200_dL1-16"]["lowest_eigenvalues"] = rowwise_spectra()


def _row_mixture(out):
    from laughlinedge_amd import Engine
    # the mixture estimator on the same workload (register sweep, FP32 tiles)
    log("rows: mixture estimator N=200 dL=1..16")
    steps = 200
    with Engine(N, M_INV, SECTORS, w=W_CONF, eps=EPS_CONF, walkers=WALKERS, samples=steps + 8, seed=SEED,
                tune_sweeps=50, step=0.5, estimator=2, mixture_rho=MIX_RHO, lanes_per_walker=8,
                sweep_registers=1, symmetry_moves=1) as e:
        e.init_walkers()
        e.tune_step()
        e.burnin(300)
        wall, ms_gram = _timed(e, 8, steps)
        out["mixture_N200_dL1-16"] = {
            "sample_updates_per_s": WALKERS * steps / wall, "ms_per_step": 1e3 * wall / steps,
            "gram_ms_per_step": ms_gram / steps, "acceptance": e.stats()["acceptance"], "rho": MIX_RHO,
            "moves": "single-particle sweeps + centre-of-mass random walk and rotation per kept sample"}
    out["mixture_N200_dL1-16"]["lowest_eigenvalues"] = mixture_spectra()


def _row_fused(out):
    from laughlinedge_amd import Engine
    # the fused step kernel on the headline workload (one launch per kept sample)
    log("rows: fused step N=200 dL=1..16")
    steps = 200
    with Engine(N, M_INV, SECTORS, w=W_CONF, eps=EPS_CONF, walkers=WALKERS, samples=steps + 8, seed=SEED,
                tune_sweeps=50, step=0.5, fused_step=1) as e:
        e.init_walkers()
        e.tune_step()
        e.burnin(300)
        e.sample(8)
        st0 = e.stats()
        t0 = time.perf_counter()
        e.sample(steps)
        wall = time.perf_counter() - t0
        st1 = e.stats()
        out["fused_step_N200_dL1-16"] = {
            "sample_updates_per_s": WALKERS * steps / wall, "ms_per_step": 1e3 * wall / steps,
            "device_ms_per_step": (st1["ms_step"] - st0["ms_step"]) / steps,
            "acceptance": st1["acceptance"],
            "note": "one launch per kept sample: register sweeps, power sums and monomials of the sample and "
                    "the S/H Gram tiles of the previous one as work items (fused_step = 1); the headline "
                    "runs the same work as separate kernels on two streams"}


def _row_occupation(out):
    from laughlinedge_amd.engine import OccupationEngine
    # occupation numbers, dL = 0
    log("rows: occupation numbers dL=0 N=200")
    steps = 40
    lcut = M_INV * (N - 1) + 1
    with OccupationEngine(N, M_INV, 0, lcut, walkers=WALKERS, samples=steps + 4, seed=SEED, tune_sweeps=50,
                          step=0.5, lanes_per_walker=-1) as e:
        e.init_walkers()
        e.tune_step()
        e.burnin(100)
        wall, ms_gram = _timed(e, 4, steps)
        out["occupation_N200_dL0"] = {"sample_updates_per_s": WALKERS * steps / wall, "ms_per_step": 1e3 * wall / steps,
                                      "orbitals": lcut, "gram_ms_per_step": ms_gram / steps,
                                      "lanes_per_walker": e.info.lanes_per_walker, "acceptance": e.stats()["acceptance"]}


def mixture_spectra(seeds=(1, 2), steps=8192):
    """Lowest eigenvalue of every sector dL = 1..16 (N = 200) from the mixture estimator (walkers
    on w |Psi_L|^2 with its centre-of-mass random walk and rotation after every kept sample, S and H
    reweighted by 1 / w), two seeds of 32768 walkers x `steps` samples in 16 blocks: the values,
    their batch sigma, the jackknife values (leading 1/n bias removed) and the seed-to-seed
    deviation in combined batch sigma."""
    from laughlinedge_amd import Engine, geneig
    from laughlinedge_amd.stats import batch_sigma, jackknife_eigenvalues
    runs = []
    for seed in seeds:
        with Engine(N, M_INV, SECTORS, w=W_CONF, eps=EPS_CONF, walkers=WALKERS, estimator=2, mixture_rho=MIX_RHO,
                    tune_sweeps=50, step=0.5, seed=seed, samples=steps, n_blocks=16, lanes_per_walker=8,
                    sweep_registers=1, symmetry_moves=1) as e:
            e.init_walkers()
            e.tune_step()
            e.burnin(600)
            for _ in range(16):
                e.sample(steps // 16)
            sp = [e.spectra(b) for b in range(16)]
            mean = e.spectra()
            blocks = [e.moments(b) for b in range(16)]
        r = {}
        for s, dl in enumerate(SECTORS):
            jk, jks, _ = jackknife_eigenvalues([bl[s] for bl in blocks], geneig)
            r[str(dl)] = (float(mean[s][0]), float(batch_sigma(np.array([b[s][0] for b in sp]))), float(jk[0]),
                          float(jks[0]))
        runs.append(r)
    dev = {k: abs(runs[0][k][0] - runs[1][k][0]) / float(np.hypot(runs[0][k][1], runs[1][k][1])) for k in runs[0]}
    return {"seeds": list(seeds), "samples": WALKERS * steps, "rho": MIX_RHO,
            "lowest": {k: [round(runs[0][k][0], 4), round(runs[0][k][1], 4)] for k in runs[0]},
            "lowest_jackknife": {k: [round(runs[0][k][2], 4), round(runs[0][k][3], 4)] for k in runs[0]},
            "seed_deviation_sigma": {k: round(v, 2) for k, v in dev.items()},
            "max_seed_deviation_sigma": round(max(dev.values()), 2)}


def rowwise_spectra(seeds=(1, 2), steps=512):
    """Lowest eigenvalue of every sector dL = 1..16 (N = 200) from the row-wise estimator, two
    seeds, 36 chains per row, `steps` samples per chain in 8 blocks: the values, their batch
    sigma and the seed-to-seed deviation in combined sigma."""
    from laughlinedge_amd import Engine
    from laughlinedge_amd.stats import batch_sigma
    runs = []
    for seed in seeds:
        with Engine(N, M_INV, SECTORS, w=W_CONF, eps=EPS_CONF, walkers=36, estimator=1, tune_sweeps=30, step=0.5,
                    seed=seed, samples=steps, n_blocks=8, target_acceptance=0.3) as e:
            e.init_walkers()
            e.tune_step()
            e.burnin(300)
            for _ in range(8):
                e.sample(steps // 8)
            sp = [e.spectra(b) for b in range(8)]
            mean = e.spectra()
        runs.append({str(dl): (float(mean[s][0]), float(batch_sigma(np.array([b[s][0] for b in sp]))))
                     for s, dl in enumerate(SECTORS)})
    dev = {k: abs(runs[0][k][0] - runs[1][k][0]) / float(np.hypot(runs[0][k][1], runs[1][k][1])) for k in runs[0]}
    return {"seeds": list(seeds), "samples_per_chain": steps, "chains_per_row": 36,
            "lowest": {k: [round(runs[0][k][0], 4), round(runs[0][k][1], 4)] for k in runs[0]},
            "seed_deviation_sigma": {k: round(v, 2) for k, v in dev.items()},
            "max_seed_deviation_sigma": round(max(dev.values()), 2)}


def forms(args, thins):
    """Sampler/accumulation forms at each thinning: ESS/s, and spectra at the best thinning.

    The plain forms tune the step to the reference's 50% acceptance; the *_wide forms tune it
    to 30%, which moves particles further per sweep (larger jump distance, shorter tau_int).
    After each kept sample's sweeps, the *_cmrot forms add the exact centre-of-mass heat bath
    and a uniform global rotation, and the *_sym forms also the exact dilation heat bath of the
    coordinates relative to the centre of mass.  learned_* train the MLP proposal first (2000
    Adam steps, lr 1e-3, 1e5 configurations, seed 99) and sample with it; the training time is
    reported, not charged to ESS/s.

    Two ESS figures per run, each at its own best thinning: ess_per_s takes the slowest of all
    traced observables (Re/Im p_1..p_4, V and the S and H entries); ess_per_s_sh the slowest S
    or H entry only.  The rotation move decorrelates Re/Im p_k at once but leaves every S and H
    entry unchanged, so only ess_per_s_sh measures what the spectra gain from it."""
    res = {}
    tw = min(args.trace_walkers, args.walkers)
    fs = args.form_steps
    nblk = 32
    # (name, precision, learned, target acceptance, symmetry moves, estimator)
    table = (("plain_fp64", 1, False, 0.5, 0, 0), ("plain_fp32", 0, False, 0.5, 0, 0),
             ("plain_fp32_wide", 0, False, 0.3, 0, 0), ("plain_bf16x6_wide", 3, False, 0.3, 0, 0),
             ("plain_fp32_wide_cmrot", 0, False, 0.3, 1, 0), ("plain_fp32_wide_sym", 0, False, 0.3, 2, 0),
             ("plain_bf16x6_wide_sym", 3, False, 0.3, 2, 0), ("learned_fp32", 0, True, 0.5, 0, 0),
             ("learned_fp32_sym", 0, True, 0.5, 2, 0), ("mixture_fp64", 1, False, 0.5, 0, 2),
             ("mixture_fp32", 0, False, 0.5, 0, 2), ("mixture_fp32_cmrot", 0, False, 0.5, 1, 2))
    for name, prec, learned, target, sym, est in table:
        eng = make_engine(args.walkers, prec, (len(thins) + 1) * (fs + 4), nblk, tw, target, sym=sym, estimator=est)
        eng.init_walkers()
        step = eng.tune_step()
        eng.burnin(args.burnin)
        r = {"precision": ["fp32", "fp64", "bf16x3", "bf16x6"][prec], "step": step, "target_acceptance": target,
             "symmetry_moves": ["none", "centre of mass + rotation", "centre of mass + rotation + dilation"][sym],
             "estimator": ["plain", "rowwise", "mixture"][est]}
        if learned:
            t0 = time.perf_counter()
            rep = eng.train_proposal(steps=2000, lr=1e-3, batch=4096, n_configs=100000, snapshot_sweeps=2, seed=99)
            r["training"] = dict(rep, wall_s=time.perf_counter() - t0)
            eng.burnin(50)
        runs = []
        for thin in thins:
            eng.set_thin(thin)
            eng.sample(4)
            wall, d = timed_sampling(eng, fs, 1)
            taus = taus_of(eng, fs, tw)
            sh = {k: v for k, v in taus.items() if k not in PK_NAMES}
            tmax, tmax_sh = max(taus.values()), tau_sh(taus)
            # mixture chains: the reweighted estimates keep kish = (sum 1/w)^2 / (n sum 1/w^2) of the
            # samples (importance weights 1/w of the walkers' current configurations)
            kish = mixture_kish(eng) if est == 2 else 1.0
            ess = kish * args.walkers * fs / (2.0 * tmax) / wall
            ess_sh = kish * args.walkers * fs / (2.0 * tmax_sh) / wall
            runs.append({"thin": thin, "ms_per_kept_sample": 1e3 * wall / fs,
                         "ms_sampler_per_kept_sample": (d["ms_sweep"] + d["ms_features"] + d["ms_sampler_graph"]) / fs,
                         "tau_max": tmax, "tau_max_obs": max(taus, key=taus.get),
                         "tau_max_sh": tmax_sh, "tau_max_sh_obs": max(sh, key=sh.get),
                         "tau_re_im_pk": tau_pk(taus), "tau_V": taus["V"], "tau_p1sq": taus["|p1|^2"],
                         "ess_per_s": ess, "ess_per_s_sh": ess_sh, "kish": kish,
                         "acceptance": eng.stats()["acceptance"]})
            log(f"form {name} thin {thin}: {1e3 * wall / fs:.3f} ms/sample tau_max {tmax:.2f} (S/H {tmax_sh:.2f}) "
                f"ESS/s {ess:.3e} (S/H {ess_sh:.3e})")
        best = max(runs, key=lambda x: x["ess_per_s"])
        best_sh = max(runs, key=lambda x: x["ess_per_s_sh"])
        r["best_thin"] = best["thin"]
        r["ess_per_s"] = best["ess_per_s"]
        r["best_thin_sh"] = best_sh["thin"]
        r["ess_per_s_sh"] = best_sh["ess_per_s_sh"]
        r["runs"] = runs
        # spectra with batch errors from nblk blocks at the best thinning for S and H
        eng.set_thin(best_sh["thin"])
        eng.sample(4)
        timed_sampling(eng, 4 * fs, nblk)
        r["_spec"], r["_sig"] = spectra_blocks(eng, eng.dims)
        jk = spectra_jackknife(eng, eng.dims)
        r["_jspec"] = [None if j is None else j[0] for j in jk]
        r["_jsig"] = [None if j is None else j[1] for j in jk]
        eng.close()
        res[name] = r
    ref = res["plain_fp64"]
    mref = res["mixture_fp64"]
    for name, *_ in table[1:]:
        r = res[name]
        r["ess_ratio_vs_plain_fp64"] = r["ess_per_s"] / ref["ess_per_s"]
        r["ess_sh_ratio_vs_plain_fp64"] = r["ess_per_s_sh"] / ref["ess_per_s_sh"]
        r["spectra_vs_plain_fp64"] = compare_spectra(r["_spec"], r["_sig"], ref["_spec"], ref["_sig"])
        r["spectra_vs_plain_fp64_jackknife"] = compare_spectra(r["_jspec"], r["_jsig"], ref["_jspec"], ref["_jsig"])
        if name.startswith("mixture_") and name != "mixture_fp64":
            r["spectra_vs_mixture_fp64"] = compare_spectra(r["_spec"], r["_sig"], mref["_spec"], mref["_sig"])
            r["spectra_vs_mixture_fp64_jackknife"] = compare_spectra(r["_jspec"], r["_jsig"], mref["_jspec"],
                                                                     mref["_jsig"])
    for name, *_ in table:
        for k in ("_spec", "_sig", "_jspec", "_jsig"):
            res[name].pop(k)
    res["note"] = ("independent chains per form: spectra differences are statistical, so about 68% "
                   "(95%) of eigenvalues are expected within 1 (2) combined sigma (the eigenvalues of one "
                   "sector move together, so these fractions scatter widely, and the high eigenvalues of "
                   "large sectors rest on heavy-tailed entries such as |p_1|^(2 dL), whose batch sigma "
                   "underestimates the error); median of a Gaussian |deviation| is 0.67; batch sigma from "
                   "32 blocks of 12 kept samples at the best thinning for S and H; the systematic error of "
                   "each accumulation on identical samples is in 'precision'; "
                   "spectra_vs_plain_fp64_jackknife compares the delete-one-block jackknife estimates "
                   "(the O(1/samples) low bias of the noisy Rayleigh-Ritz values removed) with their "
                   "jackknife sigma, over the same 32 blocks; the mixture_* forms sample w |Psi_L|^2, "
                   "whose reweighted S and H integrands are bounded, so their batch sigma is honest: "
                   "spectra_vs_mixture_fp64 compares them with the FP64 form of the same estimator, and "
                   "their ESS carries the Kish factor of the 1/w weights")
    return res


def precision_check(eng, sig, n_fixed=8192):
    """FP32 and BF16-split GPU moments vs the FP64 CPU oracle on identical configurations
    (the production walkers' current positions)."""
    from oracle.oracle import Oracle, geneig
    from laughlinedge_amd import Engine
    pos = eng.positions()[: n_fixed]
    o = Oracle(N, M_INV, SECTORS, w=W_CONF, eps=EPS_CONF)
    t0 = time.perf_counter()
    So, Ho, _ = o.accumulate_fixed(pos)
    t_cpu = time.perf_counter() - t0
    n = pos.shape[0]
    out = {"fixed_samples": int(n), "oracle_seconds": t_cpu}
    for label, prec in (("fp32", 0), ("bf16x3", 2), ("bf16x6", 3)):
        with Engine(N, M_INV, SECTORS, w=W_CONF, eps=EPS_CONF, walkers=1024, samples=1,
                    precision=prec) as eg:
            eg.accumulate_positions(pos)
            mom = eg.moments()
            spec = eg.spectra()
        rs, rh, eig_rel, eig_sig = [], [], [], []
        for s, (Sg, Hg) in enumerate(mom):
            S64, H64 = So[s] / n, Ho[s] / n
            rs.append(float(np.linalg.norm(Sg - S64) / np.linalg.norm(S64)))
            rh.append(float(np.linalg.norm(Hg - H64) / np.linalg.norm(H64)))
            try:
                e64 = geneig(S64, H64)
            except np.linalg.LinAlgError:
                continue
            if spec[s] is None:
                continue
            eig_rel.append(float(np.max(np.abs(spec[s] - e64) / np.abs(e64))))
            if sig[s] is not None:
                eig_sig.append(float(np.max(np.abs(spec[s] - e64) / np.maximum(sig[s], 1e-300))))
        out[label] = {"rel_fro_S_max": max(rs), "rel_fro_H_max": max(rh),
                      "eig_rel_err_max": max(eig_rel) if eig_rel else None,
                      "eig_shift_in_mc_sigma_max": max(eig_sig) if eig_sig else None}
    return out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(gpu_value):
    """CPU restatements of the reference's estimators (FP64, OpenMP, host cores), each on a
    bounded sample of the headline workload (N=200, m=3, dL=1..16, D=914), burn-in included:

    rowwise  the reference's own estimator (em_spectrum.cu:16-137, 318-390; oracle
             or_rowwise_sector): one |p_f Psi_L|^2 chain per basis row f accumulating g/f,
             V g/f and |g/f|^2 over the row's sector.  One kept sample of one chain fills one
             row, so a full S,H sample-update (every row of every sector sampled once) costs D
             chain samples: sample-updates/s = row-samples/s / D.  Timed on the rows of
             sector dL=10 (42 rows); the cost per row-sample does not depend on the sector (the
             sweep is O(N^2), the features O(N K + D)).
    plain    the plain |Psi_L|^2 chain with the S and H accumulation (oracle or_plain_run),
             the estimator the GPU path computes.
    `value` is the rowwise leg, the reference's method."""
    from oracle.oracle import Oracle
    threads = min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", "16")))
    o = Oracle(N, M_INV, SECTORS, w=W_CONF, eps=EPS_CONF)
    D = sum(o.dims)
    legs = {}
    sec = SECTORS.index(10)
    d = o.dims[sec]
    walkers, burn, samples = 4 * threads, 10, 60
    t0 = time.perf_counter()
    o.rowwise_sector(sec, walkers, burn, samples, 0.5, SEED, threads=threads)
    dt = time.perf_counter() - t0
    row_rate = d * walkers * samples / dt
    legs["rowwise"] = {"value": row_rate / D, "unit": "sample-updates/s", "row_samples_per_s": row_rate,
                       "seconds": dt,
                       "sample": f"sector dL=10: {d} rows x {walkers} walkers x ({burn} burn-in + {samples} kept "
                                 f"sweeps); sample-updates/s = row-samples/s / D (D={D})"}
    walkers, burn, samples = 32 * threads, 20, 400
    t0 = time.perf_counter()
    o.plain_run(walkers, burn, samples, 0.5, SEED, threads=threads)
    dt = time.perf_counter() - t0
    legs["plain"] = {"value": walkers * samples / dt, "unit": "sample-updates/s", "seconds": dt,
                     "sample": f"{walkers} walkers x ({burn} burn-in + {samples} kept sweeps)"}
    r = legs["rowwise"]
    return {"value": r["value"], "unit": "sample-updates/s", "cores": threads, "kind": "port",
            "estimator": "rowwise (the reference's)", "host_logical_cpus": os.cpu_count(),
            "cpu_model": cpu_model(), "sample": r["sample"],
            "gpu_over_cpu": gpu_value / r["value"],
            "gpu_over_cpu_plain": gpu_value / legs["plain"]["value"], "legs": legs}


def config0_full():
    """The small plain configuration in full (N=10, m=3, dL 0..6, 64 walkers, seed 1234,
    step 0.5, 2000 burn-in sweeps, 2e4 samples per walker, FP64) on the oracle port and on the
    GPU: the two chains follow the same random streams, so S and H agree to rounding."""
    from oracle.oracle import Oracle
    from laughlinedge_amd import Engine
    threads = min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", "16")))
    n10, secs = 10, list(range(7))
    o = Oracle(n10, 3, secs, w=0.05, eps=1e-4)
    t0 = time.perf_counter()
    So, Ho, _, _, _ = o.plain_run(64, 2000, 20000, 0.5, 1234, threads=threads)
    cpu_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    with Engine(n10, 3, secs, w=0.05, eps=1e-4, walkers=64, precision=1, samples=20000, seed=1234,
                tune_sweeps=0, step=0.5) as e:
        e.init_walkers()
        e.burnin(2000)
        e.sample(20000)
        mom = e.moments()
    gpu_s = time.perf_counter() - t0
    rel = max(max(np.linalg.norm(Sg - So[s]) / np.linalg.norm(So[s]),
                  np.linalg.norm(Hg - Ho[s]) / np.linalg.norm(Ho[s])) for s, (Sg, Hg) in enumerate(mom))
    return {"cpu_seconds": cpu_s, "cpu_threads": threads, "gpu_seconds": gpu_s,
            "max_rel_diff_S_H": float(rel), "samples": 64 * 20000}


if __name__ == "__main__":
    main()
