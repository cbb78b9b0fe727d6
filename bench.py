#!/usr/bin/env python3
"""Unweighted events/s at 1% relative error on one MI355X.

Procedure (per workload):
  1. adaptation: MultiChannel::Optimize semantics (include/Channel/
     MultiChannel.hh:253-268) -- iterate with training, adapt alpha and every
     channel's VEGAS grid -- for the workload's iteration count, recording
     when the pooled relative error first reaches 1%;
  2. freeze grids and alpha;
  3. `--warmup` untimed then `--steps` timed production steps, one step = one
     batch of accepted points through the whole hot path (pick, bucket,
     forward map + cut, every channel's inverse map + VEGAS density,
     multichannel weight, statistics), values kept in HBM;
  4. eps_unw = <w> / w_max over the production sample (also with w_max = the
     99.9% quantile), unweighted events/s = accepted points/s x eps_unw.
The roofline entry prices the dominant kernel (k_main, the all-channel
inverse/weight kernel) by its algorithmic FP64 flop-equivalents per point
(N_arith + 20 N_trans, counted by the op-counting CPU oracle) over its
HIP-event duration.  The CPU baseline times the oracle's scalar restatement
of the reference loop on the host on a bounded sample with the same frozen
state.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "unweighted events/s at 1% rel. err: plain VEGAS vs flow, 1 MI355X"
FP64_PEAK_TFLOPS = 78.6
SEED = 0xC0FFEE00


def workloads():
    from chili_amd import processes as P
    return {
        "gg_4g": dict(proc=lambda: P.gluon_process(4, smax=3), f=lambda: P.integrand_propprod(), bins=64, batch=1 << 20,
                      iters=10, cfg=1,
                      desc="2->4 massless gluons, sqrt(s)=1 TeV, 15 s/t channels (smax=3), 1/s propagator maps, "
                           "cuts pT>20 |y|<2.5 dR>0.4, 10 dims x 64 bins, 2^20 points/iteration"),
        "tt_6j": dict(proc=lambda: P.tt_process(smax=5), f=lambda: P.integrand_resonant(), bins=64, batch=1 << 22,
                      iters=10, cfg=4,
                      desc="2->6 g g -> t t~ -> 6 jets, t(173,1.42) -> W(80.4,2.08) j, W -> j j, sqrt(s)=1 TeV, "
                           "2 s/t channels (smax=5), BW x quadratic form in s_ij, 16 dims x 64 bins"),
    }


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def adapt(gpu, ora_lib, W, uni_seed):
    """Optimize: iterate + train (alpha, grids), pooled StatsData."""
    from chili_amd import abi, lib
    pooled = np.zeros(6)
    pooled[1], pooled[2] = np.finfo(np.float64).max, np.finfo(np.float64).tiny
    t0 = time.perf_counter()
    t_1pct = None
    hist_rel = []
    best_alpha, min_diff = gpu.alpha(), np.inf
    for it in range(W["iters"]):
        uni = abi.philox_uniforms(uni_seed, stream=it)
        st, T, hist, na = gpu.iterate(uni, W["batch"], train=True)
        if np.isnan(T).any():
            raise RuntimeError("alpha training became NaN")
        # MultiChannel::Train: alpha adapt, per-channel grid adapt, MaxDifference
        alpha = lib.alpha_adapt(gpu.alpha(), T)
        grids = gpu.grids()
        for c in range(gpu.nchannels):
            if np.any(hist[c] != 0):
                grids[c] = lib.grid_adapt(grids[c], hist[c], 1.0)
        gpu.set_grids(grids)
        gpu.set_alpha(alpha)
        md = lib.max_difference(T)
        if md < min_diff:
            best_alpha, min_diff = alpha.copy(), md
        pooled[0] += st.n; pooled[3] += st.sum; pooled[4] += st.sum2; pooled[5] += st.n_finite
        mean = pooled[3] / pooled[0]
        err = np.sqrt((pooled[4] / pooled[0] - mean * mean) / (pooled[0] - 1))
        rel = err / abs(mean)
        hist_rel.append(rel)
        el = time.perf_counter() - t0
        if t_1pct is None and rel <= 0.01:
            t_1pct = el
        log(f"  iter {it}: I={st.sum / st.n:.6e} +- {st.error():.2e}  pooled {mean:.6e} rel {rel * 100:.3f}%  "
            f"acc {W['batch']}/{na} ({W['batch'] / na * 100:.1f}%)  {el:.2f}s")
    return dict(pooled_mean=float(mean), pooled_rel_err=float(rel), time_adapt_s=time.perf_counter() - t0,
                time_to_1pct_s=t_1pct, rel_err_history=[float(x) for x in hist_rel])


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="gg_4g", choices=sorted(workloads()))
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="CPU baseline sample length")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one MI355X: this build has no multi-GPU path (each GPU would run independent "
              "point batches and all-reduce C*d*B histograms per iteration); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    from chili_amd import abi, lib
    from oracle import oracle as O

    W = workloads()[args.workload]
    if lib.device_count() < 1:
        raise SystemExit("no HIP device")
    torch.cuda.init()
    proc, f = W["proc"](), W["f"]()
    ch = lib.Channels(proc)
    gpu = lib.MultiChannelGPU(ch, W["bins"], f, device=0)
    log(f"workload {args.workload}: {gpu.nchannels} channels, {gpu.ndims} dims, {gpu.nbins} bins")
    seed = SEED + W["cfg"]
    ad = adapt(gpu, None, W, seed)

    # ---- production with frozen state ----
    n = W["batch"]
    vals = torch.empty(n * (args.steps + 1), dtype=torch.float64, device="cuda:0")
    for s in range(args.warmup):
        gpu.iterate(abi.philox_uniforms(seed, stream=1000 + s), n, train=False)
    gpu.timing(enable=True, reset=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    atts = 0
    for s in range(args.steps):
        st, _, _, na = gpu.iterate(abi.philox_uniforms(seed, stream=2000 + s), n, train=False,
                                   vals_ptr=ctypes.c_void_p(vals.data_ptr() + 8 * n * s), vals_cap=n)
        atts += na
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    tm = gpu.timing(enable=False)
    w = vals[: n * args.steps]
    wmax = float(w.max())
    wmean = float(w.mean())
    ws = torch.sort(w).values
    q999 = float(ws[int(0.999 * (ws.numel() - 1))])
    eps = wmean / wmax
    eps999 = wmean / q999
    pts_s = n * args.steps / dt
    unw_s = pts_s * eps
    ms_step = dt / args.steps * 1e3

    # ---- roofline of k_main ----
    ora = O.Oracle(proc, gpu.nbins, f, count=True)
    ora.set_grids(gpu.grids())
    ora.set_alpha(gpu.alpha())
    O.lib(True).oracle_count_reset()
    cnt = ora.count_points(abi.philox_uniforms(seed, stream=99), 0, 4000)
    f_main = (cnt["wgt_arith"] + 20 * cnt["wgt_trans"]) / max(cnt["accepted"], 1)
    f_fwd = (cnt["fwd_arith"] + 20 * cnt["fwd_trans"]) / cnt["attempts"]
    avg_ms = tm["main_ms"] / max(tm["launches"], 1)
    pts_launch = tm["points"] / max(tm["launches"], 1)
    achieved = f_main * pts_launch / (avg_ms * 1e-3) / 1e12
    roof = dict(bound="mfma", achieved=round(achieved, 4), peak=FP64_PEAK_TFLOPS, unit="TFLOP/s",
                frac=round(achieved / FP64_PEAK_TFLOPS, 5), traffic=None,
                kernel="k_main (all-channel inverse maps + VEGAS densities + multichannel weight)",
                compute="FP64 VALU; MI355X's FP64 vector and matrix peaks are both 78.6 TFLOP/s",
                flop_eq_per_point=round(f_main, 1), points_per_launch=round(pts_launch, 1),
                avg_launch_ms=round(avg_ms, 4), launches=tm["launches"],
                forward_kernel_ms_per_step=round(tm["forward_ms"] / args.steps, 4))

    # ---- CPU baseline: scalar oracle, one core, bounded sample ----
    cpu = None
    if not args.no_cpu:
        orc = O.Oracle(proc, gpu.nbins, f)
        orc.set_grids(gpu.grids())
        orc.set_alpha(gpu.alpha())
        probe = 2000
        t = time.perf_counter()
        orc.iterate(abi.philox_uniforms(seed, stream=3000), probe, train=False)
        rate = probe / (time.perf_counter() - t)
        ncpu = int(max(probe, min(rate * args.cpu_seconds, 5e6)))
        t = time.perf_counter()
        sc, _, _, nac = orc.iterate(abi.philox_uniforms(seed, stream=3001), ncpu, train=False)
        cdt = time.perf_counter() - t
        cpu_pts = ncpu / cdt
        cpu = dict(value=round(cpu_pts * eps, 3), unit="events/s", cores=1, kind="port",
                   points_per_s=round(cpu_pts, 1),
                   sample=f"{ncpu} accepted points ({nac} attempts) of the frozen {args.workload} state in {cdt:.1f} s on one "
                          f"host core; unweighted rate uses the GPU production sample's eps_unw")

    out = {
        "metric": METRIC, "value": round(unw_s, 1), "unit": "events/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_step, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (Philox4x32-10 uniforms)",
        "config": {"workload": args.workload, "description": W["desc"], "channels": gpu.nchannels, "dims": gpu.ndims,
                   "bins": gpu.nbins, "points_per_step": n, "path": "plain multichannel VEGAS"},
        "points_per_s": round(pts_s, 1), "eps_unw": eps, "eps_unw_q999": eps999, "w_max": wmax, "w_mean": wmean,
        "attempts_per_step": atts / args.steps,
        "adaptation": ad, "roofline": roof, "cpu_baseline": cpu,
        "gpu_vs_cpu": round(pts_s / cpu["points_per_s"], 1) if cpu else None,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
