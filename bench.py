#!/usr/bin/env python3
"""Unweighted events/s at 1% relative error on one MI355X.

Procedure (the 2->6 resonant workload by default):
  1. adaptation: the product's MultiChannel::Optimize (chili_mc_optimize:
     iterate with training, alpha and grid adaptation, MaxDifference; the
     workload's iteration count, fixed bin count), recording when the pooled
     relative error first reaches 1%;
  2. freeze grids and alpha;
  3. `--warmup` untimed then `--steps` timed production steps, one step = one
     batch of accepted points through the whole hot path (pick, bucket,
     forward map + cut, every channel's inverse map + VEGAS density,
     multichannel weight, statistics), values kept in HBM;
  4. eps_unw = <w> / w_max over the production sample (also with w_max = the
     99.9% quantile), derived unweighted events/s = accepted points/s x eps_unw;
  5. flow leg: a coupling-layer spline flow per channel is trained on points
     and weights the frozen plain proposal generates, the training epoch is
     chosen on a validation sample and the defensive-mixture weight on a
     selection sample (their own Philox streams), then the production steps
     run through the flow-assisted path;
  6. the headline: each leg's unweighting stage (chili_mc_generate /
     chili_mc_flow_generate, with the exact early rejection of
     chili_mc_specialize bit 3: the same events as weighting every point)
     generates unweighted events against its production w_max; the leg with
     more events/s on a selection pass (plain multichannel VEGAS the
     fallback) is timed over `--warmup` + `--steps` passes, one step = one
     pass of 2^22 events, and `value` = generated events/s.  Both legs'
     production and generation figures are reported.
The roofline entry prices the dominant plain kernel (k_main, the all-channel
inverse/weight kernel) by its algorithmic FP64 flop-equivalents per point
(N_arith + 20 N_trans, counted by the op-counting CPU oracle) over its
HIP-event duration; roofline_flow prices the flow leg's coupling-layer
kernel (k_couple) by its algorithmic BF16 conditioner flops over its
HIP-event duration.  The CPU baseline times the oracle's scalar restatement
of the reference loop on the host on a bounded sample with the same frozen
state.  Short legs for the other workloads (plain VEGAS e+e- -> mu+mu-, the
2->4 multichannel run, 32 batched jobs) are reported under `configs`.

    python bench.py [--steps K --warmup W]                 the default line
    python bench.py --workload gg_4g                       the 2->4 workload
    python bench.py --kernel couple --dims 16 --layers 8   k_couple alone (profiling)
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
BF16_PEAK_TFLOPS = 2500.0
SEED = 0xC0FFEE00
# Philox stream ranges (each sample its own): adaptation 0.., kernel-variant
# timing 900.., production 2000.., unweighting 3000.., flow training records
# 500.., flow validation 700.., selection samples 4000..
S_PROD, S_SEL = 2000, 4000


def workloads():
    from chili_amd import processes as P
    return {
        "gg_4g": dict(proc=lambda: P.gluon_process(4, smax=3), f=lambda: P.integrand_propprod(), bins=64, batch=1 << 20,
                      iters=10, cfg=1, flow_layers=6,
                      desc="2->4 massless gluons, sqrt(s)=1 TeV, 15 s/t channels (smax=3), 1/s propagator maps, "
                           "cuts pT>20 |y|<2.5 dR>0.4, 10 dims x 64 bins, 2^20 points/iteration"),
        "tt_6j": dict(proc=lambda: P.tt_process(smax=5), f=lambda: P.integrand_resonant(), bins=64, batch=1 << 24,
                      iters=10, cfg=4, flow_layers=8,
                      desc="2->6 g g -> t t~ -> 6 jets, t(173,1.42) -> W(80.4,2.08) j, W -> j j, sqrt(s)=1 TeV, "
                           "2 s/t channels (smax=5), BW x quadratic form in s_ij, 16 dims x 64 bins, 2^24 points/iteration"),
    }


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def adapt(gpu, W, seed, iters=None, ncalls=None):
    """MultiChannel::Optimize through the product (chili_mc_optimize), one
    iteration per call so the time to 1% is recorded; the workload's bin
    count is fixed (nrefine = 0) and rtol = 0 runs every iteration."""
    from chili_amd import abi, lib
    iters = iters or W["iters"]
    params = lib.mc_params(ncalls=ncalls or W["batch"], niterations=iters, rtol=0.0, nrefine=0)
    summ = lib.mc_summary()
    best = gpu.alpha().copy()
    t0 = time.perf_counter()
    t_1pct, hist_rel = None, []
    uni = abi.philox_uniforms(seed, stream=0)
    for it in range(iters):
        st = gpu.optimize(uni, params, summ, best, max_steps=1)[0]
        if np.isnan(gpu.alpha()).any():
            raise RuntimeError("alpha training became NaN")
        rel = summ.sum.error() / abs(summ.sum.mean())
        hist_rel.append(rel)
        el = time.perf_counter() - t0
        if t_1pct is None and rel <= 0.01:
            t_1pct = el
        log(f"  iter {it}: I={st.sum / st.n:.6e} +- {st.error():.2e}  pooled {summ.sum.mean():.6e} rel {rel * 100:.3f}%  "
            f"acc {params.ncalls}/{int(st.n)} ({params.ncalls / st.n * 100:.1f}%)  {el:.2f}s")
    return dict(pooled_mean=float(summ.sum.mean()), pooled_rel_err=float(hist_rel[-1]), time_adapt_s=time.perf_counter() - t0,
                time_to_1pct_s=t_1pct, rel_err_history=[float(x) for x in hist_rel], driver="chili_mc_optimize",
                params=dict(ncalls=int(params.ncalls), niterations=iters, nrefine=0, beta=params.beta, min_alpha=params.min_alpha))


def measured_traffic(workload, variant, pts_launch):
    """HBM bytes per k_main launch (variant "forward": per forward-kernel
    launch, pts_launch attempts) at this launch size, projected from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE measurement (profiles/),
    scaled by the counters' reading of a copy kernel of known size; None if
    absent."""
    if variant == "forward":
        path, key = os.path.join(ROOT, "profiles", f"k_forward_traffic_{workload}.json"), "bytes_per_attempt"
    else:
        kind = "split" if "split" in variant else ("specialized" if variant in ("specialized", "specialized-weight") else "generic")
        path, key = os.path.join(ROOT, "profiles", f"k_main_traffic_{workload}_{kind}.json"), "bytes_per_point"
    if not os.path.exists(path):
        return None, None
    with open(path) as fh:
        t = json.load(fh)
    return round(t[key] * pts_launch, 1), t[key]


def fixture_f_main(workload):
    """F_main of the committed op-count fixture (tests/golden/opcounts.json:
    uniform grids, alpha = 1/C), for reference next to the run's own count."""
    path = os.path.join(ROOT, "tests", "golden", "opcounts.json")
    try:
        with open(path) as fh:
            return json.load(fh)["workloads"][workload]["F_main_per_point"]
    except (OSError, KeyError, ValueError):
        return None


def _cpu_worker(job):
    """One host core: the oracle's MultiChannel loop over a share of the sample."""
    proc_name, grids, alpha, seed, stream, ncalls = job
    from chili_amd import abi
    from oracle import oracle as O
    W = workloads()[proc_name]
    orc = O.Oracle(W["proc"](), grids.shape[-1] - 1, W["f"]())
    orc.set_grids(grids)
    orc.set_alpha(alpha)
    t = time.perf_counter()
    orc.iterate(abi.philox_uniforms(seed, stream=stream), ncalls, train=False)
    return ncalls, time.perf_counter() - t


def cpu_baseline(proc, f, grids, alpha, seed, seconds, eps, workload):
    """The oracle's scalar restatement of the reference loop, one process per
    host core (at most 16, the box's share), each on its own stream; about
    `seconds` of work per core.  Rate = total accepted points / slowest worker."""
    import concurrent.futures as cf
    import multiprocessing as mp
    from chili_amd import abi
    from oracle import oracle as O
    orc = O.Oracle(proc, grids.shape[-1] - 1, f)
    orc.set_grids(grids)
    orc.set_alpha(alpha)
    probe = 2000
    t = time.perf_counter()
    orc.iterate(abi.philox_uniforms(seed, stream=3000), probe, train=False)
    rate1 = probe / (time.perf_counter() - t)
    ncores = max(1, min(16, len(os.sched_getaffinity(0))))
    per = int(max(probe, min(rate1 * seconds, 5e6)))
    jobs = [(workload, grids, alpha, seed, 3001 + k, per) for k in range(ncores)]
    with cf.ProcessPoolExecutor(max_workers=ncores, mp_context=mp.get_context("spawn")) as ex:
        res = list(ex.map(_cpu_worker, jobs))
    tot = sum(r[0] for r in res)
    wall = max(r[1] for r in res)
    pts = tot / wall
    return dict(value=round(pts * eps, 3), unit="events/s", cores=ncores, kind="port", points_per_s=round(pts, 1),
                points_per_s_one_core=round(rate1, 1),
                sample=f"{ncores} host processes x {per} accepted points of the frozen {workload} state "
                       f"({wall:.1f} s, oracle restatement of the reference loop, -O2); unweighted rate uses the GPU "
                       f"production sample's eps_unw")


def couple_flops_per_row(masks, d, K=16, H=64):
    """Algorithmic BF16 flops of one row through every coupling layer's
    conditioner: 2 (nA H + 2 H^2 + H nB (3K - 1)) per layer."""
    return sum(2 * ((d - len(B)) * H + 2 * H * H + H * len(B) * (3 * K - 1)) for B in masks)


# ---------------------------------------------------------------------------
# legs of the other workloads (short; reported under "configs")
# ---------------------------------------------------------------------------
def leg_vegas_eemumu():
    """Plain VEGAS e+e- -> mu+mu- (2 dims x 50 bins): Vegas::Optimize through
    the product (chili_vegas_optimize: 10 iterations of 1e4 points, refinement
    every 5 -> 200 bins, ncalls 4e4), then 1e5 production points, timed."""
    import torch
    from chili_amd import abi, lib
    from chili_amd import processes as P
    f = P.integrand_eemumu()
    v = lib.VegasGPU(2, 50, f)
    p = lib.vegas_params(ncalls=10000, nrefine=5, niterations=10)
    tot = abi.Stats(0, np.finfo(float).max, np.finfo(float).tiny, 0, 0, 0)
    t0 = time.perf_counter()
    res, n = v.optimize(abi.philox_uniforms(SEED, stream=0), p, tot, 0, 10)
    t_opt = time.perf_counter() - t0
    v.iterate(abi.philox_uniforms(SEED, stream=50), 100000)         # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    reps = 20
    for k in range(reps):
        st, _ = v.iterate(abi.philox_uniforms(SEED, stream=51 + k), 100000)
    dt = (time.perf_counter() - t0) / reps
    return dict(workload="ee_mumu: e+e- -> mu+mu- plain VEGAS, 2 dims, 50 bins refined to 200, 10 x 1e4 (x2 per refinement) "
                         "then 1e5 points", optimize_iterations=n, bins_after=v.nbins, ncalls_after=int(p.ncalls),
                optimize_s=round(t_opt, 4), integral=float(tot.sum / tot.n), rel_err=float(tot.error() / abs(tot.sum / tot.n)),
                production_points=100000, production_ms=round(dt * 1e3, 4), points_per_s=round(100000 / dt, 1),
                note="a 1e5-point iteration is launch-latency bound (two kernels)")


def timed_steps(step, n, steps, seed, vals):
    """`steps` production steps of n accepted points (values into vals),
    after one warm-up step: (seconds, attempts, <w>, w_max)."""
    import torch
    from chili_amd import abi
    step(abi.philox_uniforms(seed, stream=1000), None, 0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    atts = 0
    for s in range(steps):
        atts += step(abi.philox_uniforms(seed, stream=S_PROD + s), ctypes.c_void_p(vals.data_ptr() + 8 * n * s), n)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    w = vals[: n * steps]
    return dt, atts, float(w.mean()), float(w.max())


VARIANTS = {0: "generic", 1: "specialized-forward", 2: "specialized-weight", 3: "specialized", 4: "split-weight",
            5: "specialized-forward+split-weight"}


def choose_variant(gpu, step, seed, modes=(1, 2, 3, 4, 5), log_fn=None):
    """Times one production step (a warm-up, then the fastest of three) with
    the generic kernels and with each run-time specialised mode
    (chili_mc_specialize bits: 1 forward, 2 fused weight, 4 split weight),
    leaves the fastest selected and returns (name, compile seconds, times)."""
    import torch
    from chili_amd import abi, lib

    def one_step(stream):
        step(abi.philox_uniforms(seed, stream=stream), None, 0)
        best = None
        for k in range(3):
            torch.cuda.synchronize()
            t = time.perf_counter()
            step(abi.philox_uniforms(seed, stream=stream + 1 + k), None, 0)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t
            best = dt if best is None else min(best, dt)
        return best

    times, compile_s = {"generic": one_step(900)}, {}
    try:
        for mode in modes:
            compile_s[VARIANTS[mode]] = gpu.specialize(mode)
            times[VARIANTS[mode]] = one_step(900 + 10 * mode)
    except lib.ChiliError as e:
        if log_fn:
            log_fn(f"  specialisation unavailable: {e}")
    variant = min(times, key=times.get)
    gpu.specialize({v: k for k, v in VARIANTS.items()}[variant])
    return variant, compile_s.get(variant), times


def leg_multichannel(name, steps=5, flow_epochs=0):
    """A workload's plain leg, short: adaptation through chili_mc_optimize,
    then `steps` timed production iterations (values in HBM); with
    flow_epochs > 0 also its flow-assisted leg (flows trained for that many
    epochs on 2^20 method-generated points, mixture 0, moment-16 loss)."""
    import torch
    from chili_amd import abi, lib
    W = workloads()[name]
    gpu = lib.MultiChannelGPU(lib.Channels(W["proc"]()), W["bins"], W["f"](), device=0)
    seed = SEED + W["cfg"]
    ad = adapt(gpu, W, seed)
    n = W["batch"]
    vals = torch.empty(n * steps, dtype=torch.float64, device="cuda:0")

    def plain(uni, vp, cap):
        return gpu.iterate(uni, n, train=False, vals_ptr=vp, vals_cap=cap)[3]

    # the fused specialised weight kernel spills beyond a few channels: the
    # generic kernels, the specialised forward kernel and the split pass
    variant, compile_s, t_var = choose_variant(gpu, plain, seed, modes=(1, 4, 5))
    gpu.timing(enable=True, reset=True)
    dt, atts, wmean, wmax = timed_steps(plain, n, steps, seed, vals)
    tm = gpu.timing(enable=False)
    out = dict(workload=W["desc"], channels=gpu.nchannels, dims=gpu.ndims, adaptation_s=round(ad["time_adapt_s"], 3),
               time_to_1pct_s=ad["time_to_1pct_s"], points_per_s=round(n * steps / dt, 1), ms_per_step=round(dt / steps * 1e3, 3),
               eps_unw=wmean / wmax, unweighted_per_s=round(n * steps / dt * wmean / wmax, 1), acceptance=n * steps / atts,
               kernel_variant=variant, variant_step_ms={k: round(v * 1e3, 3) for k, v in t_var.items()},
               specialize_compile_s=compile_s,
               k_main_avg_ms=round(tm["main_ms"] / max(tm["launches"], 1), 4),
               k_main_points_per_launch=round(tm["points"] / max(tm["launches"], 1), 1))
    if flow_epochs > 0:
        C, d, cap = gpu.nchannels, gpu.ndims, 1 << 20
        r = torch.empty(C * cap * d, dtype=torch.float64, device="cuda:0")
        beta = torch.empty(C * cap, dtype=torch.float64, device="cuda:0")
        w2 = torch.empty(cap, dtype=torch.float64, device="cuda:0")
        fl = lib.FlowGPU(d, W["flow_layers"], C, seed=seed)
        fl.set_moment(16)
        t0 = time.perf_counter()
        _, nk, _ = gpu.flow_data(None, abi.philox_uniforms(seed, stream=500), cap, r.data_ptr(), beta.data_ptr(), w2.data_ptr(), cap)
        losses = fl.train(nk, r.data_ptr(), beta.data_ptr(), w2.data_ptr(), 1 << 16, flow_epochs, 1e-3, seed=seed)
        torch.cuda.synchronize()
        t_train = time.perf_counter() - t0
        del r, beta, w2

        def flow(uni, vp, cap_):
            return gpu.flow_iterate(fl, uni, n, vals_ptr=vp, vals_cap=cap_)[1]

        dt, atts, wmean, wmax = timed_steps(flow, n, steps, seed, vals)
        out["flow"] = dict(layers=W["flow_layers"], epochs=flow_epochs, train_points=int(nk), train_s=round(t_train, 2),
                           loss_first=losses[0], loss_last=losses[-1], points_per_s=round(n * steps / dt, 1),
                           ms_per_step=round(dt / steps * 1e3, 3), eps_unw=wmean / wmax,
                           unweighted_per_s=round(n * steps / dt * wmean / wmax, 1), acceptance=n * steps / atts,
                           note="short recipe (config: 200 epochs on 2^22 points, weighted-variance loss)")
        out["flow"]["vs_plain_unweighted"] = round(out["flow"]["unweighted_per_s"] / out["unweighted_per_s"], 4)
    return out


def leg_batched(iters=5, adapt_iters=3):
    """32 batched jobs (2->2..2->5, log-uniform accepted points): adaptation
    through the batch, then one iteration of all jobs timed three ways (loop
    of single-job iterations, one eager batched pass, the hipGraph replay
    after the launch-shape autotuner)."""
    import torch
    from chili_amd import abi, lib
    from chili_amd import processes as P
    spec = P.batched_jobs()
    chans = [lib.Channels(p) for p, _, _ in spec]
    jobs = [lib.MultiChannelGPU(ch, 64, f) for ch, (_, f, _) in zip(chans, spec)]
    ncalls = [n for _, _, n in spec]
    J = len(jobs)
    b = lib.BatchGPU(jobs)
    seeds = [P.BATCH_SEED + j for j in range(J)]
    stream = 0
    for it in range(adapt_iters):
        _, _, Ts, hs = b.iterate(seeds, [stream] * J, ncalls, train=True, graph=False)
        stream += 1
        for gpu, T, h in zip(jobs, Ts, hs):
            gpu.set_alpha(lib.alpha_adapt(gpu.alpha(), T))
            g = gpu.grids()
            for c in range(gpu.nchannels):
                if np.any(h[c] != 0):
                    g[c] = lib.grid_adapt(g[c], h[c], 1.0)
            gpu.set_grids(g)
    total = int(sum(ncalls))

    def timed(fn):
        nonlocal stream
        fn(stream)
        stream += 1
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(iters):
            fn(stream)
            stream += 1
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / iters

    def single_jobs(s):
        for gpu, sd, n in zip(jobs, seeds, ncalls):
            gpu.iterate(abi.philox_uniforms(sd, stream=s), n, train=False)

    t_single = timed(single_jobs)
    t_eager = timed(lambda s: b.iterate(seeds, [s] * J, ncalls, train=False, graph=False))
    tune = b.tune(seeds, [stream] * J, ncalls)
    stream += 1
    t_graph = timed(lambda s: b.iterate(seeds, [s] * J, ncalls, train=False, graph=True))
    sh = b.shape()
    return dict(workload="32 jobs, 2->2..2->5 all-gluon (2/5/15/52 channels, 4/7/10/13 dims, 64 bins), accepted points per "
                         "job log-uniform [1e4, 1e6] from seed 0xC0FFEE02, propagator products with pT/|y|/dR cuts",
                jobs=J, accepted_points=total, single_job_loop_ms=round(t_single * 1e3, 3),
                batched_eager_ms=round(t_eager * 1e3, 3), batched_graph_ms=round(t_graph * 1e3, 3),
                points_per_s=round(total / t_graph, 1), points_per_s_single=round(total / t_single, 1),
                speedup_vs_single=round(t_single / t_graph, 3), tuned_shape=tune, topups=sh["topups"],
                captures=sh["captures"], replays=sh["replays"])


def kernel_couple(args):
    """The coupling-layer kernel alone (the profiles' workload): log q of
    `rows` random hypercube points (density direction) and as many samples
    (inverse direction) through one flow of random parameters; prints rows/s
    and the algorithmic BF16 rate, with the kernel's own HIP-event time."""
    import torch
    from chili_amd import lib
    d, L, n = args.dims, args.layers, args.rows
    fl = lib.FlowGPU(d, L, 1, seed=5)
    rng = np.random.default_rng(1)
    p = fl.params(0)
    fl.set_params(0, (p + rng.normal(0, 0.02, p.shape)).astype(np.float32))
    flops = couple_flops_per_row(fl.masks(), d)
    r = torch.rand(n, d, dtype=torch.float64, device="cuda")
    lq = torch.empty(n, dtype=torch.float64, device="cuda")
    z = torch.empty_like(r)
    out = {}
    for direction in ("density", "sample"):
        def run():
            if direction == "density":
                fl.log_density(0, n, r.data_ptr(), lq.data_ptr(), z.data_ptr())
            else:
                z.copy_(r)
                fl.sample(0, n, z.data_ptr(), lq.data_ptr())
        run()
        torch.cuda.synchronize()
        fl.timing(enable=True, reset=True)
        t0 = time.perf_counter()
        for _ in range(args.reps):
            run()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / args.reps
        tm = fl.timing(enable=False, reset=True)
        avg_us = tm["ms"] / max(tm["launches"], 1) * 1e3
        out[direction] = dict(rows_per_s=n / dt, ms_per_pass=dt * 1e3, k_couple_avg_us=round(avg_us, 2),
                              k_couple_tflops=round(tm["flops"] / (tm["ms"] * 1e-3) / 1e12, 2),
                              k_couple_frac_of_bf16_peak=round(tm["flops"] / (tm["ms"] * 1e-3) / 1e12 / BF16_PEAK_TFLOPS, 4))
    print(json.dumps(dict(kernel="k_couple", dims=d, layers=L, rows=n, flops_per_row=flops, **out)))


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="tt_6j", choices=sorted(workloads()))
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="CPU baseline sample length")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-flow", action="store_true", help="skip the flow leg")
    ap.add_argument("--no-configs", action="store_true", help="skip the short legs of the other workloads")
    ap.add_argument("--sel-steps", type=int, default=3, help="steps of each selection sample (mixture weight, leg)")
    ap.add_argument("--flow-train", type=int, default=1 << 22, help="training points (method-generated)")
    ap.add_argument("--flow-epochs", type=int, default=30, help="training epochs (the config's schedule is 200)")
    ap.add_argument("--flow-batch", type=int, default=1 << 16)
    ap.add_argument("--flow-lr", type=float, default=1e-3)
    ap.add_argument("--flow-val-every", type=int, default=3, help="epochs between validation passes (model selection)")
    ap.add_argument("--gen-events", type=int, default=1 << 23,
                    help="unweighted events to generate against the production w_max (0: skip)")
    ap.add_argument("--flow-mix", default="0,0.3,0.5,0.7",
                    help="defensive mixture weights to try on the selection sample (comma list)")
    ap.add_argument("--flow-train-mix", type=float, default=0.0,
                    help="mixture weight of the proposal the flows are trained (and validated) for")
    ap.add_argument("--flow-rounds", type=int, default=1,
                    help="training rounds; rounds after the first train on records drawn by the best flows so far")
    ap.add_argument("--flow-moment", type=int, default=16,
                    help="training loss: this moment of the weights under the proposal (2: variance; higher moments "
                         "control the largest weights, which set the unweighting efficiency)")
    ap.add_argument("--no-specialize", action="store_true",
                    help="keep the generic kernels (default: the autotuner compiles kernels with the channel programs as "
                         "constants and keeps them when a production step is faster)")
    ap.add_argument("--specialize-mode", type=int, default=5,
                    help="kernel variant of --traffic-probe (chili_mc_specialize bits: 1 forward, 2 fused weight, 4 split "
                         "weight)")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="for rocprofv3 --pmc runs: adapt, run two production steps and a 1 GiB device copy (the "
                         "calibration kernel of known traffic), print nothing else")
    ap.add_argument("--kernel", choices=["couple"], help="time one kernel alone (for profiles) and exit")
    ap.add_argument("--dims", type=int, default=16)
    ap.add_argument("--layers", type=int, default=8)
    ap.add_argument("--rows", type=int, default=1 << 22)
    ap.add_argument("--reps", type=int, default=5)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one MI355X: this build has no multi-GPU path (each GPU would run independent "
              "point batches and all-reduce C*d*B histograms per iteration); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    from chili_amd import abi, lib
    from oracle import oracle as O

    if lib.device_count() < 1:
        raise SystemExit("no HIP device")
    torch.cuda.init()
    if args.kernel == "couple":
        kernel_couple(args)
        return

    W = workloads()[args.workload]
    proc, f = W["proc"](), W["f"]()
    ch = lib.Channels(proc)
    gpu = lib.MultiChannelGPU(ch, W["bins"], f, device=0)
    log(f"workload {args.workload}: {gpu.nchannels} channels, {gpu.ndims} dims, {gpu.nbins} bins")
    seed = SEED + W["cfg"]
    ad = adapt(gpu, W, seed)
    if args.traffic_probe:
        if not args.no_specialize:
            gpu.specialize(args.specialize_mode)
        for s in range(2):
            gpu.iterate(abi.philox_uniforms(seed, stream=S_PROD + s), W["batch"], train=False)
        a = torch.empty(1 << 27, dtype=torch.float64, device="cuda:0").fill_(1.0)
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        log(f"traffic probe done: {gpu.nchannels} channels, copy of {a.numel() * 8} bytes")
        return

    # ---- production with frozen state ----
    n = W["batch"]
    vals = torch.empty(n * max(args.steps, args.sel_steps), dtype=torch.float64, device="cuda:0")

    def production(step_fn, tag, stream0=S_PROD, steps=None, warmup=None):
        steps = args.steps if steps is None else steps
        warmup = args.warmup if warmup is None else warmup
        for s in range(warmup):
            step_fn(abi.philox_uniforms(seed, stream=1000 + s), None, 0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        atts = 0
        step_ms = []
        for s in range(steps):
            ts = time.perf_counter()
            atts += step_fn(abi.philox_uniforms(seed, stream=stream0 + s), ctypes.c_void_p(vals.data_ptr() + 8 * n * s), n)
            step_ms.append(round((time.perf_counter() - ts) * 1e3, 2))   # each step ends synchronised
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        w = vals[: n * steps]
        wmax, wmean = float(w.max()), float(w.mean())
        q999 = float(torch.sort(w).values[int(0.999 * (w.numel() - 1))])
        per_step = w.view(steps, n)
        eps_steps = (per_step.mean(dim=1) / per_step.max(dim=1).values).cpu().numpy()
        pts_s = n * steps / dt
        res = dict(points_per_s=round(pts_s, 1), eps_unw=wmean / wmax, eps_unw_q999=wmean / q999,
                   unweighted_per_s=round(pts_s * wmean / wmax, 1), unweighted_per_s_q999=round(pts_s * wmean / q999, 1),
                   ms_per_step=round(dt / steps * 1e3, 3), w_max=wmax, w_mean=wmean,
                   w_rel_std=float(w.std()) / wmean, acceptance=n * steps / atts,
                   integral=wmean * n * steps / atts, steps=steps, first_stream=stream0,
                   eps_unw_per_step=[round(float(e), 5) for e in eps_steps], step_ms=step_ms)
        log(f"  [{tag}] {pts_s:.3e} points/s, eps {res['eps_unw']:.4f} (q999 {res['eps_unw_q999']:.4f}, per step "
            f"{eps_steps.min():.4f}..{eps_steps.max():.4f}), {res['unweighted_per_s']:.3e} unweighted/s, rel std "
            f"{res['w_rel_std']:.3f}, {res['ms_per_step']:.2f} ms/step")
        return res

    def plain_step(uni, vp, cap):
        st, _, _, na = gpu.iterate(uni, n, train=False, vals_ptr=vp, vals_cap=cap)
        return na

    # kernel variant: generic (programs read from HBM) or specialised at run
    # time (programs as constants); the faster production step wins
    variant, compile_s, t_variant = "generic", None, {}
    if not args.no_specialize:
        variant, compile_s, t_variant = choose_variant(gpu, plain_step, seed, log_fn=log)
        log(f"  kernel variant: {variant} (step s: {t_variant}, compile {compile_s} s)")

    plain_sel = production(plain_step, "plain, selection sample", stream0=S_SEL, steps=args.sel_steps, warmup=0)
    gpu.timing(enable=True, reset=True)
    plain = production(plain_step, "plain")
    tm = gpu.timing(enable=False)
    pts_s, eps, eps999 = plain["points_per_s"], plain["eps_unw"], plain["eps_unw_q999"]
    wmax, wmean = plain["w_max"], plain["w_mean"]
    unw_s = plain["unweighted_per_s"]
    atts = n * args.steps / plain["acceptance"]

    # ---- unweighted events actually generated (device unweighting stage
    # against the production w_max): the events/s the derived figure above
    # stands for ----
    if args.gen_events > 0:
        ne = args.gen_events
        ev = torch.empty(4 * gpu.ntot * ne, dtype=torch.float64, device="cuda:0")
        ew = torch.empty(ne, dtype=torch.float64, device="cuda:0")
        gpu.generate(abi.philox_uniforms(seed, stream=3000), 1 << 16, wmax, ctypes.c_void_p(ev.data_ptr()),
                     ctypes.c_void_p(ew.data_ptr()), ne)                      # warm-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        gst, gatt = gpu.generate(abi.philox_uniforms(seed, stream=3001), ne, wmax, ctypes.c_void_p(ev.data_ptr()),
                                 ctypes.c_void_p(ew.data_ptr()), ne)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        plain["generated"] = dict(events=ne, attempts=int(gatt), seconds=round(dt, 4), events_per_s=round(ne / dt, 1),
                                  wmax=wmax, overweight=int((ew.abs() > 1).sum()),
                                  note="chili_mc_generate: the MultiChannel loop with an unweighting PostProcess, "
                                       "events (momenta, weights) written to HBM")
        log(f"  [generate] {ne} unweighted events in {dt:.3f} s: {ne / dt:.3e} events/s ({gatt} attempts)")
        # the same events with exact early rejection (chili_mc_specialize bit
        # 3): only points whose picked channel's bound can pass the accept
        # test are weighted in full
        if not args.no_specialize:
            mode = {v: k for k, v in VARIANTS.items()}.get(variant, 0)
            gpu.specialize(mode | 8)
            gpu.generate(abi.philox_uniforms(seed, stream=3000), 1 << 16, wmax, ctypes.c_void_p(ev.data_ptr()),
                         ctypes.c_void_p(ew.data_ptr()), ne)                  # warm-up
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            gst2, gatt2 = gpu.generate(abi.philox_uniforms(seed, stream=3001), ne, wmax, ctypes.c_void_p(ev.data_ptr()),
                                       ctypes.c_void_p(ew.data_ptr()), ne)
            torch.cuda.synchronize()
            dt2 = time.perf_counter() - t0
            gpu.specialize(mode)
            plain["generated"]["early_rejection"] = dict(
                seconds=round(dt2, 4), events_per_s=round(ne / dt2, 1), attempts=int(gatt2),
                same_attempts=bool(gatt2 == gatt), sum_rel_diff=abs(gst2.sum - gst.sum) / abs(gst.sum),
                note="chili_mc_specialize bit 3: a passing point is weighted in full only when the bound "
                     "f / (alpha_c / ps_c / vegas_c) of its picked channel c can pass the accept test (the "
                     "multichannel weight never exceeds it): the same events")
            log(f"  [generate, early rejection] {ne / dt2:.3e} events/s ({gatt2} attempts)")
        del ev, ew

    # ---- flow leg ----
    flow, flow_sel, roof_flow = None, None, None

    def flow_leg():
        C, d = gpu.nchannels, gpu.ndims
        cap = args.flow_train
        r = torch.empty(C * cap * d, dtype=torch.float64, device="cuda:0")
        beta = torch.empty(C * cap, dtype=torch.float64, device="cuda:0")
        w2 = torch.empty(cap, dtype=torch.float64, device="cuda:0")
        fl = lib.FlowGPU(d, W["flow_layers"], C, seed=seed)
        fl.set_mixture(args.flow_train_mix)
        fl.set_moment(args.flow_moment)
        losses = []
        done = 0
        # epoch selection: after every --flow-val-every epochs the flows'
        # unweighting efficiency <w>/w_max is measured on a validation sample
        # (its own Philox streams); the best parameters are kept
        nv = min(n, 1 << 22)
        best = None
        val_hist = []
        t_data = 0.0
        t1 = time.perf_counter()
        total = args.flow_epochs * args.flow_rounds
        for rnd in range(args.flow_rounds):
            # training records: drawn by plain VEGAS first, then (later
            # rounds) by the best flows so far
            t0 = time.perf_counter()
            if rnd > 0:
                for j in range(C):
                    fl.set_params(j, best[2][j])
            _, nk, _ = gpu.flow_data(fl if rnd > 0 else None, abi.philox_uniforms(seed, stream=500 + rnd), cap, r.data_ptr(),
                                     beta.data_ptr(), w2.data_ptr(), cap)
            torch.cuda.synchronize()
            t_data += time.perf_counter() - t0
            end = done + args.flow_epochs
            while done < end:
                e = min(args.flow_val_every, end - done)
                losses += list(fl.train(nk, r.data_ptr(), beta.data_ptr(), w2.data_ptr(), args.flow_batch, e, args.flow_lr,
                                        seed=seed + done))
                done += e
                gpu.flow_iterate(fl, abi.philox_uniforms(seed, stream=700 + done), nv, vals_ptr=ctypes.c_void_p(vals.data_ptr()),
                                 vals_cap=nv)
                v = vals[:nv]
                eps_v = float(v.mean() / v.max())
                val_hist.append(dict(epoch=done, loss=round(losses[-1], 3), eps_unw=round(eps_v, 5),
                                     w_rel_std=round(float(v.std() / v.mean()), 4)))
                if best is None or eps_v > best[0]:
                    best = (eps_v, done, [fl.params(j) for j in range(C)])
                log(f"  flow training: epoch {done}/{total}, loss {losses[-1]:.5e}, validation eps {eps_v:.4f} "
                    f"({time.perf_counter() - t1:.1f} s)")
        for j in range(C):
            fl.set_params(j, best[2][j])
        log(f"  flow parameters of epoch {best[1]} kept (validation eps {best[0]:.4f})")
        t_train = time.perf_counter() - t1 - t_data
        del r, beta, w2

        def flow_step(uni, vp, cap_):
            st, na = gpu.flow_iterate(fl, uni, n, vals_ptr=vp, vals_cap=cap_)
            return na

        # proposal (1 - lam) q + lam: the uniform component bounds every
        # weight by 1/lam times the plain one; lam is chosen on the selection
        # sample, then the production sample is timed with it
        scan = {}
        for lam in [float(v) for v in args.flow_mix.split(",")]:
            fl.set_mixture(lam)
            scan[lam] = production(flow_step, f"flow mix {lam:g}, selection sample", stream0=S_SEL, steps=args.sel_steps,
                                   warmup=1)
        lam_best = max(scan, key=lambda k: scan[k]["unweighted_per_s"])
        fl.set_mixture(lam_best)
        fl.timing(enable=True, reset=True)
        prod = production(flow_step, f"flow mix {lam_best:g}")
        tf = fl.timing(enable=False, reset=True)
        # the step's algorithmic BF16 MFMA work (k_couple's conditioners)
        # against the dense peak, and the coupling kernel's own rate
        per_row = couple_flops_per_row(fl.masks(), d)
        step_tflops = tf["flops"] / (args.steps + args.warmup) / (prod["ms_per_step"] * 1e-3) / 1e12
        prod["mfma"] = dict(mlp_flops_per_row=per_row, rows_per_step=round(tf["rows"] / (args.steps + args.warmup)),
                            achieved_tflops=round(step_tflops, 2), peak_tflops=BF16_PEAK_TFLOPS,
                            frac=round(step_tflops / BF16_PEAK_TFLOPS, 5),
                            note="whole flow step (spline transforms, VEGAS and phase-space maps included) against the "
                                 "dense BF16 MFMA peak; rows counted on the device")
        kc_ms = tf["ms"] / max(tf["launches"], 1)
        kc_tf = tf["flops"] / (tf["ms"] * 1e-3) / 1e12
        roof = dict(bound="mfma", achieved=round(kc_tf, 2), peak=BF16_PEAK_TFLOPS, unit="TFLOP/s",
                    frac=round(kc_tf / BF16_PEAK_TFLOPS, 5), traffic=None,
                    kernel="k_couple (fused coupling layer: BF16 MFMA conditioner MLP + FP32/FP64 spline)",
                    launches=tf["launches"], avg_launch_ms=round(kc_ms, 4),
                    rows_per_launch=round(tf["rows"] * W["flow_layers"] / max(tf["launches"], 1), 1),
                    flops_per_row_layer=round(per_row / W["flow_layers"], 1),
                    note="algorithmic conditioner flops of every timed launch (rows counted on the device) over the "
                         "launches' HIP-event time, production steps incl. warm-up; traffic not measured (MFMA/VALU bound)")
        prod.update(dict(mixture=lam_best,
                         mixture_selection={f"{k:g}": dict(unweighted_per_s=v["unweighted_per_s"], eps_unw=v["eps_unw"],
                                                           w_rel_std=round(v["w_rel_std"], 4), acceptance=round(v["acceptance"], 4))
                                            for k, v in scan.items()},
                         selected_epoch=best[1], validation=val_hist,
                         train_points=int(nk), train_s=round(t_train, 2), data_s=round(t_data, 3),
                         loss_first=losses[0] if losses else None, loss_last=losses[-1] if losses else None,
                         vs_plain_unweighted=round(prod["unweighted_per_s"] / unw_s, 4),
                         vs_plain_unweighted_q999=round(prod["unweighted_per_s_q999"] / plain["unweighted_per_s_q999"], 4),
                         vs_plain_acceptance=round(prod["acceptance"] / plain["acceptance"], 4)))
        # unweighted events actually generated on the learned proposal
        # against its production w_max, without and with the exact early
        # rejection (the latter needs the specialised kernels and no mixture)
        if args.gen_events > 0:
            ne = args.gen_events
            ev = torch.empty(4 * gpu.ntot * ne, dtype=torch.float64, device="cuda:0")
            ew = torch.empty(ne, dtype=torch.float64, device="cuda:0")
            mode = {v: k for k, v in VARIANTS.items()}.get(variant, 0)
            gen = {}
            for tag, md in (("generated", mode), ("early_rejection", mode | 8)):
                if md & 8 and (args.no_specialize or lam_best > 0):
                    continue
                gpu.specialize(md)
                gpu.flow_generate(fl, abi.philox_uniforms(seed, stream=3000), 1 << 16, prod["w_max"], ctypes.c_void_p(ev.data_ptr()),
                                  ctypes.c_void_p(ew.data_ptr()), ne)                    # warm-up
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                gst, gatt = gpu.flow_generate(fl, abi.philox_uniforms(seed, stream=3001), ne, prod["w_max"],
                                              ctypes.c_void_p(ev.data_ptr()), ctypes.c_void_p(ew.data_ptr()), ne)
                torch.cuda.synchronize()
                dt = time.perf_counter() - t0
                gen[tag] = dict(events=ne, attempts=int(gatt), seconds=round(dt, 4), events_per_s=round(ne / dt, 1),
                                sum=gst.sum)
                log(f"  [flow generate{', early rejection' if md & 8 else ''}] {ne / dt:.3e} events/s ({gatt} attempts)")
            gpu.specialize(mode)
            if "generated" in gen:
                g0 = gen["generated"]
                prod["generated"] = dict(events=ne, attempts=g0["attempts"], seconds=g0["seconds"],
                                         events_per_s=g0["events_per_s"], wmax=prod["w_max"],
                                         overweight=int((ew.abs() > 1).sum()),
                                         note="chili_mc_flow_generate: the unweighting stage on the learned proposal")
                if "early_rejection" in gen:
                    g1 = gen["early_rejection"]
                    prod["generated"]["early_rejection"] = dict(
                        seconds=g1["seconds"], events_per_s=g1["events_per_s"], attempts=g1["attempts"],
                        same_attempts=bool(g1["attempts"] == g0["attempts"]), sum_rel_diff=abs(g1["sum"] - g0["sum"]) / abs(g0["sum"]))
            del ev, ew

        def gen_fn(stream, ne, ev_ptr, ew_ptr):
            return gpu.flow_generate(fl, abi.philox_uniforms(seed, stream=stream), ne, prod["w_max"], ev_ptr, ew_ptr, ne)

        gen_fn.flow = fl
        gen_fn.per_row = per_row / W["flow_layers"]
        return prod, scan[lam_best], roof, gen_fn

    flow_gen = None
    if not args.no_flow:
        # a failing flow leg (e.g. training that diverged) is reported in
        # the bench line; the plain leg still stands
        try:
            flow, flow_sel, roof_flow, flow_gen = flow_leg()
        except (lib.ChiliError, RuntimeError, ValueError, ZeroDivisionError) as e:
            log(f"  flow leg failed: {e}")
            flow = dict(error=str(e))

    # ---- the headline: unweighted events actually generated ----
    # each leg's unweighting stage against its production w_max (with the
    # exact early rejection when the specialised kernels are in use: the
    # same events as weighting every point); the leg with more events/s on a
    # selection pass (its own streams) is timed over --warmup + --steps
    # passes of `ne_step` events each (production streams)
    flow_ok = bool(flow and "unweighted_per_s" in flow)
    ne_step = max(1 << 16, args.gen_events // 2)
    mode = {v: k for k, v in VARIANTS.items()}.get(variant, 0)
    if not args.no_specialize:
        gpu.specialize(mode | 8)
    legs = {"plain": lambda st, ev_p, ew_p: gpu.generate(abi.philox_uniforms(seed, stream=st), ne_step, wmax, ev_p, ew_p, ne_step)}
    if flow_ok and flow_gen is not None:
        legs["flow"] = lambda st, ev_p, ew_p: flow_gen(st, ne_step, ev_p, ew_p)
    ev = torch.empty(4 * gpu.ntot * ne_step, dtype=torch.float64, device="cuda:0")
    ew = torch.empty(ne_step, dtype=torch.float64, device="cuda:0")
    evp, ewp = ctypes.c_void_p(ev.data_ptr()), ctypes.c_void_p(ew.data_ptr())
    gen_sel = {}
    for leg, fn in legs.items():
        fn(3050, evp, ewp)                                  # warm-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fn(S_SEL + 100, evp, ewp)
        torch.cuda.synchronize()
        gen_sel[leg] = round(ne_step / (time.perf_counter() - t0), 1)
    gen_leg = max(gen_sel, key=gen_sel.get)
    fn = legs[gen_leg]
    for s_ in range(args.warmup):
        fn(S_PROD + 1100 + s_, evp, ewp)
    # the step's dominant kernel timed live with HIP events on the stream it
    # runs on: the forward map (plain leg) or the coupling kernel (flow leg)
    timer = gpu if gen_leg == "plain" else flow_gen.flow
    timer.timing(enable=True, reset=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    gen_atts, gen_ms = 0, []
    for s_ in range(args.steps):
        ts = time.perf_counter()
        _, a_ = fn(S_PROD + 1200 + s_, evp, ewp)
        gen_atts += a_
        gen_ms.append(round((time.perf_counter() - ts) * 1e3, 2))
    torch.cuda.synchronize()
    gen_dt = time.perf_counter() - t0
    tm_gen = timer.timing(enable=False, reset=True)
    gen_value = ne_step * args.steps / gen_dt
    log(f"  [headline] {gen_leg}: {gen_value:.3e} unweighted events/s generated ({args.steps} passes of {ne_step} events; "
        f"selection {gen_sel})")
    del ev, ew

    # ---- rooflines: the headline step's dominant kernel, and k_main ----
    ora = O.Oracle(proc, gpu.nbins, f, count=True)
    ora.set_grids(gpu.grids())
    ora.set_alpha(gpu.alpha())
    O.lib(True).oracle_count_reset()
    cnt = ora.count_points(abi.philox_uniforms(seed, stream=99), 0, 4000)
    f_main = (cnt["wgt_arith"] + 20 * cnt["wgt_trans"]) / max(cnt["accepted"], 1)
    avg_ms = tm["main_ms"] / max(tm["launches"], 1)
    pts_launch = tm["points"] / max(tm["launches"], 1)
    achieved = f_main * pts_launch / (avg_ms * 1e-3) / 1e12
    traffic, bpp = measured_traffic(args.workload, variant, pts_launch)
    roof = dict(bound="fp64-valu", achieved=round(achieved, 4), peak=FP64_PEAK_TFLOPS, unit="TFLOP/s",
                frac=round(achieved / FP64_PEAK_TFLOPS, 5), traffic=traffic,
                traffic_source=("projected: bytes per point from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE run "
                                f"(profiles/k_main_traffic_{args.workload}_*.md, {bpp} B/point) x this run's points per launch"
                                if traffic is not None else None),
                kernel="k_main (all-channel inverse maps + VEGAS densities + multichannel weight)",
                compute="FP64 vector ALU (no matrix work); MI355X's FP64 vector peak 78.6 TFLOP/s",
                scope="the multichannel weight kernel of the plain leg; the flow leg's coupling kernel is roofline_flow",
                flop_eq_per_point=round(f_main, 1),
                flop_eq_count="N_arith + 20 N_trans, op-counting oracle on 4000 attempts of the frozen state",
                flop_eq_fixture=fixture_f_main(args.workload),
                points_per_launch=round(pts_launch, 1), avg_launch_ms=round(avg_ms, 4), launches=tm["launches"],
                forward_kernel_ms_per_step=round(tm["forward_ms"] / (args.steps + args.warmup), 4))
    if gen_leg == "plain":
        # the specialised forward kernel: Philox uniforms, VEGAS map, channel
        # mapping, cut, packed momenta of the passing attempts
        f_fwd = (cnt["fwd_arith"] + 20 * cnt["fwd_trans"]) / max(cnt["attempts"], 1)
        nl = max(tm_gen["forward_launches"], 1)
        fwd_ms = tm_gen["forward_ms"] / nl
        att_launch = tm_gen["forward_attempts"] / nl
        ach = f_fwd * att_launch / (fwd_ms * 1e-3) / 1e12
        ftraffic, fbpa = measured_traffic(args.workload, "forward", att_launch)
        roof_head = dict(bound="fp64-valu", achieved=round(ach, 4), peak=FP64_PEAK_TFLOPS, unit="TFLOP/s",
                         frac=round(ach / FP64_PEAK_TFLOPS, 5), traffic=ftraffic,
                         traffic_source=("projected: bytes per attempt from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE "
                                         f"run (profiles/k_forward_traffic_{args.workload}.md, {fbpa} B/attempt) x this "
                                         "run's attempts per launch" if ftraffic is not None else None),
                         kernel="chili_k_forward (run-time specialised: Philox uniforms, VEGAS forward map, channel "
                                "mapping, cut, packed momenta of the passing attempts)",
                         compute="FP64 vector ALU (no matrix work); MI355X's FP64 vector peak 78.6 TFLOP/s",
                         scope="the dominant kernel of the headline step (the plain leg's unweighting stage)",
                         flop_eq_per_attempt=round(f_fwd, 1),
                         flop_eq_count="N_arith + 20 N_trans of the forward map, op-counting oracle on 4000 attempts of "
                                       "the frozen state (integer Philox work not counted)",
                         attempts_per_launch=round(att_launch, 1), avg_launch_ms=round(fwd_ms, 4), launches=nl,
                         share_of_step=round(tm_gen["forward_ms"] / (gen_dt * 1e3), 4))
    else:
        kc_ms = tm_gen["ms"] / max(tm_gen["launches"], 1)
        kc_tf = tm_gen["flops"] / (tm_gen["ms"] * 1e-3) / 1e12
        roof_head = dict(bound="mfma", achieved=round(kc_tf, 2), peak=BF16_PEAK_TFLOPS, unit="TFLOP/s",
                         frac=round(kc_tf / BF16_PEAK_TFLOPS, 5), traffic=None,
                         kernel="k_couple (fused coupling layer: BF16 MFMA conditioner MLP + FP32/FP64 spline)",
                         scope="the dominant kernel of the headline step (the flow leg's unweighting stage)",
                         launches=tm_gen["launches"], avg_launch_ms=round(kc_ms, 4),
                         rows_per_launch=round(tm_gen["rows"] / max(tm_gen["launches"], 1), 1),
                         flops_per_row_layer=round(flow_gen.per_row, 1),
                         share_of_step=round(tm_gen["ms"] / (gen_dt * 1e3), 4),
                         note="algorithmic conditioner flops of the timed launches over their HIP-event time; traffic "
                              "not measured (MFMA/VALU bound)")

    # ---- short legs of the other workloads ----
    configs = {}
    if not args.no_configs:
        for key, fn in (("ee_mumu", leg_vegas_eemumu), ("gg_4g", lambda: leg_multichannel("gg_4g", flow_epochs=6)),
                        ("batched_32_jobs", leg_batched)):
            if key == "gg_4g" and args.workload == "gg_4g":
                continue
            try:
                configs[key] = fn()
                log(f"  [{key}] {json.dumps(configs[key])[:300]}")
            except (lib.ChiliError, RuntimeError, ValueError) as e:
                configs[key] = dict(error=str(e))
                log(f"  [{key}] failed: {e}")

    # ---- CPU baseline: the scalar oracle on the host cores, bounded sample ----
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(proc, f, gpu.grids(), gpu.alpha(), seed, args.cpu_seconds, eps, args.workload)

    # the production samples' derived rates (accepted points/s x <w>/w_max)
    # name the leg the integration would pick; the headline is the leg that
    # generates more events on the selection pass (plain VEGAS the fallback)
    selected = "flow" if (flow_ok and flow_sel["unweighted_per_s"] > plain_sel["unweighted_per_s"]) else "plain"
    value = gen_value
    ms_step = gen_dt / args.steps * 1e3
    recipe = None
    if flow_ok:
        recipe = dict(layers=W["flow_layers"], hidden="3x64 ReLU", spline="rational-quadratic, 16 bins",
                      train_points=flow["train_points"], batch=args.flow_batch, lr=args.flow_lr, epochs=args.flow_epochs,
                      rounds=args.flow_rounds, loss=f"(1/p) log mean w^p, p = {args.flow_moment}",
                      train_mixture=args.flow_train_mix,
                      deviations_from_config=[f"{args.flow_epochs} epochs with epoch selection on a validation sample "
                                              "(config: 200)",
                                              f"moment-{args.flow_moment} loss (config: weighted variance, p = 2)"] +
                                             (["defensive mixture lambda chosen on a selection sample"]
                                              if len(args.flow_mix.split(",")) > 1 else []))
    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "events/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_step, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (Philox4x32-10 uniforms)",
        "config": {"workload": args.workload, "description": W["desc"], "channels": gpu.nchannels, "dims": gpu.ndims,
                   "bins": gpu.nbins, "points_per_step": n, "events_per_step": ne_step, "selected_leg": gen_leg,
                   "step": f"one pass of the selected leg's unweighting stage: {ne_step} unweighted events against its "
                           "production w_max (forward map, cut, weights of the candidates, accept test, events and "
                           "weights written to HBM)",
                   "selection": dict(rule="value: the leg with more generated events/s on a selection pass (Philox "
                                          f"streams {S_SEL + 100}), timed over the production passes (streams "
                                          f"{S_PROD + 1200}..); derived rates (accepted points/s x <w>/w_max) of the "
                                          f"production samples under plain / flow, the mixture weight chosen on a "
                                          f"selection sample ({args.sel_steps} steps, streams {S_SEL}..)",
                                     generated_selection_events_per_s=gen_sel,
                                     derived_selected_leg=selected,
                                     plain_selection_unweighted_per_s=plain_sel["unweighted_per_s"],
                                     flow_selection_unweighted_per_s=flow_sel["unweighted_per_s"] if flow_sel else None),
                   "early_rejection": not args.no_specialize,
                   "kernel_variant": variant, "specialize_compile_s": compile_s, "flow_recipe": recipe},
        # the headline leg's generation passes, then its production
        # figures (both legs in full below)
        "generated": dict(leg=gen_leg, events=ne_step * args.steps, attempts=int(gen_atts), seconds=round(gen_dt, 4),
                          pass_ms=gen_ms),
        **(dict(points_per_s=flow["points_per_s"], eps_unw=flow["eps_unw"], eps_unw_q999=flow["eps_unw_q999"],
                w_max=flow["w_max"], w_mean=flow["w_mean"], attempts_per_step=n / flow["acceptance"],
                derived_unweighted_per_s=flow["unweighted_per_s"])
           if gen_leg == "flow" else
           dict(points_per_s=round(pts_s, 1), eps_unw=eps, eps_unw_q999=eps999, w_max=wmax, w_mean=wmean,
                attempts_per_step=atts / args.steps, derived_unweighted_per_s=unw_s)),
        "adaptation": ad, "roofline": roof_head, "roofline_weight": roof, "roofline_flow": roof_flow, "cpu_baseline": cpu,
        "plain": plain, "flow": flow,
        "configs": configs,
        # the CPU baseline runs the plain algorithm: its points/s against the plain leg's
        "gpu_vs_cpu": round(pts_s / cpu["points_per_s"], 1) if cpu else None,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
