"""Batched integrations on one GPU: 32 independent multichannel jobs,
multiplicities 2->2 .. 2->5 (4, 7, 10, 13 VEGAS dimensions), accepted points
per job log-uniform in [1e4, 1e6] from the seed, propagator-product
integrand with pT / |y| / dR cuts, FP64.

    python -m chili_amd.batchbench [--iters 5] [--adapt 3]

Each job is adapted (alpha and grids) for `--adapt` iterations through the
batch, then `--iters` production iterations are timed three ways: one
single-job iteration after another (chili_mc_iterate), the batched pass
launched eagerly, and the batched pass replayed from its hipGraph after the
autotuner chose the launch shape.  Prints one JSON line.
"""
import argparse
import json
import time

import numpy as np

SEED = 0xC0FFEE02


def build_jobs(seed, njobs=32):
    from chili_amd import lib
    from chili_amd import processes as P
    rng = np.random.default_rng(seed)
    chans, jobs, ncalls = [], [], []
    for j in range(njobs):
        n_out = 2 + j % 4
        proc = P.gluon_process(n_out, smax=n_out - 1)
        ch = lib.Channels(proc)
        jobs.append(lib.MultiChannelGPU(ch, 64, P.integrand_propprod()))
        chans.append(ch)
        ncalls.append(int(10 ** (4 + 2 * rng.random())))
    return chans, jobs, ncalls


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--iters", type=int, default=5)
    ap.add_argument("--adapt", type=int, default=3)
    args = ap.parse_args()

    import torch
    from chili_amd import abi, lib

    torch.cuda.init()
    chans, jobs, ncalls = build_jobs(SEED)
    J = len(jobs)
    b = lib.BatchGPU(jobs)
    seeds = [SEED + j for j in range(J)]
    stream = 0
    for it in range(args.adapt):
        stats, att, Ts, hs = b.iterate(seeds, [stream] * J, ncalls, train=True, graph=False)
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
        for _ in range(args.iters):
            fn(stream)
            stream += 1
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / args.iters

    def single_jobs(s):
        for gpu, sd, n in zip(jobs, seeds, ncalls):
            gpu.iterate(abi.philox_uniforms(sd, stream=s), n, train=False)

    t_single = timed(single_jobs)
    t_eager = timed(lambda s: b.iterate(seeds, [s] * J, ncalls, train=False, graph=False))
    tune = b.tune(seeds, [stream] * J, ncalls)
    stream += 1
    t_graph = timed(lambda s: b.iterate(seeds, [s] * J, ncalls, train=False, graph=True))
    sh = b.shape()
    out = dict(workload="32 jobs, 2->2..2->5, accepted points log-uniform [1e4, 1e6]", jobs=J, accepted_points=total,
               dims=sorted({gpu.ndims for gpu in jobs}), channels=sorted({gpu.nchannels for gpu in jobs}),
               single_job_loop_ms=round(t_single * 1e3, 3), batched_eager_ms=round(t_eager * 1e3, 3),
               batched_graph_ms=round(t_graph * 1e3, 3), points_per_s_graph=round(total / t_graph, 1),
               points_per_s_single=round(total / t_single, 1), speedup_vs_single=round(t_single / t_graph, 3),
               tuned_shape=tune, topups=sh["topups"], captures=sh["captures"], replays=sh["replays"])
    print(json.dumps(out))


if __name__ == "__main__":
    main()
