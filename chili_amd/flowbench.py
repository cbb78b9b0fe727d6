"""Throughput of the fused coupling-layer kernels on one GPU.

    python -m chili_amd.flowbench [--dims 10] [--layers 6] [--rows 2097152] [--reps 5]

Evaluates log q of `rows` random hypercube points (density direction) and
samples as many (inverse direction) with one flow of random parameters, and
prints rows/s, row-layers/s and the algorithmic BF16 MFMA rate of the
conditioner networks (2 x (nA x 64 + 2 x 64 x 64 + 64 x nB (3K-1)) flops per
row and layer).
"""
import argparse
import json
import time

import numpy as np


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--dims", type=int, default=10)
    ap.add_argument("--layers", type=int, default=6)
    ap.add_argument("--rows", type=int, default=1 << 21)
    ap.add_argument("--reps", type=int, default=5)
    args = ap.parse_args()

    import torch
    from chili_amd import lib

    d, L, n = args.dims, args.layers, args.rows
    fl = lib.FlowGPU(d, L, 1, seed=5)
    rng = np.random.default_rng(1)
    p = fl.params(0)
    fl.set_params(0, (p + rng.normal(0, 0.02, p.shape)).astype(np.float32))
    masks = fl.masks()
    flops = 0
    for B in masks:
        nB = len(B)
        nA = d - nB
        flops += 2 * (nA * 64 + 2 * 64 * 64 + 64 * nB * 47)
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
        t0 = time.perf_counter()
        for _ in range(args.reps):
            run()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / args.reps
        out[direction] = dict(rows_per_s=n / dt, row_layers_per_s=n * L / dt, mfma_tflops=flops * n / dt / 1e12,
                              mfma_frac_of_2500=flops * n / dt / 2.5e15, ms=dt * 1e3)
    print(json.dumps(dict(dims=d, layers=L, rows=n, flops_per_row=flops, **out)))


if __name__ == "__main__":
    main()
