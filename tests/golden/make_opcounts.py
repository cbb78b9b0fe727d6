#!/usr/bin/env python3
"""Writes tests/golden/opcounts.json: the algorithmic work per point of the
benchmark workloads, counted by the op-counting build of the oracle
(oracle/_build/liboracle_count.so: every FP64 +, -, x, /, sqrt is one
N_arith, every log / exp / pow / sin / cos / tan / atan / atan2 / sinh call
one N_trans) on a fixed sample: uniform grids, alpha = 1/C, Philox seed
0xC0FFEE00 + config, stream 0, attempts 0..3999.  F = N_arith + 20 N_trans
flop-equivalents per attempt (forward map) and per accepted point (all
channels' inverse maps, VEGAS densities and the multichannel weight).

    python tests/golden/make_opcounts.py      (run from the repository root)
"""
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from chili_amd import abi  # noqa: E402
from chili_amd import processes as P  # noqa: E402
from oracle import oracle as O  # noqa: E402

WORKLOADS = {
    "tt_6j": (4, lambda: P.tt_process(smax=5), lambda: P.integrand_resonant()),
    "gg_4g": (1, lambda: P.gluon_process(4, smax=3), lambda: P.integrand_propprod()),
    "gg_2g": (2, lambda: P.gluon_process(2, smax=1), lambda: P.integrand_propprod()),
    "gg_3g": (2, lambda: P.gluon_process(3, smax=2), lambda: P.integrand_propprod()),
    "gg_5g": (2, lambda: P.gluon_process(5, smax=4), lambda: P.integrand_propprod()),
}
N = 4000


def count(name):
    cfg, proc, f = WORKLOADS[name]
    ora = O.Oracle(proc(), 64, f(), count=True)
    O.lib(True).oracle_count_reset()
    c = ora.count_points(abi.philox_uniforms(0xC0FFEE00 + cfg, stream=0), 0, N)
    acc = int(c["accepted"])
    return dict(channels=ora.nchannels, dims=ora.ndims, attempts=N, accepted=acc,
                fwd_arith=int(c["fwd_arith"]), fwd_trans=int(c["fwd_trans"]),
                wgt_arith=int(c["wgt_arith"]), wgt_trans=int(c["wgt_trans"]),
                F_fwd_per_attempt=round((c["fwd_arith"] + 20 * c["fwd_trans"]) / N, 3),
                F_main_per_point=round((c["wgt_arith"] + 20 * c["wgt_trans"]) / max(acc, 1), 3))


def main():
    out = {"convention": "F = N_arith + 20 N_trans (FP64 +,-,x,/,sqrt; log/exp/pow/sin/cos/tan/atan/atan2/sinh)",
           "sample": f"uniform grids (64 bins), alpha = 1/C, Philox seed 0xC0FFEE00 + config, stream 0, attempts 0..{N - 1}",
           "workloads": {k: count(k) for k in WORKLOADS}}
    path = os.path.join(ROOT, "tests", "golden", "opcounts.json")
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1)
        fh.write("\n")
    print(json.dumps(out["workloads"], indent=1))


if __name__ == "__main__":
    main()
