"""Batched jobs: one device pass over many independent multichannel
integrations (persistent kernels, device-side cutoff, hipGraph replay,
autotuned launch shape) against each job's single-context iteration on the
same Philox streams.  Integer results (attempts, counts) must match exactly,
sums to 1e-12 (their order differs)."""
import numpy as np
import pytest

from chili_amd import abi, lib
from chili_amd import processes as P
from tests.helpers import assert_rel

pytestmark = pytest.mark.gpu

SEED = 0xC0FFEE02


def random_state(gpu, rng):
    B = gpu.nbins
    g = np.zeros((gpu.nchannels, gpu.ndims, B + 1))
    for c in range(gpu.nchannels):
        for k in range(gpu.ndims):
            g[c, k] = np.concatenate([[0.0], np.sort(rng.uniform(0, 1, B - 1)), [1.0]])
    a = rng.uniform(0.2, 1.0, gpu.nchannels)
    gpu.set_grids(g)
    gpu.set_alpha(a / a.sum())


def make_jobs():
    specs = [(P.gluon_process(2), P.integrand_propprod()),
             (P.gluon_process(3), P.integrand_propprod()),
             (P.gluon_process(3), P.integrand_propprod()),
             (P.gluon_process(4, smax=3), P.integrand_propprod()),
             (P.tt_process(smax=5), P.integrand_resonant())]
    rng = np.random.default_rng(3)
    chans, jobs = [], []
    for proc, f in specs:
        ch = lib.Channels(proc)
        gpu = lib.MultiChannelGPU(ch, 32, f)
        random_state(gpu, rng)
        chans.append(ch)
        jobs.append(gpu)
    return chans, jobs


NCALLS = [20000, 12000, 30000, 6000, 25000]


def single(jobs, seeds, streams, train):
    out = []
    for gpu, s, st, n in zip(jobs, seeds, streams, NCALLS):
        out.append(gpu.iterate(abi.philox_uniforms(int(s), stream=int(st)), n, a0=0, train=train))
    return out


def compare(batch_res, single_res, what):
    stats, att, Ts, hs = batch_res
    for j, (st1, T1, h1, na1) in enumerate(single_res):
        st0 = stats[j]
        assert att[j] == na1, f"{what} job {j}: attempts {att[j]} vs {na1}"
        assert st0.n == st1.n and st0.n_finite == st1.n_finite, f"{what} job {j}: counts"
        assert st0.min == st1.min and st0.max == st1.max, f"{what} job {j}: min/max"
        assert_rel(np.array([st0.sum, st0.sum2]), np.array([st1.sum, st1.sum2]), 1e-12, what=f"{what} job {j} sums")
        assert_rel(Ts[j], T1, 1e-12, scale=max(np.abs(T1).max(), 1e-300), what=f"{what} job {j} T")
        assert_rel(hs[j], h1, 1e-12, scale=max(np.abs(h1).max(), 1e-300), what=f"{what} job {j} hist")


def test_batch_matches_single_jobs():
    chans, jobs = make_jobs()
    b = lib.BatchGPU(jobs)
    seeds = [SEED + j for j in range(len(jobs))]
    streams = [7 + j for j in range(len(jobs))]
    # first pass: the default acceptance estimate under-plans the cut-heavy
    # jobs, which then continue through the single-job path
    res = b.iterate(seeds, streams, NCALLS, train=True, graph=False)
    compare(res, single(jobs, seeds, streams, True), "eager pass 1")
    assert b.shape()["topups"] > 0
    # second pass plans from the measured acceptance
    streams2 = [s + 100 for s in streams]
    res = b.iterate(seeds, streams2, NCALLS, train=False, graph=False)
    compare(res, single(jobs, seeds, streams2, False), "eager pass 2")


def test_batch_graph_replays():
    chans, jobs = make_jobs()
    b = lib.BatchGPU(jobs)
    seeds = [SEED + 10 + j for j in range(len(jobs))]
    b.iterate(seeds, [1] * len(jobs), NCALLS, train=False, graph=False)    # acceptance estimates
    for rep in range(3):
        streams = [200 + 10 * rep + j for j in range(len(jobs))]
        res = b.iterate(seeds, streams, NCALLS, train=True, graph=True)
        # eager single-job work between the replays
        compare(res, single(jobs, seeds, streams, True), f"graph replay {rep}")
    sh = b.shape()
    assert sh["replays"] == 3 and sh["captures"] >= 1


def test_batch_autotune():
    chans, jobs = make_jobs()
    b = lib.BatchGPU(jobs)
    seeds = [SEED + 20 + j for j in range(len(jobs))]
    b.iterate(seeds, [2] * len(jobs), NCALLS, train=False, graph=False)
    t = b.tune(seeds, [3] * len(jobs), NCALLS)
    assert t["fwd_wpc"] in (1, 2, 4, 8) and t["main_wpc"] in (1, 2) and t["main_rnd"] in (1, 2, 4)
    assert t["forward_ms"] > 0 and t["weight_ms"] > 0
    streams = [300 + j for j in range(len(jobs))]
    res = b.iterate(seeds, streams, NCALLS, train=True, graph=True)
    compare(res, single(jobs, seeds, streams, True), "tuned graph")


def test_batch_kt_jet_cut():
    """Jobs with the kT-jet PreProcess in a batch, in the same multiplicity
    class as a job without it: attempts, statistics and training sums equal
    each job's single-context iteration (whose kT cut is pinned against the
    oracle's JetCluster)."""
    specs = [(P.gluon_process(4, smax=3), P.integrand_propprod(cuts=("kt", 0.4, 900.0))),
             (P.gluon_process(4, smax=3), P.integrand_propprod()),
             (P.gluon_process(3), P.integrand_propprod(cuts=("kt", 0.4)))]
    ncalls = [15000, 12000, 20000]
    rng = np.random.default_rng(5)
    chans, jobs = [], []
    for proc, f in specs:
        ch = lib.Channels(proc)
        gpu = lib.MultiChannelGPU(ch, 32, f)
        random_state(gpu, rng)
        chans.append(ch)
        jobs.append(gpu)
    b = lib.BatchGPU(jobs)
    seeds = [SEED + 40 + j for j in range(len(jobs))]
    b.iterate(seeds, [1] * len(jobs), ncalls, train=False, graph=False)   # acceptance estimates
    for rep, graph in enumerate((False, True, True)):
        streams = [400 + 10 * rep + j for j in range(len(jobs))]
        stats, att, Ts, hs = b.iterate(seeds, streams, ncalls, train=True, graph=graph)
        for j, (gpu, n) in enumerate(zip(jobs, ncalls)):
            st1, T1, h1, na1 = gpu.iterate(abi.philox_uniforms(int(seeds[j]), stream=int(streams[j])), n, a0=0, train=True)
            what = f"kt batch rep {rep} job {j}"
            assert att[j] == na1, f"{what}: attempts {att[j]} vs {na1}"
            assert stats[j].n == st1.n and stats[j].n_finite == st1.n_finite, what
            assert stats[j].min == st1.min and stats[j].max == st1.max, what
            assert_rel(np.array([stats[j].sum, stats[j].sum2]), np.array([st1.sum, st1.sum2]), 1e-12, what=what + " sums")
            assert_rel(Ts[j], T1, 1e-12, scale=max(np.abs(T1).max(), 1e-300), what=what + " T")
            assert_rel(hs[j], h1, 1e-12, scale=max(np.abs(h1).max(), 1e-300), what=what + " hist")
