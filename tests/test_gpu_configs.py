"""The benchmark configurations end to end on the GPU:

* the batched workload (32 jobs, 2->2 .. 2->5, log-uniform accepted points
  from its seed): the batched pass against every job's single-context
  iteration, and one job per multiplicity's single path against the oracle
  on the GPU's momenta;
* the flow-assisted 2->4 workload (gg_4g, 15 channels, 6 coupling layers)
  and the flow-assisted 2->6 resonant workload (tt_6j, 8 layers): weights of
  the flow path's accepted points against the oracle recomputing the
  multichannel weight from the hypercube points the GPU's flows drew (the
  GPU's log q staged, and the CPU flow restatement's own log q), and the
  integral of briefly trained flows against plain VEGAS (unbiasedness; tt_6j
  at 2^24 points)."""
import numpy as np
import pytest
import torch

from chili_amd import abi, lib
from chili_amd import processes as P
from oracle import flow_oracle as FO
from oracle import oracle as O
from tests.helpers import assert_rel

pytestmark = pytest.mark.gpu


# ---------------------------------------------------------------------------
# batched workload
# ---------------------------------------------------------------------------
def test_batched_32_jobs_against_single_and_oracle():
    spec = P.batched_jobs()
    assert len(spec) == 32 and [s[0].nflav - 2 for s in spec[:4]] == [2, 3, 4, 5]
    chans = [lib.Channels(p) for p, _, _ in spec]
    jobs = [lib.MultiChannelGPU(ch, 64, f) for ch, (_, f, _) in zip(chans, spec)]
    assert sorted({g.nchannels for g in jobs}) == [2, 5, 15, 52] and sorted({g.ndims for g in jobs}) == [4, 7, 10, 13]
    ncalls = [n for _, _, n in spec]
    assert min(ncalls) >= 10 ** 4 and max(ncalls) < 10 ** 6
    J = len(jobs)
    b = lib.BatchGPU(jobs)
    seeds = [P.BATCH_SEED + j for j in range(J)]
    for rnd, (graph, train) in enumerate(((False, True), (True, True), (True, False))):
        streams = [10 * rnd + 1] * J
        print(f"batched pass {rnd}: graph={graph} train={train}", flush=True)
        stats, att, Ts, hs = b.iterate(seeds, streams, ncalls, train=train, graph=graph)
        for j, gpu in enumerate(jobs):
            st1, T1, h1, na1 = gpu.iterate(abi.philox_uniforms(seeds[j], stream=streams[j]), ncalls[j], train=train)
            assert att[j] == na1 and stats[j].n == st1.n and stats[j].n_finite == st1.n_finite, f"job {j}"
            assert stats[j].min == st1.min and stats[j].max == st1.max, f"job {j}"
            assert_rel(np.array([stats[j].sum, stats[j].sum2]), np.array([st1.sum, st1.sum2]), 1e-12, what=f"job {j} sums")
            if train:
                assert_rel(Ts[j], T1, 1e-12, scale=max(np.abs(T1).max(), 1e-300), what=f"job {j} T")
                assert_rel(hs[j], h1, 1e-12, scale=max(np.abs(h1).max(), 1e-300), what=f"job {j} hist")
        if train:   # adapt every job between passes, as the benchmark does
            for gpu, T, h in zip(jobs, Ts, hs):
                gpu.set_alpha(lib.alpha_adapt(gpu.alpha(), T))
                g = gpu.grids()
                for c in range(gpu.nchannels):
                    if np.any(h[c] != 0):
                        g[c] = lib.grid_adapt(g[c], h[c], 1.0)
                gpu.set_grids(g)
    # one job per multiplicity (the smallest of its class): the single path
    # against the oracle replaying the GPU's forward momenta
    for n_out in (2, 3, 4, 5):
        j = min((k for k in range(J) if spec[k][0].nflav - 2 == n_out), key=lambda k: ncalls[k])
        gpu, (proc, f, _) = jobs[j], spec[j]
        ora = O.Oracle(proc, 64, f)
        ora.set_grids(gpu.grids())
        ora.set_alpha(gpu.alpha())
        uni = abi.philox_uniforms(seeds[j], stream=99)
        sg, Tg, hg, nag = gpu.iterate(uni, ncalls[j], train=True)
        rec = gpu.points(uni, 0, nag)
        so, To, ho, nao = ora.iterate(uni, ncalls[j], mom_in=rec["mom"])
        print(f"job {j} (2->{n_out}, {gpu.nchannels} channels, {ncalls[j]} points): attempts {nag}", flush=True)
        assert nag == nao and sg.n == so.n and sg.n_finite == so.n_finite
        assert_rel(np.array([sg.sum, sg.sum2]), np.array([so.sum, so.sum2]), 1e-12, what=f"job {j} vs oracle sums")
        assert_rel(Tg, To, 1e-12, what=f"job {j} vs oracle T")
        assert_rel(hg, ho, 1e-12, scale=np.abs(ho).max(), what=f"job {j} vs oracle histograms")


# ---------------------------------------------------------------------------
# flow-assisted workloads
# ---------------------------------------------------------------------------
def adapted(case, iters, ncalls, seed):
    proc, f = {"gg_4g": lambda: (P.gluon_process(4, smax=3), P.integrand_propprod()),
               "tt_6j": lambda: (P.tt_process(smax=5), P.integrand_resonant())}[case]()
    ch = lib.Channels(proc)
    gpu = lib.MultiChannelGPU(ch, 64, f)
    params = lib.mc_params(ncalls=ncalls, niterations=iters, rtol=0.0, nrefine=0)
    gpu.optimize(abi.philox_uniforms(seed, stream=0), params, lib.mc_summary(), gpu.alpha().copy())
    return proc, f, gpu


def trained_flows(gpu, layers, seed, epochs=3, npts=1 << 17, moment=16):
    C, d = gpu.nchannels, gpu.ndims
    r = torch.empty(C * npts * d, dtype=torch.float64, device="cuda")
    beta = torch.empty(C * npts, dtype=torch.float64, device="cuda")
    w2 = torch.empty(npts, dtype=torch.float64, device="cuda")
    _, nk, _ = gpu.flow_data(None, abi.philox_uniforms(seed, stream=500), npts, r.data_ptr(), beta.data_ptr(), w2.data_ptr(), npts)
    fl = lib.FlowGPU(d, layers, C, seed=seed)
    fl.set_moment(moment)
    losses = fl.train(nk, r.data_ptr(), beta.data_ptr(), w2.data_ptr(), 1 << 14, epochs, 1e-3, seed=seed)
    assert np.all(np.isfinite(losses))
    return fl


def staged_weights_vs_oracle(proc, f, gpu, fl, n, seed, max_points):
    ora = O.Oracle(proc, gpu.nbins, f)
    alpha, grids = gpu.alpha(), gpu.grids()
    ora.set_grids(grids)
    ora.set_alpha(alpha)
    fos = FO.FlowOracle(gpu.ndims, fl.masks(), mode="gpu")
    ps = [torch.tensor(fl.params(j), dtype=torch.float64) for j in range(gpu.nchannels)]
    uni = abi.philox_uniforms(seed, stream=5)
    g = gpu.flow_points(fl, uni, 0, n)
    ok = np.nonzero(g["status"] == 1)[0][:max_points]
    assert len(ok) > 50
    dWst, dW = [], []
    for i in ok:
        c = g["pick"][i]
        x, _ = O.grid_forward(grids[c], g["r"][i].copy())
        mom = ora.generate_point(c, x)
        fv, passed = O.integrand(f, mom, proc.sqrts)
        assert passed
        tot, tot_st = 0.0, 0.0
        for j in range(gpu.nchannels):
            psj, xj = ora.generate_weight(j, mom)
            rj, vwj = O.grid_inverse(grids[j], xj)
            with torch.no_grad():
                lo, _ = fos.density(ps[j], torch.tensor(rj[None, :]))
            tot += alpha[j] * np.exp(float(lo[0])) / psj / vwj
            tot_st += alpha[j] * np.exp(g["lq"][i, j]) / psj / vwj
        dWst.append(abs(1 / tot_st - g["w"][i]) * abs(tot_st))
        dW.append(abs(1 / tot - g["w"][i]) * abs(tot))
    dWst, dW = np.array(dWst), np.array(dW)
    print(f"  {len(ok)} accepted points: staged rel dW median {np.median(dWst):.1e} max {dWst.max():.1e}; with the "
          f"restatement's log q median {np.median(dW):.1e} max {dW.max():.1e}", flush=True)
    # staged: the weight algebra after log q, as the plain path's end-to-end weights
    assert np.median(dWst) <= 1e-10 and dWst.max() <= 1e-6
    # the restatement's own log q (bf16 MLP rounding order)
    assert np.median(dW) < 1e-5 and dW.max() < 1e-3


def mean_err(st):
    m = st.sum / st.n
    return m, np.sqrt((st.sum2 / st.n - m * m) / (st.n - 1))


@pytest.mark.parametrize("case,layers,n_unbiased", [("gg_4g", 6, 1 << 20), ("tt_6j", 8, 1 << 24)])
def test_flow_workload(case, layers, n_unbiased):
    seed = {"gg_4g": 0xC0FFEE03, "tt_6j": 0xC0FFEE04}[case]
    print(f"{case}: adapting", flush=True)
    proc, f, gpu = adapted(case, 3, 1 << 18, seed)
    print(f"{case}: training {gpu.nchannels} flows x {layers} layers", flush=True)
    fl = trained_flows(gpu, layers, seed)
    staged_weights_vs_oracle(proc, f, gpu, fl, 1200, seed, 120 if case == "gg_4g" else 300)
    print(f"{case}: unbiasedness at {n_unbiased} points", flush=True)
    st_plain, _, _, _ = gpu.iterate(abi.philox_uniforms(seed, stream=800), n_unbiased, train=False)
    st_flow, _ = gpu.flow_iterate(fl, abi.philox_uniforms(seed, stream=801), n_unbiased)
    m1, e1 = mean_err(st_plain)
    m2, e2 = mean_err(st_flow)
    print(f"  plain {m1:.6e} +- {e1:.2e}; flows {m2:.6e} +- {e2:.2e}", flush=True)
    assert abs(m1 - m2) < 4 * np.hypot(e1, e2)
