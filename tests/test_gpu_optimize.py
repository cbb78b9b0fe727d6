"""The product's optimisation drivers on the GPU against the oracle's:
chili_mc_optimize (MultiChannel::Optimize with Train and RefineChannels,
include/Channel/MultiChannel.hh:122-130,182-194,253-268) and
chili_vegas_optimize (Vegas::Optimize / Refine, src/Integrator/Vegas.cc:
54-82).  Each side trains with its own code.  Counts, bin numbers, ncalls and
the refinement counter must match exactly; statistics, alpha, best weights
and grids to 1e-12 relative (grids: 1e-12 absolute, edges lie in [0, 1]) --
the device sums its histograms in another (fixed) order than the oracle's
serial loop, so bit equality is not expected."""
import numpy as np
import pytest

from chili_amd import abi, lib
from chili_amd import processes as P
from oracle import oracle as O
from tests.helpers import assert_rel, stats_tuple

pytestmark = pytest.mark.gpu

SEED = 0xC0FFEE10


def compare_state(gpu, ora, pg, po, sg, so, bg, bo, what):
    assert (pg.ncalls, pg.iteration, gpu.nbins, sg.nresults) == (po.ncalls, po.iteration, ora.nbins, so.nresults), what
    assert_rel(gpu.alpha(), ora.alpha(), 1e-12, what=f"{what} alpha")
    assert_rel(bg, bo, 1e-12, what=f"{what} best weights")
    assert_rel(np.array([sg.min_diff]), np.array([so.min_diff]), 1e-10, what=f"{what} min_diff")
    assert_rel(gpu.grids(), ora.grids(), 1e-12, scale=1.0, what=f"{what} grids")
    a, b = stats_tuple(sg.sum), stats_tuple(so.sum)
    assert a[0] == b[0] and a[5] == b[5], what
    assert_rel(a[1:5], b[1:5], 1e-12, what=f"{what} pooled stats")


@pytest.mark.parametrize("case", ["gg_3g", "tt_6j"])
def test_mc_optimize_with_refinement_vs_oracle(case):
    """Three iterations with refinement after each (nrefine = 1): 16 -> 32 ->
    64 -> 128 bins, ncalls x 1.5 each time; the oracle replays the GPU's
    forward momenta of every iteration (its weights then agree to 1e-12)."""
    proc, f = {"gg_3g": lambda: (P.gluon_process(3, smax=2), P.integrand_propprod()),
               "tt_6j": lambda: (P.tt_process(smax=5), P.integrand_resonant())}[case]()
    ch = lib.Channels(proc)
    gpu = lib.MultiChannelGPU(ch, 16, f)
    ora = O.Oracle(proc, 16, f)
    pg, po = lib.mc_params(ncalls=4000, niterations=10, nrefine=1), lib.mc_params(ncalls=4000, niterations=10, nrefine=1)
    sg, so = lib.mc_summary(), lib.mc_summary()
    bg, bo = np.zeros(gpu.nchannels), np.zeros(gpu.nchannels)
    uni = abi.philox_uniforms(SEED, stream=100)
    for k in range(3):
        print(f"{case} iteration {k}: {pg.ncalls} points, {gpu.nbins} bins", flush=True)
        # the GPU's forward map at this iteration's state, for the oracle
        rec = gpu.points(abi.philox_uniforms(SEED, stream=100 + k), 0, 30 * pg.ncalls)
        res = gpu.optimize(uni, pg, sg, bg, max_steps=1)
        st = ora.optimize_step(uni, po, so, bo, mom_in=rec["mom"])
        assert len(res) == 1 and res[0].n == st.n and res[0].n_finite == st.n_finite
        assert_rel(stats_tuple(res[0])[1:5], stats_tuple(st)[1:5], 1e-12, what=f"iteration {k} stats")
        compare_state(gpu, ora, pg, po, sg, so, bg, bo, f"{case} after iteration {k}")
    assert gpu.nbins == 128 and pg.ncalls == int(int(int(4000 * 1.5) * 1.5) * 1.5)


def test_mc_optimize_loop_stops_like_the_oracle():
    """A whole Optimize call on each side with its own forward map: the same
    number of iterations (stopped by rtol before niterations), refinement
    schedule and pooled estimate within the weights' conditioning spread."""
    proc, f = P.gluon_process(4, smax=3), P.integrand_propprod()
    ch = lib.Channels(proc)
    gpu = lib.MultiChannelGPU(ch, 16, f)
    ora = O.Oracle(proc, 16, f)
    kw = dict(ncalls=20000, niterations=12, nrefine=2, rtol=0.005, max_bins=64)
    pg, po = lib.mc_params(**kw), lib.mc_params(**kw)
    sg, so = lib.mc_summary(), lib.mc_summary()
    bg, bo = np.zeros(gpu.nchannels), np.zeros(gpu.nchannels)
    uni = abi.philox_uniforms(SEED + 1, stream=0)
    rg = gpu.optimize(uni, pg, sg, bg)
    ro = ora.optimize(uni, po, so, bo)
    print(f"gg_4g Optimize: {len(rg)} iterations, rel err {sg.sum.error() / abs(sg.sum.mean()):.4f}, "
          f"{gpu.nbins} bins, ncalls {pg.ncalls}")
    assert len(rg) == len(ro) < 12 and sg.nresults == so.nresults
    assert (pg.ncalls, pg.iteration, gpu.nbins) == (po.ncalls, po.iteration, ora.nbins)
    assert sg.sum.n == so.sum.n
    assert abs(sg.sum.sum - so.sum.sum) <= 1e-6 * abs(so.sum.sum)
    assert sg.sum.error() / abs(sg.sum.mean()) <= 0.005


def test_vegas_optimize_vs_oracle():
    """The e+e- -> mu+mu- plain VEGAS run (10 iterations of 1e4 points,
    refinement every 5: 50 -> 100 -> 200 bins, ncalls 1e4 -> 2e4 -> 4e4),
    each side adapting with its own code."""
    f = P.integrand_eemumu()
    gpu = lib.VegasGPU(2, 50, f)
    kw = dict(ncalls=10000, nrefine=5, niterations=10, rtol=1e-4, atol=1e-4)
    pg, po = lib.vegas_params(**kw), lib.vegas_params(**kw)
    tg = abi.Stats(0, np.finfo(float).max, np.finfo(float).tiny, 0, 0, 0)
    to = abi.Stats(0, np.finfo(float).max, np.finfo(float).tiny, 0, 0, 0)
    uni = abi.philox_uniforms(SEED + 2, stream=0)
    rg, ng = gpu.optimize(uni, pg, tg, 0, 10)
    eo, ro, no = O.vegas_optimize(O.uniform_grid(2, 50), f, uni, po, to, 0, 10)
    assert ng == no == 10 and gpu.nbins == 200 and eo.shape == (2, 201)
    assert (pg.ncalls, pg.irefine) == (po.ncalls, po.irefine) == (40000, 0)
    for k in range(10):
        assert rg[k].n == ro[k].n
        assert_rel(stats_tuple(rg[k])[1:5], stats_tuple(ro[k])[1:5], 1e-12, what=f"pass {k}")
    assert_rel(gpu.grid(), eo, 1e-12, scale=1.0, what="grid")
    assert_rel(stats_tuple(tg)[1:5], stats_tuple(to)[1:5], 1e-12, what="pooled")
    print(f"ee->mumu: {tg.sum / tg.n:.6e} +- {tg.error():.2e}")
