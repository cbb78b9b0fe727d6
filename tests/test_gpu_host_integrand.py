"""The host-callback integrand (chili_mc_set_host_integrand: the reference's
std::function integrand and PreProcess, include/Channel/Integrand.hh:41-47)
against the device integrand: a context whose integrand and pT / |y| / dR cut
run on the host (here the oracle's restatement of the same definitions) gives
the device-evaluated context's picks, cut decisions, weights, values, attempt
counts, statistics and training sums."""
import numpy as np
import pytest

from chili_amd import abi, lib
from chili_amd import processes as P
from oracle import oracle as O
from tests.helpers import assert_rel, stats_tuple

pytestmark = pytest.mark.gpu

SEED = 0xC0FFEE20


def host_fn(f, sqrts, calls):
    def fn(mom):
        calls.append(len(mom))
        out = [O.integrand(f, m, sqrts) for m in mom]
        return np.array([v for v, _ in out]), np.array([ok for _, ok in out])
    return fn


@pytest.mark.parametrize("case", ["gg_3g", "tt_6j"])
def test_host_integrand_matches_device(case):
    proc, f = {"gg_3g": lambda: (P.gluon_process(3, smax=2), P.integrand_propprod()),
               "tt_6j": lambda: (P.tt_process(smax=5), P.integrand_resonant())}[case]()
    f_nocut = {"gg_3g": lambda: P.integrand_propprod(cuts=None), "tt_6j": lambda: P.integrand_resonant(cuts=None)}[case]()
    ch = lib.Channels(proc)
    dev = lib.MultiChannelGPU(ch, 32, f)
    host = lib.MultiChannelGPU(ch, 32, f_nocut)
    rng = np.random.default_rng(4)
    g = np.array([[np.concatenate([[0.0], np.sort(rng.uniform(0, 1, 31)), [1.0]]) for _ in range(dev.ndims)]
                  for _ in range(dev.nchannels)])
    a = rng.uniform(0.2, 1, dev.nchannels)
    for ctx in (dev, host):
        ctx.set_grids(g)
        ctx.set_alpha(a / a.sum())
    calls = []
    host.set_host_integrand(host_fn(f, proc.sqrts, calls))
    uni = abi.philox_uniforms(SEED, stream=1)
    n = 3000
    pd, ph = dev.points(uni, 0, n), host.points(uni, 0, n)
    assert sum(calls) > 0
    assert np.array_equal(pd["pick"], ph["pick"]) and np.array_equal(pd["status"], ph["status"])
    ok = pd["status"] == 1
    assert ok.sum() > 100
    assert_rel(ph["f"][ok], pd["f"][ok], 1e-12, what="integrand")
    assert_rel(ph["w"][ok], pd["w"][ok], 1e-12, what="weights")
    assert_rel(ph["val"][ok], pd["val"][ok], 1e-12, what="values")
    # one iteration with training: the retry loop over several sub-batches
    ncalls = 70000 if case == "gg_3g" else 40000
    sd, Td, hd, nd = dev.iterate(uni, ncalls, train=True)
    calls.clear()
    sh, Th, hh, nh = host.iterate(uni, ncalls, train=True)
    print(f"{case}: {nh} attempts, {len(calls)} callback batches of at most {max(calls)} points")
    assert nd == nh and max(calls) <= 1 << 16 and len(calls) > 1
    a_, b_ = stats_tuple(sd), stats_tuple(sh)
    assert a_[0] == b_[0] and a_[5] == b_[5]
    assert_rel(b_[1:5], a_[1:5], 1e-12, what="stats")
    assert_rel(Th, Td, 1e-12, what="T")
    assert_rel(hh, hd, 1e-12, scale=np.abs(hd).max(), what="histograms")
    # the split weight pass (its first inverse kernel hands f on) with the
    # host callback: the same values
    host.specialize(5)
    ps5 = host.points(uni, 0, n)
    assert np.array_equal(pd["status"], ps5["status"])
    assert_rel(ps5["f"][ok], pd["f"][ok], 1e-12, what="integrand, split pass")
    assert_rel(ps5["val"][ok], pd["val"][ok], 1e-12, what="values, split pass")
    host.specialize(False)
    # back to the device integrand
    host.set_host_integrand(None)
    s2, _, _, n2 = host.iterate(uni, 5000, train=False)
    assert n2 > 0


def test_host_integrand_errors():
    proc = P.gluon_process(3, smax=2)
    ctx = lib.MultiChannelGPU(lib.Channels(proc), 16, P.integrand_propprod(cuts=None))

    def bad(mom):
        raise ValueError("user integrand failed")

    ctx.set_host_integrand(bad)
    with pytest.raises(lib.ChiliError, match="callback"):
        ctx.iterate(abi.philox_uniforms(SEED, stream=2), 1000)
    ctx.set_host_integrand(lambda mom: (np.ones(len(mom)), np.ones(len(mom), bool)))
    import torch
    ev = torch.empty(4 * ctx.ntot * 16, dtype=torch.float64, device="cuda")
    ew = torch.empty(16, dtype=torch.float64, device="cuda")
    with pytest.raises(lib.ChiliError, match="host callback"):
        ctx.generate(abi.philox_uniforms(SEED, stream=3), 16, 1.0, ev.data_ptr(), ew.data_ptr(), 16)
    with pytest.raises(lib.ChiliError, match="host callback"):
        lib.BatchGPU([ctx])
    # f = 1 and every attempt accepted: the values are the weights
    st, _, _, na = ctx.iterate(abi.philox_uniforms(SEED, stream=4), 2000, train=False)
    assert st.n == na and na >= 2000
