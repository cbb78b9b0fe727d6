"""Unweighting stage (chili_mc_generate): the MultiChannel loop with an
unweighting PostProcess (MultiChannel.hh:240-245), against the oracle's
restatement on the same Philox streams; the events' distribution against the
weighted one."""
import ctypes

import numpy as np
import pytest
import torch

from chili_amd import abi, lib
from chili_amd import processes as P
from oracle import oracle as O
from tests.helpers import assert_rel, stats_tuple
from tests.test_gpu_parity import CASES, random_alpha, random_grids

pytestmark = pytest.mark.gpu

SEED = 0xC0FFEE05


def setup(case, B=50, seed=3):
    proc, f = CASES[case]()
    ch = lib.Channels(proc)
    gpu = lib.MultiChannelGPU(ch, B, f)
    ora = O.Oracle(proc, B, f)
    rng = np.random.default_rng(seed)
    g = random_grids(rng, gpu.nchannels, gpu.ndims, B)
    a = random_alpha(rng, gpu.nchannels)
    gpu.set_grids(g); ora.set_grids(g)
    gpu.set_alpha(a); ora.set_alpha(a)
    return proc, ch, gpu, ora


def generate(gpu, uni, nevents, wmax, a0=0):
    ntot = gpu.ntot
    ev = torch.empty(4 * ntot * nevents, dtype=torch.float64, device="cuda")
    ew = torch.empty(nevents, dtype=torch.float64, device="cuda")
    st, na = gpu.generate(uni, nevents, wmax, ctypes.c_void_p(ev.data_ptr()), ctypes.c_void_p(ew.data_ptr()), nevents, a0=a0)
    # [4 ntot][n] -> [n][ntot][4]
    events = ev.view(ntot, 4, nevents).permute(2, 0, 1).cpu().numpy()
    return st, na, events, ew.cpu().numpy()


@pytest.mark.parametrize("case", ["gg_3g", "gg_4g", "tt_6j"])
def test_generate_parity(case):
    proc, ch, gpu, ora = setup(case)
    uni = abi.philox_uniforms(SEED, stream=3)
    probe = gpu.points(uni, 1 << 20, 4000)
    v = np.abs(probe["val"][probe["status"] == 1])
    wmax = float(np.quantile(v, 0.9))        # overweight events included
    n, a0 = 3000, 7
    st, na, ev, ew = generate(gpu, uni, n, wmax, a0=a0)
    assert np.all(np.isin(np.abs(ew[np.abs(ew) <= 1]), [1.0])) and np.any(np.abs(ew) > 1)
    # staged: the oracle's loop on the GPU's own forward momenta of every attempt
    rec = gpu.points(uni, a0, na)
    so, nao, evo, ewo = ora.generate(uni, n, wmax, a0=a0, mom_in=rec["mom"])
    assert nao == na, f"attempts {na} vs {nao}"
    sg_, so_ = stats_tuple(st), stats_tuple(so)
    assert sg_[0] == so_[0] == na and sg_[5] == so_[5]
    assert_rel(sg_[1:5], so_[1:5], 1e-12, what=f"{case} stats")
    assert np.array_equal(ev, evo), f"{case}: events differ"
    assert_rel(ew, ewo, 1e-12, what=f"{case} event weights")
    # the oracle's own forward maps: the same attempts and decisions
    so2, nao2, evo2, _ = ora.generate(uni, n, wmax, a0=a0)
    assert nao2 == na
    sc = np.abs(evo2).max(axis=(1, 2))[:, None, None]
    assert np.all(np.abs(ev - evo2) <= 1e-12 * sc)
    print(f"{case}: {n} events from {na} attempts (eff {n / na:.4f}), {np.sum(np.abs(ew) > 1)} overweight")


def test_generate_distribution_matches_weighted():
    """Unweighted events (with their overweight factors) reproduce the
    weighted distribution of an observable: the mean leading transverse
    momentum of 2->3 events against the value-weighted mean of weighted
    points, within errors; and the event count is the kept count."""
    proc, ch, gpu, ora = setup("gg_3g", B=64, seed=9)
    n = 1 << 18
    vals = torch.empty(n, dtype=torch.float64, device="cuda")
    uni_w = abi.philox_uniforms(SEED + 1, stream=1)
    gpu.iterate(uni_w, n, train=False, vals_ptr=ctypes.c_void_p(vals.data_ptr()), vals_cap=n)
    wmax = float(torch.quantile(vals.abs()[: 1 << 16], 0.99))
    rec = gpu.points(uni_w, 0, 1 << 19)
    ok = rec["status"] == 1
    ptl = np.max(np.hypot(rec["mom"][ok, 2:, 1], rec["mom"][ok, 2:, 2]), axis=1)
    wv = rec["val"][ok]
    m_w = np.sum(wv * ptl) / np.sum(wv)
    st, na, ev, ew = generate(gpu, abi.philox_uniforms(SEED + 2, stream=1), n, wmax)
    pte = np.max(np.hypot(ev[:, 2:, 1], ev[:, 2:, 2]), axis=1)
    m_u = np.sum(ew * pte) / np.sum(ew)
    err = np.std(pte) / np.sqrt(n) + np.std(ptl) / np.sqrt(len(ptl)) * np.sqrt(np.mean(wv ** 2)) / np.mean(wv)
    print(f"mean leading pT: weighted {m_w:.4f}, unweighted {m_u:.4f} (+- {err:.4f}); eff {n / na:.4f}")
    assert abs(m_u - m_w) < 5 * err


def test_generate_errors():
    proc, ch, gpu, ora = setup("gg_2g")
    ev = torch.empty(4 * gpu.ntot * 10, dtype=torch.float64, device="cuda")
    ew = torch.empty(10, dtype=torch.float64, device="cuda")
    p_ev, p_ew = ctypes.c_void_p(ev.data_ptr()), ctypes.c_void_p(ew.data_ptr())
    uni = abi.philox_uniforms(SEED)
    with pytest.raises(lib.ChiliError):
        gpu.generate(uni, 10, 0.0, p_ev, p_ew, 10)
    with pytest.raises(lib.ChiliError):
        gpu.generate(uni, 20, 1.0, p_ev, p_ew, 10)
    hu = abi.host_uniforms(np.random.default_rng(1).random((gpu.ndims + 1, 100)))
    with pytest.raises(lib.ChiliError):
        gpu.generate(hu, 5, 1.0, p_ev, p_ew, 10)


@pytest.mark.parametrize("case,modes", [("gg_3g", (0, 8)), ("gg_4g", (5, 13)), ("tt_6j", (5, 13)), ("gg_4g_kt", (1, 9))])
def test_generate_early_rejection_same_events(case, modes):
    """Exact early rejection (chili_mc_specialize bit 3): a passing point is
    weighted in full only when its picked channel's bound can pass the accept
    test.  Attempts, the events (momenta and weights) and StatsData's counts,
    min and max equal the stage without it bit for bit, on two sub-batch sizes
    of the loop; the sums agree to rounding (the loop sizes its sub-batches
    from a running acceptance estimate that the previous call moved, and the
    sums are grouped by sub-batch)."""
    proc, ch, gpu, ora = setup(case)
    uni = abi.philox_uniforms(SEED + 1, stream=5)
    probe = gpu.points(uni, 1 << 20, 4000)
    v = np.abs(probe["val"][probe["status"] == 1])
    for q, n in ((0.9, 3000), (0.999, 20000)):
        wmax = float(np.quantile(v, q))
        out = []
        for mode in modes:
            gpu.specialize(mode)
            out.append(generate(gpu, uni, n, wmax, a0=11))
        (s0, n0, e0, w0), (s1, n1, e1, w1) = out
        assert n0 == n1, f"attempts {n0} vs {n1}"
        t0, t1 = stats_tuple(s0), stats_tuple(s1)
        assert np.array_equal(t0[[0, 1, 2, 5]], t1[[0, 1, 2, 5]])
        assert_rel(t0[3:5], t1[3:5], 1e-13, what=f"{case} sums")
        assert np.array_equal(e0, e1) and np.array_equal(w0, w1)


def test_generate_forward_timing():
    """chili_mc_timing / chili_mc_timing_forward over the unweighting stage:
    one timed forward launch per sub-batch, covering every attempt the loop
    consumed (the last sub-batch's attempts past the cutoff included)."""
    proc, ch, gpu, ora = setup("gg_4g")
    uni = abi.philox_uniforms(SEED + 2, stream=9)
    probe = gpu.points(uni, 1 << 20, 4000)
    wmax = float(np.quantile(np.abs(probe["val"][probe["status"] == 1]), 0.99))
    gpu.timing(enable=True, reset=True)
    st, na, ev, ew = generate(gpu, uni, 5000, wmax)
    tm = gpu.timing(enable=False, reset=True)
    assert tm["forward_launches"] >= 1 and tm["forward_ms"] > 0
    assert tm["forward_attempts"] >= na
    assert tm["forward_attempts"] < na + 2 * tm["forward_attempts"] / tm["forward_launches"] + 1
    # reset: nothing left to report
    tm2 = gpu.timing(enable=False, reset=False)
    assert tm2["forward_launches"] == 0 and tm2["forward_attempts"] == 0 and tm2["forward_ms"] == 0
