"""GPU parity: the HIP path through the C-ABI against the CPU oracle on the
same uniforms.  Integers (channel picks, cut decisions, bin indices, attempt
counts) must match exactly; reals to 1e-12 relative (momenta to 1e-12 of the
point's largest component)."""
import numpy as np
import pytest

from chili_amd import abi, lib
from chili_amd import processes as P
from oracle import oracle as O
from tests.helpers import assert_rel, stats_tuple

pytestmark = pytest.mark.gpu

SEED = 0xC0FFEE00


def random_grids(rng, nch, d, B):
    g = np.zeros((nch, d, B + 1))
    for c in range(nch):
        for k in range(d):
            inner = np.sort(rng.uniform(0.0, 1.0, B - 1))
            g[c, k] = np.concatenate([[0.0], inner, [1.0]])
    return g


def random_alpha(rng, n):
    a = rng.uniform(0.2, 1.0, n)
    return a / a.sum()


CASES = {
    "gg_2g": lambda: (P.gluon_process(2), P.integrand_propprod()),
    "gg_3g": lambda: (P.gluon_process(3), P.integrand_propprod()),
    "gg_4g": lambda: (P.gluon_process(4, smax=3), P.integrand_propprod()),
    "gg_5g": lambda: (P.gluon_process(5, smax=4), P.integrand_propprod()),
    "tt_6j": lambda: (P.tt_process(smax=5), P.integrand_resonant()),
    "gg_4g_nocut": lambda: (P.gluon_process(4, smax=3), P.integrand_unit(None)),
}


def compare_points(g, o, what):
    n = len(o["pick"])
    assert np.array_equal(g["pick"], o["pick"]), f"{what}: channel picks differ"
    assert np.array_equal(g["status"], o["status"]), f"{what}: cut/NaN status differs at {np.nonzero(g['status'] != o['status'])[0][:5]}"
    sc = np.abs(o["mom"]).max(axis=(1, 2))[:, None, None]
    assert_rel(g["mom"], o["mom"], 1e-12, scale=sc, what=f"{what} momenta")
    ok = o["status"] == 1
    # the inverse point of the picked channel is only defined where its
    # phase-space weight is nonzero (the reference leaves stale slots otherwise)
    pk = o["pick"]
    live = ok & (o["ps"][np.arange(n), pk] > 0)
    assert np.array_equal(g["bins"][live], o["bins"][live]), f"{what}: FindBin indices differ"
    assert_rel(g["xpick"][live], o["xpick"][live], 1e-12, scale=1.0, what=f"{what} inverse point")
    assert_rel(g["ps"][ok], o["ps"][ok], 1e-12, what=f"{what} phase-space weights")
    assert_rel(g["w"][ok], o["w"][ok], 1e-12, what=f"{what} multichannel weight")
    assert_rel(g["val"][ok], o["val"][ok], 1e-12, what=f"{what} values")
    assert_rel(g["f"][ok], o["f"][ok], 1e-12, what=f"{what} integrand")


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("grids", ["uniform", "random"])
def test_points_parity(case, grids):
    proc, f = CASES[case]()
    ch = lib.Channels(proc)
    B = 64
    gpu = lib.MultiChannelGPU(ch, B, f)
    ora = O.Oracle(proc, B, f)
    assert gpu.nchannels == ora.nchannels and gpu.ndims == ora.ndims
    if grids == "random":
        rng = np.random.default_rng(7)
        g = random_grids(rng, gpu.nchannels, gpu.ndims, B)
        a = random_alpha(rng, gpu.nchannels)
        gpu.set_grids(g); ora.set_grids(g)
        gpu.set_alpha(a); ora.set_alpha(a)
    n = 3000
    uni = abi.philox_uniforms(SEED + 1, stream=3)
    og = ora.points(uni, 17, n)
    gg = gpu.points(uni, 17, n)
    print(f"{case}/{grids}: C={gpu.nchannels} accepted {np.sum(og['status'] == 1)}/{n}")
    compare_points(gg, og, f"{case}/{grids}")


def test_points_host_uniforms():
    proc, f = CASES["gg_4g"]()
    ch = lib.Channels(proc)
    gpu = lib.MultiChannelGPU(ch, 64, f)
    ora = O.Oracle(proc, 64, f)
    rng = np.random.default_rng(11)
    n = 2000
    u = np.ascontiguousarray(rng.uniform(0, 1, (gpu.ndims + 1, n)))
    uni = abi.host_uniforms(u)
    compare_points(gpu.points(uni, 0, n), ora.points(uni, 0, n), "host uniforms")


@pytest.mark.parametrize("case", ["gg_2g", "gg_4g", "tt_6j"])
def test_iterate_parity(case):
    proc, f = CASES[case]()
    ch = lib.Channels(proc)
    B = 64
    gpu = lib.MultiChannelGPU(ch, B, f)
    ora = O.Oracle(proc, B, f)
    rng = np.random.default_rng(5)
    g = random_grids(rng, gpu.nchannels, gpu.ndims, B)
    a = random_alpha(rng, gpu.nchannels)
    gpu.set_grids(g); ora.set_grids(g)
    gpu.set_alpha(a); ora.set_alpha(a)
    uni = abi.philox_uniforms(SEED + 2, stream=1)
    ncalls = 6000
    so, To, ho, nao = ora.iterate(uni, ncalls, a0=5)
    sg, Tg, hg, nag = gpu.iterate(uni, ncalls, a0=5)
    assert nag == nao, f"attempts {nag} vs {nao}"
    sg_, so_ = stats_tuple(sg), stats_tuple(so)
    assert sg_[0] == so_[0] and sg_[5] == so_[5]
    assert_rel(sg_[1:3], so_[1:3], 1e-12, what="min/max")
    assert_rel(sg_[3:5], so_[3:5], 1e-12, what="sums")
    assert_rel(Tg, To, 1e-12, what="alpha training")
    assert_rel(hg, ho, 1e-12, scale=np.abs(ho).max(), what="histograms")


def test_vegas_parity():
    f = P.integrand_eemumu()
    B = 50
    gpu = lib.VegasGPU(2, B, f)
    edges = O.uniform_grid(2, B)
    uni = abi.philox_uniforms(SEED + 0)
    n = 20000
    for it in range(3):
        so, ho, _ = O.vegas_iterate(edges, f, uni, n, a0=it * n)
        sg, hg = gpu.iterate(uni, n, a0=it * n)
        assert stats_tuple(sg)[0] == n
        assert_rel(stats_tuple(sg)[1:3], stats_tuple(so)[1:3], 1e-12, what="min/max")
        assert_rel(stats_tuple(sg)[3:5], stats_tuple(so)[3:5], 1e-12, what="sums")
        assert_rel(hg, ho, 1e-12, scale=np.abs(ho).max(), what="vegas histogram")
        edges = O.grid_adapt(edges, ho)
        gpu.set_grid(edges)
