"""Batch evaluation op (chili_mc_evaluate), the reference's ChiliCPU.evaluate
/ CallChili on device tensors: one channel's mapping on raw randoms against
the oracle's mapper, weight and cut; the multichannel mode against the
per-attempt records of the same kernels."""
import numpy as np
import pytest
import torch

from chili_amd import abi, lib
from oracle import oracle as O
from tests.test_gpu_parity import CASES, random_alpha, random_grids

pytestmark = pytest.mark.gpu


def context(case, B=64, seed=4):
    proc, f = CASES[case]()
    ch = lib.Channels(proc)
    gpu = lib.MultiChannelGPU(ch, B, f)
    ora = O.Oracle(proc, B, f)
    rng = np.random.default_rng(seed)
    g = random_grids(rng, gpu.nchannels, gpu.ndims, B)
    a = random_alpha(rng, gpu.nchannels)
    gpu.set_grids(g); ora.set_grids(g)
    gpu.set_alpha(a); ora.set_alpha(a)
    return proc, f, gpu, ora


@pytest.mark.parametrize("case", ["gg_3g", "gg_4g", "tt_6j"])
def test_evaluate_single_channel_vs_oracle(case):
    proc, f, gpu, ora = context(case)
    rng = np.random.default_rng(21)
    n = 2000
    r = rng.random((n, gpu.ndims))
    for c in sorted({0, gpu.nchannels - 1}):
        out = gpu.evaluate(torch.tensor(r, device="cuda"), channel=c).cpu().numpy()
        want = np.zeros(n)
        for i in range(n):
            mom = ora.generate_point(c, r[i].copy())
            fv, passed = O.integrand(f, mom, proc.sqrts)
            if not passed:
                continue
            ps, _ = ora.generate_weight(c, mom)
            want[i] = 0.0 if ps == 0 else fv * ps
        assert np.array_equal(out == 0, want == 0), f"{case} channel {c}: cut decisions differ"
        nz = want != 0
        d = np.abs(out[nz] - want[nz]) / np.abs(want[nz])
        print(f"{case} channel {c}: {nz.sum()}/{n} pass; rel diff median {np.median(d):.1e} max {d.max():.1e}")
        # independent forward maps: the spread of the mapper's libm calls
        # through sqrt(|m^2|) of massless children (as in the points parity)
        assert np.median(d) <= 1e-10 and d.max() <= 1e-6


def test_evaluate_dlpack_and_row_stride():
    proc, f, gpu, ora = context("gg_3g")
    rng = np.random.default_rng(5)
    wide = torch.tensor(rng.random((3000, gpu.ndims + 3)), device="cuda")
    rows = wide[:, 1:1 + gpu.ndims]                       # row stride d + 3
    a = gpu.evaluate(rows, channel=1)
    b = gpu.evaluate(torch.utils.dlpack.to_dlpack(rows.contiguous()), channel=1)
    assert torch.equal(a, b)


def test_evaluate_multichannel_matches_points():
    proc, f, gpu, ora = context("tt_6j")
    rng = np.random.default_rng(8)
    n = 5000
    u = rng.random((n, gpu.ndims + 1))
    out = gpu.evaluate(torch.tensor(u, device="cuda"), channel=-1).cpu().numpy()
    uT = np.ascontiguousarray(u.T)        # kept alive while the library reads it
    rec = gpu.points(abi.host_uniforms(uT), 0, n)
    want = np.where(rec["status"] == 1, rec["val"], 0.0)
    assert np.array_equal(out, want)
    assert (rec["status"] == 1).sum() > 100


def test_evaluate_errors():
    proc, f, gpu, ora = context("gg_2g")
    r = torch.rand(10, gpu.ndims + 1, dtype=torch.float64, device="cuda")
    with pytest.raises(lib.ChiliError):
        gpu.evaluate(r, channel=gpu.nchannels)
    with pytest.raises(lib.ChiliError):
        gpu.evaluate(r.float(), channel=0)
    with pytest.raises(lib.ChiliError):
        gpu.evaluate(r[:, :gpu.ndims], channel=-1)
    r[3, 0] = 1.0
    with pytest.raises(lib.ChiliError):
        gpu.evaluate(r, channel=-1)
