"""chili_vegas_iterate_static (the allocation-free variant of
chili_vegas_iterate, its totals and partials in a static per-GPU device
array) against chili_vegas_iterate through the C-ABI: statistics,
histograms and per-point values bit for bit, on two calls in a row, the
second with a smaller input, and on an input that needs several launch
chunks of the static scratch."""
import sys

import numpy as np
import pytest
import torch

from chili_amd import abi, lib
from chili_amd import processes as P
from tests.helpers import stats_tuple

pytestmark = pytest.mark.gpu


def run_both(v, uni, n, a0):
    va = torch.full((n,), -1.0, dtype=torch.float64, device="cuda:0")
    vb = torch.full((n,), -2.0, dtype=torch.float64, device="cuda:0")
    torch.cuda.synchronize()
    sa, ha = v.iterate(uni, n, a0=a0, vals_ptr=va.data_ptr(), vals_cap=n)
    sb, hb = v.iterate_static(uni, n, a0=a0, vals_ptr=vb.data_ptr(), vals_cap=n)
    return (sa, ha, va.cpu().numpy()), (sb, hb, vb.cpu().numpy())


@pytest.mark.parametrize("B", [50, 256])
def test_static_matches_allocating(B):
    v = lib.VegasGPU(2, B, P.integrand_eemumu())
    rng = np.random.default_rng(B)
    g = np.array([np.concatenate([[0.0], np.sort(rng.uniform(0, 1, B - 1)), [1.0]]) for _ in range(2)])
    v.set_grid(g)
    # two calls in a row, the second smaller; then one spanning several
    # chunks of the scratch (B = 256: 258,048 points per chunk; B = 50:
    # 1,263,616)
    for n, stream in ((300_000, 1), (70_001, 2), (3_000_000, 3)):
        print(f"B={B} n={n}", flush=True)
        uni = abi.philox_uniforms(0xC0FFEE00, stream=stream)
        (sa, ha, va), (sb, hb, vb) = run_both(v, uni, n, a0=11)
        assert np.array_equal(stats_tuple(sa), stats_tuple(sb)), (stats_tuple(sa), stats_tuple(sb))
        assert np.array_equal(ha, hb)
        assert np.array_equal(va, vb)
        assert sa.n == n


def test_static_rejects_host_uniforms():
    v = lib.VegasGPU(2, 50, P.integrand_eemumu())
    u = np.ascontiguousarray(np.random.default_rng(0).uniform(0, 1, (2, 100)))
    with pytest.raises(lib.ChiliError, match="device-generated"):
        v.iterate_static(abi.host_uniforms(u), 100)
