"""Pins of the CPU flow restatement (oracle/flow_oracle.py) by mathematics:
identity at zero parameters, bijection of [0,1], inverse round trips,
log-Jacobians against central differences, normalisation of the density by
quadrature and loss gradients against central differences."""
import math

import numpy as np
import pytest
import torch

from oracle.flow_oracle import FlowOracle, loss_and_grads, training_loss, variance_loss


def masks_for(d, L, seed=1):
    rng = np.random.default_rng(seed)
    out = []
    for l in range(L):
        if l % 2 == 0:
            half = sorted(rng.permutation(d)[: d // 2].tolist())
            out.append(half)
        else:
            out.append([i for i in range(d) if i not in half])
    return out


def random_params(o, scale, seed):
    rng = np.random.default_rng(seed)
    p = rng.normal(0.0, scale, o.nparams)
    return torch.tensor(p.astype(np.float32).astype(np.float64))


@pytest.mark.parametrize("mode", ["exact", "gpu"])
def test_identity_at_zero(mode):
    d = 6
    o = FlowOracle(d, masks_for(d, 4), mode=mode)
    r = torch.rand(500, d, dtype=torch.float64, generator=torch.Generator().manual_seed(3))
    lq, z = o.density(torch.zeros(o.nparams, dtype=torch.float64), r)
    tol = 1e-13 if mode == "exact" else 1e-6
    assert float(lq.abs().max()) < tol
    assert float((z - r).abs().max()) < tol


def test_spline_bijection_and_inverse():
    o = FlowOracle(2, [[0], [1]], mode="exact")
    g = torch.Generator().manual_seed(5)
    u = torch.randn(1, 3 * 16 - 1, dtype=torch.float64, generator=g) * 2.0
    x = torch.linspace(0, 1, 20001, dtype=torch.float64)
    y, ld = o.rqs(u.expand(x.shape[0], -1), x)
    assert abs(float(y[0])) < 1e-15 and abs(float(y[-1]) - 1) < 1e-15
    assert bool((y[1:] > y[:-1]).all())
    xb, ldb = o.rqs(u.expand(x.shape[0], -1), y, inverse=True)
    assert float((xb - x).abs().max()) < 1e-12
    assert float((ldb - ld).abs().max()) < 1e-9


def test_spline_logdet_finite_difference():
    o = FlowOracle(2, [[0], [1]], mode="exact")
    g = torch.Generator().manual_seed(7)
    u = torch.randn(400, 3 * 16 - 1, dtype=torch.float64, generator=g) * 1.5
    x = torch.rand(400, dtype=torch.float64, generator=g) * 0.98 + 0.01
    hstep = 1e-7
    yp, _ = o.rqs(u, x + hstep)
    ym, _ = o.rqs(u, x - hstep)
    _, ld = o.rqs(u, x)
    fd = (yp - ym) / (2 * hstep)
    rel = ((fd - torch.exp(ld)) / torch.exp(ld)).abs()
    # points within h of a knot see two pieces; the spline is C1 there
    assert float(torch.median(rel)) < 1e-7
    assert float(rel.max()) < 1e-4


def test_density_normalised_2d():
    # q integrates to one over [0,1]^2 (midpoint rule on a fine grid)
    o = FlowOracle(2, [[0], [1], [0], [1]], mode="exact")
    p = random_params(o, 0.1, 11)
    n = 500
    c = (torch.arange(n, dtype=torch.float64) + 0.5) / n
    r = torch.stack(torch.meshgrid(c, c, indexing="ij"), -1).reshape(-1, 2)
    with torch.no_grad():
        lq, _ = o.density(p, r)
    q = torch.exp(lq)
    assert float(q.max() / q.min()) > 10.0         # a far from uniform density
    assert abs(float(q.mean()) - 1.0) < 5e-5


def test_sample_inverts_density():
    d = 5
    o = FlowOracle(d, masks_for(d, 4), mode="exact")
    p = random_params(o, 0.2, 13)
    z = torch.rand(300, d, dtype=torch.float64, generator=torch.Generator().manual_seed(2))
    with torch.no_grad():
        r, lq_s = o.sample(p, z)
        lq_d, z2 = o.density(p, r)
    assert float((z2 - z).abs().max()) < 1e-10
    assert float((lq_s - lq_d).abs().max()) < 1e-8


@pytest.mark.parametrize("mix,moment", [(0.0, 2), (0.4, 2), (0.0, 4), (0.3, 3)])
def test_loss_gradient_finite_difference(mix, moment):
    d = 4
    C = 2
    o = [FlowOracle(d, masks_for(d, 2, seed=s), mode="exact") for s in (1, 2)]
    ps = [random_params(oo, 0.2, 20 + j).numpy() for j, oo in enumerate(o)]
    g = np.random.default_rng(4)
    n = 64
    r = g.random((C, n, d))
    beta = g.random((C, n)) + 0.1
    beta /= beta.sum(0)
    w2 = g.random(n) ** 2
    L, grads = loss_and_grads(o, ps, r, beta, w2, mix, moment)
    rng = np.random.default_rng(9)
    for j in range(C):
        idx = rng.choice(o[j].nparams, 12, replace=False)
        for k in idx:
            h = 1e-6
            pp = [torch.tensor(p, dtype=torch.float64) for p in ps]
            pm = [torch.tensor(p, dtype=torch.float64) for p in ps]
            pp[j][k] += h
            pm[j][k] -= h
            with torch.no_grad():
                Lp = float(training_loss(o, pp, torch.tensor(r), torch.tensor(beta), torch.tensor(w2), mix, moment))
                Lm = float(training_loss(o, pm, torch.tensor(r), torch.tensor(beta), torch.tensor(w2), mix, moment))
            fd = (Lp - Lm) / (2 * h)
            assert abs(fd - grads[j][k]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (j, k, fd, grads[j][k])


def test_mixture_loss_limits():
    # identity flows (q = 1): the loss is mean w2 / sum_j beta_j whatever the
    # mixture weight; an all-uniform proposal (mix = 1) has zero gradient
    d, C = 3, 2
    o = [FlowOracle(d, masks_for(d, 2, seed=s), mode="exact") for s in (1, 2)]
    g = np.random.default_rng(6)
    n = 32
    r = g.random((C, n, d))
    beta = g.random((C, n)) + 0.1
    w2 = g.random(n) ** 2
    want = float(np.mean(w2 / beta.sum(0)))
    zero = [np.zeros(oo.nparams) for oo in o]
    for mix in (0.0, 0.3, 1.0):
        L, _ = loss_and_grads(o, zero, r, beta, w2, mix)
        assert abs(L - want) <= 1e-12 * want
    # the p-th moment: mean w^p / (sum_j beta_j)^(p-1) for q = 1 (trained as
    # its log over p)
    for p in (3, 4):
        wp = float(np.mean(w2 ** (p / 2) / beta.sum(0) ** (p - 1)))
        Lp = float(variance_loss(o, [torch.tensor(z) for z in zero], torch.tensor(r), torch.tensor(beta), torch.tensor(w2), 0.2, p))
        assert abs(Lp - wp) <= 1e-12 * wp
        Jp, _ = loss_and_grads(o, zero, r, beta, w2, 0.2, p)
        assert abs(Jp - np.log(wp) / p) <= 1e-12
    ps = [random_params(oo, 0.2, 30 + j).numpy() for j, oo in enumerate(o)]
    L1, g1 = loss_and_grads(o, ps, r, beta, w2, 1.0)
    assert abs(L1 - want) <= 1e-12 * want
    assert all(np.abs(gj).max() == 0.0 for gj in g1)


def test_gpu_mode_close_to_exact():
    # bf16 operands and FP32 knots move log q by far less than its spread
    d = 6
    masks = masks_for(d, 4)
    oe, og = FlowOracle(d, masks, mode="exact"), FlowOracle(d, masks, mode="gpu")
    p = random_params(oe, 0.05, 17)
    r = torch.rand(2000, d, dtype=torch.float64, generator=torch.Generator().manual_seed(8))
    with torch.no_grad():
        le, _ = oe.density(p, r)
        lg, _ = og.density(p, r)
    assert float(le.std()) > 0.05
    assert float((le - lg).abs().max()) < 0.01 * float(le.std())
