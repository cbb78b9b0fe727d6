"""GPU tests of the learned proposal: the MFMA GEMM's fragment layout, the
flow's density / sampling / gradients against the CPU restatement
(oracle/flow_oracle.py), and the multichannel path with flows against the
plain path and the oracle."""
import ctypes
import os

import numpy as np
import pytest
import torch

from chili_amd import abi, lib
from chili_amd import processes as P
from oracle import flow_oracle as FO
from oracle import oracle as O
from tests.helpers import assert_rel

pytestmark = pytest.mark.gpu

SEED = 0xC0FFEE03


def dev(a):
    # callers keep the tensor alive while the library reads its pointer
    return torch.as_tensor(np.ascontiguousarray(a), device="cuda")


def bf16_bits(x):
    t = torch.as_tensor(np.asarray(x, dtype=np.float32)).to(torch.bfloat16)
    return t.view(torch.int16).numpy().view(np.uint16), t.to(torch.float64).numpy()


def gemm_probe(A_bits, B_bits, M, N, K, transA=False, transB=True):
    """The product's BF16 MFMA GEMM kernel on host buffers, through the
    test-only probe library (tests/probes)."""
    import ctypes as C
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "probes", "_build", "libchili_probe.so")
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built (make -C tests/probes)")
    L = C.CDLL(path)
    L.chili_probe_gemm_bf16.argtypes = [C.c_int] * 3 + [C.POINTER(C.c_uint16)] * 2 + [C.POINTER(C.c_float), C.c_int, C.c_int]
    a = np.ascontiguousarray(A_bits, dtype=np.uint16)
    b = np.ascontiguousarray(B_bits, dtype=np.uint16)
    c = np.zeros((M, N), dtype=np.float32)
    rc = L.chili_probe_gemm_bf16(M, N, K, a.ctypes.data_as(C.POINTER(C.c_uint16)), b.ctypes.data_as(C.POINTER(C.c_uint16)),
                                 c.ctypes.data_as(C.POINTER(C.c_float)), int(transA), int(transB))
    assert rc == 0, f"probe rc {rc}"
    return c.astype(np.float64)


@pytest.mark.parametrize("shape", [(64, 64, 64), (96, 70, 80), (200, 235, 64), (5, 64, 8)])
def test_gemm_mfma_layout(shape):
    M, N, K = shape
    rng = np.random.default_rng(1)
    a_bits, a = bf16_bits(rng.normal(size=(M, K)))
    b_bits, b = bf16_bits(rng.normal(size=(N, K)))
    c = gemm_probe(a_bits, b_bits, M, N, K, transA=False, transB=True)
    ref = a @ b.T
    err = np.abs(c - ref) / (np.abs(a) @ np.abs(b).T)
    assert err.max() < 1e-6, f"A B^T: max scaled error {err.max():.2e}"
    # transposed-A / plain-B form (weight gradients): A stored [K][M], B [K][N]
    at_bits, at = bf16_bits(rng.normal(size=(K, M)))
    bn_bits, bn = bf16_bits(rng.normal(size=(K, N)))
    c2 = gemm_probe(at_bits, bn_bits, M, N, K, transA=True, transB=False)
    ref2 = at.T @ bn
    err2 = np.abs(c2 - ref2) / (np.abs(at.T) @ np.abs(bn))
    assert err2.max() < 1e-6, f"A^T B: max scaled error {err2.max():.2e}"


def make_flow(d, L, nflows, scale, seed=3):
    fl = lib.FlowGPU(d, L, nflows, seed=seed)
    rng = np.random.default_rng(seed)
    ps = []
    for j in range(nflows):
        p = fl.params(j)
        if scale > 0:
            p = (p + rng.normal(0, scale, p.shape)).astype(np.float32)
            fl.set_params(j, p)
        ps.append(p)
    return fl, ps


def test_flow_identity_at_init():
    fl, _ = make_flow(10, 6, 2, 0.0)
    n = 4096
    r = torch.rand(n, 10, dtype=torch.float64, device="cuda")
    lq = torch.empty(n, dtype=torch.float64, device="cuda")
    z = torch.empty_like(r)
    fl.log_density(1, n, r.data_ptr(), lq.data_ptr(), z.data_ptr())
    torch.cuda.synchronize()
    assert float(lq.abs().max()) < 1e-6
    assert float((z - r).abs().max()) < 1e-7


# odd d leaves one lane half of the last dimension pair idle; d = 22 has six
# pairs per layer (the widest coupling kernel)
@pytest.mark.parametrize("d,L", [(10, 6), (16, 8), (4, 2), (7, 4), (3, 2), (22, 2)])
def test_flow_density_vs_oracle(d, L):
    fl, ps = make_flow(d, L, 1, 0.05, seed=d)
    o = FO.FlowOracle(d, fl.masks(), mode="gpu")
    assert o.nparams == fl.nparams
    n = 4096
    rng = np.random.default_rng(5)
    r = rng.random((n, d))
    lq = torch.empty(n, dtype=torch.float64, device="cuda")
    z = torch.empty(n, d, dtype=torch.float64, device="cuda")
    rd = dev(r)
    fl.log_density(0, n, rd.data_ptr(), lq.data_ptr(), z.data_ptr())
    torch.cuda.synchronize()
    with torch.no_grad():
        lo, zo = o.density(torch.tensor(ps[0], dtype=torch.float64), torch.tensor(r))
    dl = np.abs(lq.cpu().numpy() - lo.numpy())
    dz = np.abs(z.cpu().numpy() - zo.numpy())
    spread = float(lo.std())
    print(f"d={d} L={L}: log q spread {spread:.3f}; |dlq| median {np.median(dl):.1e} p99 {np.quantile(dl, .99):.1e} "
          f"max {dl.max():.1e}; |dz| max {dz.max():.1e}")
    assert spread > 0.05
    # same bf16 operands; FP32 accumulation order differs, which can flip the
    # bf16 rounding of an activation now and then
    assert np.median(dl) < 1e-5
    assert np.quantile(dl, 0.99) < 1e-3
    assert dl.max() < 0.02 * spread


@pytest.mark.parametrize("d", [10, 7])
def test_flow_sample_inverts_density(d):
    fl, _ = make_flow(d, 6, 1, 0.05, seed=4)
    n = 8192
    z = torch.rand(n, d, dtype=torch.float64, device="cuda")
    r = z.clone()
    lqs = torch.empty(n, dtype=torch.float64, device="cuda")
    fl.sample(0, n, r.data_ptr(), lqs.data_ptr())
    lqd = torch.empty_like(lqs)
    z2 = torch.empty_like(z)
    fl.log_density(0, n, r.data_ptr(), lqd.data_ptr(), z2.data_ptr())
    torch.cuda.synchronize()
    assert float((r - z).abs().max()) > 1e-3          # not the identity
    assert float((z2 - z).abs().max()) < 1e-10
    assert float((lqs - lqd).abs().max()) < 1e-8


def test_flow_normalised_2d():
    # the density integrates to one whatever the network computes
    fl, _ = make_flow(2, 4, 1, 0.1, seed=6)
    n = 600
    c = (np.arange(n) + 0.5) / n
    r = np.stack(np.meshgrid(c, c, indexing="ij"), -1).reshape(-1, 2)
    lq = torch.empty(n * n, dtype=torch.float64, device="cuda")
    rd = dev(r)
    fl.log_density(0, n * n, rd.data_ptr(), lq.data_ptr())
    q = torch.exp(lq).cpu().numpy()
    assert q.max() / q.min() > 5
    assert abs(q.mean() - 1) < 5e-5


# C = 5: more flows than training lanes (round robin over four streams);
# mix: the loss of the defensive-mixture proposal; moment: the loss's order
@pytest.mark.parametrize("C,mix,moment", [(2, 0.0, 2), (5, 0.0, 2), (2, 0.4, 2), (2, 0.0, 4), (3, 0.3, 3), (2, 0.0, 8)])
def test_flow_gradients_vs_oracle(C, mix, moment):
    d, L = 6, 4
    fl, ps = make_flow(d, L, C, 0.05, seed=8)
    fl.set_mixture(mix)
    fl.set_moment(moment)
    o = [FO.FlowOracle(d, fl.masks(), mode="gpu") for _ in range(C)]
    rng = np.random.default_rng(11)
    n = 2048
    r = rng.random((C, n, d))
    beta = rng.random((C, n)) + 0.05
    beta /= beta.sum(0)
    w2 = rng.gamma(1.0, 1.0, n)
    rd, bd, wd = dev(r), dev(beta), dev(w2)
    loss = fl.loss_grad(n, rd.data_ptr(), bd.data_ptr(), wd.data_ptr())
    Lo, go = FO.loss_and_grads(o, ps, r, beta, w2, mix, moment)
    assert abs(loss - Lo) <= 1e-4 * max(1.0, abs(Lo))
    for j in range(C):
        g = fl.grads(j).astype(np.float64)
        rel = np.linalg.norm(g - go[j]) / np.linalg.norm(go[j])
        cos = g @ go[j] / (np.linalg.norm(g) * np.linalg.norm(go[j]))
        print(f"flow {j}: |g - g_oracle| / |g_oracle| = {rel:.2e}, cos = {cos:.6f}")
        # the backward pass rounds gradients to bf16 at each product
        assert rel < 3e-2 and cos > 0.999


def test_flow_training_reduces_loss():
    d, L, C = 4, 4, 1
    fl, _ = make_flow(d, L, C, 0.0, seed=9)
    # target density on [0,1]^4: a product of peaked profiles; samples drawn
    # uniformly, so beta = 1 and w2 = f^2 (the variance loss of q against f)
    rng = np.random.default_rng(12)
    n = 1 << 16
    r = rng.random((C, n, d))
    f = np.prod(1.0 + 4.0 * np.exp(-((r[0] - 0.3) ** 2) / 0.02), axis=1)
    f /= f.mean()
    beta = np.ones((C, n))
    w2 = f ** 2
    rd, bd, wd = dev(r), dev(beta), dev(w2)
    hist = fl.train(n, rd.data_ptr(), bd.data_ptr(), wd.data_ptr(), 4096, 12, 2e-3, seed=1)
    print("loss by epoch:", np.array2string(hist, precision=4))
    # loss = E[f^2 / q] = 1 + var(f/q) for normalised f: starts at E[f^2]
    assert hist[0] > hist[-1]
    assert hist[-1] < 0.5 * float(np.mean(w2))


# ---------------------------------------------------------------------------
# multichannel path with flows
# ---------------------------------------------------------------------------
def adapted_context(case="gg_3g", B=64, iters=3, seed=SEED):
    proc, f = (P.gluon_process(3), P.integrand_propprod()) if case == "gg_3g" else (P.gluon_process(4, smax=3), P.integrand_propprod())
    ch = lib.Channels(proc)
    gpu = lib.MultiChannelGPU(ch, B, f)
    uni = abi.philox_uniforms(seed, stream=1)
    a0 = 0
    for it in range(iters):
        st, T, hist, na = gpu.iterate(uni, 20000, a0=a0, train=True)
        a0 += na
        gpu.set_alpha(lib.alpha_adapt(gpu.alpha(), T))
        g = gpu.grids()
        for c in range(gpu.nchannels):
            if hist[c].sum() > 0:
                g[c] = lib.grid_adapt(g[c], hist[c])
        gpu.set_grids(g)
    return proc, f, ch, gpu


def test_flow_path_without_flows_is_plain_path():
    proc, f, ch, gpu = adapted_context()
    uni = abi.philox_uniforms(SEED + 9, stream=2)
    n = 20000
    a = gpu.points(uni, 5, n)
    b = gpu.flow_points(None, uni, 5, n)
    assert np.array_equal(a["status"], b["status"])
    assert np.array_equal(a["pick"], b["pick"])
    ok = a["status"] == 1
    assert np.array_equal(a["w"][ok], b["w"][ok])
    assert np.array_equal(a["val"][ok], b["val"][ok])


def test_flow_path_identity_flows():
    proc, f, ch, gpu = adapted_context()
    fl = lib.FlowGPU(gpu.ndims, 4, gpu.nchannels, seed=2)
    uni = abi.philox_uniforms(SEED + 9, stream=2)
    n = 20000
    a = gpu.points(uni, 5, n)
    b = gpu.flow_points(fl, uni, 5, n)
    assert np.array_equal(a["pick"], b["pick"])
    ok = (a["status"] == 1) & (b["status"] == 1)
    assert ok.sum() > 0.9 * (a["status"] == 1).sum()
    d = np.abs(b["w"][ok] - a["w"][ok]) / np.abs(a["w"][ok])
    print(f"identity flows: rel dW median {np.median(d):.1e} max {d.max():.1e}; |lq| max {np.abs(b['lq'][ok]).max():.1e}")
    assert np.median(d) < 1e-6 and np.quantile(d, 0.99) < 1e-4


def test_flow_points_vs_oracle():
    """End to end: the oracle recomputes the proposal's weight of every
    accepted point from the hypercube point the GPU's flow drew."""
    proc, f, ch, gpu = adapted_context()
    fl, ps = make_flow(gpu.ndims, 4, gpu.nchannels, 0.03, seed=13)
    ora = O.Oracle(proc, gpu.nbins, f)
    ora.set_grids(gpu.grids())
    ora.set_alpha(gpu.alpha())
    alpha = gpu.alpha()
    grids = gpu.grids()
    fos = FO.FlowOracle(gpu.ndims, fl.masks(), mode="gpu")
    uni = abi.philox_uniforms(SEED + 4, stream=5)
    n = 1500
    g = gpu.flow_points(fl, uni, 11, n)
    # the flow's sample of each attempt vs the oracle's inverse flow
    z = np.array([[O.philox_uniform(uni.seed, uni.stream, 11 + i, k) for k in range(gpu.ndims)] for i in range(n)])
    pick = np.array([O.pick_channel(alpha, O.philox_uniform(uni.seed, uni.stream, 11 + i, gpu.ndims)) for i in range(n)])
    assert np.array_equal(pick, g["pick"])
    for c in range(gpu.nchannels):
        sel = pick == c
        if not sel.any():
            continue
        with torch.no_grad():
            ro, _ = fos.sample(torch.tensor(ps[c], dtype=torch.float64), torch.tensor(z[sel]))
        assert np.abs(ro.numpy() - g["r"][sel]).max() < 1e-3
    ok = np.nonzero(g["status"] == 1)[0]
    assert len(ok) > 100
    dW, dWst, dlq = [], [], []
    for i in ok:
        c = g["pick"][i]
        r = g["r"][i].copy()
        x, _ = O.grid_forward(grids[c], r)
        mom = ora.generate_point(c, x)
        fv, passed = O.integrand(f, mom, proc.sqrts)
        assert passed
        tot, tot_st = 0.0, 0.0
        for j in range(gpu.nchannels):
            psj, xj = ora.generate_weight(j, mom)
            rj, vwj = O.grid_inverse(grids[j], xj)
            with torch.no_grad():
                lo, _ = fos.density(torch.tensor(ps[j], dtype=torch.float64), torch.tensor(rj[None, :]))
            lo = float(lo[0])
            dlq.append(abs(lo - g["lq"][i, j]))
            tot += alpha[j] * np.exp(lo) / psj / vwj
            tot_st += alpha[j] * np.exp(g["lq"][i, j]) / psj / vwj
        dW.append(abs(1 / tot - g["w"][i]) / abs(1 / tot))
        dWst.append(abs(1 / tot_st - g["w"][i]) / abs(1 / tot_st))
    dW, dWst, dlq = map(np.array, (dW, dWst, dlq))
    print(f"flow points: {len(ok)} accepted; staged rel dW median {np.median(dWst):.1e} max {dWst.max():.1e}; "
          f"full rel dW median {np.median(dW):.1e} max {dW.max():.1e}; |dlq| median {np.median(dlq):.1e} max {dlq.max():.1e}")
    # staged (the GPU's log q): the remaining weight algebra agrees as the plain path's end-to-end weights do
    assert np.median(dWst) <= 1e-10 and dWst.max() <= 1e-6
    # full: the oracle's own log q (bf16 MLP, FP32 accumulation order)
    assert np.median(dW) < 1e-5 and dW.max() < 1e-2


def test_flow_path_unbiased():
    """Any proposal gives the same integral: random (untrained) flows against
    the plain path, independent samples, within 4 combined sigma."""
    proc, f, ch, gpu = adapted_context()
    fl, _ = make_flow(gpu.ndims, 4, gpu.nchannels, 0.05, seed=21)
    n = 400000
    st1, _, _, _ = gpu.iterate(abi.philox_uniforms(SEED + 30, stream=1), n, train=False)
    st2, _ = gpu.flow_iterate(fl, abi.philox_uniforms(SEED + 31, stream=1), n)

    def mean_err(st):
        m = st.sum / st.n
        var = (st.sum2 / st.n - m * m) / (st.n - 1)
        return m, np.sqrt(var)

    m1, e1 = mean_err(st1)
    m2, e2 = mean_err(st2)
    print(f"plain {m1:.6e} +- {e1:.2e}; random flows {m2:.6e} +- {e2:.2e}")
    assert abs(m1 - m2) < 4 * np.hypot(e1, e2)


def test_flow_mixture_points_vs_oracle():
    """Defensive mixture (1 - lam) q + lam: attempts whose uniform number d+1
    is below lam keep their uniforms as the hypercube point, the others take
    the flow's sample; every weight uses the mixture density."""
    proc, f, ch, gpu = adapted_context()
    fl, ps = make_flow(gpu.ndims, 4, gpu.nchannels, 0.03, seed=17)
    lam = 0.4
    fl.set_mixture(lam)
    ora = O.Oracle(proc, gpu.nbins, f)
    ora.set_grids(gpu.grids())
    ora.set_alpha(gpu.alpha())
    alpha, grids = gpu.alpha(), gpu.grids()
    fos = FO.FlowOracle(gpu.ndims, fl.masks(), mode="gpu")
    uni = abi.philox_uniforms(SEED + 6, stream=7)
    n, a0, d = 1500, 23, gpu.ndims
    g = gpu.flow_points(fl, uni, a0, n)
    z = np.array([[O.philox_uniform(uni.seed, uni.stream, a0 + i, k) for k in range(d)] for i in range(n)])
    um = np.array([O.philox_uniform(uni.seed, uni.stream, a0 + i, d + 1) for i in range(n)])
    pick = np.array([O.pick_channel(alpha, O.philox_uniform(uni.seed, uni.stream, a0 + i, d)) for i in range(n)])
    assert np.array_equal(pick, g["pick"])
    uni_comp = um < lam
    assert 0.3 * n < uni_comp.sum() < 0.5 * n
    assert np.array_equal(g["r"][uni_comp], z[uni_comp])
    for c in range(gpu.nchannels):
        sel = (pick == c) & ~uni_comp
        if not sel.any():
            continue
        with torch.no_grad():
            ro, _ = fos.sample(torch.tensor(ps[c], dtype=torch.float64), torch.tensor(z[sel]))
        assert np.abs(ro.numpy() - g["r"][sel]).max() < 1e-3
    ok = np.nonzero(g["status"] == 1)[0]
    assert len(ok) > 100
    dWst = []
    for i in ok:
        c = g["pick"][i]
        x, _ = O.grid_forward(grids[c], g["r"][i].copy())
        mom = ora.generate_point(c, x)
        tot = 0.0
        for j in range(gpu.nchannels):
            psj, xj = ora.generate_weight(j, mom)
            rj, vwj = O.grid_inverse(grids[j], xj)
            tot += alpha[j] * ((1 - lam) * np.exp(g["lq"][i, j]) + lam) / psj / vwj
        dWst.append(abs(1 / tot - g["w"][i]) / abs(1 / tot))
    dWst = np.array(dWst)
    print(f"mixture lam={lam}: {uni_comp.sum()} uniform-component attempts; staged rel dW median {np.median(dWst):.1e} "
          f"max {dWst.max():.1e}")
    assert np.median(dWst) <= 1e-10 and dWst.max() <= 1e-6
    with pytest.raises(lib.ChiliError):
        fl.set_mixture(1.5)


def test_flow_mixture_unbiased():
    """The mixture proposal gives the plain path's integral (independent
    samples, 4 combined sigma)."""
    proc, f, ch, gpu = adapted_context()
    fl, _ = make_flow(gpu.ndims, 4, gpu.nchannels, 0.05, seed=23)
    fl.set_mixture(0.3)
    n = 400000
    st1, _, _, _ = gpu.iterate(abi.philox_uniforms(SEED + 40, stream=1), n, train=False)
    st2, _ = gpu.flow_iterate(fl, abi.philox_uniforms(SEED + 41, stream=1), n)

    def mean_err(st):
        m = st.sum / st.n
        return m, np.sqrt((st.sum2 / st.n - m * m) / (st.n - 1))

    m1, e1 = mean_err(st1)
    m2, e2 = mean_err(st2)
    print(f"plain {m1:.6e} +- {e1:.2e}; mixture flows {m2:.6e} +- {e2:.2e}")
    assert abs(m1 - m2) < 4 * np.hypot(e1, e2)


@pytest.mark.parametrize("mode", [4, 5])
@pytest.mark.parametrize("mix", [0.0, 0.4])
def test_flow_path_specialised_context(mix, mode):
    """A context with run-time specialised kernels gives the flow path's
    attempts, statistics and values bit for bit (two identically adapted
    contexts, one specialised): mode 4 the per-channel hypercube points and
    densities of the kept points, 5 with the specialised forward kernel of
    the uniform path set as well."""
    _, _, _, g0 = adapted_context()
    _, _, _, g1 = adapted_context()
    g1.specialize(mode)
    fl, _ = make_flow(g0.ndims, 4, g0.nchannels, 0.05, seed=29)
    fl.set_mixture(mix)
    n = 60000
    out = []
    for g in (g0, g1):
        vals = torch.empty(n, dtype=torch.float64, device="cuda")
        st, na = g.flow_iterate(fl, abi.philox_uniforms(SEED + 60, stream=3), n, vals_ptr=vals.data_ptr(), vals_cap=n)
        torch.cuda.synchronize()
        out.append((na, [st.n, st.min, st.max, st.sum, st.sum2, st.n_finite], vals.cpu().numpy()))
    assert out[0][0] == out[1][0]
    assert out[0][1] == out[1][1]
    assert np.array_equal(out[0][2], out[1][2])


def weight_spread(step, n, stream):
    """StatsData mean and error over all attempts (rejected ones add zeros),
    and the relative std and <w>/w_max of the n accepted values."""
    vals = torch.empty(n, dtype=torch.float64, device="cuda")
    st = step(abi.philox_uniforms(SEED + 50, stream=stream), vals, n)
    v = vals.cpu().numpy()
    m = st.sum / st.n
    err = np.sqrt((st.sum2 / st.n - m * m) / (st.n - 1))
    return m, err, v.std() / v.mean(), v.mean() / v.max()


def test_trained_flow_gains_over_plain_vegas():
    """What the trained network gains: on a briefly adapted 2->3 context,
    flows trained on the method's own samples lower the relative spread of
    the weights against the same context without them, and raise the
    unweighting efficiency; the integral stays within errors."""
    proc, f, ch, gpu = adapted_context(iters=2)
    C, d = gpu.nchannels, gpu.ndims
    cap = 1 << 17
    r = torch.empty(C * cap * d, dtype=torch.float64, device="cuda")
    beta = torch.empty(C * cap, dtype=torch.float64, device="cuda")
    w2 = torch.empty(cap, dtype=torch.float64, device="cuda")
    _, nk, _ = gpu.flow_data(None, abi.philox_uniforms(SEED + 51, stream=1), cap, r.data_ptr(), beta.data_ptr(), w2.data_ptr(), cap)
    fl = lib.FlowGPU(d, 4, C, seed=31)
    hist = fl.train(nk, r.data_ptr(), beta.data_ptr(), w2.data_ptr(), 4096, 8, 2e-3, seed=5)
    n = 1 << 18

    def plain(uni, vals, n_):
        return gpu.iterate(uni, n_, train=False, vals_ptr=C_ptr(vals), vals_cap=n_)[0]

    def flow(uni, vals, n_):
        return gpu.flow_iterate(fl, uni, n_, vals_ptr=C_ptr(vals), vals_cap=n_)[0]

    m0, err0, s0, e0 = weight_spread(plain, n, 60)
    m1, err1, s1, e1 = weight_spread(flow, n, 61)
    print(f"loss {hist[0]:.4g} -> {hist[-1]:.4g}; plain: {m0:.5e} +- {err0:.1e}, rel std {s0:.4f}, eps {e0:.4f}; "
          f"trained flows: {m1:.5e} +- {err1:.1e}, rel std {s1:.4f}, eps {e1:.4f}")
    assert s1 < 0.8 * s0
    assert e1 > e0
    assert err1 < err0
    assert abs(m1 - m0) < 4 * np.hypot(err0, err1)


def test_higher_moment_loss_raises_efficiency():
    """The unweighting efficiency <w>/w_max is set by the largest weights: on
    the same context and training records, flows trained on the 16th moment
    of the weights reach a higher efficiency than flows trained on the
    variance, and both integrals agree with plain VEGAS."""
    proc, f, ch, gpu = adapted_context(iters=2)
    C, d = gpu.nchannels, gpu.ndims
    cap = 1 << 17
    r = torch.empty(C * cap * d, dtype=torch.float64, device="cuda")
    beta = torch.empty(C * cap, dtype=torch.float64, device="cuda")
    w2 = torch.empty(cap, dtype=torch.float64, device="cuda")
    _, nk, _ = gpu.flow_data(None, abi.philox_uniforms(SEED + 51, stream=1), cap, r.data_ptr(), beta.data_ptr(), w2.data_ptr(), cap)
    n = 1 << 18
    res = {}
    for p in (2, 16):
        fl = lib.FlowGPU(d, 4, C, seed=31)
        fl.set_moment(p)
        fl.train(nk, r.data_ptr(), beta.data_ptr(), w2.data_ptr(), 4096, 8, 2e-3, seed=5)

        def flow(uni, vals, n_):
            return gpu.flow_iterate(fl, uni, n_, vals_ptr=C_ptr(vals), vals_cap=n_)[0]

        res[p] = weight_spread(flow, n, 62)

    def plain(uni, vals, n_):
        return gpu.iterate(uni, n_, train=False, vals_ptr=C_ptr(vals), vals_cap=n_)[0]

    m0, err0, s0, e0 = weight_spread(plain, n, 63)
    for p, (m, err, sp, e) in res.items():
        print(f"p = {p}: {m:.5e} +- {err:.1e}, rel std {sp:.4f}, eps {e:.4f} (plain: {m0:.5e} +- {err0:.1e}, eps {e0:.4f})")
        assert abs(m - m0) < 4 * np.hypot(err0, err)
    assert res[16][3] > res[2][3]
    assert res[16][3] > e0


def C_ptr(t):
    import ctypes
    return ctypes.c_void_p(t.data_ptr())


def test_flow_path_errors():
    """Argument errors of the flow entry points are reported, not run: no
    calls, flows of another shape, a mixture on host uniforms, a bad mixture
    weight."""
    _, _, _, gpu = adapted_context(iters=1)
    fl, _ = make_flow(gpu.ndims, 2, gpu.nchannels, 0.0)
    uni = abi.philox_uniforms(SEED + 70, stream=1)
    with pytest.raises(lib.ChiliError):
        gpu.flow_iterate(fl, uni, 0)
    other, _ = make_flow(gpu.ndims + 1, 2, gpu.nchannels, 0.0)
    with pytest.raises(lib.ChiliError):
        gpu.flow_iterate(other, uni, 1000)
    fewer, _ = make_flow(gpu.ndims, 2, gpu.nchannels + 1, 0.0)
    with pytest.raises(lib.ChiliError):
        gpu.flow_iterate(fewer, uni, 1000)
    fl.set_mixture(0.5)
    hu = abi.host_uniforms(np.random.default_rng(2).random((gpu.ndims + 1, 5000)))
    with pytest.raises(lib.ChiliError):
        gpu.flow_iterate(fl, hu, 100)
    for bad in (-0.1, 1.5):
        with pytest.raises(lib.ChiliError):
            fl.set_mixture(bad)
    for bad in (1, 65):
        with pytest.raises(lib.ChiliError):
            fl.set_moment(bad)
    # the context still works after the errors
    fl.set_mixture(0.0)
    st, na = gpu.flow_iterate(fl, uni, 1000)
    assert na >= 1000 and st.n == na



def flow_generate(g, fl, uni, n, wmax, a0):
    ev = torch.empty(4 * g.ntot * n, dtype=torch.float64, device="cuda")
    ew = torch.empty(n, dtype=torch.float64, device="cuda")
    st, na = g.flow_generate(fl, uni, n, wmax, ctypes.c_void_p(ev.data_ptr()), ctypes.c_void_p(ew.data_ptr()), n, a0=a0)
    torch.cuda.synchronize()
    return st, na, ev.view(g.ntot * 4, n).cpu().numpy(), ew.cpu().numpy()


@pytest.mark.parametrize("mode", [0, 13])
def test_flow_generate_vs_records(mode):
    """The unweighting stage on the learned proposal (chili_mc_flow_generate)
    is its definition applied to the flow path's per-attempt records: an
    attempt that passes is kept when its uniform d + 1 is below |v| / wmax,
    the first n kept are the events, weights sign(v) max(1, |v| / wmax), and
    StatsData sums the kept values over all attempts used.  Mode 13 adds the
    exact early rejection to the specialised kernels."""
    _, _, _, g = adapted_context()
    if mode:
        g.specialize(mode)
    fl, _ = make_flow(g.ndims, 4, g.nchannels, 0.05, seed=31)
    uni = abi.philox_uniforms(SEED + 70, stream=4)
    a0, span = 5, 120000
    rec = g.flow_points(fl, uni, a0, span)
    ok = rec["status"] == 1
    wmax = float(np.quantile(np.abs(rec["val"][ok]), 0.95))
    u = np.array([lib.philox_uniform(uni.seed, uni.stream, a0 + i, g.ndims + 1) for i in range(span)])
    keep = ok & (u < np.abs(rec["val"]) / wmax)
    n = 2000
    idx = np.flatnonzero(keep)[:n]
    assert len(idx) == n, "record span too short"
    st, na, ev, ew = flow_generate(g, fl, uni, n, wmax, a0)
    assert na == idx[-1] + 1, f"attempts {na} vs {idx[-1] + 1}"
    v = rec["val"][idx]
    assert np.array_equal(ew, np.where(np.abs(v) > wmax, v / wmax, np.copysign(1.0, v)))
    assert st.n == na and st.n_finite == np.count_nonzero(v)
    assert_rel(np.array([st.sum, st.sum2]), np.array([v.sum(), (v * v).sum()]), 1e-12, what="stats")


def test_flow_generate_early_rejection_same_events():
    """With the exact early rejection the flow path weights in full only the
    points whose picked channel's term can pass the accept test; attempts,
    StatsData's counts, min and max, event momenta and weights equal the
    stage without it bit for bit, the sums to rounding (specialised kernels in
    both, so only the rejection differs)."""
    _, _, _, g = adapted_context("gg_4g")
    fl, _ = make_flow(g.ndims, 4, g.nchannels, 0.05, seed=37)
    uni = abi.philox_uniforms(SEED + 71, stream=6)
    probe = g.flow_points(fl, uni, 1 << 20, 20000)
    v = np.abs(probe["val"][probe["status"] == 1])
    for q, n in ((0.9, 3000), (0.995, 12000)):
        wmax = float(np.quantile(v, q))
        out = []
        for mode in (5, 13):
            g.specialize(mode)
            out.append(flow_generate(g, fl, uni, n, wmax, 17))
        (s0, n0, e0, w0), (s1, n1, e1, w1) = out
        assert n0 == n1, f"attempts {n0} vs {n1}"
        assert (s0.n, s0.min, s0.max, s0.n_finite) == (s1.n, s1.min, s1.max, s1.n_finite)
        # the sums are grouped by sub-batch, whose sizes follow a running
        # acceptance estimate that the previous call moved
        assert_rel(np.array([s0.sum, s0.sum2]), np.array([s1.sum, s1.sum2]), 1e-13, what="sums")
        assert np.array_equal(e0, e1) and np.array_equal(w0, w1)
