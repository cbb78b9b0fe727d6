"""CPU restatement of the learned proposal (coupling-layer rational-quadratic
spline flows on [0,1]^d) -- TEST INFRASTRUCTURE ONLY.

The proposal has no counterpart in the reference (it is new in this path);
the algorithm follows Durkan, Bekasov, Murray, Papamakarios (2019), "Neural
Spline Flows" (rational-quadratic spline, eq. 4 and the inverse of appendix
A.3), with coupling layers as in Dinh et al. (2017), "Density estimation using
Real NVP".  Only tests, the smoke check and the CPU-baseline leg of the
benchmark may use this module, and only as the checker.

Conventions restated here (the product computes the same):
  * parameters of one flow: a flat fp32 vector, per coupling layer
    W1[nA][H], b1[H], W2[H][H], b2[H], W3[H][H], b3[H], W4[H][out], b4[out],
    out = nB (3K-1); raw block of transformed dim b at columns b (3K-1):
    K width logits, K height logits, K-1 interior derivative inputs;
  * conditioner: xa = x[A]; h = relu(xa W1 + b1), two more ReLU layers,
    raw = h W4 + b4;
  * knots: X_i = i minW + (1-K minW) sum_{j<i} softmax_j, X_0 = 0, X_K = 1
    (in FP32 as fma((1-K minW)/S_K, S_i, i minW) of the sequential prefix
    sums S_i of the exponentials), bin search exact against the FP32 knots;
    transform and log-Jacobian in FP64 from those knots; derivatives d_0 = d_K = 1,
    d_i = minD + softplus_f32(u_i + shift), shift = log(expm1(1 - minD)) so
    that the zero network is the identity;
  * log q(r) = sum over layers and transformed dims of log dy/dx.

mode "exact": FP64 everywhere (for finite-difference and quadrature pins);
mode "gpu": operands of every product rounded to bf16 as the device rounds
them (straight through for gradients), knots in FP32.

Parity of this restatement is pinned by: identity at zero parameters,
monotone bijection of [0,1] onto itself, inverse(forward(x)) = x, log dy/dx
against central differences, normalisation of q by quadrature in two
dimensions, and autograd gradients against central differences of the loss.
"""
import math

import numpy as np
import torch

K_BINS = 16
HIDDEN = 64
MIN_W = 1e-3
MIN_H = 1e-3
MIN_D = 1e-3
SHIFT32 = float(np.float32(math.log(math.expm1(1.0 - MIN_D))))


def layer_shapes(masks, d, K=K_BINS, H=HIDDEN):
    """Per layer (A, B, offsets) for the flat parameter vector."""
    out = []
    off = 0
    for B in masks:
        B = sorted(B)
        A = [i for i in range(d) if i not in B]
        nA, nB = len(A), len(B)
        o = nB * (3 * K - 1)
        offs = {}
        for name, size in (("W1", nA * H), ("b1", H), ("W2", H * H), ("b2", H), ("W3", H * H), ("b3", H),
                           ("W4", H * o), ("b4", o)):
            offs[name] = (off, size)
            off += size
        out.append(dict(A=A, B=B, nA=nA, nB=nB, out=o, offs=offs))
    return out, off


def _bf16(t):
    # straight-through rounding: forward value bf16(float32(t)), identity gradient
    r = t.to(torch.float32).to(torch.bfloat16).to(t.dtype)
    return t + (r - t).detach()


def _f32(t):
    r = t.to(torch.float32).to(t.dtype)
    return t + (r - t).detach()


class FlowOracle:
    def __init__(self, d, masks, K=K_BINS, H=HIDDEN, mode="gpu"):
        self.d, self.K, self.H, self.mode = d, K, H, mode
        self.layers, self.nparams = layer_shapes(masks, d, K, H)

    # -- parameters -------------------------------------------------------
    def unpack(self, p):
        """p: torch float64 tensor [nparams] (values of the fp32 parameters)."""
        res = []
        for ly in self.layers:
            g = {}
            for name, (o, s) in ly["offs"].items():
                g[name] = p[o:o + s]
            nA, H, out = ly["nA"], self.H, ly["out"]
            g["W1"] = g["W1"].reshape(nA, H)
            g["W2"] = g["W2"].reshape(H, H)
            g["W3"] = g["W3"].reshape(H, H)
            g["W4"] = g["W4"].reshape(H, out)
            res.append(g)
        return res

    # -- conditioner --------------------------------------------------------
    def conditioner(self, g, xa):
        rnd = _bf16 if self.mode == "gpu" else (lambda t: t)
        h = rnd(xa)
        for (W, b) in (("W1", "b1"), ("W2", "b2"), ("W3", "b3")):
            h = torch.relu(h @ rnd(g[W]) + g[b])
            h = rnd(h)
        raw = h @ rnd(g["W4"]) + g["b4"]
        if self.mode == "gpu":
            raw = _f32(raw)
        return raw

    # -- spline --------------------------------------------------------------
    def knots(self, u):
        """u: [..., 3K-1] -> X, Y [..., K+1], d [..., K+1] (FP64 tensors).
        gpu mode: knots are FP32 sums (sequential, as the device adds them)."""
        K = self.K
        if self.mode == "gpu":
            u32 = u.to(torch.float32)
            cw = torch.tensor(1.0 - K * MIN_W, dtype=torch.float32)

            def axis(e):
                # S_i = sequential FP32 prefix sums of the exponentials,
                # a = (1 - K minW) / S_K in FP32, knot i = fma(a, S_i, i minW)
                # rounded once to FP32 (evaluated in FP64, where the product
                # is exact), knot 0 = 0, knot K = 1
                S = [torch.zeros_like(e[..., 0])]
                for i in range(K):
                    S.append(S[-1] + e[..., i])
                a = cw * (1.0 / S[K])
                out = [torch.zeros_like(e[..., 0])]
                for i in range(1, K):
                    c = float(np.float32(i * MIN_W))
                    out.append((a.to(torch.float64) * S[i].to(torch.float64) + c).to(torch.float32))
                out.append(torch.ones_like(e[..., 0]))
                return torch.stack(out, -1).to(torch.float64)

            ew = torch.exp(u32[..., :K] - u32[..., :K].max(-1, keepdim=True).values.detach())
            eh = torch.exp(u32[..., K:2 * K] - u32[..., K:2 * K].max(-1, keepdim=True).values.detach())
            X, Y = axis(ew), axis(eh)
            dz = torch.nn.functional.softplus(u32[..., 2 * K:] + SHIFT32).to(torch.float64)
            one = torch.ones_like(X[..., :1])
            return X, Y, torch.cat([one, MIN_D + dz, one], -1)
        shift = math.log(math.expm1(1.0 - MIN_D))
        sw = torch.softmax(u[..., :K], dim=-1)
        sh = torch.softmax(u[..., K:2 * K], dim=-1)
        dz = torch.nn.functional.softplus(u[..., 2 * K:] + shift)
        w = MIN_W + (1 - K * MIN_W) * sw
        h = MIN_H + (1 - K * MIN_H) * sh
        zero = torch.zeros_like(w[..., :1])
        one = torch.ones_like(w[..., :1])
        X = torch.cat([zero, torch.cumsum(w[..., :K - 1], -1), one], -1)
        Y = torch.cat([zero, torch.cumsum(h[..., :K - 1], -1), one], -1)
        dd = torch.cat([one, MIN_D + dz, one], -1)
        return X, Y, dd

    def rqs(self, u, x, inverse=False):
        """Rational-quadratic spline of x (FP64 [...]) with raw params u [..., 3K-1].
        Returns (y, log dy/dx) (inverse: (x, log dy/dx at x))."""
        K = self.K
        X, Y, dd = self.knots(u)
        v = x.clamp(0.0, 1.0)
        edges = Y if inverse else X
        k = (v.detach().unsqueeze(-1) >= edges[..., 1:K].detach()).sum(-1)
        k = k.unsqueeze(-1)
        X0 = X.gather(-1, k).squeeze(-1)
        Y0 = Y.gather(-1, k).squeeze(-1)
        w = X.gather(-1, k + 1).squeeze(-1) - X0
        h = Y.gather(-1, k + 1).squeeze(-1) - Y0
        d0 = dd.gather(-1, k).squeeze(-1)
        d1 = dd.gather(-1, k + 1).squeeze(-1)
        s = h / w
        if not inverse:
            xi = (v - X0) / w
        else:
            dy = v - Y0
            E = d1 + d0 - 2 * s
            a = h * (s - d0) + dy * E
            b = h * d0 - dy * E
            c = -s * dy
            disc = torch.clamp(b * b - 4 * a * c, min=0.0)
            xi = ((2 * c) / (-b - torch.sqrt(disc))).clamp(0.0, 1.0)
        t = xi * (1 - xi)
        den = s + (d1 + d0 - 2 * s) * t
        D = d1 * xi * xi + 2 * s * t + d0 * (1 - xi) * (1 - xi)
        ld = torch.log(s * s * D / (den * den))
        if not inverse:
            return Y0 + h * (s * xi * xi + d0 * t) / den, ld
        return X0 + xi * w, ld

    # -- flow ------------------------------------------------------------------
    def density(self, p, r):
        """p: float64 tensor [nparams]; r: float64 [n, d] -> (log q(r) [n], z [n, d])."""
        gs = self.unpack(p)
        x = r
        lq = torch.zeros(r.shape[0], dtype=torch.float64)
        K = self.K
        for ly, g in zip(self.layers, gs):
            raw = self.conditioner(g, x[:, ly["A"]])
            cols = []
            xn = x.clone()
            for bi, c in enumerate(ly["B"]):
                u = raw[:, bi * (3 * K - 1):(bi + 1) * (3 * K - 1)]
                y, ld = self.rqs(u, x[:, c])
                xn = xn.index_copy(1, torch.tensor([c]), y.unsqueeze(1))
                lq = lq + ld
            x = xn
        return lq, x

    def sample(self, p, z):
        """z -> r (inverse of density's map) and log q(r)."""
        gs = self.unpack(p)
        x = z.clone()
        lq = torch.zeros(z.shape[0], dtype=torch.float64)
        K = self.K
        for ly, g in reversed(list(zip(self.layers, gs))):
            raw = self.conditioner(g, x[:, ly["A"]])
            for bi, c in enumerate(ly["B"]):
                u = raw[:, bi * (3 * K - 1):(bi + 1) * (3 * K - 1)]
                xv, ld = self.rqs(u, x[:, c], inverse=True)
                x = x.index_copy(1, torch.tensor([c]), xv.unsqueeze(1))
                lq = lq + ld
        return x, lq


def variance_loss(oracles, params, r, beta, w2, mix=0.0, moment=2):
    """L = mean_i w2_i^(p/2) / ratio_i^(p-1), ratio_i = sum_j beta_ji q'_j(r_ji),
    q' = (1 - mix) exp(lq_j) + mix (the proposal blended with the uniform
    density), p = moment (2: the variance loss); r [C][n][d], beta [C][n]."""
    ratio = 0
    for j, (o, p) in enumerate(zip(oracles, params)):
        lq, _ = o.density(p, r[j])
        ratio = ratio + beta[j] * ((1.0 - mix) * torch.exp(lq) + mix)
    return torch.mean(w2 ** (0.5 * moment) / ratio ** (moment - 1))


def training_loss(oracles, params, r, beta, w2, mix=0.0, moment=2):
    """The objective the device trains: L_2 for the variance loss, (1/p) log L_p
    for a higher moment p (same minimiser, bounded gradients)."""
    L = variance_loss(oracles, params, r, beta, w2, mix, moment)
    return L if moment == 2 else torch.log(L) / moment


def loss_and_grads(oracles, params_np, r, beta, w2, mix=0.0, moment=2):
    """numpy in/out: params_np list of fp32 arrays; returns (loss, [grads])
    of training_loss."""
    ps = [torch.tensor(np.asarray(p, dtype=np.float64), requires_grad=True) for p in params_np]
    L = training_loss(oracles, ps, torch.as_tensor(r, dtype=torch.float64), torch.as_tensor(beta, dtype=torch.float64),
                      torch.as_tensor(w2, dtype=torch.float64), mix, moment)
    L.backward()
    return float(L), [p.grad.numpy().copy() for p in ps]
