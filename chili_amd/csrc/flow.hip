// flow.hip -- learned proposal on gfx950: per-channel coupling-layer flows with
// rational-quadratic splines over the VEGAS input hypercube.  MLPs run on
// hand-written BF16 MFMA GEMMs (FP32 accumulation); spline transforms,
// log-Jacobians and the training loss run in FP64.
//
// Density direction (r -> z): layer l keeps x[A_l], maps x[B_l] through a
// spline whose parameters an MLP computes from x[A_l]; log q(r) = sum of the
// spline log-derivatives.  Sampling direction (z -> r) inverts the layers in
// reverse order.  Training minimises the weighted-variance loss of the
// multichannel weight over samples drawn by the method itself:
//   L = mean_i w_i^2 / sum_j beta_ij q_j(r_ij),
// w_i the point's weight under the sampler that drew it, beta_ij = alpha_j
// g_j(p_i) / g(p_i) that sampler's channel shares and r_ij the point's
// coordinates in channel j's hypercube.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <random>
#include <vector>

#include "common.hpp"
#include "flow.hpp"
#include "flow_kernels.hpp"

using namespace chili_nf;

namespace {

constexpr int kSplineBins = 16;
constexpr int kTh = 256;

// ---------------------------------------------------------------------------
// small kernels
// ---------------------------------------------------------------------------
// xa[i][c] = bf16(x[i][A[c]]) (padded columns zero)
__global__ void k_gather_bf16(int64_t n, int d, const double* __restrict__ x, const int* __restrict__ A, int nA, int ld,
                              bf16* __restrict__ xa) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int c = 0; c < ld; ++c) xa[i * ld + c] = f2bf(c < nA ? (float)x[i * d + A[c]] : 0.f);
}

// Forward (density direction) spline of one layer, one thread per (point,
// transformed dim); A dims are copied by the b == 0 thread, log-Jacobians go
// to ldt[i][b] (summed per point in fixed order by k_add_rows).
template <int K>
__global__ void k_spline_fwd(int64_t n, int d, const double* __restrict__ x, const float* __restrict__ raw, int ldr,
                             const int* __restrict__ Bd, int nB, const int* __restrict__ Ad, int nA, float shift,
                             double* __restrict__ y, double* __restrict__ ldt) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * nB) return;
    const int64_t i = t / nB;
    const int b = (int)(t - i * nB);
    if (b == 0)
        for (int c = 0; c < nA; ++c) y[i * d + Ad[c]] = x[i * d + Ad[c]];
    double l;
    const int c = Bd[b];
    y[i * d + c] = rqs_forward<K>(raw + i * ldr + b * (3 * K - 1), shift, x[i * d + c], l);
    ldt[t] = l;
}

// lq[i] += sum_b ldt[i][b] (b in order)
__global__ void k_add_rows(int64_t n, int nB, const double* __restrict__ ldt, double* __restrict__ lq) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0;
    for (int b = 0; b < nB; ++b) s += ldt[i * nB + b];
    lq[i] += s;
}

// Spline backward, one thread per (point, transformed dim): gx holds dL/dy on
// entry and dL/dx on exit for the B dims; draw / drawb receive dL/draw (fp32
// for the bias sums, bf16 for the next product).
template <int K>
__global__ __launch_bounds__(kTh) void k_spline_bwd(int64_t n, int d, const double* __restrict__ x, const float* __restrict__ raw, int ldr,
                             const int* __restrict__ Bd, int nB, float shift, const double* __restrict__ glq,
                             double* __restrict__ gx, float* __restrict__ draw, bf16* __restrict__ drawb) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * nB) return;
    const int64_t i = t / nB;
    const int b = (int)(t - i * nB);
    const int c = Bd[b];
    float du[3 * K - 1];                 // registers: every index is a constant
    gx[i * d + c] = rqs_backward<K>(raw + i * ldr + b * (3 * K - 1), shift, x[i * d + c], gx[i * d + c], glq[i], du);
    float* dr = draw + i * ldr + b * (3 * K - 1);
    bf16* drb = drawb + i * ldr + b * (3 * K - 1);
#pragma unroll
    for (int q = 0; q < 3 * K - 1; ++q) {
        dr[q] = du[q];
        drb[q] = f2bf(du[q]);
    }
}

// gh (bf16) = relu'(h) * g (fp32)
__global__ void k_relu_mask(int64_t n, const float* g, const bf16* __restrict__ h, bf16* __restrict__ out, float* outf) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float v = bf2f(h[i]) > 0.f ? g[i] : 0.f;
    out[i] = f2bf(v);
    outf[i] = v;
}

// gx[i][A[c]] += dxa[i][c]
__global__ void k_scatter_add(int64_t n, int d, const float* __restrict__ dxa, int ld, const int* __restrict__ A, int nA,
                              double* __restrict__ gx) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int c = 0; c < nA; ++c) gx[i * d + A[c]] += (double)dxa[i * ld + c];
}

// Column sums of g [n x m] in fixed order: block (z, cb) sums rows
// [z chunk, (z+1) chunk) of columns cb*64 .. cb*64+63, four row phases per
// column combined in order.
__global__ __launch_bounds__(256) void k_colsum_part(int64_t n, int m, const float* __restrict__ g, int64_t chunk,
                                                     float* __restrict__ part) {
    __shared__ float s[4][64];
    const int c = blockIdx.y * 64 + (threadIdx.x & 63), ph = threadIdx.x >> 6;
    const int64_t r0 = (int64_t)blockIdx.x * chunk, r1 = std::min<int64_t>(n, r0 + chunk);
    float acc = 0;
    if (c < m)
        for (int64_t r = r0 + ph; r < r1; r += 4) acc += g[r * m + c];
    s[ph][threadIdx.x & 63] = acc;
    __syncthreads();
    if (ph == 0 && c < m) part[(int64_t)blockIdx.x * m + c] = ((s[0][threadIdx.x] + s[1][threadIdx.x]) + s[2][threadIdx.x]) + s[3][threadIdx.x];
}

// dst[k] += sum_z part[z][k] in a fixed order: four interleaved sums per
// element (z mod 4, one per wave of the workgroup), then combined in order.
__global__ __launch_bounds__(256) void k_sum_parts(int64_t cnt, int nz, const float* __restrict__ part, float* __restrict__ dst) {
    __shared__ float s[4][64];
    const int t = threadIdx.x & 63, ph = threadIdx.x >> 6;
    const int64_t k = (int64_t)blockIdx.x * 64 + t;
    float acc = 0;
    if (k < cnt)
        for (int z = ph; z < nz; z += 4) acc += part[(int64_t)z * cnt + k];
    s[ph][t] = acc;
    __syncthreads();
    if (ph == 0 && k < cnt) dst[k] += ((s[0][t] + s[1][t]) + s[2][t]) + s[3][t];
}

// Weighted-variance loss over C flows; loss partial per block; dlq[j][i].
// Loss: the p-th moment of the weights under the new proposal, estimated on
// the old one's samples, L = mean w^p / ratio^(p-1) with ratio =
// sum_j beta_j q'_j (w^2 = w2); with a mixture weight mix the proposal is
// q' = (1 - mix) q + mix.  p = 2 is the variance loss; larger p weighs the
// largest weights more.
__global__ void k_loss(int64_t n, int C, const double* __restrict__ lq, const double* __restrict__ beta,
                       const double* __restrict__ w2, double inv_n, double mix, int p, double* __restrict__ dlq,
                       double* __restrict__ lpart) {
    __shared__ double s[kTh];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double li = 0;
    if (i < n) {
        double ratio = 0;
        for (int j = 0; j < C; ++j) ratio += beta[(int64_t)j * n + i] * ((1.0 - mix) * exp(lq[(int64_t)j * n + i]) + mix);
        double wp = w2[i];                       // w^p
        for (int q = 4; q <= p; q += 2) wp *= w2[i];
        if (p & 1) wp *= sqrt(w2[i]);
        double rp = ratio;                       // ratio^(p-1)
        for (int q = 3; q <= p; ++q) rp *= ratio;
        li = wp / rp;
        for (int j = 0; j < C; ++j)
            dlq[(int64_t)j * n + i] =
                -(p - 1) * inv_n * wp * beta[(int64_t)j * n + i] * (1.0 - mix) * exp(lq[(int64_t)j * n + i]) / (rp * ratio);
    }
    s[threadIdx.x] = li;
    __syncthreads();
    for (int off = kTh / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) lpart[blockIdx.x] = s[0];
}

// Higher moments (p > 2) in log space: the objective is J = (1/p) log L_p
// (same minimiser; its gradient dL_p / (p L_p) stays bounded where L_p's own
// would overflow FP32).  Per point lt_i = log w_i^p - (p - 1) log ratio_i;
// with m = max_i lt_i and s_i = exp(lt_i - m), S = sum_i s_i:
// J = (m + log(S / n)) / p and dJ/dlq_j(i) = -(p - 1) / p (s_i / S)
// beta_ji (1 - mix) q_j / ratio_i.  Three passes: terms and block maxima;
// block sums of s_i; gradients (every block combines the partials in the
// same fixed order).
__global__ void k_loss_terms(int64_t n, int C, const double* __restrict__ lq, const double* __restrict__ beta,
                             const double* __restrict__ w2, double mix, int p, double* __restrict__ lt, double* __restrict__ rat,
                             double* __restrict__ bmax) {
    __shared__ double s[kTh];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double v = -INFINITY;
    if (i < n) {
        double ratio = 0;
        for (int j = 0; j < C; ++j) ratio += beta[(int64_t)j * n + i] * ((1.0 - mix) * exp(lq[(int64_t)j * n + i]) + mix);
        v = (w2[i] > 0.0) ? 0.5 * p * log(w2[i]) - (p - 1) * log(ratio) : -INFINITY;
        lt[i] = v;
        rat[i] = ratio;
    }
    s[threadIdx.x] = v;
    __syncthreads();
    for (int off = kTh / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) s[threadIdx.x] = fmax(s[threadIdx.x], s[threadIdx.x + off]);
        __syncthreads();
    }
    if (threadIdx.x == 0) bmax[blockIdx.x] = s[0];
}

__device__ __forceinline__ double loss_max(const double* __restrict__ bmax, int nbk) {
    double m = -INFINITY;
    for (int b = 0; b < nbk; ++b) m = fmax(m, bmax[b]);
    return m;
}

__global__ void k_loss_sums(int64_t n, int nbk, const double* __restrict__ lt, const double* __restrict__ bmax,
                            double* __restrict__ bsum) {
    __shared__ double s[kTh];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const double m = loss_max(bmax, nbk);
    s[threadIdx.x] = (i < n && lt[i] > -INFINITY) ? exp(lt[i] - m) : 0.0;
    __syncthreads();
    for (int off = kTh / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) bsum[blockIdx.x] = s[0];
}

__global__ void k_loss_grad_log(int64_t n, int C, const double* __restrict__ lq, const double* __restrict__ beta, double mix, int p,
                                const double* __restrict__ lt, const double* __restrict__ rat, const double* __restrict__ bmax,
                                const double* __restrict__ bsum, int nbk, double* __restrict__ dlq, double* __restrict__ lpart) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const double m = loss_max(bmax, nbk);
    double S = 0;
    for (int b = 0; b < nbk; ++b) S += bsum[b];
    if (i < n) {
        const double c = (S > 0.0 && lt[i] > -INFINITY) ? -(double)(p - 1) / p * exp(lt[i] - m) / S * (1.0 - mix) / rat[i] : 0.0;
        for (int j = 0; j < C; ++j) dlq[(int64_t)j * n + i] = c * beta[(int64_t)j * n + i] * exp(lq[(int64_t)j * n + i]);
    }
    // the host sums lpart over blocks and divides by n
    if (threadIdx.x == 0) lpart[blockIdx.x] = (blockIdx.x == 0) ? (m + log(S / (double)n)) / p * (double)n : 0.0;
}

__global__ void k_adam(int64_t n, float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       float lr, float b1, float b2, float eps, float c1, float c2) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1 - b1) * gi;
    const float vi = b2 * v[i] + (1 - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// bf16 copies of a weight W [in x out]: Wb (same layout) and Wt [out x in]
__global__ void k_weights_bf16(int in, int out, const float* __restrict__ W, bf16* __restrict__ Wb, bf16* __restrict__ Wt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= in * out) return;
    const int r = i / out, c = i % out;
    Wb[i] = f2bf(W[i]);
    Wt[c * in + r] = f2bf(W[i]);
}

// Swizzle one flow's fp32 parameters (layer l: W1[nA][H] b1 W2[H][H] b2 W3 b3
// W4[H][out] b4[out]) into the fused layout.  One thread per output element.
__global__ void k_swizzle(const float* __restrict__ p, int64_t pW1, int64_t pb1, int64_t pW2, int64_t pb2, int64_t pW3, int64_t pb3,
                          int64_t pW4, int64_t pb4, int nA, int nB, bf16* __restrict__ w, float* __restrict__ b) {
    const FusedLayout FL = fused_layout(nB);
    const int out = nB * kRaw;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < FL.wel) {
        const int jj = (int)(i & 7), lane = (int)((i >> 3) & 63), h = lane >> 5;
        const int64_t frag = i >> 9;            // 512 elements per fragment block
        float v = 0.f;
        if (i < FL.oW2) {                       // W1: [to][lane][8], natural k
            const int to = (int)frag, rowo = 32 * to + (lane & 31), k = 8 * h + jj;
            if (k < nA) v = p[pW1 + (int64_t)k * kHid + rowo];
        } else if (i < FL.oW4) {                // W2 / W3: [to][s][lane][8]
            const bool second = i >= FL.oW3;
            const int64_t f = (i - (second ? FL.oW3 : FL.oW2)) >> 9;
            const int to = (int)(f / 4), s = (int)(f % 4), rowo = 32 * to + (lane & 31), k = phi_k(s, h, jj);
            v = p[(second ? pW3 : pW2) + (int64_t)k * kHid + rowo];
        } else {                                // W4: [3 npair tiles][s][lane][8]
            const int64_t f = (i - FL.oW4) >> 9;
            const int T = (int)(f / 4), s = (int)(f % 4), P = T / 3, tt = T % 3;
            const int fr = lane & 31, ds = (fr >> 2) & 1, q = 16 * tt + 4 * (fr >> 3) + (fr & 3), dim = 2 * P + ds;
            if (q < kRaw && dim < nB) v = p[pW4 + (int64_t)phi_k(s, h, jj) * out + dim * kRaw + q];
        }
        w[i] = f2bf(v);
    }
    if (i < FL.bel) {
        float v = 0.f;
        if (i < 64) v = p[pb1 + i];
        else if (i < 128) v = p[pb2 + i - 64];
        else if (i < 192) v = p[pb3 + i - 128];
        else {
            const int64_t k = i - 192;
            const int P = (int)(k / kPairRows), rp = (int)(k % kPairRows), ds = rp / 48, q = rp % 48, dim = 2 * P + ds;
            if (q < kRaw && dim < nB) v = p[pb4 + dim * kRaw + q];
        }
        b[i] = v;
    }
}

// Work-item table of a fused pass: flow j's rows start at start[j] (seg[j],
// or j * stride) and number count[j] (seg[j+1] - seg[j], or ucount); items[]
// prefix-sums ceil(count / chunk).
__global__ void k_items(int nfl, const int* __restrict__ seg, const int* __restrict__ seg_end, int64_t ucount, int64_t stride,
                        int64_t chunk, int64_t* start, int64_t* count, int* items, int64_t* rows_out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    int acc = 0;
    int64_t rows = 0;
    for (int j = 0; j < nfl; ++j) {
        const int64_t s0 = seg ? (int64_t)seg[j] : (int64_t)j * stride;
        const int64_t c = seg ? (int64_t)(seg_end ? seg_end[j] : seg[j + 1]) - seg[j] : ucount;
        start[j] = s0;
        count[j] = c;
        items[j] = acc;
        acc += (int)((c + chunk - 1) / chunk);
        rows += c;
    }
    items[nfl] = acc;
    if (rows_out) *rows_out = rows;
}

// out[i] = in[perm[i]] rows of width w (fp64)
__global__ void k_gather_rows(int64_t n, int w, const double* __restrict__ in, const int64_t* __restrict__ perm, int64_t base,
                              double* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t src = perm[base + i];
    for (int c = 0; c < w; ++c) out[i * w + c] = in[src * w + c];
}

inline unsigned nblk(int64_t n, int t = kTh) { return (unsigned)((n + t - 1) / t); }

}  // namespace

// ===========================================================================
// host engine
// ===========================================================================
namespace chili_nf {

void FlowEngine::build_layers(uint64_t seed) {
    // Coupling masks: pairs of layers transform a seeded random half of the
    // dimensions and then its complement, so every dimension is transformed
    // (and conditions the others) in every pair of layers.
    std::mt19937_64 rng(seed);
    layers.clear();
    std::vector<int> half;
    for (int l = 0; l < L; ++l) {
        std::vector<int> B;
        if (l % 2 == 0) {
            std::vector<int> perm(d);
            for (int i = 0; i < d; ++i) perm[i] = i;
            std::shuffle(perm.begin(), perm.end(), rng);
            half.assign(perm.begin(), perm.begin() + d / 2);
            std::sort(half.begin(), half.end());
            B = half;
        } else {
            for (int i = 0; i < d; ++i)
                if (!std::binary_search(half.begin(), half.end(), i)) B.push_back(i);
        }
        Layer ly;
        ly.B = B;
        for (int i = 0; i < d; ++i)
            if (!std::binary_search(B.begin(), B.end(), i)) ly.A.push_back(i);
        ly.nA = (int)ly.A.size();
        ly.nB = (int)ly.B.size();
        ly.out = ly.nB * (3 * kSplineBins - 1);
        ly.ldA = ((ly.nA + 7) / 8) * 8;
        layers.push_back(std::move(ly));
    }
    nparams = 0;
    for (auto& ly : layers) {
        ly.oW1 = nparams; nparams += (size_t)ly.nA * H;
        ly.ob1 = nparams; nparams += H;
        ly.oW2 = nparams; nparams += (size_t)H * H;
        ly.ob2 = nparams; nparams += H;
        ly.oW3 = nparams; nparams += (size_t)H * H;
        ly.ob3 = nparams; nparams += H;
        ly.oW4 = nparams; nparams += (size_t)H * ly.out;
        ly.ob4 = nparams; nparams += ly.out;
    }
    for (auto& ly : layers) {
        std::vector<int> a = ly.A, b = ly.B;
        ly.dA.upload(a, stream);
        ly.dB.upload(b, stream);
    }
    fwstride = 0;
    fbstride = 0;
    for (auto& ly : layers) {
        const FusedLayout FL = fused_layout(ly.nB);
        fwstride = std::max<int64_t>(fwstride, ((FL.wel + 7) / 8) * 8);
        fbstride = std::max<int64_t>(fbstride, ((FL.bel + 3) / 4) * 4);
    }
}

void FlowEngine::init_params(uint64_t seed) {
    std::mt19937_64 rng(seed ^ 0x9E3779B97F4A7C15ull);
    for (int f = 0; f < nflows; ++f) {
        std::vector<float> p(nparams, 0.f);
        for (auto& ly : layers) {
            auto fill = [&](size_t off, int in, int out) {
                const double bnd = 1.0 / std::sqrt((double)in);
                std::uniform_real_distribution<double> u(-bnd, bnd);
                for (size_t k = 0; k < (size_t)in * out; ++k) p[off + k] = (float)u(rng);
            };
            fill(ly.oW1, ly.nA, H);
            fill(ly.oW2, H, H);
            fill(ly.oW3, H, H);
            // last layer zero: every spline starts as the identity
        }
        flows[f].params.upload(p, stream);
        std::vector<float> z(nparams, 0.f);
        flows[f].grads.upload(z, stream);
        flows[f].m.upload(z, stream);
        flows[f].v.upload(z, stream);
        flows[f].steps = 0;
        refresh_bf16(f);
    }
}

void FlowEngine::refresh_bf16(int f) {
    Flow& F = flows[f];
    fusedW.resize((size_t)nflows * L * fwstride);
    fusedB.resize((size_t)nflows * L * fbstride);
    for (int l = 0; l < L; ++l) {
        const Layer& ly = layers[l];
        const FusedLayout FL = fused_layout(ly.nB);
        const int64_t cnt = std::max(FL.wel, FL.bel);
        const int64_t blk = (int64_t)f * L + l;
        hipLaunchKernelGGL(k_swizzle, dim3(nblk(cnt)), dim3(kTh), 0, stream, F.params.ptr, (int64_t)ly.oW1, (int64_t)ly.ob1,
                           (int64_t)ly.oW2, (int64_t)ly.ob2, (int64_t)ly.oW3, (int64_t)ly.ob3, (int64_t)ly.oW4, (int64_t)ly.ob4, ly.nA,
                           ly.nB, fusedW.ptr + blk * fwstride, fusedB.ptr + blk * fbstride);
    }
    size_t wt = 0;
    for (auto& ly : layers) wt += (size_t)ly.nA * H + 2 * (size_t)H * H + (size_t)H * ly.out;
    F.wb.resize(wt);
    F.wt.resize(wt);
    size_t o = 0;
    for (auto& ly : layers) {
        const int ins[4] = {ly.nA, H, H, H}, outs[4] = {H, H, H, ly.out};
        const size_t offs[4] = {ly.oW1, ly.oW2, ly.oW3, ly.oW4};
        for (int q = 0; q < 4; ++q) {
            ly.wo[q] = o;
            hipLaunchKernelGGL(k_weights_bf16, dim3(nblk((int64_t)ins[q] * outs[q])), dim3(kTh), 0, stream, ins[q], outs[q],
                               F.params.ptr + offs[q], F.wb.ptr + o, F.wt.ptr + o);
            o += (size_t)ins[q] * outs[q];
        }
    }
}

void FlowEngine::ensure_ws(int64_t nb, bool keep) {
    // keep: activations of every layer stay for the backward pass; otherwise
    // the layers share one slot (and two for x)
    const size_t Ls = keep ? (size_t)L : 1;
    const size_t Lx = keep ? (size_t)L + 1 : 2;
    const int maxout = ws_out(), maxA = ws_lda();
    ws.x.resize(Lx * nb * d);
    ws.xa.resize(Ls * nb * maxA);
    ws.h.resize(Ls * 3 * nb * H);
    ws.raw.resize(Ls * nb * maxout);
    if (keep) {
        ws.draw.resize((size_t)nb * maxout);
        ws.drawb.resize((size_t)nb * maxout);
        ws.gf.resize((size_t)nb * std::max(H, maxout));
        ws.gb.resize((size_t)nb * std::max(H, maxout));
        ws.gx.resize((size_t)nb * d);
        ws.dxa.resize((size_t)nb * maxA);
    }
}

// GEMM launch helpers ------------------------------------------------------
void FlowEngine::gemm_fwd(int64_t M, int N, int K, const bf16* A, int lda, const bf16* Bt, int ldb, const float* bias, int relu,
                          float* C, int ldc, bf16* Cb, int ldcb) {
    dim3 g((N + kGT - 1) / kGT, (unsigned)((M + kGT - 1) / kGT), 1);
    hipLaunchKernelGGL((k_gemm<false, true>), g, dim3(kGThreads), 0, stream, (int)M, N, K, A, lda, Bt, ldb, C, ldc, bias, relu, Cb,
                       ldcb, K, (int64_t)0);
}
// dW [in x out] += A^T G with A [n x in] (bf16), G [n x out] (bf16), split over n
void FlowEngine::gemm_wgrad(int64_t n, int in, int out, const bf16* A, int lda, const bf16* G, int ldg, float* dW) {
    const int chunk = 512;     // points per split-K partial (a multiple of the 32-point K step)
    const int nz = (int)((n + chunk - 1) / chunk);
    part.resize((size_t)nz * in * out);
    dim3 g((out + kGT - 1) / kGT, (in + kGT - 1) / kGT, nz);
    hipLaunchKernelGGL((k_gemm<true, false>), g, dim3(kGThreads), 0, stream, in, out, (int)n, A, lda, G, ldg, part.ptr, out,
                       (const float*)nullptr, 0, (bf16*)nullptr, 0, chunk, (int64_t)in * out);
    hipLaunchKernelGGL(k_sum_parts, dim3(nblk((int64_t)in * out, 64)), dim3(256), 0, stream, (int64_t)in * out, nz, part.ptr, dW);
}
void FlowEngine::colsum(int64_t n, int m, const float* g, float* db) {
    const int64_t chunk = 256;
    const int nz = (int)((n + chunk - 1) / chunk);
    part2.resize((size_t)nz * m);
    hipLaunchKernelGGL(k_colsum_part, dim3(nz, (m + 63) / 64), dim3(256), 0, stream, n, m, g, chunk, part2.ptr);
    hipLaunchKernelGGL(k_sum_parts, dim3(nblk(m, 64)), dim3(256), 0, stream, (int64_t)m, nz, part2.ptr, db);
}

// Density direction over nb points: x0 = r (fp64 [nb x d]) -> z = x_L; lq.
// keep = true stores every layer's activations for backward().
void FlowEngine::density(int f, int64_t nb, const double* r, double* lq, double* zout, bool keep) {
    ensure_ws(nb, keep);
    Flow& F = flows[f];
    HIP_CHECK(hipMemcpyAsync(ws.x.ptr, r, (size_t)nb * d * sizeof(double), hipMemcpyDeviceToDevice, stream));
    hipLaunchKernelGGL(k_fill<double>, dim3(nblk(nb)), dim3(kTh), 0, stream, nb, 0.0, lq);
    const int lda = ws_lda(), ldo = ws_out();
    for (int l = 0; l < L; ++l) {
        Layer& ly = layers[l];
        const size_t sl = keep ? (size_t)l : 0;
        double* xl = ws.x.ptr + (keep ? (size_t)l : (size_t)(l & 1)) * nb * d;
        double* xn = ws.x.ptr + (keep ? (size_t)l + 1 : (size_t)((l + 1) & 1)) * nb * d;
        bf16* xa = ws.xa.ptr + sl * nb * lda;
        bf16* h = ws.h.ptr + sl * 3 * nb * H;
        float* raw = ws.raw.ptr + sl * nb * ldo;
        mlp_forward(F, ly, nb, xl, xa, h, raw);
        ws.ldt.resize((size_t)nb * ly.nB);
        hipLaunchKernelGGL(k_spline_fwd<kSplineBins>, dim3(nblk(nb * ly.nB)), dim3(kTh), 0, stream, nb, d, xl, raw, ly.out, ly.dB.ptr,
                           ly.nB, ly.dA.ptr, ly.nA, shift, xn, ws.ldt.ptr);
        hipLaunchKernelGGL(k_add_rows, dim3(nblk(nb)), dim3(kTh), 0, stream, nb, ly.nB, ws.ldt.ptr, lq);
    }
    const double* zl = ws.x.ptr + (keep ? (size_t)L : (size_t)(L & 1)) * nb * d;
    if (zout) HIP_CHECK(hipMemcpyAsync(zout, zl, (size_t)nb * d * sizeof(double), hipMemcpyDeviceToDevice, stream));
}

// Conditioner of one coupling layer: xa = bf16(x[A]); three ReLU layers; raw
// spline parameters (fp32).
void FlowEngine::mlp_forward(Flow& F, Layer& ly, int64_t nb, const double* x, bf16* xa, bf16* h, float* raw) {
    hipLaunchKernelGGL(k_gather_bf16, dim3(nblk(nb)), dim3(kTh), 0, stream, nb, d, x, ly.dA.ptr, ly.nA, ly.ldA, xa);
    gemm_fwd(nb, H, ly.nA, xa, ly.ldA, F.wt.ptr + ly.wo[0], ly.nA, F.params.ptr + ly.ob1, 1, nullptr, 0, h, H);
    gemm_fwd(nb, H, H, h, H, F.wt.ptr + ly.wo[1], H, F.params.ptr + ly.ob2, 1, nullptr, 0, h + (size_t)nb * H, H);
    gemm_fwd(nb, H, H, h + (size_t)nb * H, H, F.wt.ptr + ly.wo[2], H, F.params.ptr + ly.ob3, 1, nullptr, 0, h + 2 * (size_t)nb * H, H);
    gemm_fwd(nb, ly.out, H, h + 2 * (size_t)nb * H, H, F.wt.ptr + ly.wo[3], H, F.params.ptr + ly.ob4, 0, raw, ly.out, nullptr, 0);
}

// Backward of density(): dlq (fp64 [nb]) -> gradients added to flow f's grads.
void FlowEngine::backward(int f, int64_t nb, const double* dlq) {
    Flow& F = flows[f];
    HIP_CHECK(hipMemsetAsync(ws.gx.ptr, 0, (size_t)nb * d * sizeof(double), stream));
    for (int l = L - 1; l >= 0; --l) {
        Layer& ly = layers[l];
        double* xl = ws.x.ptr + (size_t)l * nb * d;
        bf16* xa = ws.xa.ptr + (size_t)l * nb * ws_lda();
        bf16* h = ws.h.ptr + (size_t)l * 3 * nb * H;
        float* raw = ws.raw.ptr + (size_t)l * nb * ws_out();
        hipLaunchKernelGGL(k_spline_bwd<kSplineBins>, dim3(nblk(nb * ly.nB)), dim3(kTh), 0, stream, nb, d, xl, raw, ly.out, ly.dB.ptr, ly.nB,
                           shift, dlq, ws.gx.ptr, ws.draw.ptr, ws.drawb.ptr);
        // last linear layer
        gemm_wgrad(nb, H, ly.out, h + 2 * (size_t)nb * H, H, ws.drawb.ptr, ly.out, F.grads.ptr + ly.oW4);
        colsum(nb, ly.out, ws.draw.ptr, F.grads.ptr + ly.ob4);
        const bf16* gin = ws.drawb.ptr;
        int gin_ld = ly.out, gin_w = ly.out;
        const size_t wq[3] = {ly.wo[3], ly.wo[2], ly.wo[1]};
        const size_t ow[3] = {ly.oW3, ly.oW2, ly.oW1};
        const size_t ob[3] = {ly.ob3, ly.ob2, ly.ob1};
        for (int q = 0; q < 3; ++q) {
            // g_h = (gin W^T) * relu'(h), h = hidden activation 2 - q
            const bf16* hq = h + (size_t)(2 - q) * nb * H;
            dim3 g((H + kGT - 1) / kGT, (unsigned)((nb + kGT - 1) / kGT), 1);
            hipLaunchKernelGGL((k_gemm<false, true>), g, dim3(kGThreads), 0, stream, (int)nb, H, gin_w, gin, gin_ld,
                               F.wb.ptr + wq[q], gin_w, ws.gf.ptr, H, (const float*)nullptr, 0, (bf16*)nullptr, 0, gin_w, (int64_t)0);
            hipLaunchKernelGGL(k_relu_mask, dim3(nblk(nb * H)), dim3(kTh), 0, stream, nb * H, ws.gf.ptr, hq, ws.gb.ptr, ws.gf.ptr);
            // weight / bias grads of the layer feeding h_{2-q}
            const bf16* ain = (q < 2) ? h + (size_t)(1 - q) * nb * H : xa;
            const int ain_w = (q < 2) ? H : ly.nA, ain_ld = (q < 2) ? H : ly.ldA;
            gemm_wgrad(nb, ain_w, H, ain, ain_ld, ws.gb.ptr, H, F.grads.ptr + ow[q]);
            colsum(nb, H, ws.gf.ptr, F.grads.ptr + ob[q]);
            gin = ws.gb.ptr;
            gin_ld = H;
            gin_w = H;
        }
        // input gradient of the first linear layer: dxa = g_h1 W1^T
        {
            dim3 g((ly.nA + kGT - 1) / kGT, (unsigned)((nb + kGT - 1) / kGT), 1);
            hipLaunchKernelGGL((k_gemm<false, true>), g, dim3(kGThreads), 0, stream, (int)nb, ly.nA, H, ws.gb.ptr, H, F.wb.ptr + ly.wo[0],
                               H, ws.dxa.ptr, ly.ldA, (const float*)nullptr, 0, (bf16*)nullptr, 0, H, (int64_t)0);
            hipLaunchKernelGGL(k_scatter_add, dim3(nblk(nb)), dim3(kTh), 0, stream, nb, d, ws.dxa.ptr, ly.ldA, ly.dA.ptr, ly.nA, ws.gx.ptr);
        }
    }
}

// Waves per workgroup: 8, or 4 where that keeps more waves resident per CU
// given the kernel's registers and this layer's LDS (the weights are loaded
// once per workgroup, so larger groups are preferred on ties; multiples of
// four waves spread evenly over a CU's four SIMDs).
// The choice depends only on the kernel and the LDS size, so it is made once
// per (device, LDS size) and kept (the attribute calls cost host time on
// every one of a step's many launches otherwise).
template <bool INV, int NP>
static void launch_couple_np(dim3 g, size_t lds, hipStream_t st, const CoupleArgs& a) {
    struct Shape { int device; size_t lds; int nw; };
    static thread_local std::vector<Shape> known;
    int dev = 0;
    HIP_CHECK(hipGetDevice(&dev));
    int nw = 0;
    for (const Shape& k : known)
        if (k.device == dev && k.lds == lds) nw = k.nw;
    if (nw == 0) {
        const void* fn = (const void*)k_couple<kSplineBins, INV, NP>;
        HIP_CHECK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        int best = 0;
        nw = 8;
        for (int w : {8, 4}) {
            int nb = 0;
            HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, 64 * w, lds));
            if (nb * w > best) {
                best = nb * w;
                nw = w;
            }
        }
        if (best == 0) throw chili_error(CHILI_ERANGE, "coupling layer does not fit one CU");
        known.push_back(Shape{dev, lds, nw});
    }
    hipLaunchKernelGGL((k_couple<kSplineBins, INV, NP>), g, dim3(64 * nw), lds, st, a);
}

void FlowEngine::launch_couple(bool inverse, int npair, dim3 g, size_t lds, const CoupleArgs& a) {
#define CHILI_COUPLE_CASE(NPV)                                                        \
    case NPV:                                                                         \
        if (inverse) launch_couple_np<true, NPV>(g, lds, stream, a);                 \
        else launch_couple_np<false, NPV>(g, lds, stream, a);                        \
        break;
    switch (npair) {
        CHILI_COUPLE_CASE(1)
        CHILI_COUPLE_CASE(2)
        CHILI_COUPLE_CASE(3)
        CHILI_COUPLE_CASE(4)
        CHILI_COUPLE_CASE(5)
        CHILI_COUPLE_CASE(6)
        default: throw chili_error(CHILI_ERANGE, "too many transformed dimensions per coupling layer");
    }
#undef CHILI_COUPLE_CASE
}

void FlowEngine::fused_pass(bool inverse, int nfl, int flow0, const int* seg, int64_t ucount, int64_t stride, int64_t rows_bound,
                            double* x, double* lq, const int* seg_end, int64_t rows_hint) {
    d_start.resize(nfl);
    d_count.resize(nfl);
    d_items.resize(nfl + 1);
    // rows per work item: large enough to amortise the workgroup's weight
    // load, small enough to give every CU several items at the expected row
    // count (rows_hint, e.g. the flow component of a mixture; the grid still
    // covers rows_bound)
    const int64_t rows_exp = (rows_hint > 0) ? std::min(rows_hint, rows_bound) : rows_bound;
    int64_t chunk = ((rows_exp / 1024 + 255) / 256) * 256;
    chunk = std::max<int64_t>(256, std::min<int64_t>(kChunk, chunk));
    int slot = -1;
    if (timing) {
        if (tslots == kTimingSlots) harvest_timing();
        if (!d_prows.ptr) d_prows.resize(kTimingSlots);
        slot = tslots++;
    }
    hipLaunchKernelGGL(k_items, dim3(1), dim3(1), 0, stream, nfl, seg, seg_end, ucount, stride, chunk, d_start.ptr, d_count.ptr,
                       d_items.ptr, slot >= 0 ? d_prows.ptr + slot : (int64_t*)nullptr);
    const unsigned grid = (unsigned)(rows_bound / chunk + nfl + 1);
    for (int q = 0; q < L; ++q) {
        const int l = inverse ? L - 1 - q : q;
        const Layer& ly = layers[l];
        const FusedLayout FL = fused_layout(ly.nB);
        const size_t wbytes = ((FL.wel * 2 + 15) / 16) * 16, bbytes = ((FL.bel + 3) / 4) * 16;
        const size_t lds = wbytes + bbytes;
        if (lds > 163840) throw chili_error(CHILI_ERANGE, "coupling layer does not fit the LDS of one workgroup");
        CoupleArgs a{};
        a.x = x;
        a.lq = lq;
        a.start = d_start.ptr;
        a.count = d_count.ptr;
        a.items = d_items.ptr;
        a.nflows = nfl;
        a.flow0 = flow0;
        a.chunk = chunk;
        a.w = fusedW.ptr;
        a.b = fusedB.ptr;
        a.wstride = fwstride;
        a.bstride = fbstride;
        a.L = L;
        a.layer = l;
        a.d = d;
        a.nA = ly.nA;
        a.nB = ly.nB;
        for (int i = 0; i < 16; ++i) {
            a.A[i] = i < ly.nA ? ly.A[i] : 0;
            a.B[i] = i < ly.nB ? ly.B[i] : 0;
        }
        a.shift = shift;
        if (slot >= 0) {
            CoupleEv e{};
            if (!tpool.empty()) {
                e.a = tpool.back().first;
                e.b = tpool.back().second;
                tpool.pop_back();
            } else {
                HIP_CHECK(hipEventCreate(&e.a));
                HIP_CHECK(hipEventCreate(&e.b));
            }
            e.slot = slot;
            // algorithmic BF16 work of the conditioner per row: nA -> 64 -> 64
            // -> 64 -> nB (3K - 1) spline parameters, 2 flops per multiply-add
            e.flops_row = 2.0 * ((double)ly.nA * H + 2.0 * H * H + (double)H * ly.nB * (3 * kSplineBins - 1));
            HIP_CHECK(hipEventRecord(e.a, stream));
            launch_couple(inverse, FL.npair, dim3(grid), lds, a);
            HIP_CHECK(hipEventRecord(e.b, stream));
            tev.push_back(e);
        } else {
            launch_couple(inverse, FL.npair, dim3(grid), lds, a);
        }
    }
}

// Sums the timed launches (synchronises the device first: timed passes may
// have run on several streams).
void FlowEngine::harvest_timing() {
    if (tev.empty() && tslots == 0) return;
    HIP_CHECK(hipDeviceSynchronize());
    std::vector<int64_t> rows((size_t)tslots);
    if (tslots) HIP_CHECK(hipMemcpy(rows.data(), d_prows.ptr, sizeof(int64_t) * (size_t)tslots, hipMemcpyDeviceToHost));
    std::vector<char> counted((size_t)tslots, 0);
    for (auto& e : tev) {
        float ms = 0;
        HIP_CHECK(hipEventElapsedTime(&ms, e.a, e.b));
        couple_ms += ms;
        couple_launches += 1;
        couple_flops += e.flops_row * (double)rows[(size_t)e.slot];
        if (!counted[(size_t)e.slot]) {
            counted[(size_t)e.slot] = 1;
            couple_rows += rows[(size_t)e.slot];
        }
        tpool.emplace_back(e.a, e.b);
    }
    tev.clear();
    tslots = 0;
}

void FlowEngine::release_timing() {
    for (auto& e : tev) { (void)hipEventDestroy(e.a); (void)hipEventDestroy(e.b); }
    for (auto& e : tpool) { (void)hipEventDestroy(e.first); (void)hipEventDestroy(e.second); }
    tev.clear();
    tpool.clear();
}

// log q_f(r) of n points (device [n][d]); z (optional) receives the image.
void FlowEngine::log_density(int f, int64_t n, const double* r, double* lq, double* z) {
    double* x = z;
    if (!x) {
        xwork.resize((size_t)n * d);
        x = xwork.ptr;
    }
    HIP_CHECK(hipMemcpyAsync(x, r, (size_t)n * d * sizeof(double), hipMemcpyDeviceToDevice, stream));
    hipLaunchKernelGGL(k_fill<double>, dim3(nblk(n)), dim3(kTh), 0, stream, n, 0.0, lq);
    fused_pass(false, 1, f, nullptr, n, 0, n, x, lq);
}

// z -> r in place for n points of flow f; lq = log q_f(r).
void FlowEngine::sample(int f, int64_t n, double* zr, double* lq) {
    hipLaunchKernelGGL(k_fill<double>, dim3(nblk(n)), dim3(kTh), 0, stream, n, 0.0, lq);
    fused_pass(true, 1, f, nullptr, n, 0, n, zr, lq);
}

double FlowEngine::loss_grad(int64_t nb, const double* r, const double* beta, const double* w2) {
    // r: [C][nb][d], beta: [C][nb], w2: [nb]
    lqb.resize((size_t)nflows * nb);
    dlqb.resize((size_t)nflows * nb);
    const unsigned nbk = nblk(nb);
    lpart.resize(nbk);
    for (int f = 0; f < nflows; ++f) {
        HIP_CHECK(hipMemsetAsync(flows[f].grads.ptr, 0, nparams * sizeof(float), stream));
    }
    // the loss needs every flow's log-density first (fused pass over all
    // flows); the backward pass then recomputes one flow at a time, keeping
    // its activations
    xwork.resize((size_t)nflows * nb * d);
    HIP_CHECK(hipMemcpyAsync(xwork.ptr, r, (size_t)nflows * nb * d * sizeof(double), hipMemcpyDeviceToDevice, stream));
    hipLaunchKernelGGL(k_fill<double>, dim3(nblk((int64_t)nflows * nb)), dim3(kTh), 0, stream, (int64_t)nflows * nb, 0.0, lqb.ptr);
    fused_pass(false, nflows, 0, nullptr, nb, nb, (int64_t)nflows * nb, xwork.ptr, lqb.ptr);
    if (moment == 2) {
        hipLaunchKernelGGL(k_loss, dim3(nbk), dim3(kTh), 0, stream, nb, nflows, lqb.ptr, beta, w2, 1.0 / (double)nb, mix, moment,
                           dlqb.ptr, lpart.ptr);
    } else {
        ltb.resize((size_t)2 * nb);
        lred.resize((size_t)2 * nbk);
        double* lt = ltb.ptr;
        double* rat = ltb.ptr + nb;
        double* bmax = lred.ptr;
        double* bsum = lred.ptr + nbk;
        hipLaunchKernelGGL(k_loss_terms, dim3(nbk), dim3(kTh), 0, stream, nb, nflows, lqb.ptr, beta, w2, mix, moment, lt, rat, bmax);
        hipLaunchKernelGGL(k_loss_sums, dim3(nbk), dim3(kTh), 0, stream, nb, (int)nbk, lt, bmax, bsum);
        hipLaunchKernelGGL(k_loss_grad_log, dim3(nbk), dim3(kTh), 0, stream, nb, nflows, lqb.ptr, beta, mix, moment, lt, rat, bmax, bsum,
                           (int)nbk, dlqb.ptr, lpart.ptr);
    }
    std::vector<double> lp(nbk);
    lpart.download(lp.data(), nbk, stream);
    // per flow (independent: its own slices of lqb / dlqb and its own
    // gradients): density keeping the activations, then backward; with
    // several flows, round robin over the training lanes
    const int K = std::min(kTrainLanes, nflows);
    if (K > 1) {
        ensure_lanes(K);
        HIP_CHECK(hipEventRecord(ev_ready, stream));
        for (int k = 0; k < K; ++k) HIP_CHECK(hipStreamWaitEvent(lanes[k]->stream, ev_ready, 0));
        for (int f = 0; f < nflows; ++f) {
            LaneSwap sw(*this, *lanes[f % K]);
            density(f, nb, r + (size_t)f * nb * d, lqb.ptr + (size_t)f * nb, nullptr, true);
            backward(f, nb, dlqb.ptr + (size_t)f * nb);
        }
        for (int k = 0; k < K; ++k) {
            HIP_CHECK(hipEventRecord(lanes[k]->done, lanes[k]->stream));
            HIP_CHECK(hipStreamWaitEvent(stream, lanes[k]->done, 0));
        }
    } else {
        density(0, nb, r, lqb.ptr, nullptr, true);
        backward(0, nb, dlqb.ptr);
    }
    HIP_CHECK(hipStreamSynchronize(stream));
    double s = 0;
    for (double v : lp) s += v;
    return s / (double)nb;
}

void FlowEngine::ensure_lanes(int k) {
    if (!ev_ready) HIP_CHECK(hipEventCreateWithFlags(&ev_ready, hipEventDisableTiming));
    while ((int)lanes.size() < k) {
        auto ln = std::make_unique<Lane>();
        HIP_CHECK(hipStreamCreateWithFlags(&ln->stream, hipStreamNonBlocking));
        HIP_CHECK(hipEventCreateWithFlags(&ln->done, hipEventDisableTiming));
        lanes.push_back(std::move(ln));
    }
}

void FlowEngine::release_lanes() {
    for (auto& ln : lanes) {
        if (ln->stream) (void)hipStreamSynchronize(ln->stream);
        if (ln->done) (void)hipEventDestroy(ln->done);
        if (ln->stream) (void)hipStreamDestroy(ln->stream);
    }
    lanes.clear();
    if (ev_ready) (void)hipEventDestroy(ev_ready);
    ev_ready = nullptr;
}

void FlowEngine::adam_step(double lr) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int f = 0; f < nflows; ++f) {
        Flow& F = flows[f];
        F.steps += 1;
        const float c1 = 1.f - std::pow(b1, (float)F.steps), c2 = 1.f - std::pow(b2, (float)F.steps);
        hipLaunchKernelGGL(k_adam, dim3(nblk((int64_t)nparams)), dim3(kTh), 0, stream, (int64_t)nparams, F.params.ptr, F.grads.ptr, F.m.ptr,
                           F.v.ptr, (float)lr, b1, b2, eps, c1, c2);
        refresh_bf16(f);
    }
}

}  // namespace chili_nf

// ===========================================================================
// C-ABI
// ===========================================================================
extern "C" {

int chili_flow_create(int ndims, int nlayers, int hidden, int nflows, uint64_t seed, int device, chili_flow** out) {
    return guarded([&] {
        if (!out) throw chili_error(CHILI_EINVAL, "null argument");
        if (ndims < 2 || ndims > CHILI_MAX_DIMS || nlayers < 1 || nlayers > 32 || hidden != 64 || nflows < 1 || nflows > 256)
            throw chili_error(CHILI_ERANGE, "flow sizes out of range (hidden width must be 64)");
        select_device(device);
        auto F = std::make_unique<chili_flow>();
        auto& E = F->eng;
        E.device = device;
        HIP_CHECK(hipStreamCreateWithFlags(&E.stream, hipStreamNonBlocking));
        E.d = ndims;
        E.L = nlayers;
        E.H = hidden;
        E.nflows = nflows;
        E.shift = (float)std::log(std::expm1(1.0 - kMinD));
        E.build_layers(seed);
        E.flows.resize(nflows);
        E.init_params(seed);
        HIP_CHECK(hipStreamSynchronize(E.stream));
        *out = F.release();
    });
}

int chili_flow_destroy(chili_flow* f) {
    return guarded([&] {
        if (!f) return;
        select_device(f->eng.device);
        hipStream_t s = f->eng.stream;
        // teardown is best effort: a sticky error of earlier work must not
        // keep the context from being freed
        if (s) (void)hipStreamSynchronize(s);
        f->eng.release_lanes();
        f->eng.release_timing();
        delete f;
        if (s) (void)hipStreamDestroy(s);
    });
}

int chili_flow_info(const chili_flow* f, int64_t* nparams, int* ndims, int* nlayers, int* nflows) {
    return guarded([&] {
        if (!f) throw chili_error(CHILI_EINVAL, "null flow");
        if (nparams) *nparams = (int64_t)f->eng.nparams;
        if (ndims) *ndims = f->eng.d;
        if (nlayers) *nlayers = f->eng.L;
        if (nflows) *nflows = f->eng.nflows;
    });
}

int chili_flow_set_mixture(chili_flow* f, double lambda) {
    return guarded([&] {
        if (!f) throw chili_error(CHILI_EINVAL, "null flow");
        if (!(lambda >= 0.0 && lambda <= 1.0)) throw chili_error(CHILI_EINVAL, "mixture weight must lie in [0, 1]");
        f->eng.mix = lambda;
    });
}

int chili_flow_set_moment(chili_flow* f, int p) {
    return guarded([&] {
        if (!f) throw chili_error(CHILI_EINVAL, "null flow");
        if (p < 2 || p > 64) throw chili_error(CHILI_ERANGE, "loss moment must lie in [2, 64]");
        f->eng.moment = p;
    });
}

// k_couple timing: enable 1 on / 0 off, reset clears the totals; outputs
// (each may be NULL): summed launch milliseconds, launch count, rows of the
// timed passes (each row passes every layer) and algorithmic BF16 flops.
int chili_flow_timing(chili_flow* f, int enable, int reset, double* ms, int64_t* launches, int64_t* rows, double* flops) {
    return guarded([&] {
        if (!f) throw chili_error(CHILI_EINVAL, "null flow");
        select_device(f->eng.device);
        f->eng.harvest_timing();
        if (ms) *ms = f->eng.couple_ms;
        if (launches) *launches = f->eng.couple_launches;
        if (rows) *rows = f->eng.couple_rows;
        if (flops) *flops = f->eng.couple_flops;
        if (reset) {
            f->eng.couple_ms = 0;
            f->eng.couple_flops = 0;
            f->eng.couple_launches = 0;
            f->eng.couple_rows = 0;
        }
        f->eng.timing = enable != 0;
    });
}

// Coupling masks: transformed dimensions of layer l (nB entries).
int chili_flow_mask(const chili_flow* f, int layer, int32_t* B, int* nB) {
    return guarded([&] {
        if (!f || layer < 0 || layer >= f->eng.L) throw chili_error(CHILI_EINVAL, "bad layer");
        const auto& ly = f->eng.layers[layer];
        if (nB) *nB = ly.nB;
        if (B) for (int i = 0; i < ly.nB; ++i) B[i] = ly.B[i];
    });
}

int chili_flow_get_params(chili_flow* f, int flow, float* out) {
    return guarded([&] {
        if (!f || !out || flow < 0 || flow >= f->eng.nflows) throw chili_error(CHILI_EINVAL, "bad argument");
        select_device(f->eng.device);
        f->eng.flows[flow].params.download(out, f->eng.nparams, f->eng.stream);
        HIP_CHECK(hipStreamSynchronize(f->eng.stream));
    });
}

int chili_flow_set_params(chili_flow* f, int flow, const float* in) {
    return guarded([&] {
        if (!f || !in || flow < 0 || flow >= f->eng.nflows) throw chili_error(CHILI_EINVAL, "bad argument");
        select_device(f->eng.device);
        std::vector<float> p(in, in + f->eng.nparams);
        f->eng.flows[flow].params.upload(p, f->eng.stream);
        f->eng.refresh_bf16(flow);
        HIP_CHECK(hipStreamSynchronize(f->eng.stream));
    });
}

int chili_flow_get_grads(chili_flow* f, int flow, float* out) {
    return guarded([&] {
        if (!f || !out || flow < 0 || flow >= f->eng.nflows) throw chili_error(CHILI_EINVAL, "bad argument");
        select_device(f->eng.device);
        f->eng.flows[flow].grads.download(out, f->eng.nparams, f->eng.stream);
        HIP_CHECK(hipStreamSynchronize(f->eng.stream));
    });
}

// log q(r) (and z) for n points of flow `flow`; device pointers, r/z [n][d].
int chili_flow_log_density(chili_flow* f, int flow, int64_t n, const double* r_dev, double* lq_dev, double* z_dev) {
    return guarded([&] {
        if (!f || !r_dev || !lq_dev || flow < 0 || flow >= f->eng.nflows) throw chili_error(CHILI_EINVAL, "bad argument");
        if (n <= 0 || n > (int64_t(1) << 24)) throw chili_error(CHILI_ERANGE, "n out of range");
        select_device(f->eng.device);
        f->eng.log_density(flow, n, r_dev, lq_dev, z_dev);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipStreamSynchronize(f->eng.stream));
    });
}

// z -> r in place for n points (device [n][d]); lq = log q(r).
int chili_flow_sample(chili_flow* f, int flow, int64_t n, double* zr_dev, double* lq_dev) {
    return guarded([&] {
        if (!f || !zr_dev || !lq_dev || flow < 0 || flow >= f->eng.nflows) throw chili_error(CHILI_EINVAL, "bad argument");
        if (n <= 0 || n > (int64_t(1) << 24)) throw chili_error(CHILI_ERANGE, "n out of range");
        select_device(f->eng.device);
        f->eng.sample(flow, n, zr_dev, lq_dev);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipStreamSynchronize(f->eng.stream));
    });
}

// Weighted-variance loss and parameter gradients (all flows) on one batch
// (device: r [C][n][d], beta [C][n], w2 [n]).  Gradients stay on the device
// (chili_flow_get_grads); nothing is updated.
int chili_flow_loss_grad(chili_flow* f, int64_t n, const double* r_dev, const double* beta_dev, const double* w2_dev, double* loss) {
    return guarded([&] {
        if (!f || !r_dev || !beta_dev || !w2_dev || !loss) throw chili_error(CHILI_EINVAL, "null argument");
        if (n <= 0 || n > (int64_t(1) << 20)) throw chili_error(CHILI_ERANGE, "batch out of range");
        select_device(f->eng.device);
        *loss = f->eng.loss_grad(n, r_dev, beta_dev, w2_dev);
        HIP_CHECK(hipGetLastError());
    });
}

// Training: epochs over n points in batches (Adam); loss_hist (host, optional)
// receives the mean batch loss of every epoch.
int chili_flow_train(chili_flow* f, int64_t n, const double* r_dev, const double* beta_dev, const double* w2_dev, int64_t batch,
                     int epochs, double lr, uint64_t seed, double* loss_hist) {
    return guarded([&] {
        if (!f || !r_dev || !beta_dev || !w2_dev) throw chili_error(CHILI_EINVAL, "null argument");
        if (n <= 0 || batch <= 0 || batch > (int64_t(1) << 20) || epochs < 0) throw chili_error(CHILI_ERANGE, "sizes out of range");
        auto& E = f->eng;
        select_device(E.device);
        const int C = E.nflows, d = E.d;
        batch = std::min(batch, n);
        const int64_t nbat = n / batch;
        // one permutation of the points, fixed by the seed
        std::vector<int64_t> perm(n);
        for (int64_t i = 0; i < n; ++i) perm[i] = i;
        std::mt19937_64 rng(seed);
        std::shuffle(perm.begin(), perm.end(), rng);
        DevBuf<int64_t> dperm;
        dperm.upload(perm, E.stream);
        DevBuf<double> br, bb, bw;
        br.resize((size_t)C * batch * d);
        bb.resize((size_t)C * batch);
        bw.resize(batch);
        std::vector<int64_t> order(nbat);
        for (int64_t b = 0; b < nbat; ++b) order[b] = b;
        for (int ep = 0; ep < epochs; ++ep) {
            std::shuffle(order.begin(), order.end(), rng);
            double ls = 0;
            for (int64_t bi = 0; bi < nbat; ++bi) {
                const int64_t base = order[bi] * batch;
                for (int c = 0; c < C; ++c) {
                    hipLaunchKernelGGL(k_gather_rows, dim3(nblk(batch)), dim3(kTh), 0, E.stream, batch, d, r_dev + (size_t)c * n * d,
                                       dperm.ptr, base, br.ptr + (size_t)c * batch * d);
                    hipLaunchKernelGGL(k_gather_rows, dim3(nblk(batch)), dim3(kTh), 0, E.stream, batch, 1, beta_dev + (size_t)c * n,
                                       dperm.ptr, base, bb.ptr + (size_t)c * batch);
                }
                hipLaunchKernelGGL(k_gather_rows, dim3(nblk(batch)), dim3(kTh), 0, E.stream, batch, 1, w2_dev, dperm.ptr, base, bw.ptr);
                ls += E.loss_grad(batch, br.ptr, bb.ptr, bw.ptr);
                E.adam_step(lr);
            }
            if (loss_hist) loss_hist[ep] = nbat ? ls / (double)nbat : 0.0;
        }
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipStreamSynchronize(E.stream));
    });
}

}  // extern "C"
