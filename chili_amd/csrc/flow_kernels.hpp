// flow_kernels.hpp -- device kernels of the learned proposal: a coupling-layer
// normalizing flow with rational-quadratic spline transforms (Durkan, Bekasov,
// Murray, Papamakarios 2019, "Neural Spline Flows") on [0,1]^d.
//
//  * k_gemm: C = A x B on MFMA (v_mfma_f32_32x32x16_bf16): BF16 operands,
//    FP32 accumulation, 64x64 output tile per workgroup (four waves of 32x32),
//    K in chunks of 32 staged through LDS; optional transposed operands,
//    bias + ReLU epilogue, BF16 copy of the result, split-K partials.
//  * spline kernels: FP64 rational-quadratic transforms, their log-Jacobians,
//    inverses and hand-derived reverse-mode gradients.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace chili_nf {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kGT = 64;     // output tile (M and N)
constexpr int kGK = 32;     // K chunk staged in LDS
constexpr int kGThreads = 256;

__device__ __forceinline__ float bf2f(bf16 v) { return (float)v; }
__device__ __forceinline__ bf16 f2bf(float v) { return (bf16)v; }

// MFMA 32x32x16 operand fragments: lane l holds row (A) / column (B) l % 32
// and the 8 consecutive k values starting at 8 * (l / 32).
__device__ __forceinline__ int frag_k0(int lane) { return 8 * (lane >> 5); }

// C[m][n] = sum_k A(m, k) B(k, n) (+ bias[n], ReLU), fp32 out (ldc), optional
// bf16 copy (ldcb).  A(m, k) = A[m * lda + k] (A_T = false) or A[k * lda + m];
// B(k, n) = B[n * ldb + k] (B_T = true) or B[k * ldb + n].  Split-K: blockIdx.z
// covers K range [z * kchunk, (z+1) * kchunk) and writes partial z of C.
template <bool A_T, bool B_T>
__global__ __launch_bounds__(kGThreads) void k_gemm(int M, int N, int K, const bf16* __restrict__ A, int lda,
                                                    const bf16* __restrict__ B, int ldb, float* __restrict__ C, int ldc,
                                                    const float* __restrict__ bias, int relu, bf16* __restrict__ Cb, int ldcb,
                                                    int kchunk, int64_t part_stride) {
    __shared__ bf16 sA[kGT][kGK + 8];
    __shared__ bf16 sB[kGT][kGK + 8];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m0 = blockIdx.y * kGT, n0 = blockIdx.x * kGT;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    const int kbeg = blockIdx.z * kchunk;
    const int kend = min(K, kbeg + kchunk);
    f32x16 acc;
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] = 0.f;
    for (int k0 = kbeg; k0 < kend; k0 += kGK) {
        // stage A tile [64 m][32 k] and B tile [64 n][32 k] (zero outside)
        for (int i = tid; i < kGT * kGK; i += kGThreads) {
            int r, c;
            if (!A_T) { r = i / kGK; c = i % kGK; } else { c = i / kGT; r = i % kGT; }
            const int m = m0 + r, k = k0 + c;
            bf16 v = f2bf(0.f);
            if (m < M && k < kend) v = A_T ? A[(int64_t)k * lda + m] : A[(int64_t)m * lda + k];
            sA[r][c] = v;
        }
        for (int i = tid; i < kGT * kGK; i += kGThreads) {
            int r, c;
            if (B_T) { r = i / kGK; c = i % kGK; } else { c = i / kGT; r = i % kGT; }
            const int n = n0 + r, k = k0 + c;
            bf16 v = f2bf(0.f);
            if (n < N && k < kend) v = B_T ? B[(int64_t)n * ldb + k] : B[(int64_t)k * ldb + n];
            sB[r][c] = v;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < kGK; kk += 16) {
            bf16x8 a, b;
            const int kb = kk + frag_k0(lane);
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                a[j] = sA[wm + (lane & 31)][kb + j];
                b[j] = sB[wn + (lane & 31)][kb + j];
            }
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        }
        __syncthreads();
    }
    // D layout: lane l, register r -> row 8 (r / 4) + 4 (l / 32) + r % 4, column l % 32
    float* Cz = C + (int64_t)blockIdx.z * part_stride;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int row = m0 + wm + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        const int col = n0 + wn + (lane & 31);
        if (row < M && col < N) {
            float v = acc[r];
            if (bias) v += bias[col];
            if (relu) v = v > 0.f ? v : 0.f;
            if (Cz) Cz[(int64_t)row * ldc + col] = v;
            if (Cb) Cb[(int64_t)row * ldcb + col] = f2bf(v);
        }
    }
}

// ---------------------------------------------------------------------------
// Rational-quadratic spline on [0, 1] with K bins.  Raw parameters per
// transformed dimension (3K-1 floats): K width logits, K height logits, K-1
// interior-derivative pre-activations.
//   * knots: sm = softmax(logits) in FP32; w_i = minW + (1 - K minW) sm_i in
//     FP64; X_0 = 0, X_{i+1} = X_i + w_i, X_K = 1 (same for heights / Y);
//   * derivatives: d_0 = d_K = 1, d_i = minD + softplus(u_i + shift) (FP32
//     softplus), shift chosen so that u = 0 gives slope 1;
//   * transform and log-Jacobian in FP64 from those knots.
// The map is a monotone bijection of [0,1] whatever the FP32 rounding of the
// knots, and its log-derivative is evaluated from the same knots, so the
// density stays exactly normalised.
// ---------------------------------------------------------------------------
constexpr double kMinW = 1e-3, kMinH = 1e-3, kMinD = 1e-3;

__device__ __forceinline__ float softplusf(float z) { return z > 20.f ? z : log1pf(expf(z)); }
__device__ __forceinline__ float sigmoidf(float z) { return 1.f / (1.f + expf(-z)); }

// Bin of the spline containing v (search over X when by_y is false, else Y).
struct SplineBin {
    int k;
    double X0, Y0, w, h, d0, d1;
};

template <int K>
__device__ __forceinline__ SplineBin spline_bin(const float* __restrict__ u, float shift, double v, bool by_y, float* ew, float* eh,
                                                float& sw, float& sh) {
    float mw = u[0], mh = u[K];
#pragma unroll
    for (int i = 1; i < K; ++i) { mw = fmaxf(mw, u[i]); mh = fmaxf(mh, u[K + i]); }
    sw = 0.f; sh = 0.f;
#pragma unroll
    for (int i = 0; i < K; ++i) {
        ew[i] = expf(u[i] - mw); sw += ew[i];
        eh[i] = expf(u[K + i] - mh); sh += eh[i];
    }
    const double cw = 1.0 - K * kMinW, ch = 1.0 - K * kMinH;
    SplineBin b{0, 0.0, 0.0, 0.0, 0.0, 1.0, 1.0};
    double X = 0, Y = 0, Xn = 0, Yn = 0;
#pragma unroll
    for (int i = 0; i < K; ++i) {
        Xn = (i == K - 1) ? 1.0 : X + (kMinW + cw * (double)(ew[i] / sw));
        Yn = (i == K - 1) ? 1.0 : Y + (kMinH + ch * (double)(eh[i] / sh));
        const bool in = (i == 0) || (by_y ? (v >= Y) : (v >= X));
        if (in) { b.k = i; b.X0 = X; b.Y0 = Y; b.w = Xn - X; b.h = Yn - Y; }
        X = Xn;
        Y = Yn;
    }
    b.d0 = (b.k == 0) ? 1.0 : kMinD + (double)softplusf(u[2 * K + b.k - 1] + shift);
    b.d1 = (b.k == K - 1) ? 1.0 : kMinD + (double)softplusf(u[2 * K + b.k] + shift);
    return b;
}

// Forward transform x -> y, logdet = log dy/dx.
template <int K>
__device__ __forceinline__ double rqs_forward(const float* __restrict__ u, float shift, double x, double& logdet) {
    float ew[K], eh[K], sw, sh;
    x = fmin(fmax(x, 0.0), 1.0);
    const SplineBin b = spline_bin<K>(u, shift, x, false, ew, eh, sw, sh);
    const double xi = (x - b.X0) / b.w, s = b.h / b.w, t = xi * (1 - xi);
    const double den = s + (b.d1 + b.d0 - 2 * s) * t;
    const double D = b.d1 * xi * xi + 2 * s * t + b.d0 * (1 - xi) * (1 - xi);
    logdet = log(s * s * D / (den * den));
    return b.Y0 + b.h * (s * xi * xi + b.d0 * t) / den;
}

// Inverse transform y -> x; logdet receives log dy/dx at that x.
template <int K>
__device__ __forceinline__ double rqs_inverse(const float* __restrict__ u, float shift, double y, double& logdet) {
    float ew[K], eh[K], sw, sh;
    y = fmin(fmax(y, 0.0), 1.0);
    const SplineBin b = spline_bin<K>(u, shift, y, true, ew, eh, sw, sh);
    const double s = b.h / b.w, dy = y - b.Y0, E = b.d1 + b.d0 - 2 * s;
    const double a = b.h * (s - b.d0) + dy * E;
    const double bq = b.h * b.d0 - dy * E;
    const double c = -s * dy;
    const double disc = fmax(bq * bq - 4 * a * c, 0.0);
    double xi = (2 * c) / (-bq - sqrt(disc));
    xi = fmin(fmax(xi, 0.0), 1.0);
    const double t = xi * (1 - xi);
    const double den = s + E * t;
    const double D = b.d1 * xi * xi + 2 * s * t + b.d0 * (1 - xi) * (1 - xi);
    logdet = log(s * s * D / (den * den));
    return b.X0 + xi * b.w;
}

// Reverse-mode gradient of (y, logdet) = rqs_forward(x): given gy = dL/dy and
// gl = dL/dlogdet, returns dL/dx and writes dL/du for the 3K-1 raw inputs.
template <int K>
__device__ __forceinline__ double rqs_backward(const float* __restrict__ u, float shift, double x, double gy, double gl,
                                               float* __restrict__ du) {
    float ew[K], eh[K], sw, sh;
    const bool inside = x > 0.0 && x < 1.0;
    x = fmin(fmax(x, 0.0), 1.0);
    const SplineBin bn = spline_bin<K>(u, shift, x, false, ew, eh, sw, sh);
    const int k = bn.k;
    const double w = bn.w, h = bn.h, dk = bn.d0, dk1 = bn.d1;
    const double xi = (x - bn.X0) / w, s = h / w, t = xi * (1 - xi), dt = 1 - 2 * xi;
    const double E = dk1 + dk - 2 * s;
    const double num = h * (s * xi * xi + dk * t), den = s + E * t;
    const double D = dk1 * xi * xi + 2 * s * t + dk * (1 - xi) * (1 - xi);
    const double den2 = den * den;
    // partials of y = Y_k + num / den
    const double y_xi = (h * (2 * s * xi + dk * dt) * den - num * E * dt) / den2;
    const double y_s = (h * xi * xi * den - num * (1 - 2 * t)) / den2;
    const double y_h = (s * xi * xi + dk * t) / den;
    const double y_dk = (h * t * den - num * t) / den2;
    const double y_dk1 = -num * t / den2;
    // partials of logdet = 2 log s + log D - 2 log den
    const double l_xi = (2 * dk1 * xi + 2 * s * dt - 2 * dk * (1 - xi)) / D - 2 * E * dt / den;
    const double l_s = 2 / s + 2 * t / D - 2 * (1 - 2 * t) / den;
    const double l_dk = (1 - xi) * (1 - xi) / D - 2 * t / den;
    const double l_dk1 = xi * xi / D - 2 * t / den;
    const double g_xi = gy * y_xi + gl * l_xi;
    const double g_s = gy * y_s + gl * l_s;
    const double g_dk = gy * y_dk + gl * l_dk;
    const double g_dk1 = gy * y_dk1 + gl * l_dk1;
    // xi = (x - X_k) / w, s = h / w; X_k = sum_{i<k} w_i, Y_k = sum_{i<k} h_i
    const double gx = inside ? g_xi / w : 0.0;
    const double gXk = -g_xi / w;
    const double gwk = g_xi * (-xi / w) + g_s * (-s / w);
    const double ghk = g_s / w + gy * y_h;
    // softmax backward: du_i = c sm_i (g_i - sum_j sm_j g_j)
    double sgw = 0, sgh = 0;
#pragma unroll
    for (int i = 0; i < K; ++i) {
        const double gw = (i < k) ? gXk : (i == k ? gwk : 0.0);
        const double gh = (i < k) ? gy : (i == k ? ghk : 0.0);
        sgw += (double)(ew[i] / sw) * gw;
        sgh += (double)(eh[i] / sh) * gh;
    }
    const double cw = 1.0 - K * kMinW, ch = 1.0 - K * kMinH;
#pragma unroll
    for (int i = 0; i < K; ++i) {
        const double gw = (i < k) ? gXk : (i == k ? gwk : 0.0);
        const double gh = (i < k) ? gy : (i == k ? ghk : 0.0);
        du[i] = (float)(cw * (double)(ew[i] / sw) * (gw - sgw));
        du[K + i] = (float)(ch * (double)(eh[i] / sh) * (gh - sgh));
    }
#pragma unroll
    for (int i = 1; i < K; ++i) {
        double g = 0;
        if (i == k) g += g_dk;
        if (i == k + 1) g += g_dk1;
        du[2 * K + i - 1] = (g != 0.0) ? (float)(g * (double)sigmoidf(u[2 * K + i - 1] + shift)) : 0.f;
    }
    return gx;
}

}  // namespace chili_nf
