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

#include <algorithm>

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
    __shared__ __align__(16) bf16 sA[kGT][kGK + 8];
    __shared__ __align__(16) bf16 sB[kGT][kGK + 8];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m0 = blockIdx.y * kGT, n0 = blockIdx.x * kGT;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    const int kbeg = blockIdx.z * kchunk;
    const int kend = min(K, kbeg + kchunk);
    // 16-byte loads need 8-element aligned rows and base pointers
    const bool vecA = (lda % 8 == 0) && ((reinterpret_cast<uintptr_t>(A) & 15) == 0);
    const bool vecB = (ldb % 8 == 0) && ((reinterpret_cast<uintptr_t>(B) & 15) == 0);
    f32x16 acc;
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] = 0.f;
    for (int k0 = kbeg; k0 < kend; k0 += kGK) {
        // stage A tile [64 m][32 k] and B tile [64 n][32 k] (zero outside);
        // the k-major layouts (A_T, !B_T) load 8 consecutive m / n at once
        if (A_T && vecA) {
            const int r8 = (tid & 7) * 8, c = tid >> 3;        // 32 k rows x 8 groups of 8 m
            const int m = m0 + r8, k = k0 + c;
            if (k < kend && m + 8 <= M) {
                const uint4 q = *reinterpret_cast<const uint4*>(A + (int64_t)k * lda + m);
                const bf16* e = reinterpret_cast<const bf16*>(&q);
#pragma unroll
                for (int j = 0; j < 8; ++j) sA[r8 + j][c] = e[j];
            } else {
#pragma unroll
                for (int j = 0; j < 8; ++j) sA[r8 + j][c] = (k < kend && m + j < M) ? A[(int64_t)k * lda + m + j] : f2bf(0.f);
            }
        } else if (!A_T && vecA) {
            // m-major: each thread loads 8 consecutive k of one row
            const int r = tid >> 2, c8 = (tid & 3) * 8;        // 64 rows x 4 groups of 8 k
            const int m = m0 + r, k = k0 + c8;
            if (m < M && k + 8 <= kend) {
                *reinterpret_cast<uint4*>(&sA[r][c8]) = *reinterpret_cast<const uint4*>(A + (int64_t)m * lda + k);
            } else {
#pragma unroll
                for (int j = 0; j < 8; ++j) sA[r][c8 + j] = (m < M && k + j < kend) ? A[(int64_t)m * lda + k + j] : f2bf(0.f);
            }
        } else {
            for (int i = tid; i < kGT * kGK; i += kGThreads) {
                int r, c;
                if (!A_T) { r = i / kGK; c = i % kGK; } else { c = i / kGT; r = i % kGT; }
                const int m = m0 + r, k = k0 + c;
                bf16 v = f2bf(0.f);
                if (m < M && k < kend) v = A_T ? A[(int64_t)k * lda + m] : A[(int64_t)m * lda + k];
                sA[r][c] = v;
            }
        }
        if (!B_T && vecB) {
            const int r8 = (tid & 7) * 8, c = tid >> 3;
            const int n = n0 + r8, k = k0 + c;
            if (k < kend && n + 8 <= N) {
                const uint4 q = *reinterpret_cast<const uint4*>(B + (int64_t)k * ldb + n);
                const bf16* e = reinterpret_cast<const bf16*>(&q);
#pragma unroll
                for (int j = 0; j < 8; ++j) sB[r8 + j][c] = e[j];
            } else {
#pragma unroll
                for (int j = 0; j < 8; ++j) sB[r8 + j][c] = (k < kend && n + j < N) ? B[(int64_t)k * ldb + n + j] : f2bf(0.f);
            }
        } else if (B_T && vecB) {
            const int r = tid >> 2, c8 = (tid & 3) * 8;
            const int n = n0 + r, k = k0 + c8;
            if (n < N && k + 8 <= kend) {
                *reinterpret_cast<uint4*>(&sB[r][c8]) = *reinterpret_cast<const uint4*>(B + (int64_t)n * ldb + k);
            } else {
#pragma unroll
                for (int j = 0; j < 8; ++j) sB[r][c8 + j] = (n < N && k + j < kend) ? B[(int64_t)n * ldb + k + j] : f2bf(0.f);
            }
        } else {
            for (int i = tid; i < kGT * kGK; i += kGThreads) {
                int r, c;
                if (B_T) { r = i / kGK; c = i % kGK; } else { c = i / kGT; r = i % kGT; }
                const int n = n0 + r, k = k0 + c;
                bf16 v = f2bf(0.f);
                if (n < N && k < kend) v = B_T ? B[(int64_t)n * ldb + k] : B[(int64_t)k * ldb + n];
                sB[r][c] = v;
            }
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
//   * knots (FP32): sm = softmax(logits), w_i = minW + (1 - K minW) sm_i,
//     X_0 = 0, X_{i+1} = X_i + w_i, X_K = 1 (same for heights / Y);
//   * derivatives: d_0 = d_K = 1, d_i = minD + softplus(u_i + shift) (FP32),
//     shift chosen so that u = 0 gives slope 1;
//   * transform and log-Jacobian in FP64 from those knots (exact in FP64).
// The map is a monotone bijection of [0,1] whatever the FP32 rounding of the
// knots, and its log-derivative is evaluated from the same knots, so the
// density stays exactly normalised.
// ---------------------------------------------------------------------------
constexpr double kMinW = 1e-3, kMinH = 1e-3, kMinD = 1e-3;

// softplus and its derivative on the hardware exp2 / log2 (the knot
// derivatives need a few significant digits; the map's Jacobian is exact in
// FP64 for whatever values they take)
__device__ __forceinline__ float softplusf(float z) {
    constexpr float kLog2e = 1.4426950408889634f, kLn2 = 0.6931471805599453f;
    return z > 20.f ? z : kLn2 * __builtin_amdgcn_logf(1.f + __builtin_amdgcn_exp2f(z * kLog2e));
}
__device__ __forceinline__ float sigmoidf(float z) {
    constexpr float kLog2e = 1.4426950408889634f;
    return 1.f / (1.f + __builtin_amdgcn_exp2f(-z * kLog2e));
}

// 1 / x in FP64 from the hardware's reciprocal estimate and two Newton
// steps, e = 1 - x r, r += r e (each step squares the relative error: within
// an ulp or two of the correctly rounded quotient, for 5 instead of ~11
// instructions); x is a spline width, slope or denominator, positive and far
// from the range's ends.
__device__ __forceinline__ double rcp64(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

struct SplineBin {
    int k;
    double X0, Y0, w, h, d0, d1;
};

static_assert(kMinW == kMinH, "one knot offset table serves both axes");

// i minW in FP32, the same value whether i is a constant or a register
__device__ __forceinline__ float knot_offset(int i) { return (float)((double)i * kMinW); }

// Largest float not above v (v in [0, 1]).  For a float knot X,
// v >= X holds exactly when f32_down(v) >= X, so the bin search compares in
// FP32 and still agrees with the FP64 transform.
__device__ __forceinline__ float f32_down(double v) {
    const float f = (float)v;
    return ((double)f > v) ? __builtin_bit_cast(float, __builtin_bit_cast(int, f) - 1) : f;
}

// Bin containing v (searched over X, or over Y when by_y); ew/eh receive the
// exponentials of the logits (max subtracted), rw/rh the reciprocals of their sums.
// u is the 3K - 1 raw parameters (a pointer, or an array held in registers:
// every index below is a compile-time constant).
template <int K, class U>
__device__ __forceinline__ SplineBin spline_bin(const U& u, float shift, double v, bool by_y, float* ew, float* eh,
                                                float& rw, float& rh) {
    float mw = u[0], mh = u[K];
#pragma unroll
    for (int i = 1; i < K; ++i) { mw = fmaxf(mw, u[i]); mh = fmaxf(mh, u[K + i]); }
    // softmax exponentials on the hardware exp2 (arguments <= 0, no overflow;
    // relative error a few ulp, which only moves the knots: the density stays
    // normalised), with their prefix sums S_i = sum_{j<i} e_j
    constexpr float kLog2e = 1.4426950408889634f;
    const float mwl = -mw * kLog2e, mhl = -mh * kLog2e;
    float SW[K + 1], SH[K + 1];
    SW[0] = 0.f;
    SH[0] = 0.f;
#pragma unroll
    for (int i = 0; i < K; ++i) {
        ew[i] = __builtin_amdgcn_exp2f(fmaf(u[i], kLog2e, mwl));
        eh[i] = __builtin_amdgcn_exp2f(fmaf(u[K + i], kLog2e, mhl));
        SW[i + 1] = SW[i] + ew[i];
        SH[i + 1] = SH[i] + eh[i];
    }
    // hardware reciprocal (about one ulp): the knots only need to be the same
    // wherever they are computed, which they are (one expression)
    rw = __builtin_amdgcn_rcpf(SW[K]);
    rh = __builtin_amdgcn_rcpf(SH[K]);
    // knots X_i = fma((1 - K minW) / sum, S_i, i minW) for 0 < i < K (the
    // same expression wherever a knot is needed), X_0 = 0, X_K = 1; Y alike
    const float aw = (float)(1.0 - K * kMinW) * rw, ah = (float)(1.0 - K * kMinH) * rh;
    const float as = by_y ? ah : aw, ao = by_y ? aw : ah;
    // exact: v >= X  <=>  vd >= X for every float X, so that the bin of v
    // agrees with the FP64 transform (forward and inverse round-trip)
    const float vd = f32_down(v);
    float E0 = 0.f, E1 = 1.f, P0 = 0.f, P1 = 0.f;
    float ud0 = 0.f, ud1 = 0.f;          // derivative logits at the bin's left / right knot
    int k = 0;
    // bin k = the last i with i == 0 or vd >= E_i (the knots E_i of the
    // searched axis do not decrease, so those i form a prefix); the
    // selections below are selects, not branches
    if constexpr (K == 16) {
        // two levels: the group of four bins from E_4, E_8, E_12, then the
        // bin within it from the group's own knots
        float Ek[K + 1];
        Ek[0] = 0.f;
        Ek[K] = 1.f;
#pragma unroll
        for (int i = 1; i < K; ++i) Ek[i] = fmaf(as, by_y ? SH[i] : SW[i], knot_offset(i));
        const bool b1 = vd >= Ek[8];
        const bool b0 = vd >= (b1 ? Ek[12] : Ek[4]);
        auto sel = [&](float x0, float x1, float x2, float x3) { return b1 ? (b0 ? x3 : x2) : (b0 ? x1 : x0); };
        // derivative logit slot j, clamped to the last real one (the clamped
        // values only ever serve bins whose outer derivative is fixed to 1)
        auto ulog = [&](int j) { return u[j < 3 * K - 2 ? j : 3 * K - 2]; };
        float Eg[5], Pg[5], Ug[5];
#pragma unroll
        for (int i = 0; i < 5; ++i) {
            Eg[i] = sel(Ek[i], Ek[4 + i], Ek[8 + i], Ek[12 + i]);
            Pg[i] = by_y ? sel(SW[i], SW[4 + i], SW[8 + i], SW[12 + i]) : sel(SH[i], SH[4 + i], SH[8 + i], SH[12 + i]);
            Ug[i] = sel(ulog(2 * K + i - 1), ulog(2 * K + 3 + i), ulog(2 * K + 7 + i), ulog(2 * K + 11 + i));
        }
        int kl = 0;
        E0 = Eg[0];
        E1 = Eg[1];
        P0 = Pg[0];
        P1 = Pg[1];
        ud0 = Ug[0];
        ud1 = Ug[1];
#pragma unroll
        for (int i = 1; i < 4; ++i) {
            const bool in = vd >= Eg[i];
            kl = in ? i : kl;
            E0 = in ? Eg[i] : E0;
            E1 = in ? Eg[i + 1] : E1;
            P0 = in ? Pg[i] : P0;
            P1 = in ? Pg[i + 1] : P1;
            ud0 = in ? Ug[i] : ud0;
            ud1 = in ? Ug[i + 1] : ud1;
        }
        k = 4 * (2 * (int)b1 + (int)b0) + kl;
    } else {
        float E = 0.f;
#pragma unroll
        for (int i = 0; i < K; ++i) {
            // searched axis: knot i is E, knot i + 1 is En
            const float En = (i == K - 1) ? 1.f : fmaf(as, by_y ? SH[i + 1] : SW[i + 1], knot_offset(i + 1));
            const bool in = (i == 0) || (vd >= E);
            k = in ? i : k;
            E0 = in ? E : E0;
            E1 = in ? En : E1;
            P0 = in ? (by_y ? SW[i] : SH[i]) : P0;
            P1 = in ? (by_y ? SW[i + 1] : SH[i + 1]) : P1;
            if (i >= 1) ud0 = in ? u[2 * K + i - 1] : ud0;
            if (i <= K - 2) ud1 = in ? u[2 * K + i] : ud1;
            E = En;
        }
    }
    // the other axis' knots k and k + 1 from their prefix sums
    const float O0 = (k == 0) ? 0.f : fmaf(ao, P0, knot_offset(k));
    const float O1 = (k == K - 1) ? 1.f : fmaf(ao, P1, knot_offset(k + 1));
    const float X0 = by_y ? O0 : E0, X1 = by_y ? O1 : E1, Y0 = by_y ? E0 : O0, Y1 = by_y ? E1 : O1;
    SplineBin b;
    b.k = k;
    b.X0 = (double)X0;
    b.Y0 = (double)Y0;
    b.w = (double)X1 - b.X0;
    b.h = (double)Y1 - b.Y0;
    // both softplus values are computed on every lane and then selected, so
    // that the end bins (fixed slope 1) add no branch to the caller's code
    const float sp0 = softplusf(ud0 + shift), sp1 = softplusf(ud1 + shift);
    b.d0 = (k == 0) ? 1.0 : kMinD + (double)sp0;
    b.d1 = (k == K - 1) ? 1.0 : kMinD + (double)sp1;
    return b;
}

// Forward transform x -> y, jac = dy/dx.
// The transforms are this build's own arithmetic (no reference operation
// order to keep): multiply-adds may fuse, divisions are reciprocals.
template <int K, class U>
__device__ __forceinline__ double rqs_forward_j(const U& u, float shift, double x, double& jac) {
#pragma clang fp contract(fast)
    float ew[K], eh[K], rw, rh;
    x = fmin(fmax(x, 0.0), 1.0);
    const SplineBin b = spline_bin<K>(u, shift, x, false, ew, eh, rw, rh);
    const double iw = rcp64(b.w);
    const double xi = (x - b.X0) * iw, s = b.h * iw, t = xi * (1 - xi);
    const double iden = rcp64(s + (b.d1 + b.d0 - 2 * s) * t);
    const double D = b.d1 * xi * xi + 2 * s * t + b.d0 * (1 - xi) * (1 - xi);
    jac = s * s * D * iden * iden;
    return b.Y0 + b.h * (s * xi * xi + b.d0 * t) * iden;
}

// Forward transform x -> y, logdet = log dy/dx.
template <int K, class U>
__device__ __forceinline__ double rqs_forward(const U& u, float shift, double x, double& logdet) {
    double jac;
    const double y = rqs_forward_j<K>(u, shift, x, jac);
    logdet = log(jac);
    return y;
}

// Inverse transform y -> x; jac receives dy/dx at that x.
template <int K, class U>
__device__ __forceinline__ double rqs_inverse_j(const U& u, float shift, double y, double& jac) {
#pragma clang fp contract(fast)
    float ew[K], eh[K], rw, rh;
    y = fmin(fmax(y, 0.0), 1.0);
    const SplineBin b = spline_bin<K>(u, shift, y, true, ew, eh, rw, rh);
    const double s = b.h * rcp64(b.w), dy = y - b.Y0, E = b.d1 + b.d0 - 2 * s;
    const double a = b.h * (s - b.d0) + dy * E;
    const double bq = b.h * b.d0 - dy * E;
    const double c = -s * dy;
    const double disc = fmax(bq * bq - 4 * a * c, 0.0);
    double xi = (2 * c) * rcp64(-bq - sqrt(disc));
    xi = fmin(fmax(xi, 0.0), 1.0);
    const double t = xi * (1 - xi);
    const double iden = rcp64(s + E * t);
    const double D = b.d1 * xi * xi + 2 * s * t + b.d0 * (1 - xi) * (1 - xi);
    jac = s * s * D * iden * iden;
    return b.X0 + xi * b.w;
}

// Inverse transform y -> x; logdet receives log dy/dx at that x.
template <int K, class U>
__device__ __forceinline__ double rqs_inverse(const U& u, float shift, double y, double& logdet) {
    double jac;
    const double x = rqs_inverse_j<K>(u, shift, y, jac);
    logdet = log(jac);
    return x;
}

// Reverse-mode gradient of (y, logdet) = rqs_forward(x): given gy = dL/dy and
// gl = dL/dlogdet, returns dL/dx and writes dL/du for the 3K-1 raw inputs.
template <int K>
__device__ __forceinline__ double rqs_backward(const float* __restrict__ u, float shift, double x, double gy, double gl,
                                               float* __restrict__ du) {
    float ew[K], eh[K], rw, rh;
    const bool inside = x > 0.0 && x < 1.0;
    x = fmin(fmax(x, 0.0), 1.0);
    const SplineBin bn = spline_bin<K>(u, shift, x, false, ew, eh, rw, rh);
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
        sgw += (double)(ew[i] * rw) * gw;
        sgh += (double)(eh[i] * rh) * gh;
    }
    const double cw = 1.0 - K * kMinW, ch = 1.0 - K * kMinH;
#pragma unroll
    for (int i = 0; i < K; ++i) {
        const double gw = (i < k) ? gXk : (i == k ? gwk : 0.0);
        const double gh = (i < k) ? gy : (i == k ? ghk : 0.0);
        du[i] = (float)(cw * (double)(ew[i] * rw) * (gw - sgw));
        du[K + i] = (float)(ch * (double)(eh[i] * rh) * (gh - sgh));
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


// ===========================================================================
// Fused coupling layer (inference): one kernel per layer evaluates the
// conditioner MLP and the splines of a tile of 32 points per wave.
//
// The MLP runs transposed, D[feature][point] = W^T[feature][k] X^T[k][point],
// on v_mfma_f32_32x32x16_bf16: weights are the A operand (read from LDS in a
// pre-swizzled fragment order), activations the B operand.  A 32x32 result
// tile leaves lane l holding point l % 32 and features
//     f(r) = 8 (r / 4) + 4 (l / 32) + r % 4,   r = 0..15,
// so the next layer's B fragment for k-step s (16 features) is registers
// 8 (s % 2) .. 8 (s % 2) + 7 of tile s / 2, in the feature order
//     phi(s, h, j) = 16 s + 8 (j / 4) + 4 h + j % 4     (h = l / 32);
// the swizzled weights use the same k order, so activations never leave
// registers between hidden layers.  The output layer is evaluated per pair of
// transformed dimensions (three 32-row tiles, 2 x 48 rows, 47 used per
// dimension); its weight rows are ordered so that row f of tile t holds
// parameter 16 t + 4 (f / 8) + f % 4 of dimension (f / 4) % 2, which leaves
// lane half h with all parameters of dimension 2P + h in its own registers
// (parameter q in register q % 16 of tile q / 16): each lane then runs the
// spline of one (point, dimension) with no exchange between lanes.
// ===========================================================================
constexpr int kHid = 64;
constexpr int kPairRows = 96;
constexpr int kRaw = 3 * 16 - 1;        // raw parameters per dimension (K = 16)

__host__ __device__ constexpr int phi_k(int s, int h, int j) { return 16 * s + 8 * (j / 4) + 4 * h + (j % 4); }

// Swizzled block of one (flow, layer): bf16 fragments then fp32 biases.
struct FusedLayout {
    int npair;
    int64_t oW1, oW2, oW3, oW4, wel;     // bf16 element offsets / count
    int64_t ob1, ob2, ob3, ob4, bel;     // fp32 offsets / count
};
__host__ __device__ inline FusedLayout fused_layout(int nB) {
    FusedLayout f;
    f.npair = (nB + 1) / 2;
    f.oW1 = 0;
    f.oW2 = f.oW1 + 2 * 64 * 8;
    f.oW3 = f.oW2 + 2 * 4 * 64 * 8;
    f.oW4 = f.oW3 + 2 * 4 * 64 * 8;
    f.wel = f.oW4 + (int64_t)3 * f.npair * 4 * 64 * 8;
    f.ob1 = 0; f.ob2 = 64; f.ob3 = 128; f.ob4 = 192;
    f.bel = f.ob4 + (int64_t)kPairRows * f.npair;
    return f;
}

struct CoupleArgs {
    double* x;                 // points [rows][d] (in place)
    double* lq;                // [rows] accumulated log-Jacobian
    const int64_t* start;      // [nflows] first row of flow j's points
    const int64_t* count;      // [nflows]
    const int* items;          // [nflows + 1] prefix sums of work items (chunks of `chunk` rows)
    int nflows, flow0;
    int64_t chunk;
    const bf16* w;             // swizzled blocks, flow j layer l at (j * L + l) * wstride
    const float* b;            // biases, at (j * L + l) * bstride
    int64_t wstride, bstride;
    int L, layer, d, nA, nB;
    int A[16], B[16];
    float shift;
};

__device__ __forceinline__ bf16x8 lds_frag(const bf16* base, int lane) {
    return *reinterpret_cast<const bf16x8*>(base + lane * 8);
}

// Accumulators of a 32-feature tile starting at the layer's biases: lane half
// h holds features 8 m + 4 h + i (m, i = 0..3) of the tile in register 4 m + i.
__device__ __forceinline__ f32x16 bias_tile(const float* __restrict__ bias, int h) {
    f32x16 c;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const float4 b4 = *reinterpret_cast<const float4*>(bias + 8 * m + 4 * h);
        c[4 * m] = b4.x;
        c[4 * m + 1] = b4.y;
        c[4 * m + 2] = b4.z;
        c[4 * m + 3] = b4.w;
    }
    return c;
}

// ReLU of two 32-feature tiles (biases already in), packed as the four B
// fragments of the next layer.  Rounding to bf16 keeps the sign, so the ReLU
// is applied after it, on pairs of bf16 bit patterns as signed 16-bit
// integers (every negative value and -0 has the sign bit set and becomes +0).
typedef short i16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void relu_pack(const f32x16& a0, const f32x16& a1, bf16x8* hb) {
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const f32x16& a = (s < 2) ? a0 : a1;
        u32x4 w;
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const f32x2 f = {a[8 * (s & 1) + 2 * m], a[8 * (s & 1) + 2 * m + 1]};
            const i16x2 v = __builtin_bit_cast(i16x2, __builtin_convertvector(f, bf16x2));
            w[m] = __builtin_bit_cast(unsigned, __builtin_elementwise_max(v, (i16x2)0));
        }
        hb[s] = __builtin_bit_cast(bf16x8, w);
    }
}

// NP = pairs of transformed dimensions (ceil(nB / 2)), a compile-time
// constant so that the per-lane operands of a tile live in registers.
//
// Software pipeline: the matrix work of the next unit -- the output rows of
// pair P + 1, or after the last pair the next tile's hidden layers and the
// output rows of its pair 0 -- is issued in the same scheduling region as the
// spline of pair P, which does not depend on it, so a wave keeps the matrix
// core and the vector ALU busy at once.  The spline runs on every lane
// (lanes past the chunk or without a transformed dimension compute on zero
// inputs and keep nothing), so that no branch splits the region.
template <int K, bool INV, int NP>
__global__ __launch_bounds__(512) void k_couple(CoupleArgs a) {
    extern __shared__ __align__(16) unsigned char s_raw[];
    // work item -> flow (last j with items[j] <= item) and chunk of its rows
    const int item = blockIdx.x;
    if (item >= a.items[a.nflows]) return;
    int lo = 0, hi = a.nflows - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (a.items[mid] <= item) lo = mid; else hi = mid - 1;
    }
    const int flow = a.flow0 + lo;
    const int64_t c0 = (int64_t)(item - a.items[lo]) * a.chunk;
    const int64_t cnt = std::min<int64_t>(a.chunk, a.count[lo] - c0);
    if (cnt <= 0) return;
    const FusedLayout FL = fused_layout(a.nB);
    bf16* sW = reinterpret_cast<bf16*>(s_raw);
    float* sB = reinterpret_cast<float*>(s_raw + ((FL.wel * 2 + 15) / 16) * 16);
    const int nw = blockDim.x >> 6;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, h = lane >> 5, pl = lane & 31;
    {
        const int64_t blk = (int64_t)flow * a.L + a.layer;
        const uint4* gw = reinterpret_cast<const uint4*>(a.w + blk * a.wstride);
        uint4* lw = reinterpret_cast<uint4*>(sW);
        for (int64_t i = threadIdx.x; i < FL.wel / 8; i += blockDim.x) lw[i] = gw[i];
        const float* gb = a.b + blk * a.bstride;
        for (int64_t i = threadIdx.x; i < FL.bel; i += blockDim.x) sB[i] = gb[i];
    }
    __syncthreads();
    const int64_t start = a.start[lo] + c0;
    const int64_t ntile = (cnt + 31) / 32;
    if (wave >= ntile) return;
    int aidx[8];                         // conditioning dims this lane feeds to layer 1
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const int k = 8 * h + j;
        aidx[j] = k < a.nA ? a.A[k] : -1;
    }
    int bidx[NP];                        // transformed dim of this lane in pair P (or -1)
#pragma unroll
    for (int P = 0; P < NP; ++P) bidx[P] = (2 * P + h < a.nB) ? a.B[2 * P + h] : -1;
    // raw inputs of the next tile: conditioning values, transformed values,
    // running log q.  The loads are unconditional -- a row past the chunk
    // reads the chunk's first row, a missing dimension column 0 -- and the
    // values are masked where they are used (npv: the row exists), so that no
    // branch splits the loop body, a fetch past the last tile is safe, and the
    // loads stay in flight until their first use.
    double rxa[8], rxb[NP], rlq;
    bool npv;
    auto fetch = [&](int64_t t) {
        npv = t * 32 + pl < cnt;
        const int64_t row = npv ? start + t * 32 + pl : start;
        const double* xp = a.x + row * a.d;
#pragma unroll
        for (int j = 0; j < 8; ++j) rxa[j] = xp[aidx[j] >= 0 ? aidx[j] : 0];
#pragma unroll
        for (int P = 0; P < NP; ++P) rxb[P] = xp[bidx[P] >= 0 ? bidx[P] : 0];
        rlq = a.lq[row];
    };
    // hidden layers of the fetched tile: lanes of half h feed conditioning
    // features 8h..8h+7 to layer 1; hb receives the last hidden layer as the
    // output layer's B fragments
    auto hidden = [&](bf16x8* hb) {
        bf16x8 b0;
#pragma unroll
        for (int j = 0; j < 8; ++j) b0[j] = f2bf((npv & (aidx[j] >= 0)) ? (float)rxa[j] : 0.f);
        f32x16 g0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(lds_frag(sW + FL.oW1, lane), b0, bias_tile(sB + FL.ob1, h), 0, 0, 0);
        f32x16 g1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(lds_frag(sW + FL.oW1 + 512, lane), b0, bias_tile(sB + FL.ob1 + 32, h),
                                                            0, 0, 0);
        relu_pack(g0, g1, hb);
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const bf16* Wq = sW + (q == 0 ? FL.oW2 : FL.oW3);
            const float* bq = sB + (q == 0 ? FL.ob2 : FL.ob3);
            g0 = bias_tile(bq, h);
            g1 = bias_tile(bq + 32, h);
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                g0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(lds_frag(Wq + (0 * 4 + s) * 512, lane), hb[s], g0, 0, 0, 0);
                g1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(lds_frag(Wq + (1 * 4 + s) * 512, lane), hb[s], g1, 0, 0, 0);
            }
            relu_pack(g0, g1, hb);
        }
    };
    // output rows of pair P: lane half h receives the 48 raw parameters of
    // dimension 2P + h in c[q / 16][q % 16] (the swizzled W4 rows put them
    // there), the accumulators starting at their biases
    auto outp = [&](int P, const bf16x8* hb, f32x16* c) {
        const float4* bq = reinterpret_cast<const float4*>(sB + FL.ob4 + kPairRows * P + 48 * h);
#pragma unroll
        for (int tt = 0; tt < 3; ++tt)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const float4 b4 = bq[4 * tt + v];
                c[tt][4 * v] = b4.x;
                c[tt][4 * v + 1] = b4.y;
                c[tt][4 * v + 2] = b4.z;
                c[tt][4 * v + 3] = b4.w;
            }
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
            for (int tt = 0; tt < 3; ++tt)
                c[tt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(lds_frag(sW + FL.oW4 + ((int64_t)(3 * P + tt) * 4 + s) * 512, lane),
                                                                hb[s], c[tt], 0, 0, 0);
    };
    fetch(wave);
    bf16x8 hb[4];
    hidden(hb);
    f32x16 cc[3];
    outp(0, hb, cc);
    for (int64_t t = wave; t < ntile; t += nw) {
        const int64_t row = start + t * 32 + pl;
        const bool pv = t * 32 + pl < cnt;
        double* xp = a.x + row * a.d;
        double xb[NP];
#pragma unroll
        for (int P = 0; P < NP; ++P) xb[P] = (pv & (bidx[P] >= 0)) ? rxb[P] : 0.0;
        const double lq0 = (pv & (h == 0)) ? rlq : 0.0;
        fetch(t + nw);                   // the next tile's inputs, in flight during this tile
        double jprod = 1.0;              // product of this lane's dy/dx (one log per tile)
#pragma unroll
        for (int P = 0; P < NP; ++P) {
            f32x16 cn[3];
            if (P + 1 < NP) {
                outp(P + 1, hb, cn);
            } else {
                hidden(hb);
                outp(0, hb, cn);
            }
            float u[48];
#pragma unroll
            for (int q = 0; q < 48; ++q) u[q] = cc[q >> 4][q & 15];
            double jac;
            const double y = INV ? rqs_inverse_j<K>(u, a.shift, xb[P], jac) : rqs_forward_j<K>(u, a.shift, xb[P], jac);
            const bool keep = pv && bidx[P] >= 0;
            if (keep) xp[bidx[P]] = y;
            jprod *= keep ? jac : 1.0;
#pragma unroll
            for (int tt = 0; tt < 3; ++tt) cc[tt] = cn[tt];
            // one step's operands live at a time
            __builtin_amdgcn_sched_barrier(0);
        }
        double ldsum = log(jprod);
        ldsum += __shfl_xor(ldsum, 32);
        if (pv && h == 0) a.lq[row] = lq0 + ldsum;
    }
}

template <class T>
__global__ void k_fill(int64_t n, T v, T* __restrict__ x) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = v;
}

}  // namespace chili_nf
