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
    __shared__ bf16 sA[kGT][kGK + 8];
    __shared__ bf16 sB[kGT][kGK + 8];
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

struct SplineBin {
    int k;
    double X0, Y0, w, h, d0, d1;
};

// Bin containing v (searched over X, or over Y when by_y); ew/eh receive the
// exponentials of the logits (max subtracted), rw/rh the reciprocals of their sums.
template <int K>
__device__ __forceinline__ SplineBin spline_bin(const float* __restrict__ u, float shift, double v, bool by_y, float* ew, float* eh,
                                                float& rw, float& rh) {
    float mw = u[0], mh = u[K];
#pragma unroll
    for (int i = 1; i < K; ++i) { mw = fmaxf(mw, u[i]); mh = fmaxf(mh, u[K + i]); }
    // softmax exponentials on the hardware exp2 (arguments <= 0, no overflow;
    // relative error a few ulp, which only moves the knots: the density stays
    // normalised)
    constexpr float kLog2e = 1.4426950408889634f;
    float sw = 0.f, sh = 0.f;
#pragma unroll
    for (int i = 0; i < K; ++i) {
        ew[i] = __builtin_amdgcn_exp2f((u[i] - mw) * kLog2e); sw += ew[i];
        eh[i] = __builtin_amdgcn_exp2f((u[K + i] - mh) * kLog2e); sh += eh[i];
    }
    rw = 1.f / sw;
    rh = 1.f / sh;
    const float cw = (float)(1.0 - K * kMinW), ch = (float)(1.0 - K * kMinH);
    float X = 0.f, Y = 0.f, X0 = 0.f, Y0 = 0.f, X1 = 1.f, Y1 = 1.f;
    int k = 0;
#pragma unroll
    for (int i = 0; i < K; ++i) {
        const float Xn = (i == K - 1) ? 1.f : X + ((float)kMinW + cw * (ew[i] * rw));
        const float Yn = (i == K - 1) ? 1.f : Y + ((float)kMinH + ch * (eh[i] * rh));
        // compared in FP64 (the knots are exact there), so that the bin of v
        // agrees with the FP64 transform: forward and inverse round-trip
        const bool in = (i == 0) || (by_y ? (v >= (double)Y) : (v >= (double)X));
        k = in ? i : k;                  // selects, not branches
        X0 = in ? X : X0;
        Y0 = in ? Y : Y0;
        X1 = in ? Xn : X1;
        Y1 = in ? Yn : Y1;
        X = Xn;
        Y = Yn;
    }
    SplineBin b;
    b.k = k;
    b.X0 = (double)X0;
    b.Y0 = (double)Y0;
    b.w = (double)X1 - b.X0;
    b.h = (double)Y1 - b.Y0;
    b.d0 = (k == 0) ? 1.0 : kMinD + (double)softplusf(u[2 * K + k - 1] + shift);
    b.d1 = (k == K - 1) ? 1.0 : kMinD + (double)softplusf(u[2 * K + k] + shift);
    return b;
}

// Forward transform x -> y, logdet = log dy/dx.
template <int K>
__device__ __forceinline__ double rqs_forward(const float* __restrict__ u, float shift, double x, double& logdet) {
    float ew[K], eh[K], rw, rh;
    x = fmin(fmax(x, 0.0), 1.0);
    const SplineBin b = spline_bin<K>(u, shift, x, false, ew, eh, rw, rh);
    const double iw = 1.0 / b.w;
    const double xi = (x - b.X0) * iw, s = b.h * iw, t = xi * (1 - xi);
    const double iden = 1.0 / (s + (b.d1 + b.d0 - 2 * s) * t);
    const double D = b.d1 * xi * xi + 2 * s * t + b.d0 * (1 - xi) * (1 - xi);
    logdet = log(s * s * D * iden * iden);
    return b.Y0 + b.h * (s * xi * xi + b.d0 * t) * iden;
}

// Inverse transform y -> x; logdet receives log dy/dx at that x.
template <int K>
__device__ __forceinline__ double rqs_inverse(const float* __restrict__ u, float shift, double y, double& logdet) {
    float ew[K], eh[K], rw, rh;
    y = fmin(fmax(y, 0.0), 1.0);
    const SplineBin b = spline_bin<K>(u, shift, y, true, ew, eh, rw, rh);
    const double s = b.h / b.w, dy = y - b.Y0, E = b.d1 + b.d0 - 2 * s;
    const double a = b.h * (s - b.d0) + dy * E;
    const double bq = b.h * b.d0 - dy * E;
    const double c = -s * dy;
    const double disc = fmax(bq * bq - 4 * a * c, 0.0);
    double xi = (2 * c) / (-bq - sqrt(disc));
    xi = fmin(fmax(xi, 0.0), 1.0);
    const double t = xi * (1 - xi);
    const double iden = 1.0 / (s + E * t);
    const double D = b.d1 * xi * xi + 2 * s * t + b.d0 * (1 - xi) * (1 - xi);
    logdet = log(s * s * D * iden * iden);
    return b.X0 + xi * b.w;
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
// transformed dimensions (96 rows = 2 x 48, 47 used), staged through LDS
// ([point][row], pitch 97 floats) and each lane then runs the FP64 spline of
// one (point, dimension).
// ===========================================================================
constexpr int kHid = 64;
constexpr int kPairRows = 96;
constexpr int kStagePitch = 97;
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

// bias + ReLU of two 32-feature tiles, packed as the four B fragments of the
// next layer
__device__ __forceinline__ void relu_pack(const f32x16& a0, const f32x16& a1, const float* __restrict__ bias, int h, bf16x8* hb) {
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const f32x16& a = (s < 2) ? a0 : a1;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int r = 8 * (s & 1) + j;
            float v = a[r] + bias[phi_k(s, h, j)];
            hb[s][j] = f2bf(v > 0.f ? v : 0.f);
        }
    }
}

// NP = pairs of transformed dimensions (ceil(nB / 2)), a compile-time
// constant so that the per-lane operands of a tile live in registers; the
// next tile's inputs are fetched while the current one computes.
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
    float* sStage = sB + ((FL.bel + 3) / 4) * 4 + wave * 32 * kStagePitch;
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
    int aidx[8];                         // conditioning dims this lane feeds to layer 1
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const int k = 8 * h + j;
        aidx[j] = k < a.nA ? a.A[k] : -1;
    }
    int bidx[NP];                        // transformed dim of this lane in pair P (or -1)
#pragma unroll
    for (int P = 0; P < NP; ++P) bidx[P] = (2 * P + h < a.nB) ? a.B[2 * P + h] : -1;
    // inputs of a tile: conditioning values, transformed values, running log q
    double nxa[8], nxb[NP], nlq = 0.0;
    auto fetch = [&](int64_t t) {
        const int64_t row = start + t * 32 + pl;
        const bool pv = t * 32 + pl < cnt;
        const double* xp = a.x + row * a.d;
#pragma unroll
        for (int j = 0; j < 8; ++j) nxa[j] = (pv && aidx[j] >= 0) ? xp[aidx[j]] : 0.0;
#pragma unroll
        for (int P = 0; P < NP; ++P) nxb[P] = (pv && bidx[P] >= 0) ? xp[bidx[P]] : 0.0;
        nlq = (pv && h == 0) ? a.lq[row] : 0.0;
    };
    if (wave < ntile) fetch(wave);
    f32x16 z;
#pragma unroll
    for (int i = 0; i < 16; ++i) z[i] = 0.f;
    for (int64_t t = wave; t < ntile; t += nw) {
        const int64_t row = start + t * 32 + pl;
        const bool pv = t * 32 + pl < cnt;
        double* xp = a.x + row * a.d;
        double xb[NP];
#pragma unroll
        for (int P = 0; P < NP; ++P) xb[P] = nxb[P];
        const double lq0 = nlq;
        // layer-1 operand: lanes of half h hold conditioning features 8h..8h+7
        bf16x8 b0;
#pragma unroll
        for (int j = 0; j < 8; ++j) b0[j] = f2bf((float)nxa[j]);
        if (t + nw < ntile) fetch(t + nw);
        f32x16 c0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(lds_frag(sW + FL.oW1, lane), b0, z, 0, 0, 0);
        f32x16 c1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(lds_frag(sW + FL.oW1 + 512, lane), b0, z, 0, 0, 0);
        bf16x8 hb[4];
        relu_pack(c0, c1, sB + FL.ob1, h, hb);
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const bf16* Wq = sW + (q == 0 ? FL.oW2 : FL.oW3);
            c0 = z;
            c1 = z;
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                c0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(lds_frag(Wq + (0 * 4 + s) * 512, lane), hb[s], c0, 0, 0, 0);
                c1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(lds_frag(Wq + (1 * 4 + s) * 512, lane), hb[s], c1, 0, 0, 0);
            }
            relu_pack(c0, c1, sB + (q == 0 ? FL.ob2 : FL.ob3), h, hb);
        }
        // output layer and splines, one pair of transformed dimensions at a time
        double ldsum = 0.0;
#pragma unroll
        for (int P = 0; P < NP; ++P) {
            f32x16 c[3] = {z, z, z};
#pragma unroll
            for (int s = 0; s < 4; ++s)
#pragma unroll
                for (int tt = 0; tt < 3; ++tt)
                    c[tt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(lds_frag(sW + FL.oW4 + ((int64_t)(3 * P + tt) * 4 + s) * 512, lane),
                                                                    hb[s], c[tt], 0, 0, 0);
#pragma unroll
            for (int tt = 0; tt < 3; ++tt)
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const int rr = 32 * tt + 8 * (r >> 2) + 4 * h + (r & 3);
                    sStage[pl * kStagePitch + rr] = c[tt][r] + sB[FL.ob4 + kPairRows * P + rr];
                }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            if (pv && bidx[P] >= 0) {
                double ld;
                const float* u = sStage + pl * kStagePitch + 48 * h;
                xp[bidx[P]] = INV ? rqs_inverse<K>(u, a.shift, xb[P], ld) : rqs_forward<K>(u, a.shift, xb[P], ld);
                ldsum += ld;
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
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
