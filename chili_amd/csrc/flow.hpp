// flow.hpp -- host engine of the learned proposal (coupling-layer spline flows).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <utility>
#include <vector>

#include "common.hpp"
#include "flow_kernels.hpp"

namespace chili_nf {

// One coupling layer: x[A] conditions an MLP  nA -> H -> H -> H -> nB (3K-1)
// whose output parametrises the splines of x[B].  Parameter offsets index the
// flat fp32 parameter vector; W stored [in][out].
struct Layer {
    std::vector<int> A, B;
    int nA = 0, nB = 0, out = 0, ldA = 0;
    size_t oW1 = 0, ob1 = 0, oW2 = 0, ob2 = 0, oW3 = 0, ob3 = 0, oW4 = 0, ob4 = 0;
    size_t wo[4] = {0, 0, 0, 0};   // offsets of the four weights in the bf16 copies
    DevBuf<int> dA, dB;
};

struct Flow {
    DevBuf<float> params, grads, m, v;
    DevBuf<bf16> wb, wt;           // bf16 weights [in][out] and transposed [out][in]
    int64_t steps = 0;
};

struct Workspace {
    DevBuf<double> x;              // layer inputs x_0..x_L  [L+1][n][d]
    DevBuf<bf16> xa;               // conditioning inputs    [L][n][ldA]
    DevBuf<bf16> h;                // hidden activations     [L][3][n][H]
    DevBuf<float> raw;             // spline parameters      [L][n][out]
    DevBuf<float> draw, gf, dxa;
    DevBuf<bf16> drawb, gb;
    DevBuf<double> gx;
    DevBuf<double> ldt;            // per (point, transformed dim) log-Jacobians
};

// A side stream with its own workspace: a training step runs the per-flow
// passes (density with activations kept, backward) of several flows at once.
struct Lane {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    Workspace ws;
    DevBuf<float> part, part2;
};

struct FlowEngine {
    int device = 0;
    hipStream_t stream = nullptr;
    int d = 0, L = 0, H = 64, nflows = 0;
    float shift = 0;   // softplus offset: slope 1 at zero input
    double mix = 0;    // proposal q' = (1 - mix) q + mix (uniform density on the hypercube)
    int moment = 2;    // training loss: the moment-th moment of the weights under q' (2: variance)
    size_t nparams = 0;
    std::vector<Layer> layers;
    std::vector<Flow> flows;
    Workspace ws;
    DevBuf<float> part, part2;
    DevBuf<double> lqb, dlqb, lpart;
    DevBuf<double> ltb, lred;      // higher-moment loss: per-point terms and ratios, block maxima and sums
    // fused inference: swizzled weight blocks [flow][layer] and the work table
    DevBuf<bf16> fusedW;
    DevBuf<float> fusedB;
    int64_t fwstride = 0, fbstride = 0;
    DevBuf<int64_t> d_start, d_count;
    DevBuf<int> d_items;
    DevBuf<double> xwork;
    static constexpr int64_t kChunk = 4096;   // largest rows per work item of a fused pass
    // training lanes (created on first use, released with the engine)
    static constexpr int kTrainLanes = 4;
    std::vector<std::unique_ptr<Lane>> lanes;
    hipEvent_t ev_ready = nullptr;
    // k_couple timing (chili_flow_timing): HIP events around every launch,
    // on the stream it runs on; the rows of each timed fused pass are summed
    // on the device (no host synchronisation) into d_prows[slot]
    struct CoupleEv { hipEvent_t a, b; int slot; double flops_row; };
    bool timing = false;
    std::vector<CoupleEv> tev;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> tpool;
    DevBuf<int64_t> d_prows;
    int tslots = 0;
    static constexpr int kTimingSlots = 4096;
    double couple_ms = 0, couple_flops = 0;
    int64_t couple_launches = 0, couple_rows = 0;
    void harvest_timing();
    void release_timing();

    int ws_lda() const { int m = 0; for (auto& l : layers) m = m > l.ldA ? m : l.ldA; return m; }
    int ws_out() const { int m = 0; for (auto& l : layers) m = m > l.out ? m : l.out; return m; }

    void build_layers(uint64_t seed);
    void init_params(uint64_t seed);
    void refresh_bf16(int f);
    void ensure_ws(int64_t nb, bool keep);
    void mlp_forward(Flow& F, Layer& ly, int64_t nb, const double* x, bf16* xa, bf16* h, float* raw);
    void gemm_fwd(int64_t M, int N, int K, const bf16* A, int lda, const bf16* Bt, int ldb, const float* bias, int relu, float* C,
                  int ldc, bf16* Cb, int ldcb);
    void gemm_wgrad(int64_t n, int in, int out, const bf16* A, int lda, const bf16* G, int ldg, float* dW);
    void colsum(int64_t n, int m, const float* g, float* db);
    void density(int f, int64_t nb, const double* r, double* lq, double* zout, bool keep);
    void backward(int f, int64_t nb, const double* dlq);
    void log_density(int f, int64_t n, const double* r, double* lq, double* z);
    void sample(int f, int64_t n, double* zr, double* lq);
    // Fused passes over nfl flows (flow0 ...): rows of flow j at seg[j] ..
    // seg[j+1] (seg device ints) or j * stride .. + ucount; x in place, lq
    // accumulated; rows_bound >= total rows.
    void launch_couple(bool inverse, int npair, dim3 g, size_t lds, const CoupleArgs& a);
    void fused_pass(bool inverse, int nfl, int flow0, const int* seg, int64_t ucount, int64_t stride, int64_t rows_bound, double* x,
                    double* lq, const int* seg_end = nullptr, int64_t rows_hint = -1);
    double loss_grad(int64_t nb, const double* r, const double* beta, const double* w2);
    void adam_step(double lr);
    void ensure_lanes(int k);
    void release_lanes();
};

// Runs a flow engine's per-flow passes on a lane (its stream, workspace and
// split-K buffers) for a scope.
struct LaneSwap {
    FlowEngine& e;
    Lane& l;
    LaneSwap(FlowEngine& eng, Lane& lane) : e(eng), l(lane) { swap_all(); }
    ~LaneSwap() { swap_all(); }
    void swap_all() {
        std::swap(e.stream, l.stream);
        std::swap(e.ws, l.ws);
        std::swap(e.part, l.part);
        std::swap(e.part2, l.part2);
    }
};

// Runs a flow engine's work on another stream (same device) for a scope.
struct StreamSwap {
    FlowEngine& e;
    hipStream_t old;
    StreamSwap(FlowEngine& eng, hipStream_t s) : e(eng), old(eng.stream) { e.stream = s; }
    ~StreamSwap() { e.stream = old; }
};

}  // namespace chili_nf

struct chili_flow {
    chili_nf::FlowEngine eng;
};
