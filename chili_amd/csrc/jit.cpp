// jit.cpp -- run-time specialisation of the multichannel kernels.
//
// The generic kernels read each channel's program (slot indices, masses,
// cut constants) from device memory, so the per-point momentum slots are
// indexed at run time and live in private memory.  For one context the
// programs are fixed, so this file generates a kernel source in which every
// program is a compile-time constant (one constexpr function per channel),
// compiles it with hipRTC for gfx950 against the same device headers the
// library is built from, and loads the code object.  The arithmetic is the
// generic kernels' statement for statement (same -ffp-contract=off build);
// the compiler can then keep the slots in registers and fold the constants.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <type_traits>
#include <sstream>
#include <string>
#include <vector>

#include "chanprog.hpp"
#include "common.hpp"
#include "jit.hpp"

// Header texts embedded at build time (chili_amd/build.py writes
// jit_headers.inc): { include name, text } pairs.
#include "jit_headers.inc"

namespace chili_jit {

using chili_host::ChanProg;
using chili_host::kMaxDec;
using chili_host::kMaxT;

namespace {

std::string lit(double v) {
    if (std::isnan(v)) return "__builtin_nan(\"\")";
    if (std::isinf(v)) return v > 0 ? "__builtin_inf()" : "(-__builtin_inf())";
    char buf[64];
    std::snprintf(buf, sizeof buf, "%a", v);
    return buf;
}
template <class T, size_t N>
std::string arr(const T (&a)[N]) {
    std::ostringstream o;
    o << "{";
    for (size_t i = 0; i < N; ++i) {
        if (i) o << ", ";
        if constexpr (std::is_floating_point<T>::value) o << lit(a[i]);
        else o << a[i];
    }
    o << "}";
    return o.str();
}

}  // namespace

std::string kernel_source(const std::vector<ChanProg>& progs, int NT) {
    std::ostringstream s;
    s << "// generated: channel programs as compile-time constants\n"
      << "#include \"mc_device.hpp\"\n"
      << "using namespace chili_dev;\n"
      << "using chili_host::ChanProg;\n"
      << "namespace {\n"
      << "constexpr int NT = " << NT << ";\n";
    for (size_t j = 0; j < progs.size(); ++j) {
        const ChanProg& p = progs[j];
        s << "__host__ __device__ constexpr ChanProg chili_prog_" << j << "() { return ChanProg{" << p.nd << ", " << p.nt << ", 0, 0, "
          << arr(p.dec_par) << ", " << arr(p.dec_c1) << ", " << arr(p.dec_c2) << ", " << arr(p.dec_massive) << ", "
          << arr(p.dec_mass) << ", " << arr(p.dec_width) << ", " << arr(p.dec_smin2) << ", " << arr(p.dec_smincut) << ", "
          << arr(p.dec_smax) << ", " << arr(p.t_slot) << ", " << arr(p.t_ptmin) << ", " << arr(p.t_etacut) << ", "
          << arr(p.s_par) << ", " << arr(p.s_a) << ", " << arr(p.s_b) << ", 0}; }\n";
    }
    s << "#ifndef CHILI_INV_WAVES\n#define CHILI_INV_WAVES 2\n#endif\n";
    s << "struct JitLoop {\n"
      << "    static constexpr bool kPre = false;\n"
      << "    template <class F>\n"
      << "    __device__ __forceinline__ static void each(const MCArgs& A, F&& f) {\n";
    for (size_t j = 0; j < progs.size(); ++j) s << "        f(" << j << ", chili_prog_" << j << "());\n";
    s << "    }\n};\n}  // namespace\n\n";
    for (int ls = 0; ls < 2; ++ls)
        s << "extern \"C\" __global__ __launch_bounds__(kTile) void chili_k_main" << (ls ? "_lds" : "")
          << "(MCArgs A, const int* __restrict__ seg2,\n"
             "        const int* __restrict__ tile2, const uint32_t* __restrict__ comp, const double* __restrict__ mom, int64_t nmax,\n"
             "        double* __restrict__ psbuf, int64_t psstride, int train, double* __restrict__ part, int64_t rec,\n"
             "        double* __restrict__ vals_out, int64_t vals_off, int64_t vals_cap, const uint32_t* __restrict__ sorted, PointRec R) {\n"
             "    extern __shared__ double s_mem[];\n"
             "    int ch, e0, e1;\n"
             "    if (!locate_tile(A.C, seg2, tile2, blockIdx.x, ch, e0, e1, kMainTile)) return;\n"
             "    main_tile<NT, kRounds, JitLoop, " << (ls ? "true" : "false") << ">(A, ch, e0, e1, s_mem, blockIdx.x, psbuf, psstride,\n"
             "        train, comp, mom, nmax, part + (int64_t)blockIdx.x * rec, vals_out, vals_off, vals_cap, sorted, R);\n"
             "}\n\n";
    // split weight pass: one kernel per channel (its inverse map alone, so
    // that its momentum slots stay in registers)
    for (size_t j = 0; j < progs.size(); ++j)
        s << "extern \"C\" __global__ __launch_bounds__(kTile, CHILI_INV_WAVES) void chili_k_inv_" << j
          << "(MCArgs A, int64_t kept, const int* __restrict__ seg2,\n"
             "        const uint32_t* __restrict__ comp, const double* __restrict__ mom, int64_t nmax, double* __restrict__ pv,\n"
             "        short* __restrict__ bins, const uint32_t* __restrict__ sorted, PointRec R, int lds_search) {\n"
             "    extern __shared__ double s_srch[];\n"
             "    if (lds_search) {\n"
             "        const double* g = A.search + (int64_t)" << j << " * A.d * A.P2;\n"
             "        for (int i = threadIdx.x; i < A.d * A.P2; i += kTile) s_srch[i] = g[i];\n"
             "        __syncthreads();\n"
             "    }\n"
             "    const int64_t e = (int64_t)blockIdx.x * kTile + threadIdx.x;\n"
             "    if (e >= kept) return;\n"
             "    // two inlined copies, so that the LDS table's address space is known\n"
             "    // (ds_read) rather than resolved at run time (flat loads)\n"
             "    if (lds_search)\n"
             "        inverse_one<NT, true>(A, chili_prog_" << j << "(), " << j << ", e, seg2, comp, mom, nmax, pv, bins, sorted, R, s_srch);\n"
             "    else\n"
             "        inverse_one<NT, false>(A, chili_prog_" << j << "(), " << j << ", e, seg2, comp, mom, nmax, pv, bins, sorted, R, nullptr);\n"
             "}\n\n";
    // forward map of the attempts' uniforms (packed tiles when sorted_out is
    // given, as in the generic k_forward: passing attempts first, momenta at
    // the packed position).  The learned-proposal path keeps the generic
    // forward kernel: a specialised variant reading the proposal's points
    // measured slower in the production step.
    s << "extern \"C\" __global__ __launch_bounds__(kTile) void chili_k_forward(MCArgs A, const int* __restrict__ seg_start,\n"
         "        const int* __restrict__ tile_start, const uint32_t* __restrict__ sorted, double* __restrict__ mom, int64_t nmax,\n"
         "        uint8_t* __restrict__ acc, PointRec R, const double* __restrict__ zsrc, uint32_t* __restrict__ sorted_out) {\n"
         "    __shared__ int s_cnt[2 * kWaves];\n"
         "    int ch, e0, e1;\n"
         "    if (!locate_tile(A.C, seg_start, tile_start, blockIdx.x, ch, e0, e1)) return;\n"
         "    const int e = e0 + threadIdx.x;\n"
         "    const bool live = e < e1;\n"
         "    MomRF<2 * NT> P;\n"
         "    bool ok = false;\n"
         "    uint32_t a = 0;\n"
         "    if (live) {\n"
         "        a = sorted[e];\n"
         "        switch (ch) {\n";
    for (size_t j = 0; j < progs.size(); ++j)
        s << "            case " << j << ": ok = forward_point<NT, false>(A, chili_prog_" << j << "(), " << j
          << ", e, a, A.a0 + a, acc, R, zsrc, P); break;\n";
    s << "            default: break;\n"
         "        }\n"
         "    }\n"
         "    if (!sorted_out) {\n"
         "        if (live && ok) store_mom<NT>(mom, nmax, e, P);\n"
         "        return;\n"
         "    }\n"
         "    const int ep = tile_pack_pos(live, ok, e0, s_cnt);\n"
         "    if (live) {\n"
         "        sorted_out[ep] = a;\n"
         "        if (ok) store_mom<NT>(mom, nmax, ep, P);\n"
         "    }\n"
         "}\n\n";
    // exact early rejection of the unweighting stage: every passing point's
    // integrand and picked channel's term (switch on the channel; the
    // channels' search tables in LDS when they fit), the bound test on acc
    s << "extern \"C\" __global__ __launch_bounds__(kTile) void chili_k_own(MCArgs A, int64_t kept, const int* __restrict__ seg2,\n"
         "        const uint32_t* __restrict__ comp, const double* __restrict__ mom, int64_t nmax, const uint32_t* __restrict__ sorted,\n"
         "        uint8_t* __restrict__ acc, double wmax, const double* __restrict__ lqs, double mix, int lds_search) {\n"
         "    extern __shared__ double s_srch[];\n"
         "    if (lds_search) {\n"
         "        for (int i = threadIdx.x; i < A.C * A.d * A.P2; i += kTile) s_srch[i] = A.search[i];\n"
         "        __syncthreads();\n"
         "    }\n"
         "    const int64_t k = (int64_t)blockIdx.x * kTile + threadIdx.x;\n"
         "    if (k >= kept) return;\n"
         "    int pc = 0;\n"
         "    for (int c = 1; c < A.C; ++c) pc = (seg2[c] <= k) ? c : pc;\n"
         "    switch (pc) {\n";
    for (size_t j = 0; j < progs.size(); ++j)
        s << "        case " << j << ":\n"
             "            if (lds_search) own_bound_one<NT, true>(A, chili_prog_" << j << "(), " << j
          << ", k, comp, mom, nmax, sorted, acc, wmax, lqs, mix, s_srch + (int64_t)" << j << " * A.d * A.P2);\n"
             "            else own_bound_one<NT, false>(A, chili_prog_" << j << "(), " << j
          << ", k, comp, mom, nmax, sorted, acc, wmax, lqs, mix, nullptr);\n"
             "            break;\n";
    s << "        default: break;\n    }\n}\n\n";
    // learned-proposal path: one kernel per channel writes r_j and a_j of
    // every kept point (channel 0 also f), rows staged in LDS (pitch D + 1)
    // and written out contiguously, as the generic k_geom does per channel
    for (size_t j = 0; j < progs.size(); ++j)
        s << "extern \"C\" __global__ __launch_bounds__(kTile, CHILI_INV_WAVES) void chili_k_geom_" << j
          << "(MCArgs A, int64_t kept, const uint32_t* __restrict__ comp,\n"
             "        const double* __restrict__ mom, int64_t nmax, double* __restrict__ rall, double* __restrict__ abuf,\n"
             "        double* __restrict__ fbuf, int64_t stride, int lds_search) {\n"
             "    constexpr int D = 3 * (NT - 2) - 2;\n"
             "    constexpr int PR = D + 1;\n"
             "    extern __shared__ double s_mem[];   // rows [kTile][PR], then this channel's search table [D][P2]\n"
             "    double* s_r = s_mem;\n"
             "    double* s_srch = s_mem + kTile * PR;\n"
             "    if (lds_search) {\n"
             "        const double* g = A.search + (int64_t)" << j << " * A.d * A.P2;\n"
             "        for (int i = threadIdx.x; i < A.d * A.P2; i += kTile) s_srch[i] = g[i];\n"
             "        __syncthreads();\n"
             "    }\n"
             "    const int64_t k0 = (int64_t)blockIdx.x * kTile, k = k0 + threadIdx.x;\n"
             "    const int nrow = (int)(kept - k0 < kTile ? kept - k0 : kTile);\n"
             "    double* row = s_r + threadIdx.x * PR;\n"
             "    if (k < kept) {\n"
             "        MomRF<2 * NT> P;\n"
             "        load_mom<NT>(mom, nmax, comp[k], P);\n"
          << (j == 0 ? "        fbuf[k] = integrand<NT>(A.f, P);\n" : "")
          << "        if (lds_search)\n"
             "            geom_one<NT, true>(A, chili_prog_" << j << "(), " << j << ", k, P, row, abuf, stride, s_srch);\n"
             "        else\n"
             "            geom_one<NT, false>(A, chili_prog_" << j << "(), " << j << ", k, P, row, abuf, stride, nullptr);\n"
             "    }\n"
             "    __syncthreads();\n"
             "    double* dst = rall + ((int64_t)" << j << " * stride + k0) * D;\n"
             "    for (int e = threadIdx.x; e < nrow * D; e += kTile) dst[e] = s_r[(e / D) * PR + e % D];\n"
             "}\n\n";
    return s.str();
}

Module::~Module() {
    if (mod) (void)hipModuleUnload(mod);
}

void compile(const std::string& src, Module& out, double* seconds, bool load, int nchannels) {
    std::vector<const char*> hdr, names;
    for (int i = 0; i < kJitHeaderCount; ++i) {
        names.push_back(kJitHeaders[i].name);
        hdr.push_back(kJitHeaders[i].text);
    }
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, src.c_str(), "chili_specialised.hip", (int)hdr.size(), hdr.data(), names.data()) != HIPRTC_SUCCESS)
        throw chili_error(CHILI_EHIP, "hiprtcCreateProgram failed");
    const char* opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
                          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
    const auto t0 = std::chrono::steady_clock::now();
    const hiprtcResult rc = hiprtcCompileProgram(prog, (int)(sizeof opts / sizeof opts[0]), opts);
    if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (rc != HIPRTC_SUCCESS) {
        size_t n = 0;
        hiprtcGetProgramLogSize(prog, &n);
        std::string log(n, '\0');
        if (n) hiprtcGetProgramLog(prog, &log[0]);
        hiprtcDestroyProgram(&prog);
        throw chili_error(CHILI_EHIP, "run-time compilation failed: " + log.substr(0, 2000));
    }
    size_t n = 0;
    hiprtcGetCodeSize(prog, &n);
    std::vector<char> code(n);
    hiprtcGetCode(prog, code.data());
    hiprtcDestroyProgram(&prog);
    if (!load) return;
    HIP_CHECK(hipModuleLoadData(&out.mod, code.data()));
    HIP_CHECK(hipModuleGetFunction(&out.main, out.mod, "chili_k_main"));
    HIP_CHECK(hipModuleGetFunction(&out.main_lds, out.mod, "chili_k_main_lds"));
    HIP_CHECK(hipModuleGetFunction(&out.forward, out.mod, "chili_k_forward"));
    HIP_CHECK(hipModuleGetFunction(&out.own, out.mod, "chili_k_own"));
    out.inv.assign((size_t)nchannels, nullptr);
    out.geom.assign((size_t)nchannels, nullptr);
    for (int j = 0; j < nchannels; ++j) {
        HIP_CHECK(hipModuleGetFunction(&out.inv[(size_t)j], out.mod, ("chili_k_inv_" + std::to_string(j)).c_str()));
        HIP_CHECK(hipModuleGetFunction(&out.geom[(size_t)j], out.mod, ("chili_k_geom_" + std::to_string(j)).c_str()));
    }
}

}  // namespace chili_jit
