// gemm_probe.hip -- test-only library (tests/probes/_build/libchili_probe.so):
// runs the product's BF16 MFMA GEMM kernel (chili_amd/csrc/flow_kernels.hpp,
// k_gemm) on host buffers, so that a GPU test can pin the MFMA fragment
// layout the flow kernels rely on.  Not part of the product C-ABI.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../chili_amd/csrc/flow_kernels.hpp"

using namespace chili_nf;

extern "C" {

// C [M][N] fp32 = A B^T (A [M][K], B [N][K], transA = 0, transB = 1) or
// A^T B (A [K][M], B [K][N], transA = 1, transB = 0); bf16 bits in.
// Returns 0, -1 for a bad argument, -3 for a HIP error.
int chili_probe_gemm_bf16(int M, int N, int K, const uint16_t* A, const uint16_t* B, float* C, int transA, int transB) {
    if (M <= 0 || N <= 0 || K <= 0 || M > 4096 || N > 4096 || K > (1 << 16)) return -1;
    if (!((!transA && transB) || (transA && !transB))) return -1;
    const size_t na = (size_t)M * K, nb = (size_t)N * K, nc = (size_t)M * N;
    bf16 *a = nullptr, *b = nullptr;
    float* c = nullptr;
    int rc = 0;
    if (hipMalloc((void**)&a, na * sizeof(bf16)) != hipSuccess || hipMalloc((void**)&b, nb * sizeof(bf16)) != hipSuccess ||
        hipMalloc((void**)&c, nc * sizeof(float)) != hipSuccess)
        rc = -3;
    if (!rc && (hipMemcpy(a, A, na * sizeof(bf16), hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(b, B, nb * sizeof(bf16), hipMemcpyHostToDevice) != hipSuccess))
        rc = -3;
    if (!rc) {
        // grid covers the M x N output in kGT tiles; lda / ldb as stored
        dim3 g((N + kGT - 1) / kGT, (M + kGT - 1) / kGT, 1);
        const int lda = transA ? M : K, ldb = transB ? K : N;
        if (!transA)
            hipLaunchKernelGGL((k_gemm<false, true>), g, dim3(kGThreads), 0, 0, M, N, K, a, lda, b, ldb, c, N,
                               (const float*)nullptr, 0, (bf16*)nullptr, 0, K, (int64_t)0);
        else
            hipLaunchKernelGGL((k_gemm<true, false>), g, dim3(kGThreads), 0, 0, M, N, K, a, lda, b, ldb, c, N,
                               (const float*)nullptr, 0, (bf16*)nullptr, 0, K, (int64_t)0);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) rc = -3;
    }
    if (!rc && hipMemcpy(C, c, nc * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) rc = -3;
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (c) (void)hipFree(c);
    return rc;
}

}  // extern "C"
