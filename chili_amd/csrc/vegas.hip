// vegas.hip -- plain VEGAS (Vegas::operator(), src/Integrator/Vegas.cc:18-52)
// on gfx950: uniforms -> grid map and Jacobian -> f x jac -> StatsData and the
// per-dimension f^2 histogram, partials per workgroup, fixed-order reduce.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

#include "common.hpp"
#include "mc_kernels.hpp"

using namespace chili_dev;

namespace {

constexpr int kVTile = 256;
constexpr int kVMaxDims = 16;

// e+e- -> mu+mu- line shape on [0,1]^2; params [MZ, GZ, sin^2 thetaW, alpha,
// smin, smax]: x0 -> s-hat by the Breit-Wigner map, x1 -> cos theta.
__device__ __forceinline__ double eemumu(const chili_integrand* __restrict__ f, const double* x) {
    const double MZ = f->params[0], GZ = f->params[1], sw2 = f->params[2], al = f->params[3], smin = f->params[4], smax = f->params[5];
    const double s = massive_prop_momenta(MZ, GZ, smin, smax, x[0]);
    double dummy;
    const double wbw = massive_prop_weight(MZ, GZ, smin, smax, s, dummy);
    const double ct = 2. * x[1] - 1.;
    const double cw2 = 1. - sw2;
    const double kap = 1. / (4. * sw2 * cw2);
    const double ve = -0.5 + 2. * sw2, ae = -0.5;
    const double dm = s - MZ * MZ;
    const double den = dm * dm + MZ * MZ * GZ * GZ;
    const double rechi = kap * s * dm / den;
    const double abschi2 = kap * kap * s * s / den;
    const double A0 = 1. + 2. * ve * ve * rechi + (ae * ae + ve * ve) * (ae * ae + ve * ve) * abschi2;
    const double A1 = 4. * ae * ae * rechi + 8. * ve * ae * ve * ae * abschi2;
    const double gev2pb = 0.3893793721e9;
    const double dsig = kPi * al * al / (2. * s) * (A0 * (1. + ct * ct) + A1 * ct) * gev2pb;
    return wbw * 2. * dsig;
}

struct VArgs {
    const double* edges;    // [d][B+1]
    const double* search;   // [d][P2]
    const chili_integrand* f;
    const double* u;        // host-fed uniforms on device or null
    int64_t ustride;
    uint64_t ua0, seed, a0;
    uint32_t stream;
    int d, B, P2;
};

__device__ __forceinline__ double rlane(double v, int src) {
    const uint64_t u = __double_as_longlong(v);
    const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)u, src);
    const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(u >> 32), src);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}
__device__ __forceinline__ double wsum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// Partial record: [n, min, max, sum, sum2, nf, hist d x B]
template <int D>
__global__ __launch_bounds__(kVTile) void k_vegas(VArgs A, int64_t n, double* __restrict__ part, int64_t rec,
                                                  double* __restrict__ vals, int64_t vals_cap) {
    extern __shared__ double s_mem[];
    double* s_hist = s_mem;                    // [D][B]
    double* s_small = s_hist + D * A.B;        // [4][8]
    short* s_bins = (short*)(s_small + 32);    // [D][kVTile]
    const int64_t i = (int64_t)blockIdx.x * kVTile + threadIdx.x;
    const bool valid = i < n;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    double val = 0;
    if (valid) {
        const uint64_t att = A.a0 + (uint64_t)i;
        double r[D + 1];
        if (A.u) {
#pragma unroll
            for (int k = 0; k < D; ++k) r[k] = A.u[(int64_t)k * A.ustride + (int64_t)(att - A.ua0)];
        } else {
#pragma unroll
            for (int k = 0; k < D; k += 2) {
                double u0, u1;
                philox_pair(A.seed, A.stream, att, (uint32_t)k, u0, u1);
                r[k] = u0;
                if (k + 1 < D) r[k + 1] = u1;
            }
        }
        // AdaptiveMap::operator(), AdaptiveMap.cc:40-56
        double jac = 1.0;
        const double Bd = (double)A.B;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const double pos = r[k] * Bd;
            const int idx = (int)pos;
            const double loc = pos - (double)idx;
            const double* e = A.edges + k * (A.B + 1);
            const double left = e[idx];
            const double size = e[idx + 1] - e[idx];
            r[k] = left + loc * size;
            jac *= size * Bd;
        }
        val = eemumu(A.f, r) * jac;
#pragma unroll
        for (int k = 0; k < D; ++k) s_bins[k * kVTile + threadIdx.x] = (short)find_bin(A.search + k * A.P2, A.P2, r[k]);
        if (vals && i < vals_cap) vals[i] = val;
    }
    const double v2 = val * val;
    {
        double nn = wsum(valid ? 1.0 : 0.0), s1 = wsum(val), s2 = wsum(v2), nf = wsum((valid && val != 0) ? 1.0 : 0.0);
        double mn = valid ? val : std::numeric_limits<double>::max(), mx = valid ? val : std::numeric_limits<double>::min();
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            double a = __shfl_xor(mn, off), b = __shfl_xor(mx, off);
            mn = (a < mn) ? a : mn;
            mx = (mx < b) ? b : mx;
        }
        if (lane == 0) {
            double* q = s_small + w * 8;
            q[0] = nn; q[1] = mn; q[2] = mx; q[3] = s1; q[4] = s2; q[5] = nf;
        }
    }
    __syncthreads();
    double* out = part + (int64_t)blockIdx.x * rec;
    if (threadIdx.x == 0) {
        double a[6] = {0, std::numeric_limits<double>::max(), std::numeric_limits<double>::min(), 0, 0, 0};
        for (int ww = 0; ww < 4; ++ww) {
            const double* q = s_small + ww * 8;
            a[0] += q[0]; a[1] = (q[1] < a[1]) ? q[1] : a[1]; a[2] = (a[2] < q[2]) ? q[2] : a[2]; a[3] += q[3]; a[4] += q[4]; a[5] += q[5];
        }
        for (int k = 0; k < 6; ++k) out[k] = a[k];
    }
    // f^2 histogram: lane l owns bins l, l+64, ..., wave values in lane order,
    // waves added one after the other.
    const int B = A.B;
    for (int k = 0; k < D; ++k) {
        const int bk = valid ? s_bins[k * kVTile + threadIdx.x] : 0;
        const double vk = valid ? v2 : 0.0;
        double h[4] = {0, 0, 0, 0};
        for (int src = 0; src < 64; ++src) {
            const int b = __builtin_amdgcn_readlane(bk, src);
            const double v = rlane(vk, src);
            if ((b & 63) == lane) {
                const int hi = b >> 6;
                if (hi == 0) h[0] += v; else if (hi == 1) h[1] += v; else if (hi == 2) h[2] += v; else h[3] += v;
            }
        }
        for (int ww = 0; ww < 4; ++ww) {
            if (w == ww) {
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int b = lane + 64 * q;
                    if (b < B) s_hist[k * B + b] = (ww == 0) ? h[q] : s_hist[k * B + b] + h[q];
                }
            }
            __syncthreads();
        }
    }
    for (int j = threadIdx.x; j < D * B; j += kVTile) out[6 + j] = s_hist[j];
}

__global__ __launch_bounds__(256) void k_vreduce(int ntile, int DB, const double* __restrict__ part, int64_t rec, double* __restrict__ tot) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        double a[6] = {tot[0], tot[1], tot[2], tot[3], tot[4], tot[5]};
        for (int t = 0; t < ntile; ++t) {
            const double* q = part + (int64_t)t * rec;
            a[0] += q[0]; a[1] = (q[1] < a[1]) ? q[1] : a[1]; a[2] = (a[2] < q[2]) ? q[2] : a[2]; a[3] += q[3]; a[4] += q[4]; a[5] += q[5];
        }
        for (int k = 0; k < 6; ++k) tot[k] = a[k];
        return;
    }
    const int64_t h = i - 1;
    if (h >= DB) return;
    double s = tot[6 + h];
    for (int t = 0; t < ntile; ++t) s += part[(int64_t)t * rec + 6 + h];
    tot[6 + h] = s;
}

// Allocation-free iteration (chili_vegas_iterate_static): the running
// totals and the per-workgroup partials live in this file-scope device array
// (one copy per GPU, held until the process ends); every call on a GPU shares
// it, so calls that overlap on one GPU race.  Layout: totals [0, kStaticTot),
// partials from kStaticTot on, ntile x rec doubles per launch chunk.
constexpr int64_t kStaticDoubles = int64_t(1) << 19;      // 4 MiB
constexpr int64_t kStaticTot = 1024;                      // >= 6 + d B
static __device__ __align__(256) double scratch[kStaticDoubles];

// totals to StatsData's start values (no memset: callers may capture the
// launch in a graph)
__global__ void k_vtot_init(int64_t n, double* __restrict__ tot) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    tot[i] = (i == 1) ? std::numeric_limits<double>::max() : (i == 2) ? std::numeric_limits<double>::min() : 0.0;
}

}  // namespace

struct chili_vegas {
    int device = 0;
    hipStream_t stream = nullptr;
    int d = 0, B = 0, P2 = 0;
    chili_integrand f{};
    std::vector<double> edges;
    DevBuf<double> d_edges, d_search, d_part, d_tot, d_u;
    DevBuf<chili_integrand> d_f;
    int64_t cap_tiles = 0;
};

namespace {
void vupload(chili_vegas* v) {
    std::vector<double> s((size_t)v->d * v->P2, std::numeric_limits<double>::infinity());
    for (int k = 0; k < v->d; ++k)
        for (int b = 1; b < v->B; ++b) s[(size_t)k * v->P2 + b - 1] = v->edges[(size_t)k * (v->B + 1) + b];
    v->d_edges.upload(v->edges, v->stream);
    v->d_search.upload(s, v->stream);
}
// One launch over n points from attempt A.a0: partials per workgroup into
// part, then added tile by tile, in tile order, to the totals in tot.
template <int D>
void vlaunch(chili_vegas* v, const VArgs& A, int64_t n, double* vals, int64_t vals_cap, double* part, double* tot) {
    const int grid = (int)((n + kVTile - 1) / kVTile);
    const int64_t rec = 6 + (int64_t)D * v->B;
    const size_t lds = sizeof(double) * (D * v->B + 32) + sizeof(short) * D * kVTile;
    hipLaunchKernelGGL(k_vegas<D>, dim3(grid), dim3(kVTile), lds, v->stream, A, n, part, rec, vals, vals_cap);
    const int64_t nout = 1 + (int64_t)D * v->B;
    hipLaunchKernelGGL(k_vreduce, dim3((unsigned)((nout + 255) / 256)), dim3(256), 0, v->stream, grid, D * v->B, part, rec, tot);
}

// Device address of this GPU's scratch array, looked up once per device.
double* static_scratch(int device) {
    static std::mutex mu;
    static std::map<int, double*> ptrs;
    std::lock_guard<std::mutex> lk(mu);
    auto it = ptrs.find(device);
    if (it != ptrs.end()) return it->second;
    void* p = nullptr;
    HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    ptrs[device] = (double*)p;
    return (double*)p;
}
}  // namespace

extern "C" {

int chili_vegas_create(int ndims, int nbins, const chili_integrand* f, int device, chili_vegas** out) {
    return guarded([&] {
        if (!f || !out) throw chili_error(CHILI_EINVAL, "null argument");
        if (f->kind != CHILI_F_EEMUMU) throw chili_error(CHILI_EINVAL, "plain VEGAS takes a hypercube integrand (CHILI_F_EEMUMU)");
        if (ndims != 2) throw chili_error(CHILI_ERANGE, "CHILI_F_EEMUMU is two-dimensional");
        if (nbins < 2 || nbins > 256) throw chili_error(CHILI_ERANGE, "nbins must be in [2, 256]");
        select_device(device);
        auto v = std::make_unique<chili_vegas>();
        v->device = device;
        HIP_CHECK(hipStreamCreateWithFlags(&v->stream, hipStreamNonBlocking));
        v->d = ndims;
        v->B = nbins;
        int p = 1;
        while (p < nbins) p <<= 1;
        v->P2 = p;
        v->f = *f;
        v->edges.resize((size_t)ndims * (nbins + 1));
        for (int k = 0; k < ndims; ++k)
            for (int b = 0; b <= nbins; ++b) v->edges[(size_t)k * (nbins + 1) + b] = (double)b / (double)nbins;
        std::vector<chili_integrand> fv(1, *f);
        v->d_f.upload(fv, v->stream);
        vupload(v.get());
        HIP_CHECK(hipStreamSynchronize(v->stream));
        *out = v.release();
    });
}

int chili_vegas_destroy(chili_vegas* v) {
    return guarded([&] {
        if (!v) return;
        select_device(v->device);
        hipStream_t s = v->stream;
        if (s) hipStreamSynchronize(s);
        delete v;
        if (s) hipStreamDestroy(s);
    });
}

int chili_vegas_set_grid(chili_vegas* v, const double* edges, int nbins) {
    return guarded([&] {
        if (!v || !edges) throw chili_error(CHILI_EINVAL, "null argument");
        if (nbins < 2 || nbins > 256) throw chili_error(CHILI_ERANGE, "nbins must be in [2, 256]");
        // validate before any state changes: a rejected grid leaves the
        // context as it was
        for (int k = 0; k < v->d; ++k)
            for (int b = 0; b < nbins; ++b)
                if (!(edges[(size_t)k * (nbins + 1) + b] <= edges[(size_t)k * (nbins + 1) + b + 1]))
                    throw chili_error(CHILI_EINVAL, "grid edges must be non-decreasing");
        select_device(v->device);
        v->B = nbins;
        int p = 1;
        while (p < nbins) p <<= 1;
        v->P2 = p;
        v->edges.assign(edges, edges + (size_t)v->d * (nbins + 1));
        v->cap_tiles = 0;
        vupload(v);
    });
}

int chili_vegas_get_grid(const chili_vegas* v, double* edges, int* nbins) {
    return guarded([&] {
        if (!v || !edges) throw chili_error(CHILI_EINVAL, "null argument");
        std::copy(v->edges.begin(), v->edges.end(), edges);
        if (nbins) *nbins = v->B;
    });
}

// ncalls points from attempt a0 (Vegas.cc:30-42, no retries).  hist (host,
// d x B) receives the f^2 histogram; vals_dev (device, optional) the values.
int chili_vegas_iterate(chili_vegas* v, const chili_uniforms* uni, uint64_t a0, int64_t ncalls, chili_stats* stats, double* hist,
                        double* vals_dev, int64_t vals_cap) {
    return guarded([&] {
        if (!v || !uni || !stats) throw chili_error(CHILI_EINVAL, "null argument");
        if (ncalls <= 0 || ncalls > (int64_t(1) << 31)) throw chili_error(CHILI_ERANGE, "ncalls must be in [1, 2^31]");
        select_device(v->device);
        VArgs A{};
        A.edges = v->d_edges.ptr;
        A.search = v->d_search.ptr;
        A.f = v->d_f.ptr;
        A.seed = uni->seed;
        A.stream = uni->stream;
        A.a0 = a0;
        A.ua0 = a0;
        A.d = v->d;
        A.B = v->B;
        A.P2 = v->P2;
        if (uni->u) {
            if (uni->stride < ncalls) throw chili_error(CHILI_ERANGE, "host uniforms shorter than ncalls");
            const size_t cnt = (size_t)v->d * uni->stride;
            for (size_t i = 0; i < cnt; ++i)
                if (!(uni->u[i] >= 0.0 && uni->u[i] < 1.0)) throw chili_error(CHILI_EINVAL, "host uniforms must lie in [0, 1)");
            v->d_u.resize(cnt);
            HIP_CHECK(hipMemcpyAsync(v->d_u.ptr, uni->u, cnt * sizeof(double), hipMemcpyHostToDevice, v->stream));
            A.u = v->d_u.ptr;
            A.ustride = uni->stride;
        }
        const int64_t ntile = (ncalls + kVTile - 1) / kVTile;
        const int64_t rec = 6 + (int64_t)v->d * v->B;
        if (ntile > v->cap_tiles) {
            v->d_part.resize((size_t)ntile * rec);
            v->cap_tiles = ntile;
        }
        std::vector<double> tot(6 + (size_t)v->d * v->B, 0.0);
        tot[1] = std::numeric_limits<double>::max();
        tot[2] = std::numeric_limits<double>::min();
        v->d_tot.upload(tot, v->stream);
        if (v->d == 2) vlaunch<2>(v, A, ncalls, vals_dev, vals_cap, v->d_part.ptr, v->d_tot.ptr);
        else throw chili_error(CHILI_ERANGE, "unsupported dimension");
        HIP_CHECK(hipGetLastError());
        v->d_tot.download(tot.data(), tot.size(), v->stream);
        HIP_CHECK(hipStreamSynchronize(v->stream));
        stats->n = tot[0]; stats->min = tot[1]; stats->max = tot[2]; stats->sum = tot[3]; stats->sum2 = tot[4]; stats->n_finite = tot[5];
        if (hist) std::copy(tot.begin() + 6, tot.end(), hist);
    });
}

// chili_vegas_iterate without any device allocation: device-generated
// uniforms only; totals and partials in this GPU's static scratch array, the
// points processed in chunks of at most (kStaticDoubles - kStaticTot) / rec
// workgroups (the totals are summed in the same tile order as one launch, so
// the results are chili_vegas_iterate's bit for bit).
int chili_vegas_iterate_static(chili_vegas* v, const chili_uniforms* uni, uint64_t a0, int64_t ncalls, chili_stats* stats,
                               double* hist, double* vals_dev, int64_t vals_cap) {
    return guarded([&] {
        if (!v || !uni || !stats) throw chili_error(CHILI_EINVAL, "null argument");
        if (uni->u) throw chili_error(CHILI_EINVAL, "chili_vegas_iterate_static takes device-generated (Philox) uniforms only");
        if (ncalls <= 0 || ncalls > (int64_t(1) << 31)) throw chili_error(CHILI_ERANGE, "ncalls must be in [1, 2^31]");
        const int64_t rec = 6 + (int64_t)v->d * v->B;
        if (rec > kStaticTot) throw chili_error(CHILI_ERANGE, "dims x bins too large for the static scratch (6 + d B <= 1024)");
        select_device(v->device);
        double* base = static_scratch(v->device);
        double* tot = base;
        double* part = base + kStaticTot;
        const int64_t chunk = ((kStaticDoubles - kStaticTot) / rec) * kVTile;   // points per launch
        VArgs A{};
        A.edges = v->d_edges.ptr;
        A.search = v->d_search.ptr;
        A.f = v->d_f.ptr;
        A.seed = uni->seed;
        A.stream = uni->stream;
        A.d = v->d;
        A.B = v->B;
        A.P2 = v->P2;
        hipLaunchKernelGGL(k_vtot_init, dim3(4), dim3(256), 0, v->stream, rec, tot);
        for (int64_t off = 0; off < ncalls; off += chunk) {
            const int64_t n = std::min(chunk, ncalls - off);
            A.a0 = a0 + (uint64_t)off;
            A.ua0 = A.a0;
            double* vp = (vals_dev && off < vals_cap) ? vals_dev + off : nullptr;
            const int64_t vc = vp ? vals_cap - off : 0;
            if (v->d == 2) vlaunch<2>(v, A, n, vp, vc, part, tot);
            else throw chili_error(CHILI_ERANGE, "unsupported dimension");
        }
        HIP_CHECK(hipGetLastError());
        double head[6];
        HIP_CHECK(hipMemcpyAsync(head, tot, sizeof head, hipMemcpyDeviceToHost, v->stream));
        if (hist) HIP_CHECK(hipMemcpyAsync(hist, tot + 6, sizeof(double) * (size_t)(rec - 6), hipMemcpyDeviceToHost, v->stream));
        HIP_CHECK(hipStreamSynchronize(v->stream));
        stats->n = head[0]; stats->min = head[1]; stats->max = head[2]; stats->sum = head[3]; stats->sum2 = head[4];
        stats->n_finite = head[5];
    });
}

void chili_vegas_params_default(chili_vegas_params* p) {
    if (!p) return;
    *p = chili_vegas_params{};
    p->ncalls = 10000;          // VegasParams defaults, Vegas.hh:35-36
    p->nrefine = 5;
    p->rtol = 1e-4;
    p->atol = 1e-4;
    p->alpha = 1.0;
    p->niterations = 10;
    p->irefine = 0;
}

// Vegas::Optimize (src/Integrator/Vegas.cc:54-77) with Vegas::operator()'s
// Adapt (Vegas.cc:49) and Refine (Vegas.cc:79-82); iteration k draws Philox
// stream uni->stream + k.
int chili_vegas_optimize(chili_vegas* v, const chili_uniforms* uni, chili_vegas_params* p, chili_stats* sum, int64_t* nresults,
                         chili_stats* results, int64_t results_cap, int64_t max_steps) {
    return guarded([&] {
        if (!v || !uni || !p || !sum || !nresults) throw chili_error(CHILI_EINVAL, "null argument");
        if (uni->u) throw chili_error(CHILI_EINVAL, "chili_vegas_optimize draws Philox uniforms (uni->u must be NULL)");
        if (max_steps < 1) throw chili_error(CHILI_EINVAL, "max_steps must be >= 1 (Vegas::Optimize need not end)");
        double abs_err = std::numeric_limits<double>::max(), rel_err = std::numeric_limits<double>::max();
        int64_t steps = 0;
        std::vector<double> hist, edges;
        while (((abs_err > p->atol && rel_err > p->rtol) || *nresults < p->niterations) && steps < max_steps) {
            const int B = v->B, d = v->d;
            if (p->irefine + 1 == p->nrefine && 2 * B > 256)
                throw chili_error(CHILI_ERANGE, "refinement would split the grid beyond 256 bins");
            const int64_t k = *nresults;
            chili_uniforms u = *uni;
            u.stream = uni->stream + (uint32_t)k;
            chili_stats st{};
            hist.assign((size_t)d * B, 0.0);
            int rc = chili_vegas_iterate(v, &u, 0, p->ncalls, &st, hist.data(), nullptr, 0);
            if (rc != CHILI_OK) throw chili_error(rc, chili_last_error());
            // grid.Adapt(alpha, train_data), unconditional (Vegas.cc:49)
            edges = v->edges;
            rc = chili_grid_adapt(d, B, edges.data(), p->alpha, hist.data());
            if (rc != CHILI_OK) throw chili_error(rc, chili_last_error());
            rc = chili_vegas_set_grid(v, edges.data(), B);
            if (rc != CHILI_OK) throw chili_error(rc, chili_last_error());
            // summary (Vegas.cc:50-51): StatsData::operator+=(StatsData)
            sum->n += st.n;
            sum->n_finite += st.n_finite;
            sum->sum += st.sum;
            sum->sum2 += st.sum2;
            sum->min = std::min(sum->min, st.min);
            sum->max = std::max(sum->max, st.max);
            if (results && k < results_cap) results[k] = st;
            *nresults = k + 1;
            const double mean = sum->sum / sum->n;
            abs_err = std::sqrt((sum->sum2 / sum->n - mean * mean) / (sum->n - 1));
            rel_err = abs_err / std::abs(mean);
            ++steps;
            if (++p->irefine == p->nrefine) {      // Refine: Split in two, ncalls *= 2
                std::vector<double> split((size_t)d * (2 * B + 1));
                rc = chili_grid_split(d, B, v->edges.data(), 2, split.data());
                if (rc != CHILI_OK) throw chili_error(rc, chili_last_error());
                rc = chili_vegas_set_grid(v, split.data(), 2 * B);
                if (rc != CHILI_OK) throw chili_error(rc, chili_last_error());
                p->ncalls *= 2;
                p->irefine = 0;
            }
        }
    });
}

}  // extern "C"
