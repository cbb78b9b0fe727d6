// mc_device.hpp -- device building blocks of the multichannel kernels: the
// per-attempt forward map and the fused all-channel weight tile.  The
// channel programs reach them either from device memory (GenericLoop) or as
// compile-time constants of a run-time specialised kernel (generated source,
// jit.cpp); both compile from this file.
#pragma once
#include "mc_kernels.hpp"

namespace chili_dev {

// Record layout of one workgroup partial: [n, min, max, sum, sum2, nf, nnan,
// T_0..T_{C-1}, hist d x B of the workgroup's channel].
constexpr int kStatW = 7;

__device__ __forceinline__ bool locate_tile(int C, const int* __restrict__ seg_start, const int* __restrict__ tile_start, int tile,
                                            int& ch, int& e0, int& e1, int tsize = kTile) {
    if (tile >= tile_start[C]) return false;
    int lo = 0, hi = C - 1;
    while (lo < hi) {            // last c with tile_start[c] <= tile
        int mid = (lo + hi + 1) >> 1;
        if (tile_start[mid] <= tile) lo = mid; else hi = mid - 1;
    }
    ch = lo;
    e0 = seg_start[lo] + (tile - tile_start[lo]) * tsize;
    e1 = min(e0 + tsize, seg_start[lo + 1]);
    return true;
}


__device__ __forceinline__ double readlane_f64(double v, int src) {
    const uint64_t u = __double_as_longlong(v);
    const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)u, src);
    const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(u >> 32), src);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}
__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) { double o = __shfl_xor(v, off); v = (o < v) ? o : v; }
    return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) { double o = __shfl_xor(v, off); v = (v < o) ? o : v; }
    return v;
}


// One attempt: uniforms (or the proposal's hypercube point), VEGAS forward
// map, FSMapper::GeneratePoint and the cut; returns the cut's decision and
// leaves the external momenta in P.  e: sorted position (row of zsrc), a:
// attempt slot of acc / per-attempt records, att: the attempt's uniform
// counter.
template <int NT, bool ZS>
__device__ __forceinline__ bool forward_point(const MCArgs& A, const ChanProg& p, int ch, int e, uint32_t a, uint64_t att,
                                              uint8_t* __restrict__ acc, const PointRec& R, const double* __restrict__ zsrc,
                                              MomRF<2 * NT>& P) {
    constexpr int D = 3 * (NT - 2) - 2;
    double r[D + 1];
    if (ZS) {
        // hypercube point drawn by the channel's learned proposal ([e][D])
#pragma unroll
        for (int i = 0; i < D; ++i) r[i] = zsrc[(int64_t)e * D + i];
        if (R.x)
#pragma unroll
            for (int i = 0; i < D; ++i) R.x[(int64_t)i * R.n + a] = r[i];
    } else {
        load_uniforms<D>(A, att, r);
    }
    // VEGAS forward map on the picked channel's grid (AdaptiveMap.cc:40-56);
    // the multichannel path drops its Jacobian (Integrand.hh:107-110).
    const double* __restrict__ edges = A.edges + (int64_t)ch * D * (A.B + 1);
    const double Bd = (double)A.B;
#pragma unroll
    for (int i = 0; i < D; ++i) {
        double pos = r[i] * Bd;
        int idx = min((int)pos, A.B - 1);   // r = 1 (reachable from a proposal) stays in the last bin
        double loc = pos - (double)idx;
        const double* ed = edges + i * (A.B + 1);
        double left = ed[idx];
        double size = ed[idx + 1] - ed[idx];
        r[i] = left + loc * size;
    }
    double m2[2 * NT];
    generate_point<NT>(A, p, r, P, m2);
    bool ok = cut_pass<NT>(A.f, P, A.sqrts);
    acc[a] = ok ? 1 : 0;
    if (R.status) {
        R.status[a] = ok ? 1 : 0;
        R.pick[a] = ch;
#pragma unroll
        for (int i = 0; i < NT; ++i) {
            R.mom[(int64_t)(i * 4 + 0) * R.n + a] = P.v[4 * i + 0];
            R.mom[(int64_t)(i * 4 + 1) * R.n + a] = P.v[4 * i + 1];
            R.mom[(int64_t)(i * 4 + 2) * R.n + a] = P.v[4 * i + 2];
            R.mom[(int64_t)(i * 4 + 3) * R.n + a] = P.v[4 * i + 3];
        }
    }
    return ok;
}

// External momenta of one point to column col of the SoA array [4 NT][nmax].
template <int NT>
__device__ __forceinline__ void store_mom(double* __restrict__ mom, int64_t nmax, int64_t col, const MomRF<2 * NT>& P) {
#pragma unroll
    for (int i = 0; i < NT; ++i) {
        mom[(int64_t)(i * 4 + 0) * nmax + col] = P.v[4 * i + 0];
        mom[(int64_t)(i * 4 + 1) * nmax + col] = P.v[4 * i + 1];
        mom[(int64_t)(i * 4 + 2) * nmax + col] = P.v[4 * i + 2];
        mom[(int64_t)(i * 4 + 3) * nmax + col] = P.v[4 * i + 3];
    }
}

// forward_point with the momenta of a passing attempt stored at its sorted
// position e.
template <int NT, bool ZS>
__device__ __forceinline__ void forward_one(const MCArgs& A, const ChanProg& p, int ch, int e, uint32_t a, uint64_t att,
                                            double* __restrict__ mom,
                                            int64_t nmax, uint8_t* __restrict__ acc, const PointRec& R,
                                            const double* __restrict__ zsrc) {
    MomRF<2 * NT> P;
    if (forward_point<NT, ZS>(A, p, ch, e, a, att, acc, R, zsrc, P)) store_mom<NT>(mom, nmax, e, P);
}

// Packed order of a forward tile [e0, e1) (one workgroup of kTile threads,
// thread t holding position e0 + t, live when inside the tile): the attempts
// that pass the cut first, then the others, each group in its sorted order.
// Returns the thread's new position; every thread of the workgroup must call
// it (it synchronises the workgroup).  s_cnt: LDS, 2 kWaves ints.
__device__ __forceinline__ int tile_pack_pos(bool live, bool ok, int e0, int* s_cnt) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint64_t mo = __ballot(live && ok), mn = __ballot(live && !ok);
    if (lane == 0) {
        s_cnt[w] = __popcll(mo);
        s_cnt[kWaves + w] = __popcll(mn);
    }
    __syncthreads();
    int before_o = 0, before_n = 0, tot_o = 0;
#pragma unroll
    for (int ww = 0; ww < kWaves; ++ww) {
        if (ww < w) {
            before_o += s_cnt[ww];
            before_n += s_cnt[kWaves + ww];
        }
        tot_o += s_cnt[ww];
    }
    const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
    return e0 + (ok ? before_o + __popcll(mo & lt) : tot_o + before_n + __popcll(mn & lt));
}


// ---------------------------------------------------------------------------
// k_main: for every kept point of the workgroup's channel: f, all C inverse
// maps and VEGAS densities, W = 1 / sum_j alpha_j / ps_j / vegas_j, the value
// val = W == 0 ? 0 : f W (NaN -> 0, counted), and the workgroup partials of
// StatsData, alpha-training T_j += val^2 W / ps_j and the picked channel's
// f^2 histogram (Integrand.hh:89-123, MultiChannel.hh:167-180, 224-247).
// ---------------------------------------------------------------------------
constexpr int kRounds = 1;                 // sub-tiles of kTile points per k_main workgroup
constexpr int kMainTile = kTile * kRounds;
constexpr size_t kLdsSearchMax = 32 * 1024;   // largest search-table set staged in LDS (bytes)

// One kept tile [e0, e1) of channel ch (RND sub-tiles of kTile points):
// psbuf column slot * kTile + thread, partial record `out`.  Every thread of
// the workgroup calls it (it synchronises the workgroup).  LS: the search
// tables of all channels ([C][d][P2], a few KB) are first copied into LDS, so
// that the dependent loads of the bin searches are LDS reads.
template <int NT, int RND, class Loop, bool LS = false>
__device__ __forceinline__ void main_tile(const MCArgs& A, int ch, int e0, int e1, double* s_mem, int64_t slot,
                                          double* __restrict__ psbuf, int64_t psstride, int train, const uint32_t* __restrict__ comp,
                                          const double* __restrict__ mom, int64_t nmax, double* __restrict__ out,
                                          double* __restrict__ vals_out, int64_t vals_off, int64_t vals_cap,
                                          const uint32_t* __restrict__ sorted, const PointRec& R) {
    constexpr int D = 3 * (NT - 2) - 2;
    // LDS: hist [D][B], per-wave sums [kWaves][8 + C], bins [D][kTile] shorts,
    // with LS the search tables [C][D][P2]
    const int B = A.B, C = A.C;
    double* s_hist = s_mem;
    double* s_small = s_hist + D * B;
    short* s_bins = (short*)(s_small + kWaves * (8 + C));
    double* s_srch = (double*)(s_bins + D * kTile);
    if (LS)
        for (int i = threadIdx.x; i < C * D * A.P2; i += kTile) s_srch[i] = A.search[i];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t gt = slot * kTile + threadIdx.x;
    __syncthreads();                     // a previous tile's record may still be reading LDS
    for (int i = threadIdx.x; i < D * B; i += kTile) s_hist[i] = 0.0;
    for (int i = threadIdx.x; i < kWaves * (8 + C); i += kTile) s_small[i] = 0.0;
    // per-lane StatsData moments, in round order
    double st_n = 0, st_s1 = 0, st_s2 = 0, st_nf = 0, st_nn = 0;
    double st_mn = __DBL_MAX__, st_mx = __DBL_MIN__;
    __syncthreads();
    for (int rnd = 0; rnd < RND; ++rnd) {
        const int e = e0 + rnd * kTile + threadIdx.x;
        if (e0 + rnd * kTile >= e1) break;           // uniform: no points left in this super-tile
        const bool valid = e < e1;
        MomRF<2 * NT> P;
        double val = 0, W = 0, fv = 0;
        bool isnanv = false;
        uint32_t att = 0;
        if (valid) {
            if constexpr (Loop::kPre) {
                // split pass: the first channel's inverse kernel, which holds
                // the point's momenta, left f; no momentum is read here
                if (R.status) att = sorted[comp[e]];
                fv = A.pre_pv[(int64_t)2 * C * A.pre_stride + e];
            } else {
                const uint32_t sp = comp[e];
                if (R.status) att = sorted[sp];
#pragma unroll
                for (int i = 0; i < NT; ++i) {
                    P.v[4 * i + 0] = mom[(int64_t)(i * 4 + 0) * nmax + sp];
                    P.v[4 * i + 1] = mom[(int64_t)(i * 4 + 1) * nmax + sp];
                    P.v[4 * i + 2] = mom[(int64_t)(i * 4 + 2) * nmax + sp];
                    P.v[4 * i + 3] = mom[(int64_t)(i * 4 + 3) * nmax + sp];
                }
                fv = A.fhost ? A.fhost[sp] : integrand<NT>(A.f, P);
            }
        } else if constexpr (!Loop::kPre) {
#pragma unroll
            for (int i = 0; i < NT; ++i) P.set(i, mk(1., 0., 0., 1.));
        }
        // Integrand::GenerateWeight (Integrand.hh:112-123)
        double wsum = 0;
        Loop::each(A, [&](int j, const ChanProg& pj) {
            double vw, ps;
            if constexpr (Loop::kPre) {
                // split pass: every channel's inverse ran in its own kernel
                const int64_t ek = valid ? e : 0;
                ps = A.pre_pv[(int64_t)j * A.pre_stride + ek];
                vw = A.pre_pv[(int64_t)(A.C + j) * A.pre_stride + ek];
                if (j == ch)
                    for (int k = 0; k < D; ++k) s_bins[k * kTile + threadIdx.x] = A.pre_bins[(int64_t)k * A.pre_stride + ek];
            } else {
                short* bo = (j == ch) ? (s_bins + threadIdx.x) : nullptr;
                const bool recx = R.status && valid && j == ch;
                ps = generate_weight<NT>(A, pj, j, P, &vw, bo, kTile, recx ? R.x + att : nullptr, recx ? R.bins + att : nullptr, R.n,
                                         nullptr, LS ? s_srch : nullptr);
                if (R.status && valid) R.ps[(int64_t)j * R.n + att] = ps;
            }
            wsum += A.alpha[j] / ps / vw;
            psbuf[(int64_t)j * psstride + gt] = ps;
        });
        if (valid) {
            W = 1.0 / wsum;
            val = (W == 0) ? 0.0 : fv * W;
            isnanv = isnan(val);
            if (isnanv) { val = 0; W = 0; }
            if (vals_out && vals_off + e < vals_cap) vals_out[vals_off + e] = val;
            if (R.status) {
                R.status[att] = isnanv ? 2 : 1;
                R.w[att] = W;
                R.f[att] = fv;
                R.val[att] = val;
            }
            // StatsData::operator+= (Statistics.hh:92-102)
            st_n += 1.0; st_s1 += val; st_s2 += val * val;
            if (val != 0) st_nf += 1.0;
            if (isnanv) st_nn += 1.0;
            st_mn = (val < st_mn) ? val : st_mn;
            st_mx = (st_mx < val) ? val : st_mx;
        }
        if (!train) continue;
        const double v2 = val * val;
        // alpha training T_j (MultiChannel.hh:174-177): wave sums added per round
        for (int j = 0; j < C; ++j) {
            double t = valid ? v2 * W / psbuf[(int64_t)j * psstride + gt] : 0.0;
            t = wave_sum(t);
            if (lane == 0) s_small[w * (8 + C) + 8 + j] += t;
        }
        // f^2 histogram of the picked channel (Integrand.hh:89-96): lane l of a
        // wave owns bins l, l+64, ...; wave values in lane order, waves in turn.
        __syncthreads();
        for (int k = 0; k < D; ++k) {
            const int bk = valid ? s_bins[k * kTile + threadIdx.x] : 0;
            const double vk = valid ? v2 : 0.0;
            double h0 = 0, h1 = 0, h2 = 0, h3 = 0;
            for (int src = 0; src < 64; ++src) {
                const int b = __builtin_amdgcn_readlane(bk, src);
                const double v = readlane_f64(vk, src);
                if ((b & 63) == lane) {
                    const int hi = b >> 6;
                    if (hi == 0) h0 += v; else if (hi == 1) h1 += v; else if (hi == 2) h2 += v; else h3 += v;
                }
            }
            for (int ww = 0; ww < kWaves; ++ww) {
                if (w == ww) {
                    double hq[4] = {h0, h1, h2, h3};
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const int b = lane + 64 * q;
                        if (b < B) s_hist[k * B + b] += hq[q];
                    }
                }
                __syncthreads();
            }
        }
    }
    // ---- StatsData partial of the workgroup ----
    {
        double n = wave_sum(st_n), s1 = wave_sum(st_s1), s2 = wave_sum(st_s2), nf = wave_sum(st_nf), nn = wave_sum(st_nn);
        double mn = wave_min(st_mn), mx = wave_max(st_mx);
        if (lane == 0) {
            double* q = s_small + w * (8 + C);
            q[0] = n; q[1] = mn; q[2] = mx; q[3] = s1; q[4] = s2; q[5] = nf; q[6] = nn;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double acc[7] = {0, __DBL_MAX__, __DBL_MIN__, 0, 0, 0, 0};
        for (int ww = 0; ww < kWaves; ++ww) {
            const double* q = s_small + ww * (8 + C);
            acc[0] += q[0]; acc[1] = (q[1] < acc[1]) ? q[1] : acc[1]; acc[2] = (acc[2] < q[2]) ? q[2] : acc[2];
            acc[3] += q[3]; acc[4] += q[4]; acc[5] += q[5]; acc[6] += q[6];
        }
        for (int k = 0; k < kStatW; ++k) out[k] = acc[k];
    }
    if (!train) return;
    for (int j = threadIdx.x; j < C; j += kTile) {
        double sj = 0;
        for (int ww = 0; ww < kWaves; ++ww) sj += s_small[ww * (8 + C) + 8 + j];
        out[kStatW + j] = sj;
    }
    for (int i = threadIdx.x; i < D * B; i += kTile) out[kStatW + C + i] = s_hist[i];
}


// Channel loop over the programs in device memory.
struct GenericLoop {
    static constexpr bool kPre = false;
    template <class F>
    __device__ __forceinline__ static void each(const MCArgs& A, F&& f) {
        for (int j = 0; j < A.C; ++j) f(j, A.progs[j]);
    }
};

// Channel loop of the split weight pass: the per-channel inverse kernels
// (inverse_one, run-time specialised) left ps, the VEGAS weight and the
// picked channel's bins in A.pre_*; no program is evaluated here.
struct PreLoop {
    static constexpr bool kPre = true;
    template <class F>
    __device__ __forceinline__ static void each(const MCArgs& A, F&& f) {
        for (int j = 0; j < A.C; ++j) f(j, A.progs[0]);
    }
};

// One kept point e of the split weight pass for channel j (program p): its
// inverse map and VEGAS weight (FSMapper::GenerateWeight, Channel.cc:40-61;
// AdaptiveMap::GenerateWeight, AdaptiveMap.cc:58-65), and, when j is the
// point's picked channel, its bins; channel 0 also evaluates the integrand
// (row 2 C of pv), so that the combination kernel reads no momenta; kLds:
// search is this channel's table in LDS, else null (device memory).
template <int NT, bool kLds>
__device__ __forceinline__ void inverse_one(const MCArgs& A, const ChanProg& p, int j, int64_t e, const int* __restrict__ seg2,
                                            const uint32_t* __restrict__ comp, const double* __restrict__ mom, int64_t nmax,
                                            double* __restrict__ pv, short* __restrict__ bins, const uint32_t* __restrict__ sorted,
                                            const PointRec& R, const double* search) {
    const uint32_t sp = comp[e];
    MomRF<2 * NT> P;
#pragma unroll
    for (int i = 0; i < NT; ++i) {
        P.v[4 * i + 0] = mom[(int64_t)(i * 4 + 0) * nmax + sp];
        P.v[4 * i + 1] = mom[(int64_t)(i * 4 + 1) * nmax + sp];
        P.v[4 * i + 2] = mom[(int64_t)(i * 4 + 2) * nmax + sp];
        P.v[4 * i + 3] = mom[(int64_t)(i * 4 + 3) * nmax + sp];
    }
    if (j == 0) pv[(int64_t)2 * A.C * A.pre_stride + e] = A.fhost ? A.fhost[sp] : integrand<NT>(A.f, P);
    int pc = 0;                         // picked channel: the kept segment holding e
    for (int c = 1; c < A.C; ++c) pc = (seg2[c] <= e) ? c : pc;
    const bool own = pc == j;
    const uint32_t att = R.status ? sorted[sp] : 0;
    const bool recx = R.status && own;
    double vw;
    const double ps = generate_weight<NT, kLds>(A, p, j, P, &vw, own ? bins + e : nullptr, (int)A.pre_stride,
                                                recx ? R.x + att : nullptr, recx ? R.bins + att : nullptr, R.n, nullptr,
                                                kLds ? search : nullptr, kLds);
    pv[(int64_t)j * A.pre_stride + e] = ps;
    pv[(int64_t)(A.C + j) * A.pre_stride + e] = vw;
    if (R.status) R.ps[(int64_t)j * R.n + att] = ps;
}

// Exact early rejection of the unweighting stage, one passing point k whose
// picked channel pc has program p: the multichannel weight
// W = 1 / sum_j alpha_j / ps_j / vegas_j is at most the picked channel's
// term's inverse, W_up = 1 / (alpha_pc / ps_pc / vegas_pc), computed with the
// same expressions as the full pass (so fl(sum) >= that term and, the
// operations being monotone, |f W| <= |f W_up| holds in floating point too).
// On the learned-proposal path the picked channel's term carries its
// proposal density, (alpha_pc / ps_pc / vegas_pc) ((1 - mix) exp(lqs) + mix)
// with lqs the sampling pass's log-density of the point (the same
// expression and value the combination uses for it).  The point stays a
// candidate (acc[a] kept at 1) only when the accept test u < |f W_up| / wmax
// can pass; otherwise acc[a] = 0 -- the full test would reject it with the
// same uniform u.  NaN bounds reject, as NaN values do.
template <int NT, bool kLds>
__device__ __forceinline__ void own_bound_one(const MCArgs& A, const ChanProg& p, int pc, int64_t k, const uint32_t* __restrict__ comp,
                                              const double* __restrict__ mom, int64_t nmax, const uint32_t* __restrict__ sorted,
                                              uint8_t* __restrict__ acc, double wmax, const double* __restrict__ lqs, double mix,
                                              const double* search) {
    const uint32_t sp = comp[k];
    MomRF<2 * NT> P;
#pragma unroll
    for (int i = 0; i < NT; ++i) {
        P.v[4 * i + 0] = mom[(int64_t)(i * 4 + 0) * nmax + sp];
        P.v[4 * i + 1] = mom[(int64_t)(i * 4 + 1) * nmax + sp];
        P.v[4 * i + 2] = mom[(int64_t)(i * 4 + 2) * nmax + sp];
        P.v[4 * i + 3] = mom[(int64_t)(i * 4 + 3) * nmax + sp];
    }
    const double fv = integrand<NT>(A.f, P);
    double vw;
    const double ps = generate_weight<NT, kLds>(A, p, pc, P, &vw, nullptr, 0, nullptr, nullptr, 0, nullptr, kLds ? search : nullptr, kLds);
    const double term = A.alpha[pc] / ps / vw;
    const double wup = 1.0 / (lqs ? term * ((1.0 - mix) * exp(lqs[sp]) + mix) : term);
    const double vup = (wup == 0) ? 0.0 : fv * wup;
    const uint32_t a = sorted[sp];
    const bool cand = aux_uniform(A, A.a0 + a) < fabs(vup) / wmax;      // false for NaN
    if (!cand) acc[a] = 0;
}

// Momenta of sorted position sp from the SoA columns [4 NT][nmax].
template <int NT>
__device__ __forceinline__ void load_mom(const double* __restrict__ mom, int64_t nmax, uint32_t sp, MomRF<2 * NT>& P) {
#pragma unroll
    for (int i = 0; i < NT; ++i) {
        P.v[4 * i + 0] = mom[(int64_t)(i * 4 + 0) * nmax + sp];
        P.v[4 * i + 1] = mom[(int64_t)(i * 4 + 1) * nmax + sp];
        P.v[4 * i + 2] = mom[(int64_t)(i * 4 + 2) * nmax + sp];
        P.v[4 * i + 3] = mom[(int64_t)(i * 4 + 3) * nmax + sp];
    }
}

// Learned-proposal path, one kept point k with momenta P for channel j
// (program p): the hypercube point r_j of its inverse map (written to `row`,
// D doubles) and a_j = alpha_j / ps_j / vegas_j (abuf[j][k]).  The generic
// k_geom loops over every channel; the run-time specialised kernels run one
// channel each (search: this channel's table in LDS when kLds).
template <int NT, bool kLds>
__device__ __forceinline__ void geom_one(const MCArgs& A, const ChanProg& p, int j, int64_t k, MomRF<2 * NT>& P, double* row,
                                         double* __restrict__ abuf, int64_t stride, const double* search) {
    constexpr int D = 3 * (NT - 2) - 2;
    // every slot is written by the channel program; 0.5 stands in for any a
    // degenerate point leaves out
#pragma unroll
    for (int q = 0; q < D; ++q) row[q] = 0.5;
    double vw;
    const double ps = generate_weight<NT, kLds>(A, p, j, P, &vw, nullptr, 0, nullptr, nullptr, 0, row, kLds ? search : nullptr, kLds);
    abuf[(int64_t)j * stride + k] = A.alpha[j] / ps / vw;
}

}  // namespace chili_dev
