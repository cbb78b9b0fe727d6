// mc.hip -- multichannel VEGAS on gfx950: kernels, device context and the
// chili_mc_* / chili_construct_channels entry points of include/chili_gpu.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "channels.hpp"
#include "common.hpp"
#include "flow.hpp"
#include "jit.hpp"
#include "mc_kernels.hpp"
#include "mc_device.hpp"

using namespace chili_dev;
using chili_host::ChanProg;

namespace {

// ---------------------------------------------------------------------------
// k_pick: channel of every attempt; per-(channel, block) counts, channel-major.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kTile) void k_pick(MCArgs A, int64_t n, uint16_t* __restrict__ pick, int* __restrict__ counts, int nblk) {
    extern __shared__ int s_cnt[];
    for (int c = threadIdx.x; c < A.C; c += blockDim.x) s_cnt[c] = 0;
    __syncthreads();
    int64_t i = (int64_t)blockIdx.x * kTile + threadIdx.x;
    if (i < n) {
        int c = pick_channel(A, channel_uniform(A, A.a0 + (uint64_t)i));
        pick[i] = (uint16_t)c;
        atomicAdd(&s_cnt[c], 1);
    }
    __syncthreads();
    for (int c = threadIdx.x; c < A.C; c += blockDim.x) counts[(int64_t)c * nblk + blockIdx.x] = s_cnt[c];
}

// Attempts per sub-batch of the learned-proposal path: its buffers hold
// C x attempts x d doubles twice (the r_j and the flows' working copy), kept
// within 24 GiB, between 2^23 and 2^25 attempts
int64_t flow_sub_max(int C, int d) {
    const int64_t fit = (int64_t(24) << 30) / ((int64_t)C * d * 16);
    return std::min<int64_t>(int64_t(1) << 25, std::max<int64_t>(int64_t(1) << 23, fit));
}
// Attempts per sub-batch of the plain path (iterate and the unweighting
// stage): each sub-batch synchronises with the host a few times, so large
// sub-batches keep the device busy (2^25 + 2^22: one sub-batch holds a
// 2^24-point iteration down to a cut acceptance of 0.45); momenta
// [4 NT][attempts] stay well inside HBM
constexpr int64_t kSubMax = (int64_t(1) << 25) + (int64_t(1) << 22);

// ---------------------------------------------------------------------------
// Exclusive scan of an int array in three passes: sums of 1024-element blocks,
// a single-workgroup scan of those sums, then each block's scan plus its
// offset.  out may alias in; total (device) receives the sum.
// ---------------------------------------------------------------------------
constexpr int kScanBlock = 1024;

__device__ __forceinline__ int64_t block_scan_1024(int64_t v, int64_t* s_tmp, int64_t& total) {
    // inclusive Hillis-Steele scan over 1024 threads, returns the exclusive value
    const int t = threadIdx.x;
    s_tmp[t] = v;
    __syncthreads();
    for (int off = 1; off < kScanBlock; off <<= 1) {
        int64_t x = t >= off ? s_tmp[t - off] : 0;
        __syncthreads();
        s_tmp[t] += x;
        __syncthreads();
    }
    total = s_tmp[kScanBlock - 1];
    int64_t ex = s_tmp[t] - v;
    __syncthreads();
    return ex;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_sums(const int* __restrict__ in, int64_t len, int64_t* __restrict__ bsum) {
    __shared__ int64_t s[kScanBlock / 64];
    const int64_t i = (int64_t)blockIdx.x * kScanBlock + threadIdx.x;
    int64_t v = i < len ? in[i] : 0;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t t = 0;
        for (int w = 0; w < kScanBlock / 64; ++w) t += s[w];
        bsum[blockIdx.x] = t;
    }
}

__global__ __launch_bounds__(kScanBlock) void k_scan_top(int64_t* __restrict__ bsum, int64_t nb, int64_t* total) {
    __shared__ int64_t s_tmp[kScanBlock];
    int64_t carry = 0;
    for (int64_t base = 0; base < nb; base += kScanBlock) {
        const int64_t i = base + threadIdx.x;
        int64_t v = i < nb ? bsum[i] : 0, tot;
        int64_t ex = block_scan_1024(v, s_tmp, tot);
        if (i < nb) bsum[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0 && total) *total = carry;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_apply(const int* in, int* out, int64_t len, const int64_t* __restrict__ bsum) {
    __shared__ int64_t s_tmp[kScanBlock];
    const int64_t i = (int64_t)blockIdx.x * kScanBlock + threadIdx.x;
    int64_t v = i < len ? in[i] : 0, tot;
    int64_t ex = block_scan_1024(v, s_tmp, tot);
    if (i < len) out[i] = (int)(bsum[blockIdx.x] + ex);
}

// Per-wave stable ranks of equal keys: rank among earlier lanes with the same
// key, plus the per-key lane count written by the key's first lane.
__device__ __forceinline__ int wave_rank(int key, bool valid, int* wave_count_row) {
    const int lane = threadIdx.x & 63;
    uint64_t rem = __ballot(valid);
    int rank = 0;
    const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
    while (rem) {
        int leader = __ffsll((unsigned long long)rem) - 1;
        int k0 = __shfl(key, leader);
        uint64_t m = __ballot(valid && key == k0);
        if (valid && key == k0) rank = __popcll(m & lt);
        if (lane == leader) wave_count_row[k0] = __popcll(m);
        rem &= ~m;
    }
    return rank;
}

// ---------------------------------------------------------------------------
// k_scatter: stable counting sort of attempt indices by channel.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kTile) void k_scatter(int C, int64_t n, const uint16_t* __restrict__ pick, const int* __restrict__ offs,
                                                   int nblk, uint32_t* __restrict__ sorted) {
    extern __shared__ int s_wc[];   // [kWaves][C]
    for (int i = threadIdx.x; i < kWaves * C; i += blockDim.x) s_wc[i] = 0;
    __syncthreads();
    const int w = threadIdx.x >> 6;
    int64_t i = (int64_t)blockIdx.x * kTile + threadIdx.x;
    bool valid = i < n;
    int c = valid ? pick[i] : 0;
    int rank = wave_rank(c, valid, s_wc + w * C);
    __syncthreads();
    if (valid) {
        int before = 0;
        for (int ww = 0; ww < w; ++ww) before += s_wc[ww * C + c];
        sorted[offs[(int64_t)c * nblk + blockIdx.x] + before + rank] = (uint32_t)i;
    }
}

// Channel segment table: seg_start[c] (entries), tile_start[c] (tiles of kTile).
__global__ void k_segments(int C, const int* __restrict__ offs, int nblk, int64_t total, int* seg_start, int* tile_start) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    int t = 0;
    for (int c = 0; c < C; ++c) {
        int s = offs[(int64_t)c * nblk];
        int e = (c + 1 < C) ? offs[(int64_t)(c + 1) * nblk] : (int)total;
        seg_start[c] = s;
        tile_start[c] = t;
        t += (e - s + kTile - 1) / kTile;
    }
    seg_start[C] = (int)total;
    tile_start[C] = t;
}

// Mixture proposal: bucket 2c + u of every attempt (c: channel, u = 1 when
// its mixture uniform is below mix, the uniform component), so that after
// the counting sort each channel's segment holds its flow attempts first.
__global__ __launch_bounds__(kTile) void k_pick_mix(MCArgs A, int64_t n, double mix, uint16_t* __restrict__ pick,
                                                    int* __restrict__ counts, int nblk) {
    extern __shared__ int s_cnt[];
    for (int c = threadIdx.x; c < 2 * A.C; c += blockDim.x) s_cnt[c] = 0;
    __syncthreads();
    int64_t i = (int64_t)blockIdx.x * kTile + threadIdx.x;
    if (i < n) {
        const uint64_t att = A.a0 + (uint64_t)i;
        const int c = pick_channel(A, channel_uniform(A, att));
        const int b = 2 * c + (aux_uniform(A, att) < mix ? 1 : 0);
        pick[i] = (uint16_t)b;
        atomicAdd(&s_cnt[b], 1);
    }
    __syncthreads();
    for (int c = threadIdx.x; c < 2 * A.C; c += blockDim.x) counts[(int64_t)c * nblk + blockIdx.x] = s_cnt[c];
}

// Channel segments of the bucketed sort: seg_start / tile_start as
// k_segments, flow_end[c]: end of channel c's flow attempts.
__global__ void k_segments_mix(int C, const int* __restrict__ offs, int nblk, int64_t total, int* seg_start, int* tile_start,
                               int* flow_end) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    int t = 0;
    for (int c = 0; c < C; ++c) {
        int s = offs[(int64_t)(2 * c) * nblk];
        int e = (c + 1 < C) ? offs[(int64_t)(2 * c + 2) * nblk] : (int)total;
        seg_start[c] = s;
        flow_end[c] = offs[(int64_t)(2 * c + 1) * nblk];
        tile_start[c] = t;
        t += (e - s + kTile - 1) / kTile;
    }
    seg_start[C] = (int)total;
    tile_start[C] = t;
}

// ---------------------------------------------------------------------------
// k_forward: VEGAS forward map, FSMapper::GeneratePoint and the cut for the
// attempts of one channel tile per workgroup.  Momenta of external legs go
// to mom[(i*4+k)*nmax + e'].  Packed (sorted_out non-null): within the tile
// the passing attempts move to the front (tile_pack_pos), sorted_out[e']
// receives the attempt of new position e' and e' is where the momenta go,
// so that the weight pass reads the kept points' momenta from contiguous
// columns; the set and order of kept points are unchanged.  Otherwise
// e' = e.  lq_in / lq_out (optional): a per-position FP64 row (the
// learned proposal's sampling log-density) moved to the packed order.
// ---------------------------------------------------------------------------
template <int NT, bool ZS>
__global__ __launch_bounds__(kTile) void k_forward(MCArgs A, const int* __restrict__ seg_start, const int* __restrict__ tile_start,
                                                   const uint32_t* __restrict__ sorted, double* __restrict__ mom, int64_t nmax,
                                                   uint8_t* __restrict__ acc, PointRec R, const double* __restrict__ zsrc,
                                                   uint32_t* __restrict__ sorted_out, const double* __restrict__ lq_in,
                                                   double* __restrict__ lq_out) {
    __shared__ int s_cnt[2 * kWaves];
    int ch, e0, e1;
    if (!locate_tile(A.C, seg_start, tile_start, blockIdx.x, ch, e0, e1)) return;
    const int e = e0 + threadIdx.x;
    const bool live = e < e1;
    if (!sorted_out) {
        if (live) forward_one<NT, ZS>(A, A.progs[ch], ch, e, sorted[e], A.a0 + sorted[e], mom, nmax, acc, R, zsrc);
        return;
    }
    MomRF<2 * NT> P;
    bool ok = false;
    uint32_t a = 0;
    if (live) {
        a = sorted[e];
        ok = forward_point<NT, ZS>(A, A.progs[ch], ch, e, a, A.a0 + a, acc, R, zsrc, P);
    }
    const int ep = tile_pack_pos(live, ok, e0, s_cnt);
    if (live) {
        sorted_out[ep] = a;
        if (ok) store_mom<NT>(mom, nmax, ep, P);
        if (lq_out) lq_out[ep] = lq_in[e];      // per-position rows move with their attempt
    }
}

// kT-jet PreProcess (CHILI_CUTS_KT_JETS), after the forward kernel's own
// checks: the s-hat limit and the clustering of the final state of an
// attempt that passed (momenta at column e).
template <int NT>
__device__ __forceinline__ bool kt_cut_pass(const chili_integrand* __restrict__ f, const double* __restrict__ mom, int64_t nmax,
                                            int64_t e) {
    V4 p[NT];
#pragma unroll
    for (int i = 0; i < NT; ++i)
        p[i] = mk(mom[(int64_t)(4 * i) * nmax + e], mom[(int64_t)(4 * i + 1) * nmax + e], mom[(int64_t)(4 * i + 2) * nmax + e],
                  mom[(int64_t)(4 * i + 3) * nmax + e]);
    const double smax = f->ptmin;
    if (smax > 0 && vmass2(vadd(p[0], p[1])) > smax * smax) return false;      // example.cc:29-32
    return kt_all_jets<NT - 2>(p + 2, f->drmin);
}

// The kT-jet cut of every attempt that passed the forward kernel; a failure
// clears its flag (and record).
template <int NT>
__global__ __launch_bounds__(kTile) void k_ktcut(const chili_integrand* __restrict__ f, int64_t n, const uint32_t* __restrict__ sorted,
                                                 const double* __restrict__ mom, int64_t nmax, uint8_t* __restrict__ acc, PointRec R) {
    const int64_t e = (int64_t)blockIdx.x * kTile + threadIdx.x;
    if (e >= n) return;
    const uint32_t a = sorted[e];
    if (!acc[a]) return;
    if (!kt_cut_pass<NT>(f, mom, nmax, e)) {
        acc[a] = 0;
        if (R.status) R.status[a] = 0;
    }
}

// Accepted attempts per block of kTile attempts (attempt order).// Accepted attempts per block of kTile attempts (attempt order).
__global__ __launch_bounds__(kTile) void k_acc_count(const uint8_t* __restrict__ acc, int64_t n, int* __restrict__ out) {
    __shared__ int s[kWaves];
    int64_t i = (int64_t)blockIdx.x * kTile + threadIdx.x;
    bool v = i < n && acc[i];
    uint64_t m = __ballot(v);
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = __popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int w = 0; w < kWaves; ++w) t += s[w];
        out[blockIdx.x] = t;
    }
}

// Kept entries (accepted and below the cutoff) per forward tile.
__global__ __launch_bounds__(kTile) void k_keep_count(int C, const int* __restrict__ seg_start, const int* __restrict__ tile_start,
                                                      const uint32_t* __restrict__ sorted, const uint8_t* __restrict__ acc, uint32_t cut,
                                                      int* __restrict__ kept) {
    __shared__ int s[kWaves];
    int ch, e0, e1;
    if (!locate_tile(C, seg_start, tile_start, blockIdx.x, ch, e0, e1)) {
        if (threadIdx.x == 0) kept[blockIdx.x] = 0;
        return;
    }
    const int e = e0 + threadIdx.x;
    bool k = false;
    if (e < e1) { uint32_t a = sorted[e]; k = a < cut && acc[a]; }
    uint64_t m = __ballot(k);
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = __popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int w = 0; w < kWaves; ++w) t += s[w];
        kept[blockIdx.x] = t;
    }
}

// Stable compaction of the kept entries (their sorted positions).
__global__ __launch_bounds__(kTile) void k_keep_scatter(int C, const int* __restrict__ seg_start, const int* __restrict__ tile_start,
                                                        const uint32_t* __restrict__ sorted, const uint8_t* __restrict__ acc, uint32_t cut,
                                                        const int* __restrict__ kept_off, uint32_t* __restrict__ comp) {
    __shared__ int s[kWaves];
    int ch, e0, e1;
    if (!locate_tile(C, seg_start, tile_start, blockIdx.x, ch, e0, e1)) return;
    const int e = e0 + threadIdx.x;
    bool k = false;
    if (e < e1) { uint32_t a = sorted[e]; k = a < cut && acc[a]; }
    uint64_t m = __ballot(k);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) s[w] = __popcll(m);
    __syncthreads();
    if (k) {
        int before = 0;
        for (int ww = 0; ww < w; ++ww) before += s[ww];
        const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
        comp[kept_off[blockIdx.x] + before + __popcll(m & lt)] = (uint32_t)e;
    }
}

// Segments of the compacted list: per channel, kept_off at its first tile.
__global__ void k_keep_segments(int C, const int* __restrict__ tile_start, const int* __restrict__ kept_off, int64_t kept_total,
                                int* seg2, int* tile2, int tsize) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    int t = 0;
    for (int c = 0; c < C; ++c) {
        // tiles at or past tile_start[C] belong to no channel: the segment runs to the end
        int s = (tile_start[c] >= tile_start[C]) ? (int)kept_total : kept_off[tile_start[c]];
        int e = (c + 1 < C && tile_start[c + 1] < tile_start[C]) ? kept_off[tile_start[c + 1]] : (int)kept_total;
        seg2[c] = s;
        tile2[c] = t;
        t += (e - s + tsize - 1) / tsize;
    }
    seg2[C] = (int)kept_total;
    tile2[C] = t;
}

template <int NT, bool LS>
__global__ __launch_bounds__(kTile) void k_main(MCArgs A, const int* __restrict__ seg2, const int* __restrict__ tile2,
                                                const uint32_t* __restrict__ comp, const double* __restrict__ mom, int64_t nmax,
                                                double* __restrict__ psbuf, int64_t psstride, int train,
                                                double* __restrict__ part, int64_t rec, double* __restrict__ vals_out, int64_t vals_off,
                                                int64_t vals_cap, const uint32_t* __restrict__ sorted, PointRec R) {
    extern __shared__ double s_mem[];
    int ch, e0, e1;
    if (!locate_tile(A.C, seg2, tile2, blockIdx.x, ch, e0, e1, kMainTile)) return;
    main_tile<NT, kRounds, GenericLoop, LS>(A, ch, e0, e1, s_mem, blockIdx.x, psbuf, psstride, train, comp, mom, nmax, part + (int64_t)blockIdx.x * rec,
                           vals_out, vals_off, vals_cap, sorted, R);
}

// The combination step of the split weight pass (PreLoop: ps, VEGAS weights
// and bins come from the per-channel inverse kernels).
template <int NT>
__global__ __launch_bounds__(kTile) void k_main_pre(MCArgs A, const int* __restrict__ seg2, const int* __restrict__ tile2,
                                                    const uint32_t* __restrict__ comp, const double* __restrict__ mom, int64_t nmax,
                                                    double* __restrict__ psbuf, int64_t psstride, int train,
                                                    double* __restrict__ part, int64_t rec, double* __restrict__ vals_out,
                                                    int64_t vals_off, int64_t vals_cap, const uint32_t* __restrict__ sorted, PointRec R) {
    extern __shared__ double s_mem[];
    int ch, e0, e1;
    if (!locate_tile(A.C, seg2, tile2, blockIdx.x, ch, e0, e1, kMainTile)) return;
    main_tile<NT, kRounds, PreLoop, false>(A, ch, e0, e1, s_mem, blockIdx.x, psbuf, psstride, train, comp, mom, nmax,
                                           part + (int64_t)blockIdx.x * rec, vals_out, vals_off, vals_cap, sorted, R);
}

// ---------------------------------------------------------------------------
// Fixed-order reduction of the workgroup partials into the running totals
// tot = [stats(7), T(C), hist(C x d x B)], in two levels so that no thread
// walks every partial: k_reduce_groups sums groups of kRedGroup consecutive
// partials (stats and T); k_reduce_final adds the group sums in order and sums
// each histogram entry over its channel's partials.
// ---------------------------------------------------------------------------
constexpr int kRedGroup = 64;

__device__ __forceinline__ void stat_merge(double* a, const double* q) {
    a[0] += q[0];
    a[1] = (q[1] < a[1]) ? q[1] : a[1];
    a[2] = (a[2] < q[2]) ? q[2] : a[2];
    a[3] += q[3]; a[4] += q[4]; a[5] += q[5]; a[6] += q[6];
}

__global__ __launch_bounds__(256) void k_reduce_groups(int C, const int* __restrict__ tile2, const double* __restrict__ part,
                                                       int64_t rec, double* __restrict__ gsum, int train) {
    const int ntile = tile2[C];
    const int g = blockIdx.x;
    const int t0 = g * kRedGroup;
    if (t0 >= ntile) return;
    const int t1 = min(ntile, t0 + kRedGroup);
    const int W = kStatW + C;
    double* out = gsum + (int64_t)g * W;
    if (threadIdx.x == 0) {
        double a[7] = {0, std::numeric_limits<double>::max(), std::numeric_limits<double>::min(), 0, 0, 0, 0};
        for (int t = t0; t < t1; ++t) stat_merge(a, part + (int64_t)t * rec);
        for (int k = 0; k < kStatW; ++k) out[k] = a[k];
    }
    if (!train) return;
    for (int j = threadIdx.x; j < C; j += blockDim.x) {
        double s = 0;
        for (int t = t0; t < t1; ++t) s += part[(int64_t)t * rec + kStatW + j];
        out[kStatW + j] = s;
    }
}

// Histogram groups: per channel, runs of kRedGroupH consecutive partials.
constexpr int kRedGroupH = 16;

__global__ void k_hist_groups_table(int C, const int* __restrict__ tile2, int* __restrict__ grp2) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    int g = 0;
    for (int c = 0; c < C; ++c) {
        grp2[c] = g;
        g += (tile2[c + 1] - tile2[c] + kRedGroupH - 1) / kRedGroupH;
    }
    grp2[C] = g;
}

__global__ __launch_bounds__(256) void k_hist_groups(int C, int DB, const int* __restrict__ tile2, const int* __restrict__ grp2,
                                                     const double* __restrict__ part, int64_t rec, double* __restrict__ hsum) {
    const int g = blockIdx.x;
    if (g >= grp2[C]) return;
    int lo = 0, hi = C - 1;
    while (lo < hi) {            // last c with grp2[c] <= g
        int mid = (lo + hi + 1) >> 1;
        if (grp2[mid] <= g) lo = mid; else hi = mid - 1;
    }
    const int c = lo;
    const int t0 = tile2[c] + (g - grp2[c]) * kRedGroupH;
    const int t1 = min(tile2[c + 1], t0 + kRedGroupH);
    for (int k = blockIdx.y * 256 + threadIdx.x; k < DB && k < (blockIdx.y + 1) * 256; k += 256) {
        double s = 0;
        for (int t = t0; t < t1; ++t) s += part[(int64_t)t * rec + kStatW + C + k];
        hsum[(int64_t)g * DB + k] = s;
    }
}

__global__ __launch_bounds__(256) void k_reduce_final(int C, int DB, const int* __restrict__ tile2, const double* __restrict__ part,
                                                      int64_t rec, const double* __restrict__ gsum, double* __restrict__ tot, int train,
                                                      const int* __restrict__ grp2, const double* __restrict__ hsum) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int ntile = tile2[C];
    const int ng = (ntile + kRedGroup - 1) / kRedGroup;
    const int W = kStatW + C;
    if (i == 0) {
        double a[7] = {tot[0], tot[1], tot[2], tot[3], tot[4], tot[5], tot[6]};
        for (int g = 0; g < ng; ++g) stat_merge(a, gsum + (int64_t)g * W);
        for (int k = 0; k < kStatW; ++k) tot[k] = a[k];
        return;
    }
    if (!train) return;
    const int64_t j = i - 1;
    if (j < C) {
        double s = tot[kStatW + j];
        for (int g = 0; g < ng; ++g) s += gsum[(int64_t)g * W + kStatW + j];
        tot[kStatW + j] = s;
        return;
    }
    const int64_t h = j - C;
    if (h >= (int64_t)C * DB) return;
    const int c = (int)(h / DB), k = (int)(h % DB);
    double s = tot[kStatW + C + h];
    for (int g = grp2[c]; g < grp2[c + 1]; ++g) s += hsum[(int64_t)g * DB + k];
    tot[kStatW + C + h] = s;
}

}  // namespace

// ---------------------------------------------------------------------------
// Learned-proposal path.  Per channel c a flow q_c on the VEGAS hypercube:
// the attempt's d uniforms z are mapped to r = T_c^{-1}(z) before the VEGAS
// forward map, and the point's density becomes
//   g(p) = sum_j alpha_j q_j(r_j(p)) / (ps_j(p) vegas_j(x_j(p))),
// r_j the VEGAS inverse of channel j's inverse point x_j.  W = 1 / g.
// ---------------------------------------------------------------------------
// z[e][k] = uniform k of attempt sorted[e]
template <int D>
__global__ __launch_bounds__(kTile) void k_zgen(MCArgs A, int64_t n, const uint32_t* __restrict__ sorted, double* __restrict__ z) {
    const int64_t e = (int64_t)blockIdx.x * kTile + threadIdx.x;
    if (e >= n) return;
    double r[D + 1];
    load_uniforms<D>(A, A.a0 + sorted[e], r);
#pragma unroll
    for (int k = 0; k < D; ++k) z[e * D + k] = r[k];
}

// All-channel inverse of every kept point: f, a_j = alpha_j / ps_j / vegas_j
// ([j][stride]) and the hypercube points r_j ([j][stride][D]).  Each channel's
// rows of the workgroup are staged in LDS (pitch D + 1 doubles, no bank
// conflicts) and written out contiguously.
template <int NT>
__global__ __launch_bounds__(kTile) void k_geom(MCArgs A, int64_t kept, const uint32_t* __restrict__ comp,
                                                const double* __restrict__ mom, int64_t nmax, double* __restrict__ rall,
                                                double* __restrict__ abuf, double* __restrict__ fbuf, int64_t stride) {
    constexpr int D = 3 * (NT - 2) - 2;
    constexpr int PR = D + 1;
    __shared__ double s_r[kTile * PR];
    const int64_t k0 = (int64_t)blockIdx.x * kTile;
    const int64_t k = k0 + threadIdx.x;
    const bool valid = k < kept;
    const int nrow = (int)std::min<int64_t>(kTile, kept - k0);
    MomRF<2 * NT> P;
    if (valid) {
        load_mom<NT>(mom, nmax, comp[k], P);
        fbuf[k] = integrand<NT>(A.f, P);
    }
    double* row = s_r + threadIdx.x * PR;
    for (int j = 0; j < A.C; ++j) {
        if (valid) geom_one<NT, false>(A, A.progs[j], j, k, P, row, abuf, stride, nullptr);
        __syncthreads();
        double* dst = rall + ((int64_t)j * stride + k0) * D;
        for (int e = threadIdx.x; e < nrow * D; e += kTile) dst[e] = s_r[(e / D) * PR + e % D];
        __syncthreads();
    }
}

// Training records of the kept points (optional outputs of k_combine):
// r [C][cap][d], beta [C][cap] = alpha_j h_j / g, w2 [cap] = value^2.
struct FlowData {
    double* r;
    double* beta;
    double* w2;
    int64_t off, cap;
};

// Own-channel ranges of the density pass: a kept point whose attempt took
// channel j's flow component (sorted positions [seg[j], fend[j]); without a
// mixture fend = seg[j + 1]) already carries log q_j from the sampling pass.
// comp is increasing, so those points are the kept indices own[2j] ..
// own[2j+1]; channel j's density rows are the other kept points in order
// (row i is kept point i below own[2j], else i + own[2j+1] - own[2j]), at
// rows dseg[j] .. dend[j] of the [C][stride] layout.
__global__ void k_own_ranges(int C, int64_t kept, const uint32_t* __restrict__ comp, const int* __restrict__ seg,
                             const int* __restrict__ fend, int64_t stride, int64_t* __restrict__ own, int* __restrict__ dseg,
                             int* __restrict__ dend) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= C) return;
    auto lower = [&](uint32_t v) {
        int64_t lo = 0, hi = kept;
        while (lo < hi) {
            const int64_t mid = (lo + hi) >> 1;
            if (comp[mid] < v) lo = mid + 1; else hi = mid;
        }
        return lo;
    };
    const int64_t a = lower((uint32_t)seg[j]), b = lower((uint32_t)(fend ? fend[j] : seg[j + 1]));
    own[2 * j] = a;
    own[2 * j + 1] = b;
    dseg[j] = (int)((int64_t)j * stride);
    dend[j] = (int)((int64_t)j * stride + kept - (b - a));
}

// Density rows of every channel (see k_own_ranges) from the r_j records:
// channel blockIdx.y, one double per thread (both sides contiguous).
__global__ void k_dens_rows(int64_t kept, int d, int64_t stride, const int64_t* __restrict__ own, const double* __restrict__ rall,
                            double* __restrict__ xw) {
    const int j = blockIdx.y;
    const int64_t lo = own[2 * j], skip = own[2 * j + 1] - lo;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (kept - skip) * d) return;
    const int64_t base = (int64_t)j * stride * d;
    xw[base + e] = rall[base + ((e < lo * d) ? e : e + skip * d)];
}

// W = 1 / sum_j a_j q_j, value = W == 0 ? 0 : f W (NaN -> 0, counted);
// per-workgroup StatsData partials [n, min, max, sum, sum2, nf, nnan].
// log q_j of kept point k: lq[j][k], or with own ranges (k_own_ranges) the
// sampling pass's lqs for the picked channel's flow component and the
// compacted density rows for the rest.
__global__ __launch_bounds__(kTile) void k_combine(int C, int d, int64_t kept, int64_t stride, const double* __restrict__ abuf,
                                                   const double* __restrict__ lq, double mix, const double* __restrict__ fbuf,
                                                   const double* __restrict__ rall, double* __restrict__ part,
                                                   double* __restrict__ vals_out, int64_t vals_off, int64_t vals_cap, FlowData FD,
                                                   const uint32_t* __restrict__ comp, const uint32_t* __restrict__ sorted, PointRec R,
                                                   const int64_t* __restrict__ own, const double* __restrict__ lqs) {
    __shared__ double s_part[kWaves][kStatW];
    const int64_t k = (int64_t)blockIdx.x * kTile + threadIdx.x;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    double val = 0, W = 0;
    bool valid = k < kept, isnanv = false;
    auto lqv = [&](int j) -> double {
        if (!own) return lq[(int64_t)j * stride + k];
        const int64_t lo = own[2 * j], hi = own[2 * j + 1];
        if (k >= lo && k < hi) return lqs[comp[k]];
        return lq[(int64_t)j * stride + (k < lo ? k : k - (hi - lo))];
    };
    if (valid) {
        double wsum = 0;
        for (int j = 0; j < C; ++j) {
            // the proposal's density: the flow's, blended with the uniform one
            const double q = lq ? (1.0 - mix) * exp(lqv(j)) + mix : 1.0;
            wsum += abuf[(int64_t)j * stride + k] * q;
        }
        W = 1.0 / wsum;
        const double fv = fbuf[k];
        val = (W == 0) ? 0.0 : fv * W;
        isnanv = isnan(val);
        if (isnanv) { val = 0; W = 0; }
        if (vals_out && vals_off + k < vals_cap) vals_out[vals_off + k] = val;
        if (FD.r && FD.off + k < FD.cap) {
            const int64_t o = FD.off + k;
            for (int j = 0; j < C; ++j) {
                FD.beta[(int64_t)j * FD.cap + o] = isnanv ? 0.0 : abuf[(int64_t)j * stride + k] / wsum;
                for (int i = 0; i < d; ++i) FD.r[((int64_t)j * FD.cap + o) * d + i] = rall[((int64_t)j * stride + k) * d + i];
            }
            FD.w2[o] = val * val;
        }
        if (R.status) {
            const uint32_t att = sorted[comp[k]];
            R.status[att] = isnanv ? 2 : 1;
            R.w[att] = W;
            R.f[att] = fv;
            R.val[att] = val;
            if (R.ps)
                for (int j = 0; j < C; ++j) R.ps[(int64_t)j * R.n + att] = lq ? lqv(j) : 0.0;
        }
    }
    double n = valid ? 1.0 : 0.0, s1 = val, s2 = val * val, nf = (valid && val != 0) ? 1.0 : 0.0, nn = isnanv ? 1.0 : 0.0;
    double mn = valid ? val : std::numeric_limits<double>::max(), mx = valid ? val : std::numeric_limits<double>::min();
    n = wave_sum(n); s1 = wave_sum(s1); s2 = wave_sum(s2); nf = wave_sum(nf); nn = wave_sum(nn);
    mn = wave_min(mn); mx = wave_max(mx);
    if (lane == 0) {
        s_part[w][0] = n; s_part[w][1] = mn; s_part[w][2] = mx; s_part[w][3] = s1; s_part[w][4] = s2; s_part[w][5] = nf; s_part[w][6] = nn;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a[kStatW] = {0, std::numeric_limits<double>::max(), std::numeric_limits<double>::min(), 0, 0, 0, 0};
        for (int ww = 0; ww < kWaves; ++ww) stat_merge(a, s_part[ww]);
        for (int q = 0; q < kStatW; ++q) part[(int64_t)blockIdx.x * kStatW + q] = a[q];
    }
}

// tot[0..6] += partials 0..np-1 in index order (one workgroup, fixed tree).
__global__ __launch_bounds__(256) void k_stats_final(int64_t np, const double* __restrict__ part, double* __restrict__ tot) {
    __shared__ double s[256][kStatW];
    double a[kStatW] = {0, std::numeric_limits<double>::max(), std::numeric_limits<double>::min(), 0, 0, 0, 0};
    const int64_t per = (np + 255) / 256;
    const int64_t b0 = threadIdx.x * per, b1 = std::min<int64_t>(np, b0 + per);
    for (int64_t b = b0; b < b1; ++b) stat_merge(a, part + b * kStatW);
    for (int q = 0; q < kStatW; ++q) s[threadIdx.x][q] = a[q];
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) stat_merge(s[threadIdx.x], s[threadIdx.x + off]);
        __syncthreads();
    }
    if (threadIdx.x == 0) stat_merge(tot, s[0]);
}

// ---------------------------------------------------------------------------
// Unweighting stage: MultiChannel::operator() with a PostProcess that keeps a
// point of value v with probability min(1, |v| / wmax) (MultiChannel.hh:
// 240-245: a rejected point adds a zero and is retried).  The decision uses
// the attempt's uniform number d+1.
// ---------------------------------------------------------------------------
// Per kept point (compaction order): accept flag, value and momentum column,
// scattered to attempt order.
__global__ __launch_bounds__(kTile) void k_unw(MCArgs A, int64_t kept, const uint32_t* __restrict__ comp,
                                               const uint32_t* __restrict__ sorted, const double* __restrict__ val, double wmax,
                                               uint8_t* __restrict__ flag, double* __restrict__ gv, uint32_t* __restrict__ gpos) {
    const int64_t k = (int64_t)blockIdx.x * kTile + threadIdx.x;
    if (k >= kept) return;
    const uint32_t sp = comp[k];
    const uint32_t a = sorted[sp];
    const double v = val[k];
    flag[a] = (aux_uniform(A, A.a0 + a) < fabs(v) / wmax) ? 1 : 0;
    gv[a] = v;
    gpos[a] = sp;
}

// Accepted events among attempts [0, used) in attempt order, one workgroup
// per kGenSub consecutive blocks of kTile attempts: momenta of the external
// legs, SoA ev[(i*4+k)][cap] from column base + blkoff[first block] + the
// event's rank in the workgroup, the event weight sign(v) max(1, |v| / wmax),
// and the workgroup's StatsData partial of the accepted values [n, min, max,
// sum, sum2, nf, nnan] (each lane accumulates its attempts in order, then
// lanes and waves in order).  The momenta are copied by the whole workgroup
// (one double per thread and step, rows of consecutive columns).
constexpr int kGenSub = 16;
template <int NT>
__global__ __launch_bounds__(kTile) void k_gen_out(int64_t used, int nblk, const uint8_t* __restrict__ flag,
                                                   const int* __restrict__ blkoff, const uint32_t* __restrict__ gpos,
                                                   const double* __restrict__ gv, double wmax, const double* __restrict__ mom,
                                                   int64_t nmax, double* __restrict__ ev, double* __restrict__ ew, int64_t cap,
                                                   int64_t base, double* __restrict__ part) {
    __shared__ int s_cnt[kGenSub][kWaves];
    __shared__ int s_pre[kGenSub + 1];
    __shared__ uint32_t s_sp[kGenSub * kTile];
    __shared__ double s_part[kWaves][kStatW];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint64_t below = (uint64_t(1) << lane) - 1;
    const int b0 = blockIdx.x * kGenSub;
    const int nb = min(kGenSub, nblk - b0);
    uint64_t masks[kGenSub];
#pragma unroll
    for (int t = 0; t < kGenSub; ++t) {
        const int64_t i = (int64_t)(b0 + t) * kTile + threadIdx.x;
        masks[t] = __ballot(t < nb && i < used && flag[i]);
        if (lane == 0) s_cnt[t][w] = __popcll(masks[t]);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int run = 0;
        for (int t = 0; t < kGenSub; ++t) {
            s_pre[t] = run;
            for (int ww = 0; ww < kWaves; ++ww) run += s_cnt[t][ww];
        }
        s_pre[kGenSub] = run;
    }
    __syncthreads();
    const int64_t col0 = base + blkoff[b0];
    double a[kStatW] = {0, __DBL_MAX__, __DBL_MIN__, 0, 0, 0, 0};
#pragma unroll
    for (int t = 0; t < kGenSub; ++t) {
        if (!((masks[t] >> lane) & 1)) continue;
        const int64_t i = (int64_t)(b0 + t) * kTile + threadIdx.x;
        int r = s_pre[t] + __popcll(masks[t] & below);
        for (int ww = 0; ww < w; ++ww) r += s_cnt[t][ww];
        s_sp[r] = gpos[i];
        const double v = gv[i];
        ew[col0 + r] = (fabs(v) > wmax) ? v / wmax : copysign(1.0, v);
        a[0] += 1.0;
        a[1] = (v < a[1]) ? v : a[1];
        a[2] = (a[2] < v) ? v : a[2];
        a[3] += v;
        a[4] += v * v;
        if (v != 0) a[5] += 1.0;
    }
    __syncthreads();
    const int cnt = s_pre[kGenSub];
    for (int j = threadIdx.x; j < cnt * 4 * NT; j += kTile) {
        const int q = j / cnt, e = j - q * cnt;
        ev[(int64_t)q * cap + col0 + e] = mom[(int64_t)q * nmax + s_sp[e]];
    }
    const double n = wave_sum(a[0]), s1 = wave_sum(a[3]), s2 = wave_sum(a[4]), nf = wave_sum(a[5]);
    const double mn = wave_min(a[1]), mx = wave_max(a[2]);
    if (lane == 0) {
        double* q = s_part[w];
        q[0] = n; q[1] = mn; q[2] = mx; q[3] = s1; q[4] = s2; q[5] = nf; q[6] = 0.0;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double acc[kStatW] = {0, __DBL_MAX__, __DBL_MIN__, 0, 0, 0, 0};
        for (int ww = 0; ww < kWaves; ++ww) stat_merge(acc, s_part[ww]);
        for (int q = 0; q < kStatW; ++q) part[(int64_t)blockIdx.x * kStatW + q] = acc[q];
    }
}

// Sum of the nb per-block counts into *total (one workgroup, fixed order).
__global__ __launch_bounds__(1024) void k_count_total(const int* __restrict__ cnt, int64_t nb, int64_t* __restrict__ total) {
    __shared__ int64_t s[1024 / 64];
    int64_t v = 0;
    for (int64_t i = threadIdx.x; i < nb; i += 1024) v += cnt[i];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t t = 0;
        for (int w = 0; w < 1024 / 64; ++w) t += s[w];
        *total = t;
    }
}

// ---------------------------------------------------------------------------
// Batch evaluation (the reference's ChiliCPU.evaluate / CallChili,
// nanobind.hh:23-50, Tensorflow.cc:22-25): per row of randoms, one channel's
// FSMapper::GeneratePoint on the randoms as they are (no VEGAS map), the
// PreProcess cut (0 when it fails) and val = f * GenerateWeight (that
// channel's phase-space weight; 0 when it is 0).
// ---------------------------------------------------------------------------
template <int NT>
__global__ __launch_bounds__(kTile) void k_eval_single(MCArgs A, int ch, const double* __restrict__ rans, int64_t n, int64_t ld,
                                                       double* __restrict__ out) {
    constexpr int D = 3 * (NT - 2) - 2;
    const int64_t i = (int64_t)blockIdx.x * kTile + threadIdx.x;
    if (i >= n) return;
    double r[D + 1];
#pragma unroll
    for (int k = 0; k < D; ++k) r[k] = rans[i * ld + k];
    r[D] = 0.0;
    const ChanProg& p = A.progs[ch];
    MomRF<2 * NT> P;
    double m2[2 * NT];
    generate_point<NT>(A, p, r, P, m2);
    if (!cut_pass<NT>(A.f, P, A.sqrts)) { out[i] = 0.0; return; }
    double vw;
    const double ps = generate_weight<NT>(A, p, ch, P, &vw, nullptr, 0);
    const double fv = integrand<NT>(A.f, P);
    out[i] = (ps == 0) ? 0.0 : fv * ps;
}

// Rows [n][ld] of randoms -> SoA uniforms [k][n] (k < nk); counts entries
// outside [0, 1) into bad.
__global__ __launch_bounds__(kTile) void k_rows_to_soa(const double* __restrict__ rans, int64_t n, int64_t ld, int nk,
                                                       double* __restrict__ u, int* __restrict__ bad) {
    const int64_t i = (int64_t)blockIdx.x * kTile + threadIdx.x;
    if (i >= n) return;
    int nb = 0;
    for (int k = 0; k < nk; ++k) {
        const double v = rans[i * ld + k];
        nb += (v >= 0.0 && v < 1.0) ? 0 : 1;
        u[(int64_t)k * n + i] = v;
    }
    if (nb) atomicAdd(bad, nb);
}

// Values of the kept points (compaction order) to attempt order; attempts
// that failed the cut keep the 0 they were filled with.
__global__ __launch_bounds__(kTile) void k_vals_to_rows(int64_t kept, const uint32_t* __restrict__ comp, const uint32_t* __restrict__ sorted,
                                                        const double* __restrict__ val, double* __restrict__ out) {
    const int64_t k = (int64_t)blockIdx.x * kTile + threadIdx.x;
    if (k >= kept) return;
    out[sorted[comp[k]]] = val[k];
}

// ===========================================================================
// host context
// ===========================================================================
struct chili_mc {
    int device = 0;
    hipStream_t stream = nullptr;
    chili_host::ChannelSet set;
    chili_integrand f{};
    int C = 0, d = 0, NT = 0, B = 0, P2 = 0;
    std::vector<double> grids;   // [C][d][B+1]
    std::vector<double> alpha;   // [C]
    bool have_grids = false;
    // device buffers
    DevBuf<ChanProg> d_progs;
    DevBuf<double> d_edges, d_search, d_alpha, d_cum, d_sext, d_tot, d_part, d_mom, d_ps, d_u, d_gsum, d_hsum;
    DevBuf<uint16_t> d_cells;      // cell index of the search tables [C][d][kCells + 1]
    DevBuf<chili_integrand> d_f;
    DevBuf<uint16_t> d_pick;
    DevBuf<int> d_fend;          // flow path with a mixture: end of each channel's flow attempts
    // unweighting stage: values of the kept points, per-attempt flag / value /
    // momentum column, block offsets and partials
    DevBuf<double> d_gval, d_gv, d_gpart;
    DevBuf<uint8_t> d_gflag;
    DevBuf<uint32_t> d_gpos;
    DevBuf<int> d_gblkoff;
    DevBuf<uint32_t> d_sorted, d_comp;
    DevBuf<uint32_t> d_sorted2;     // packed order written by the forward kernel (swapped with d_sorted)
    DevBuf<uint8_t> d_acc;
    DevBuf<int> d_counts, d_offs, d_seg, d_tiles, d_seg2, d_tiles2, d_blkacc, d_kept, d_keptoff, d_grp2;
    DevBuf<int64_t> d_total, d_bsum;
    int64_t* h_pin = nullptr;      // pinned host words for the per-sub-batch counts
    // learned-proposal path: z/r [fmax][d], r_j [C][fmax][d], a_j and log q_j
    // [C][fmax], f [fmax], sampling log-density [fmax], StatsData partials
    DevBuf<double> d_z, d_rall, d_abuf, d_lq, d_fbuf, d_lqs, d_lqs2, d_fpart, d_xw;
    DevBuf<int64_t> d_own;         // own-channel kept ranges [C][2] of the density pass
    DevBuf<int> d_dseg, d_dend;    // density rows of each channel
    int64_t fmax = 0;
    int64_t nmax = 0;            // attempts per sub-batch
    double acc_rate = 0.25;      // running estimate of the cut acceptance
    // run-time specialised forward / weight kernels (channel programs as constants)
    std::unique_ptr<chili_jit::Module> jit;
    bool jit_fwd = false, jit_main = false;
    bool jit_split = false;      // split weight pass: per-channel specialised inverse kernels + combine
    bool jit_early = false;      // unweighting stage: exact early rejection by the picked channel's bound
    DevBuf<double> d_pre;        // split pass: ps and VEGAS weights [2 C][kept], then f [kept]
    DevBuf<short> d_prebins;     // split pass: the picked channel's bins [d][kept]
    bool lds_search = true;      // weight kernel: search tables in LDS when they fit
    double gen_rate = 0.01;      // unweighting stage: kept events per attempt (running estimate)
    double flow_gen_rate = 0.01; // the same on the learned-proposal path
    double jit_compile_s = 0;
    // k_main timing with HIP events on the context's stream
    bool timing = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_live, ev_pool;
    std::vector<int64_t> ev_points;
    double main_ms = 0, fwd_ms = 0;
    int64_t main_launches = 0, main_points = 0;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> fev_live;
    std::vector<int64_t> fev_points;       // attempts of each timed forward launch
    int64_t fwd_launches = 0, fwd_points = 0;
    hipEvent_t producer_ev = nullptr;      // chili_mc_evaluate: the caller's stream -> the context's
    // host integrand / PreProcess callback (the reference's std::function path)
    chili_host_integrand_fn host_fn = nullptr;
    void* host_user = nullptr;
    DevBuf<double> d_fhost;                // callback values by sorted position
    ~chili_mc() {
        if (h_pin) hipHostFree(h_pin);
        if (producer_ev) hipEventDestroy(producer_ev);
        for (auto& e : ev_live) { hipEventDestroy(e.first); hipEventDestroy(e.second); }
        for (auto& e : fev_live) { hipEventDestroy(e.first); hipEventDestroy(e.second); }
        for (auto& e : ev_pool) { hipEventDestroy(e.first); hipEventDestroy(e.second); }
    }
};

namespace {

int next_pow2(int x) { int p = 1; while (p < x) p <<= 1; return p; }

void scan_int(chili_mc* m, const int* in, int* out, int64_t len);

// Grid edges must be non-decreasing (AdaptiveMap's Adapt and Split keep
// them so); the device bin search relies on it.
void check_edges(const double* e, size_t rows, int B) {
    for (size_t r = 0; r < rows; ++r)
        for (int b = 0; b < B; ++b)
            if (!(e[r * (B + 1) + b] <= e[r * (B + 1) + b + 1])) throw chili_error(CHILI_EINVAL, "grid edges must be non-decreasing");
}

// Host-fed uniforms must lie in [0, 1): 1.0 would map to bin B (past the
// grid) exactly as it would index past the reference's edge array.
void check_uniforms(const double* u, size_t cnt) {
    for (size_t i = 0; i < cnt; ++i)
        if (!(u[i] >= 0.0 && u[i] < 1.0)) throw chili_error(CHILI_EINVAL, "host uniforms must lie in [0, 1)");
}

void upload_state(chili_mc* m) {
    const int C = m->C, d = m->d, B = m->B, P2 = m->P2;
    std::vector<double> srch((size_t)C * d * P2, std::numeric_limits<double>::infinity());
    for (int c = 0; c < C; ++c)
        for (int k = 0; k < d; ++k)
            for (int b = 1; b < B; ++b) srch[((size_t)c * d + k) * P2 + b - 1] = m->grids[((size_t)c * d + k) * (B + 1) + b];
    m->d_edges.upload(m->grids, m->stream);
    m->d_search.upload(srch, m->stream);
    // cell index: t[j] = interior edges strictly below j / kCells, t[kCells]
    // = those below 1 (exact: j / kCells is a dyadic double)
    std::vector<uint16_t> cells((size_t)C * d * (kCells + 1));
    for (size_t r = 0; r < (size_t)C * d; ++r) {
        const double* e = m->grids.data() + r * (B + 1);
        int cnt = 0;
        for (int j = 0; j <= kCells; ++j) {
            const double y = (double)j / kCells;
            while (cnt < B - 1 && e[cnt + 1] < y) ++cnt;
            cells[r * (kCells + 1) + j] = (uint16_t)cnt;
        }
    }
    m->d_cells.upload(cells, m->stream);
    // cumulative alpha as std::discrete_distribution builds it (random.tcc:2656-2677)
    std::vector<double> cum(C, 1.0);
    if (C >= 2) {
        double sum = 0.0;
        for (double a : m->alpha) sum += a;
        double acc = 0;
        for (int i = 0; i < C; ++i) { double p = m->alpha[i] / sum; acc = (i == 0) ? p : acc + p; cum[i] = acc; }
        cum[C - 1] = 1.0;
    }
    m->d_alpha.upload(m->alpha, m->stream);
    m->d_cum.upload(cum, m->stream);
}

void ensure_capacity(chili_mc* m, int64_t n) {
    if (n <= m->nmax) return;
    const int64_t nblk = (n + kTile - 1) / kTile;
    const int64_t ntiles = nblk + m->C + 1;
    m->d_pick.resize(n);
    m->d_sorted.resize(n);
    m->d_sorted2.resize(n);
    m->d_comp.resize(n);
    m->d_acc.resize(n);
    m->d_mom.resize((size_t)n * m->NT * 4);
    m->d_counts.resize((size_t)nblk * m->C);
    m->d_offs.resize((size_t)nblk * m->C);
    m->d_blkacc.resize(nblk);
    m->d_kept.resize(ntiles);
    m->d_keptoff.resize(ntiles);
    const int64_t rec = kStatW + m->C + (int64_t)m->d * m->B;
    m->d_part.resize((size_t)ntiles * rec);
    m->d_ps.resize((size_t)ntiles * kTile * m->C);
    m->nmax = n;
}

MCArgs make_args(chili_mc* m, const double* u, int64_t ustride, uint64_t seed, uint32_t stream, uint64_t a0) {
    MCArgs A{};
    A.progs = m->d_progs.ptr;
    A.edges = m->d_edges.ptr;
    A.search = m->d_search.ptr;
    A.cells = m->d_cells.ptr;
    A.alpha = m->d_alpha.ptr;
    A.cum = m->d_cum.ptr;
    A.sext = m->d_sext.ptr;
    A.f = m->d_f.ptr;
    A.u = u;
    A.ustride = ustride;
    A.seed = seed;
    A.a0 = a0;
    A.stream = stream;
    A.C = m->C;
    A.d = m->d;
    A.B = m->B;
    A.P2 = m->P2;
    A.sqrts = m->set.model.sqrts;
    A.ptmax = m->set.model.sqrts / 2.0;
    A.fhost = m->host_fn ? m->d_fhost.ptr : nullptr;
    return A;
}

// a PreProcess rejection of the host callback, in the per-attempt records
__global__ void k_host_reject(int64_t n, const uint8_t* __restrict__ acc, int32_t* __restrict__ status) {
    const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (a < n && !acc[a] && status[a] == 1) status[a] = 0;
}

// The host integrand and PreProcess (Integrand.hh:41-47, MultiChannel.hh:
// 219-229) of the attempts of a sub-batch that passed the device cut: their
// momenta go to the host in attempt order, the callback returns f and the
// PreProcess decision, the decisions join the attempt flags and the values
// wait, by sorted position, for the weight kernel.
void host_pass(chili_mc* m, int64_t n, const PointRec& R) {
    const int NT = m->NT;
    std::vector<uint8_t> acc((size_t)n);
    std::vector<uint32_t> sorted((size_t)n);
    std::vector<double> cols((size_t)4 * NT * n);
    HIP_CHECK(hipMemcpyAsync(acc.data(), m->d_acc.ptr, (size_t)n, hipMemcpyDeviceToHost, m->stream));
    HIP_CHECK(hipMemcpyAsync(sorted.data(), m->d_sorted.ptr, sizeof(uint32_t) * (size_t)n, hipMemcpyDeviceToHost, m->stream));
    for (int r = 0; r < 4 * NT; ++r)
        HIP_CHECK(hipMemcpyAsync(cols.data() + (size_t)r * n, m->d_mom.ptr + (size_t)r * m->nmax, sizeof(double) * (size_t)n,
                                 hipMemcpyDeviceToHost, m->stream));
    HIP_CHECK(hipStreamSynchronize(m->stream));
    std::vector<int64_t> pos((size_t)n);                 // attempt -> sorted position
    for (int64_t e = 0; e < n; ++e) pos[sorted[(size_t)e]] = e;
    std::vector<int64_t> live;
    for (int64_t a = 0; a < n; ++a)
        if (acc[(size_t)a]) live.push_back(a);
    const int64_t nl = (int64_t)live.size();
    std::vector<double> mom((size_t)nl * NT * 4), fv((size_t)nl, 0.0), fpos((size_t)n, 0.0);
    std::vector<int32_t> pass((size_t)nl, 1);
    for (int64_t k = 0; k < nl; ++k) {
        const int64_t e = pos[(size_t)live[(size_t)k]];
        for (int r = 0; r < 4 * NT; ++r) mom[(size_t)k * NT * 4 + r] = cols[(size_t)r * n + e];
    }
    if (nl > 0 && m->host_fn(m->host_user, nl, NT, mom.data(), fv.data(), pass.data()) != 0)
        throw chili_error(CHILI_EINVAL, "the host integrand callback reported an error");
    for (int64_t k = 0; k < nl; ++k) {
        const int64_t a = live[(size_t)k];
        if (!pass[(size_t)k]) acc[(size_t)a] = 0;
        fpos[(size_t)pos[(size_t)a]] = fv[(size_t)k];
    }
    if (m->d_fhost.n < (size_t)n) throw chili_error(CHILI_EINVAL, "host value buffer smaller than the sub-batch");
    HIP_CHECK(hipMemcpyAsync(m->d_acc.ptr, acc.data(), (size_t)n, hipMemcpyHostToDevice, m->stream));
    HIP_CHECK(hipMemcpyAsync(m->d_fhost.ptr, fpos.data(), sizeof(double) * (size_t)n, hipMemcpyHostToDevice, m->stream));
    if (R.status)
        hipLaunchKernelGGL(k_host_reject, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, m->stream, n, m->d_acc.ptr, R.status);
    // the staging vectors are freed on return: finish the copies first
    HIP_CHECK(hipStreamSynchronize(m->stream));
}

void no_host_callback(const chili_mc* m, const char* what) {
    if (m->host_fn) throw chili_error(CHILI_EINVAL, std::string(what) + " evaluates the integrand on the device: clear the host callback first");
}

template <int NT>
void launch_forward(chili_mc* m, const MCArgs& A, int grid, const PointRec& R, const double* z = nullptr, bool pack = true) {
    // packed tiles (the kept points' momenta contiguous): the packed order
    // replaces the sorted order for the rest of the sub-batch, and on the
    // learned-proposal path the sampling log-densities (d_lqs) move with it.
    // A mixture proposal keeps the sorted order (pack = false): its tiles
    // can hold both components of a channel, whose ranges must stay put.
    uint32_t* out = pack ? m->d_sorted2.ptr : nullptr;
    const double* lqi = (pack && z) ? m->d_lqs.ptr : nullptr;
    double* lqo = (pack && z) ? m->d_lqs2.ptr : nullptr;
    if (m->jit_fwd && !z) {
        MCArgs a = A;
        const int* seg = m->d_seg.ptr;
        const int* tile = m->d_tiles.ptr;
        const uint32_t* sorted = m->d_sorted.ptr;
        double* mom = m->d_mom.ptr;
        int64_t nmax = m->nmax;
        uint8_t* acc = m->d_acc.ptr;
        PointRec r = R;
        const double* zs = nullptr;
        void* args[] = {&a, &seg, &tile, &sorted, &mom, &nmax, &acc, &r, &zs, &out};
        HIP_CHECK(hipModuleLaunchKernel(m->jit->forward, grid, 1, 1, kTile, 1, 1, 0, m->stream, args, nullptr));
    } else if (z) {
        hipLaunchKernelGGL((k_forward<NT, true>), dim3(grid), dim3(kTile), 0, m->stream, A, m->d_seg.ptr, m->d_tiles.ptr, m->d_sorted.ptr,
                           m->d_mom.ptr, m->nmax, m->d_acc.ptr, R, z, out, lqi, lqo);
    } else {
        hipLaunchKernelGGL((k_forward<NT, false>), dim3(grid), dim3(kTile), 0, m->stream, A, m->d_seg.ptr, m->d_tiles.ptr, m->d_sorted.ptr,
                           m->d_mom.ptr, m->nmax, m->d_acc.ptr, R, z, out, lqi, lqo);
    }
    if (out) std::swap(m->d_sorted, m->d_sorted2);
    if (lqo) std::swap(m->d_lqs, m->d_lqs2);
}
template <int NT>
void launch_main(chili_mc* m, const MCArgs& A, int grid, int train, double* vals, int64_t vals_off, int64_t vals_cap,
                 const PointRec& R, int64_t kept) {
    constexpr int D = 3 * (NT - 2) - 2;
    if (m->jit_split) {
        // split weight pass: every channel's inverse in its own specialised
        // kernel (ps, VEGAS weight, the picked channel's bins by kept index),
        // then the combination (W, value, statistics, training sums)
        m->d_pre.resize((size_t)(2 * m->C + 1) * kept);
        m->d_prebins.resize((size_t)D * kept);
        MCArgs a = A;
        a.pre_pv = m->d_pre.ptr;
        a.pre_bins = m->d_prebins.ptr;
        a.pre_stride = kept;
        const size_t one_bytes = sizeof(double) * (size_t)D * m->P2;
        int lds_search = (m->lds_search && one_bytes <= kLdsSearchMax) ? 1 : 0;
        const int* seg2 = m->d_seg2.ptr;
        const uint32_t* comp = m->d_comp.ptr;
        const double* mom = m->d_mom.ptr;
        int64_t nmax = m->nmax, kp = kept;
        double* pv = m->d_pre.ptr;
        short* bins = m->d_prebins.ptr;
        const uint32_t* sorted = m->d_sorted.ptr;
        PointRec r = R;
        void* args[] = {&a, &kp, &seg2, &comp, &mom, &nmax, &pv, &bins, &sorted, &r, &lds_search};
        const unsigned g = (unsigned)((kept + kTile - 1) / kTile);
        for (int j = 0; j < m->C; ++j)
            HIP_CHECK(hipModuleLaunchKernel(m->jit->inv[(size_t)j], g, 1, 1, kTile, 1, 1, lds_search ? (unsigned)one_bytes : 0u, m->stream,
                                            args, nullptr));
        const size_t lds = sizeof(double) * (D * m->B + kWaves * (8 + m->C)) + sizeof(short) * D * kTile;
        hipLaunchKernelGGL((k_main_pre<NT>), dim3(grid), dim3(kTile), lds, m->stream, a, m->d_seg2.ptr, m->d_tiles2.ptr, m->d_comp.ptr,
                           m->d_mom.ptr, m->nmax, m->d_ps.ptr, (int64_t)grid * kTile, train, m->d_part.ptr,
                           (int64_t)(kStatW + m->C + (int64_t)m->d * m->B), vals, vals_off, vals_cap, m->d_sorted.ptr, R);
        return;
    }
    size_t lds = sizeof(double) * (D * m->B + kWaves * (8 + m->C)) + sizeof(short) * D * kTile;
    if (lds > 160 * 1024) throw chili_error(CHILI_ERANGE, "channel count x bins exceed the LDS of one workgroup");
    // small search tables go to LDS with the rest of the workgroup's state
    const size_t srch_bytes = sizeof(double) * (size_t)m->C * D * m->P2;
    const bool ls = m->lds_search && srch_bytes <= kLdsSearchMax;
    if (ls) lds += srch_bytes;
    const int64_t rec = kStatW + m->C + (int64_t)m->d * m->B;
    if (m->jit_main) {
        MCArgs a = A;
        const int* seg2 = m->d_seg2.ptr;
        const int* tile2 = m->d_tiles2.ptr;
        const uint32_t* comp = m->d_comp.ptr;
        const double* mom = m->d_mom.ptr;
        int64_t nmax = m->nmax;
        double* ps = m->d_ps.ptr;
        int64_t psstride = (int64_t)grid * kTile;
        int tr = train;
        double* part = m->d_part.ptr;
        int64_t rc = rec;
        double* v = vals;
        int64_t voff = vals_off, vcap = vals_cap;
        const uint32_t* sorted = m->d_sorted.ptr;
        PointRec r = R;
        void* args[] = {&a, &seg2, &tile2, &comp, &mom, &nmax, &ps, &psstride, &tr, &part, &rc, &v, &voff, &vcap, &sorted, &r};
        HIP_CHECK(hipModuleLaunchKernel(ls ? m->jit->main_lds : m->jit->main, grid, 1, 1, kTile, 1, 1, (unsigned)lds, m->stream,
                                        args, nullptr));
        return;
    }
    if (ls)
        hipLaunchKernelGGL((k_main<NT, true>), dim3(grid), dim3(kTile), lds, m->stream, A, m->d_seg2.ptr, m->d_tiles2.ptr, m->d_comp.ptr,
                           m->d_mom.ptr, m->nmax, m->d_ps.ptr, (int64_t)grid * kTile, train, m->d_part.ptr, rec, vals, vals_off,
                           vals_cap, m->d_sorted.ptr, R);
    else
        hipLaunchKernelGGL((k_main<NT, false>), dim3(grid), dim3(kTile), lds, m->stream, A, m->d_seg2.ptr, m->d_tiles2.ptr, m->d_comp.ptr,
                           m->d_mom.ptr, m->nmax, m->d_ps.ptr, (int64_t)grid * kTile, train, m->d_part.ptr, rec, vals, vals_off,
                           vals_cap, m->d_sorted.ptr, R);
}

#define CHILI_DISPATCH_NT(NTV, CALL)           \
    switch (NTV) {                             \
        case 4: { constexpr int NT = 4; CALL; } break;  \
        case 5: { constexpr int NT = 5; CALL; } break;  \
        case 6: { constexpr int NT = 6; CALL; } break;  \
        case 7: { constexpr int NT = 7; CALL; } break;  \
        case 8: { constexpr int NT = 8; CALL; } break;  \
        case 9: { constexpr int NT = 9; CALL; } break;  \
        case 10: { constexpr int NT = 10; CALL; } break; \
        default: throw chili_error(CHILI_EINVAL, "unsupported number of external particles"); \
    }

void scan_int(chili_mc* m, const int* in, int* out, int64_t len) {
    const int64_t nb = (len + kScanBlock - 1) / kScanBlock;
    m->d_bsum.resize(nb);
    hipLaunchKernelGGL(k_scan_sums, dim3((unsigned)nb), dim3(kScanBlock), 0, m->stream, in, len, m->d_bsum.ptr);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kScanBlock), 0, m->stream, m->d_bsum.ptr, nb, m->d_total.ptr);
    hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nb), dim3(kScanBlock), 0, m->stream, in, out, len, m->d_bsum.ptr);
}

std::pair<hipEvent_t, hipEvent_t> take_events(chili_mc* m) {
    if (!m->ev_pool.empty()) {
        auto e = m->ev_pool.back();
        m->ev_pool.pop_back();
        return e;
    }
    std::pair<hipEvent_t, hipEvent_t> e;
    HIP_CHECK(hipEventCreate(&e.first));
    HIP_CHECK(hipEventCreate(&e.second));
    return e;
}
// Folds finished k_main / k_forward event pairs into the totals (stream synchronised).
void harvest_events(chili_mc* m) {
    for (size_t i = 0; i < m->ev_live.size(); ++i) {
        float ms = 0;
        HIP_CHECK(hipEventElapsedTime(&ms, m->ev_live[i].first, m->ev_live[i].second));
        m->main_ms += ms;
        m->main_launches += 1;
        m->main_points += m->ev_points[i];
        m->ev_pool.push_back(m->ev_live[i]);
    }
    for (size_t i = 0; i < m->fev_live.size(); ++i) {
        float ms = 0;
        HIP_CHECK(hipEventElapsedTime(&ms, m->fev_live[i].first, m->fev_live[i].second));
        m->fwd_ms += ms;
        m->fwd_launches += 1;
        m->fwd_points += m->fev_points[i];
        m->ev_pool.push_back(m->fev_live[i]);
    }
    m->ev_live.clear();
    m->ev_points.clear();
    m->fev_live.clear();
    m->fev_points.clear();
}

// Cutoff (the attempt at which `need` accepted attempts are reached; host
// sync) and stable compaction of the kept entries of the forward tiles.
// The attempt at which `need` flags (attempt order) are reached: kept =
// flags up to it, used = attempts up to it (all n when fewer are set); the
// per-block flag counts stay in d_blkacc.
void find_cutoff(chili_mc* m, const uint8_t* flags, int64_t n, int nblk, int64_t need, int64_t& kept, int64_t& used) {
    hipLaunchKernelGGL(k_acc_count, dim3(nblk), dim3(kTile), 0, m->stream, flags, n, m->d_blkacc.ptr);
    // the total first (8 bytes to pinned memory): the block counts are only
    // read when the cutoff lies in this sub-batch
    m->d_total.resize(1);
    if (!m->h_pin) HIP_CHECK(hipHostMalloc((void**)&m->h_pin, 8 * sizeof(int64_t), hipHostMallocDefault));
    hipLaunchKernelGGL(k_count_total, dim3(1), dim3(1024), 0, m->stream, m->d_blkacc.ptr, (int64_t)nblk, m->d_total.ptr);
    HIP_CHECK(hipMemcpyAsync(m->h_pin, m->d_total.ptr, sizeof(int64_t), hipMemcpyDeviceToHost, m->stream));
    HIP_CHECK(hipStreamSynchronize(m->stream));
    if (m->h_pin[0] < need) {
        kept = m->h_pin[0];
        used = n;
        return;
    }
    std::vector<int> blkacc(nblk);
    m->d_blkacc.download(blkacc.data(), nblk, m->stream);
    HIP_CHECK(hipStreamSynchronize(m->stream));
    int64_t run = 0;
    int64_t cut = n;
    for (int b = 0; b < nblk; ++b) {
        if (run + blkacc[b] >= need) {
            std::vector<uint8_t> fl(kTile);
            const int64_t base = (int64_t)b * kTile;
            const int64_t cnt = std::min<int64_t>(kTile, n - base);
            HIP_CHECK(hipMemcpyAsync(fl.data(), flags + base, cnt, hipMemcpyDeviceToHost, m->stream));
            HIP_CHECK(hipStreamSynchronize(m->stream));
            for (int64_t i = 0; i < cnt; ++i) {
                run += fl[i];
                if (run == need) { cut = base + i + 1; break; }
            }
            break;
        }
        run += blkacc[b];
    }
    kept = std::min<int64_t>(run, need);
    used = cut;
}

void cutoff_and_compact(chili_mc* m, int64_t n, int nblk, int tgrid, int64_t need, int64_t& kept, int64_t& used) {
    const int C = m->C;
    find_cutoff(m, m->d_acc.ptr, n, nblk, need, kept, used);
    const int64_t cut = used;
    hipLaunchKernelGGL(k_keep_count, dim3(tgrid), dim3(kTile), 0, m->stream, C, m->d_seg.ptr, m->d_tiles.ptr, m->d_sorted.ptr, m->d_acc.ptr,
                       (uint32_t)cut, m->d_kept.ptr);
    scan_int(m, m->d_kept.ptr, m->d_keptoff.ptr, (int64_t)tgrid);
    hipLaunchKernelGGL(k_keep_scatter, dim3(tgrid), dim3(kTile), 0, m->stream, C, m->d_seg.ptr, m->d_tiles.ptr, m->d_sorted.ptr, m->d_acc.ptr,
                       (uint32_t)cut, m->d_keptoff.ptr, m->d_comp.ptr);
    hipLaunchKernelGGL(k_keep_segments, dim3(1), dim3(1), 0, m->stream, C, m->d_tiles.ptr, m->d_keptoff.ptr, kept, m->d_seg2.ptr,
                       m->d_tiles2.ptr, kMainTile);
}

// The kT-jet PreProcess pass (no-op for the other cut kinds).
void launch_ktcut(chili_mc* m, int64_t n, const PointRec& R) {
    if (m->f.cuts != CHILI_CUTS_KT_JETS) return;
    const unsigned g = (unsigned)((n + kTile - 1) / kTile);
    CHILI_DISPATCH_NT(m->NT, hipLaunchKernelGGL((k_ktcut<NT>), dim3(g), dim3(kTile), 0, m->stream, m->d_f.ptr, n, m->d_sorted.ptr,
                                                m->d_mom.ptr, m->nmax, m->d_acc.ptr, R));
}

// Channel picks, the stable sort by channel and the forward map + cut of n
// attempts; returns the forward grid (tiles).
int pick_sort_forward(chili_mc* m, const MCArgs& A, int64_t n, const PointRec& R) {
    const int C = m->C;
    const int nblk = (int)((n + kTile - 1) / kTile);
    hipLaunchKernelGGL(k_pick, dim3(nblk), dim3(kTile), sizeof(int) * C, m->stream, A, n, m->d_pick.ptr, m->d_counts.ptr, nblk);
    scan_int(m, m->d_counts.ptr, m->d_offs.ptr, (int64_t)nblk * C);
    hipLaunchKernelGGL(k_scatter, dim3(nblk), dim3(kTile), sizeof(int) * kWaves * C, m->stream, C, n, m->d_pick.ptr, m->d_offs.ptr, nblk,
                       m->d_sorted.ptr);
    hipLaunchKernelGGL(k_segments, dim3(1), dim3(1), 0, m->stream, C, m->d_offs.ptr, nblk, n, m->d_seg.ptr, m->d_tiles.ptr);
    const int tgrid = nblk + C;
    std::pair<hipEvent_t, hipEvent_t> fev{};
    if (m->timing) {
        fev = take_events(m);
        HIP_CHECK(hipEventRecord(fev.first, m->stream));
    }
    CHILI_DISPATCH_NT(m->NT, launch_forward<NT>(m, A, tgrid, R));
    if (m->timing) {
        HIP_CHECK(hipEventRecord(fev.second, m->stream));
        m->fev_live.push_back(fev);
        m->fev_points.push_back(n);
    }
    launch_ktcut(m, n, R);
    return tgrid;
}

void gen_tail(chili_mc* m, const MCArgs& A, int64_t n, int nblk, int64_t need, double wmax, double* ev, double* ew, int64_t cap,
              int64_t got, int64_t passed, int64_t& kept, int64_t& used);

// Exact early rejection (own_bound_one) over the passing points of the
// current compaction: a point whose picked channel's bound cannot pass the
// accept test loses its flag.  lqs / mix: the learned proposal's sampling
// log-densities and mixture weight (null / 0 on the plain path).
void launch_early_reject(chili_mc* m, const MCArgs& A, int64_t passed, double wmax, const double* lqs, double mix) {
    constexpr size_t kOwnLdsMax = 32 * 1024;
    const size_t tbl = sizeof(double) * (size_t)m->C * m->d * m->P2;
    int lds_search = (m->lds_search && tbl <= kOwnLdsMax) ? 1 : 0;
    MCArgs a = A;
    int64_t kp = passed, nmax = m->nmax;
    const int* seg2 = m->d_seg2.ptr;
    const uint32_t* comp = m->d_comp.ptr;
    const double* mom = m->d_mom.ptr;
    const uint32_t* sorted = m->d_sorted.ptr;
    uint8_t* acc = m->d_acc.ptr;
    double wm = wmax, mx = mix;
    const double* lq = lqs;
    void* args[] = {&a, &kp, &seg2, &comp, &mom, &nmax, &sorted, &acc, &wm, &lq, &mx, &lds_search};
    HIP_CHECK(hipModuleLaunchKernel(m->jit->own, (unsigned)((passed + kTile - 1) / kTile), 1, 1, kTile, 1, 1,
                                    lds_search ? (unsigned)tbl : 0u, m->stream, args, nullptr));
}

// One sub-batch of the unweighting stage over attempts [a0, a0 + n): every
// point passing the cut is weighted (production: no training sums; with
// early rejection only the candidates), the accept / reject decisions are
// made per attempt, the cutoff is the attempt of the need-th accepted event;
// StatsData of the accepted events goes to d_tot, the events to ev / ew
// from column got.
void run_subbatch_gen(chili_mc* m, uint64_t seed, uint32_t stream, uint64_t a0, int64_t n, int64_t need, double wmax, double* ev,
                      double* ew, int64_t cap, int64_t got, int64_t& kept, int64_t& used) {
    ensure_capacity(m, n);
    const int C = m->C;
    const int nblk = (int)((n + kTile - 1) / kTile);
    MCArgs A = make_args(m, nullptr, 0, seed, stream, a0);
    const int tgrid = pick_sort_forward(m, A, n, PointRec{});
    int64_t passed = 0, all = 0;
    cutoff_and_compact(m, n, nblk, tgrid, std::numeric_limits<int64_t>::max(), passed, all);
    if (m->jit_early && passed > 0) {
        // the points whose bound cannot pass the accept test lose their
        // flag; the rest are compacted again and weighted in full (the same
        // decisions as weighting every point)
        launch_early_reject(m, A, passed, wmax, nullptr, 0.0);
        cutoff_and_compact(m, n, nblk, tgrid, std::numeric_limits<int64_t>::max(), passed, all);
    }
    if (passed > 0) {
        m->d_gval.resize(passed);
        const int mgrid = (int)((passed + kMainTile - 1) / kMainTile) + C;
        std::pair<hipEvent_t, hipEvent_t> evt{};
        if (m->timing) {
            evt = take_events(m);
            HIP_CHECK(hipEventRecord(evt.first, m->stream));
        }
        CHILI_DISPATCH_NT(m->NT, launch_main<NT>(m, A, mgrid, 0, m->d_gval.ptr, 0, passed, PointRec{}, passed));
        if (m->timing) {
            HIP_CHECK(hipEventRecord(evt.second, m->stream));
            m->ev_live.push_back(evt);
            m->ev_points.push_back(passed);
        }
    }
    gen_tail(m, A, n, nblk, need, wmax, ev, ew, cap, got, passed, kept, used);
}

// The unweighting decisions of the passing points of the current compaction
// (values in d_gval, compaction order), the cutoff at the need-th accepted
// event, StatsData of the accepted events into d_tot and the events to
// ev / ew from column got.
void gen_tail(chili_mc* m, const MCArgs& A, int64_t n, int nblk, int64_t need, double wmax, double* ev, double* ew, int64_t cap,
              int64_t got, int64_t passed, int64_t& kept, int64_t& used) {
    m->d_gflag.resize(n);
    m->d_gv.resize(n);
    m->d_gpos.resize(n);
    hipLaunchKernelGGL(chili_nf::k_fill<uint8_t>, dim3(nblk), dim3(kTile), 0, m->stream, n, (uint8_t)0, m->d_gflag.ptr);
    if (passed > 0)
        hipLaunchKernelGGL(k_unw, dim3((unsigned)((passed + kTile - 1) / kTile)), dim3(kTile), 0, m->stream, A, passed, m->d_comp.ptr,
                           m->d_sorted.ptr, m->d_gval.ptr, wmax, m->d_gflag.ptr, m->d_gv.ptr, m->d_gpos.ptr);
    find_cutoff(m, m->d_gflag.ptr, n, nblk, need, kept, used);
    if (got + kept > cap) throw chili_error(CHILI_ERANGE, "event buffer too small");
    const int ublk = (int)((used + kTile - 1) / kTile);
    if (ublk == 0) return;
    const int oblk = (ublk + kGenSub - 1) / kGenSub;
    m->d_gpart.resize((size_t)oblk * kStatW);
    m->d_gblkoff.resize(ublk);
    scan_int(m, m->d_blkacc.ptr, m->d_gblkoff.ptr, (int64_t)ublk);
    CHILI_DISPATCH_NT(m->NT, hipLaunchKernelGGL((k_gen_out<NT>), dim3(oblk), dim3(kTile), 0, m->stream, used, ublk, m->d_gflag.ptr,
                                                m->d_gblkoff.ptr, m->d_gpos.ptr, m->d_gv.ptr, wmax, m->d_mom.ptr, m->nmax, ev, ew,
                                                cap, got, m->d_gpart.ptr));
    hipLaunchKernelGGL(k_stats_final, dim3(1), dim3(256), 0, m->stream, (int64_t)oblk, m->d_gpart.ptr, m->d_tot.ptr);
}

// One sub-batch of attempts [a0, a0 + n): returns kept (accepted below the
// cutoff), the cutoff (attempts consumed) and NaN count, accumulating into d_tot.
void run_subbatch(chili_mc* m, const double* u, int64_t ustride, uint64_t ua0, uint64_t seed, uint32_t stream, uint64_t a0,
                  int64_t n, int64_t need, int train, double* vals, int64_t vals_off, int64_t vals_cap, int64_t& kept, int64_t& used,
                  const PointRec& R = PointRec{}) {
    ensure_capacity(m, n);
    const int C = m->C;
    const int nblk = (int)((n + kTile - 1) / kTile);
    if (m->host_fn) m->d_fhost.resize((size_t)m->nmax);
    MCArgs A = make_args(m, u, ustride, seed, stream, a0);
    A.ua0 = ua0;
    const int tgrid = pick_sort_forward(m, A, n, R);
    if (m->host_fn) host_pass(m, n, R);
    cutoff_and_compact(m, n, nblk, tgrid, need, kept, used);
    if (kept > 0) {
        const int mgrid = (int)((kept + kMainTile - 1) / kMainTile) + C;
        std::pair<hipEvent_t, hipEvent_t> ev{};
        if (m->timing) {
            ev = take_events(m);
            HIP_CHECK(hipEventRecord(ev.first, m->stream));
        }
        CHILI_DISPATCH_NT(m->NT, launch_main<NT>(m, A, mgrid, train, vals, vals_off, vals_cap, R, kept));
        if (m->timing) {
            HIP_CHECK(hipEventRecord(ev.second, m->stream));
            m->ev_live.push_back(ev);
            m->ev_points.push_back(kept);
        }
        const int64_t rec = kStatW + C + (int64_t)m->d * m->B;
        const int64_t nout = 1 + (int64_t)C + (int64_t)C * m->d * m->B;
        const int ngroups = (mgrid + kRedGroup - 1) / kRedGroup;
        m->d_gsum.resize((size_t)ngroups * (kStatW + C));
        hipLaunchKernelGGL(k_reduce_groups, dim3(ngroups), dim3(256), 0, m->stream, C, m->d_tiles2.ptr, m->d_part.ptr, rec,
                           m->d_gsum.ptr, train);
        if (train) {
            const int DB = m->d * m->B;
            const int hgroups = mgrid / kRedGroupH + C + 1;
            m->d_hsum.resize((size_t)hgroups * DB);
            hipLaunchKernelGGL(k_hist_groups_table, dim3(1), dim3(1), 0, m->stream, C, m->d_tiles2.ptr, m->d_grp2.ptr);
            hipLaunchKernelGGL(k_hist_groups, dim3(hgroups, (DB + 255) / 256), dim3(256), 0, m->stream, C, DB, m->d_tiles2.ptr,
                               m->d_grp2.ptr, m->d_part.ptr, rec, m->d_hsum.ptr);
        }
        hipLaunchKernelGGL(k_reduce_final, dim3((unsigned)((nout + 255) / 256)), dim3(256), 0, m->stream, C, m->d * m->B,
                           m->d_tiles2.ptr, m->d_part.ptr, rec, m->d_gsum.ptr, m->d_tot.ptr, train, m->d_grp2.ptr, m->d_hsum.ptr);
    }
}


// Host uniforms of a call on the device ([d+1][stride]); null for Philox.
const double* stage_uniforms(chili_mc* m, const chili_uniforms* uni, int64_t& ustride) {
    ustride = 0;
    if (!uni->u) return nullptr;
    if (uni->stride <= 0) throw chili_error(CHILI_EINVAL, "uniform stride must be positive");
    const size_t cnt = (size_t)(m->d + 1) * uni->stride;
    check_uniforms(uni->u, cnt);
    m->d_u.resize(cnt);
    HIP_CHECK(hipMemcpyAsync(m->d_u.ptr, uni->u, cnt * sizeof(double), hipMemcpyHostToDevice, m->stream));
    ustride = uni->stride;
    return m->d_u.ptr;
}

// The retry loop of one iteration (MultiChannel.hh:214-248): sub-batches of
// attempts from a0 until ncalls points pass the cut with a finite value.
// sub(u, ustride, first attempt, n, need, kept so far, kept, used) runs one
// sub-batch and accumulates into d_tot; tot receives the totals and stats
// the iteration's StatsData with a zero per rejected attempt.
template <class F>
void iterate_loop(chili_mc* m, const chili_uniforms* uni, uint64_t a0, int64_t ncalls, int64_t cap_max, std::vector<double>& tot,
                  int64_t& att, chili_stats* stats, F&& sub, double* rate_ptr = nullptr, double rate_floor = 0.02) {
    // running estimate of accepted points per attempt (the loop's own, or
    // one the caller keeps for a different acceptance such as unweighting)
    double& rate = rate_ptr ? *rate_ptr : m->acc_rate;
    int64_t ustride = 0;
    const double* ud = stage_uniforms(m, uni, ustride);
    const size_t ntot = kStatW + m->C + (size_t)m->C * m->d * m->B;
    tot.assign(ntot, 0.0);
    tot[1] = std::numeric_limits<double>::max();
    tot[2] = std::numeric_limits<double>::min();
    m->d_tot.upload(tot, m->stream);
    int64_t need = ncalls, rejected = 0, got = 0;
    att = 0;
    double prev_nan = 0;
    const int64_t cap = std::max<int64_t>(int64_t(1) << 16, std::min<int64_t>(cap_max, 8 * ncalls + 4096));
    // the reference loops until ncalls points pass; stop instead of spinning
    // when the cut (or NaN values) leave essentially nothing
    const int64_t max_attempts = 100000 * ncalls + (int64_t(1) << 24);
    int all_nan = 0;   // consecutive sub-batches whose every kept value was NaN
    while (need > 0) {
        if (att > max_attempts) throw chili_error(CHILI_ERANGE, "no attempts pass the cut: retry limit reached");
        if (all_nan >= 8) throw chili_error(CHILI_ERANGE, "every value is NaN: retry limit reached");
        // attempts expected to give `need` points, with a margin of ~1% plus
        // five binomial standard deviations; a shortfall runs another round
        const double expect = (double)need / std::max(rate, rate_floor);
        int64_t n = (int64_t)(expect * 1.01 + 5.0 * std::sqrt(expect)) + 2048;
        n = std::min(n, cap);
        if (ud) {
            n = std::min<int64_t>(n, ustride - att);
            if (n <= 0) throw chili_error(CHILI_ERANGE, "host uniforms exhausted before ncalls points passed");
        }
        int64_t kept = 0, used = 0;
        sub(ud, ustride, a0 + (uint64_t)att, n, need, got, kept, used);
        double nanc = 0;
        HIP_CHECK(hipMemcpyAsync(&nanc, m->d_tot.ptr + 6, sizeof(double), hipMemcpyDeviceToHost, m->stream));
        HIP_CHECK(hipStreamSynchronize(m->stream));
        const int64_t nan_new = (int64_t)(nanc - prev_nan);
        prev_nan = nanc;
        all_nan = (kept > 0 && nan_new == kept) ? all_nan + 1 : 0;
        rejected += used - kept;
        need -= kept - nan_new;
        got += kept;
        att += used;
        if (used > 0) rate = 0.5 * rate + 0.5 * std::max(0.5 * rate_floor, (double)kept / (double)used);
    }
    HIP_CHECK(hipMemcpyAsync(tot.data(), m->d_tot.ptr, ntot * sizeof(double), hipMemcpyDeviceToHost, m->stream));
    HIP_CHECK(hipStreamSynchronize(m->stream));
    if (m->timing) harvest_events(m);
    // rejected attempts add a zero each (MultiChannel.hh:219-223)
    double n = tot[0] + (double)rejected;
    double mn = tot[1], mx = tot[2];
    if (rejected > 0) { mn = std::min(mn, 0.0); mx = std::max(mx, 0.0); }
    stats->n = n; stats->min = mn; stats->max = mx; stats->sum = tot[3]; stats->sum2 = tot[4]; stats->n_finite = tot[5];
}

void ensure_flow_capacity(chili_mc* m, int64_t n) {
    if (n <= m->fmax) return;
    m->d_z.resize((size_t)n * m->d);
    m->d_rall.resize((size_t)m->C * n * m->d);
    m->d_abuf.resize((size_t)m->C * n);
    m->d_lq.resize((size_t)m->C * n);
    m->d_fbuf.resize(n);
    m->d_lqs.resize(n);
    m->d_lqs2.resize(n);
    m->d_fpart.resize((size_t)((n + kTile - 1) / kTile) * kStatW);
    m->fmax = n;
}

template <int NT>
void launch_zgen(chili_mc* m, const MCArgs& A, int64_t n) {
    constexpr int D = 3 * (NT - 2) - 2;
    hipLaunchKernelGGL(k_zgen<D>, dim3((unsigned)((n + kTile - 1) / kTile)), dim3(kTile), 0, m->stream, A, n, m->d_sorted.ptr,
                       m->d_z.ptr);
}
template <int NT>
void launch_geom(chili_mc* m, const MCArgs& A, int64_t kept) {
    const unsigned g = (unsigned)((kept + kTile - 1) / kTile);
    if (m->jit_split) {
        // run-time specialised: one kernel per channel (the split weight
        // pass's structure), this channel's search table in LDS when small
        constexpr int D = 3 * (NT - 2) - 2;
        const size_t one_bytes = sizeof(double) * (size_t)D * m->P2;
        int lds_search = (m->lds_search && one_bytes <= kLdsSearchMax) ? 1 : 0;
        const size_t lds = sizeof(double) * (size_t)kTile * (D + 1) + (lds_search ? one_bytes : 0);
        MCArgs a = A;
        int64_t kp = kept, nmax = m->nmax, stride = m->fmax;
        const uint32_t* comp = m->d_comp.ptr;
        const double* mom = m->d_mom.ptr;
        double* rall = m->d_rall.ptr;
        double* abuf = m->d_abuf.ptr;
        double* fbuf = m->d_fbuf.ptr;
        void* args[] = {&a, &kp, &comp, &mom, &nmax, &rall, &abuf, &fbuf, &stride, &lds_search};
        for (int j = 0; j < m->C; ++j)
            HIP_CHECK(hipModuleLaunchKernel(m->jit->geom[(size_t)j], g, 1, 1, kTile, 1, 1, (unsigned)lds, m->stream, args, nullptr));
        return;
    }
    hipLaunchKernelGGL(k_geom<NT>, dim3(g), dim3(kTile), 0, m->stream, A, kept, m->d_comp.ptr, m->d_mom.ptr, m->nmax, m->d_rall.ptr,
                       m->d_abuf.ptr, m->d_fbuf.ptr, m->fmax);
}

// One sub-batch of the learned-proposal path (fl == null: every q_j = 1,
// the plain multichannel weight through the same kernels).
// The unweighting stage's outputs on the learned-proposal path (see
// run_subbatch_gen): accept test against wmax, events to ev / ew from column
// got.
struct GenOut {
    double wmax;
    double* ev;
    double* ew;
    int64_t cap, got;
};

void run_subbatch_flow(chili_mc* m, chili_flow* fl, const double* u, int64_t ustride, uint64_t ua0, uint64_t seed, uint32_t stream,
                       uint64_t a0, int64_t n, int64_t need, double* vals, int64_t vals_off, int64_t vals_cap, FlowData FD,
                       int64_t& kept, int64_t& used, const PointRec& R = PointRec{}, const GenOut* gen = nullptr) {
    ensure_capacity(m, n);
    ensure_flow_capacity(m, n);
    const int C = m->C, d = m->d;
    const int nblk = (int)((n + kTile - 1) / kTile);
    MCArgs A = make_args(m, u, ustride, seed, stream, a0);
    A.ua0 = ua0;
    // with a mixture proposal the sort key is (channel, component): the
    // uniform component's attempts close each channel's segment and keep
    // their uniforms as the hypercube point
    const bool mixing = fl && fl->eng.mix > 0;
    if (mixing && u) throw chili_error(CHILI_EINVAL, "a mixture proposal needs device-generated (Philox) uniforms");
    const int nb = mixing ? 2 * C : C;
    m->d_counts.resize((size_t)nblk * nb);
    m->d_offs.resize((size_t)nblk * nb);
    if (mixing)
        hipLaunchKernelGGL(k_pick_mix, dim3(nblk), dim3(kTile), sizeof(int) * nb, m->stream, A, n, fl->eng.mix, m->d_pick.ptr,
                           m->d_counts.ptr, nblk);
    else
        hipLaunchKernelGGL(k_pick, dim3(nblk), dim3(kTile), sizeof(int) * C, m->stream, A, n, m->d_pick.ptr, m->d_counts.ptr, nblk);
    scan_int(m, m->d_counts.ptr, m->d_offs.ptr, (int64_t)nblk * nb);
    hipLaunchKernelGGL(k_scatter, dim3(nblk), dim3(kTile), sizeof(int) * kWaves * nb, m->stream, nb, n, m->d_pick.ptr, m->d_offs.ptr,
                       nblk, m->d_sorted.ptr);
    if (mixing) {
        m->d_fend.resize(C);
        hipLaunchKernelGGL(k_segments_mix, dim3(1), dim3(1), 0, m->stream, C, m->d_offs.ptr, nblk, n, m->d_seg.ptr, m->d_tiles.ptr,
                           m->d_fend.ptr);
    } else {
        hipLaunchKernelGGL(k_segments, dim3(1), dim3(1), 0, m->stream, C, m->d_offs.ptr, nblk, n, m->d_seg.ptr, m->d_tiles.ptr);
    }
    const int tgrid = nblk + C;
    if (fl) {
        // z -> r through each channel's flow, on the channel-sorted segments
        // (with mixing: only their flow parts)
        CHILI_DISPATCH_NT(m->NT, launch_zgen<NT>(m, A, n));
        chili_nf::StreamSwap sw(fl->eng, m->stream);
        hipLaunchKernelGGL(chili_nf::k_fill<double>, dim3((unsigned)((n + kTile - 1) / kTile)), dim3(kTile), 0, m->stream, n, 0.0,
                           m->d_lqs.ptr);
        fl->eng.fused_pass(true, C, 0, m->d_seg.ptr, 0, 0, n, m->d_z.ptr, m->d_lqs.ptr, mixing ? m->d_fend.ptr : nullptr,
                           mixing ? (int64_t)((1.0 - fl->eng.mix) * (double)n) + 1 : n);
    }
    std::pair<hipEvent_t, hipEvent_t> fev{};
    if (m->timing) {
        fev = take_events(m);
        HIP_CHECK(hipEventRecord(fev.first, m->stream));
    }
    CHILI_DISPATCH_NT(m->NT, launch_forward<NT>(m, A, tgrid, R, fl ? m->d_z.ptr : nullptr, !mixing));
    if (m->timing) {
        HIP_CHECK(hipEventRecord(fev.second, m->stream));
        m->fev_live.push_back(fev);
        m->fev_points.push_back(n);
    }
    launch_ktcut(m, n, R);
    // own ranges need the [C][fmax] row layout within the int offsets of a pass
    const bool own = fl && (int64_t)C * m->fmax < (int64_t)INT32_MAX;
    if (gen) {
        // unweighting stage: every passing point is a candidate, or with
        // early rejection those whose picked channel's bound (its sampling
        // density, which the combination uses as is when the own ranges are
        // in use and no mixture is) can pass the accept test
        int64_t all = 0;
        cutoff_and_compact(m, n, nblk, tgrid, std::numeric_limits<int64_t>::max(), kept, all);
        if (m->jit_early && own && !mixing && kept > 0) {
            launch_early_reject(m, A, kept, gen->wmax, m->d_lqs.ptr, 0.0);
            cutoff_and_compact(m, n, nblk, tgrid, std::numeric_limits<int64_t>::max(), kept, all);
        }
        m->d_gval.resize(std::max<int64_t>(kept, 1));
        vals = m->d_gval.ptr;
        vals_off = 0;
        vals_cap = kept;
        if (kept == 0) {
            const int64_t passed = 0;
            gen_tail(m, A, n, nblk, need, gen->wmax, gen->ev, gen->ew, gen->cap, gen->got, passed, kept, used);
            return;
        }
    } else {
        cutoff_and_compact(m, n, nblk, tgrid, need, kept, used);
        if (kept == 0) return;
    }
    CHILI_DISPATCH_NT(m->NT, launch_geom<NT>(m, A, kept));
    if (fl) {
        // log q_j(r_j) of every kept point and channel; the flows move a copy
        // of r_j (the training records keep the originals).  The picked
        // channel's flow-component points are skipped (their log q is the
        // sampling pass's).
        chili_nf::StreamSwap sw(fl->eng, m->stream);
        const int64_t rows = (int64_t)C * m->fmax;
        m->d_xw.resize((size_t)rows * d);
        hipLaunchKernelGGL(chili_nf::k_fill<double>, dim3((unsigned)((rows + kTile - 1) / kTile)), dim3(kTile), 0, m->stream, rows, 0.0,
                           m->d_lq.ptr);
        if (own) {
            m->d_own.resize((size_t)2 * C);
            m->d_dseg.resize(C);
            m->d_dend.resize(C);
            hipLaunchKernelGGL(k_own_ranges, dim3((C + 63) / 64), dim3(64), 0, m->stream, C, kept, m->d_comp.ptr, m->d_seg.ptr,
                               mixing ? m->d_fend.ptr : nullptr, m->fmax, m->d_own.ptr, m->d_dseg.ptr, m->d_dend.ptr);
            hipLaunchKernelGGL(k_dens_rows, dim3((unsigned)((kept * d + kTile - 1) / kTile), (unsigned)C), dim3(kTile), 0, m->stream, kept, d,
                               m->fmax, m->d_own.ptr, m->d_rall.ptr, m->d_xw.ptr);
            const int64_t expect = (int64_t)((double)C * kept - (1.0 - fl->eng.mix) * (double)kept) + 1;
            fl->eng.fused_pass(false, C, 0, m->d_dseg.ptr, 0, 0, (int64_t)C * kept, m->d_xw.ptr, m->d_lq.ptr, m->d_dend.ptr, expect);
        } else {
            HIP_CHECK(hipMemcpyAsync(m->d_xw.ptr, m->d_rall.ptr, (size_t)rows * d * sizeof(double), hipMemcpyDeviceToDevice, m->stream));
            fl->eng.fused_pass(false, C, 0, nullptr, kept, m->fmax, (int64_t)C * kept, m->d_xw.ptr, m->d_lq.ptr);
        }
    }
    const int cg = (int)((kept + kTile - 1) / kTile);
    hipLaunchKernelGGL(k_combine, dim3(cg), dim3(kTile), 0, m->stream, C, d, kept, m->fmax, m->d_abuf.ptr, fl ? m->d_lq.ptr : nullptr,
                       fl ? fl->eng.mix : 0.0,
                       m->d_fbuf.ptr, m->d_rall.ptr, m->d_fpart.ptr, vals, vals_off, vals_cap, FD, m->d_comp.ptr, m->d_sorted.ptr, R,
                       own ? m->d_own.ptr : nullptr, m->d_lqs.ptr);
    if (gen) {
        // the candidates' values (compaction order) decide; StatsData is the
        // accepted events'
        const int64_t passed = kept;
        gen_tail(m, A, n, nblk, need, gen->wmax, gen->ev, gen->ew, gen->cap, gen->got, passed, kept, used);
        return;
    }
    hipLaunchKernelGGL(k_stats_final, dim3(1), dim3(256), 0, m->stream, (int64_t)cg, m->d_fpart.ptr, m->d_tot.ptr);
}

void check_flow(const chili_mc* m, const chili_flow* fl) {
    if (!fl) return;
    if (fl->eng.device != m->device) throw chili_error(CHILI_EINVAL, "flow and context live on different devices");
    if (fl->eng.d != m->d || fl->eng.nflows != m->C)
        throw chili_error(CHILI_EINVAL, "flow dimensions / count do not match the channels");
}

}  // namespace

namespace {

// ===========================================================================
// Batched jobs: many independent multichannel integrations (each its own
// chili_mc: channels, grids, alpha, integrand) of one multiplicity class in
// one pass.  Attempts of job j occupy the blocks [blk0, blk0 + nblk) of one
// concatenated attempt space; segments are (job, channel) pairs flattened to
// g = g0(job) + channel; the forward and weight kernels are persistent
// (fixed grids looping over device-counted tiles) and the cutoff is found on
// the device, so a whole pass is host-synchronisation free and is captured
// once in a hipGraph and replayed.
// ===========================================================================
struct BJob {
    MCArgs A;           // the job's context (A.a0: first attempt counter of the pass)
    int64_t att0;       // first attempt slot (multiple of kTile)
    int64_t n;          // attempts of this pass
    int64_t need;       // accepted points wanted
    int64_t cnt0;       // offset of the job's count matrix [C][nblk]
    int blk0, nblk;     // attempt blocks
    int g0;             // first flattened segment
    int kt;             // the job's cut is the kT-jet PreProcess
    double* tot;        // running totals [7 + C + C d B]
};

__device__ __forceinline__ int block_job(const int* __restrict__ blkjob, int b) { return blkjob[b]; }

__global__ __launch_bounds__(kTile) void k_bpick(const BJob* __restrict__ J, const int* __restrict__ blkjob, uint16_t* __restrict__ pick,
                                                 int* __restrict__ counts) {
    extern __shared__ int s_cnt[];
    const BJob& jb = J[block_job(blkjob, blockIdx.x)];
    const int C = jb.A.C;
    for (int c = threadIdx.x; c < C; c += blockDim.x) s_cnt[c] = 0;
    __syncthreads();
    const int lb = blockIdx.x - jb.blk0;
    const int64_t local = (int64_t)lb * kTile + threadIdx.x;
    if (local < jb.n) {
        const int c = pick_channel(jb.A, channel_uniform(jb.A, jb.A.a0 + (uint64_t)local));
        pick[jb.att0 + local] = (uint16_t)c;
        atomicAdd(&s_cnt[c], 1);
    }
    __syncthreads();
    for (int c = threadIdx.x; c < C; c += blockDim.x) counts[jb.cnt0 + (int64_t)c * jb.nblk + lb] = s_cnt[c];
}

__global__ __launch_bounds__(kTile) void k_bscatter(const BJob* __restrict__ J, const int* __restrict__ blkjob,
                                                    const uint16_t* __restrict__ pick, const int* __restrict__ offs,
                                                    uint32_t* __restrict__ sorted) {
    extern __shared__ int s_wc[];   // [kWaves][C]
    const BJob& jb = J[block_job(blkjob, blockIdx.x)];
    const int C = jb.A.C;
    for (int i = threadIdx.x; i < kWaves * C; i += blockDim.x) s_wc[i] = 0;
    __syncthreads();
    const int w = threadIdx.x >> 6;
    const int lb = blockIdx.x - jb.blk0;
    const int64_t local = (int64_t)lb * kTile + threadIdx.x;
    const bool valid = local < jb.n;
    const int c = valid ? pick[jb.att0 + local] : 0;
    const int rank = wave_rank(c, valid, s_wc + w * C);
    __syncthreads();
    if (valid) {
        int before = 0;
        for (int ww = 0; ww < w; ++ww) before += s_wc[ww * C + c];
        sorted[offs[jb.cnt0 + (int64_t)c * jb.nblk + lb] + before + rank] = (uint32_t)(jb.att0 + local);
    }
}

// Segment table of the flattened (job, channel) list; total = attempts of the pass.
__global__ void k_bsegments(int G, const BJob* __restrict__ J, const int* __restrict__ gjob, const int* __restrict__ offs,
                            const int64_t* __restrict__ total, int* seg, int* tile) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    int t = 0;
    for (int g = 0; g < G; ++g) {
        const BJob& jb = J[gjob[g]];
        const int c = g - jb.g0;
        const int s0 = offs[jb.cnt0 + (int64_t)c * jb.nblk];
        int e;
        if (g + 1 < G) {
            const BJob& jn = J[gjob[g + 1]];
            e = offs[jn.cnt0 + (int64_t)(g + 1 - jn.g0) * jn.nblk];
        } else {
            e = (int)*total;
        }
        seg[g] = s0;
        tile[g] = t;
        t += (e - s0 + kTile - 1) / kTile;
    }
    seg[G] = (int)*total;
    tile[G] = t;
}

template <int NT>
__global__ __launch_bounds__(kTile) void k_bforward(int G, const BJob* __restrict__ J, const int* __restrict__ gjob,
                                                    const int* __restrict__ seg, const int* __restrict__ tile,
                                                    const uint32_t* __restrict__ sorted, double* __restrict__ mom, int64_t nmax,
                                                    uint8_t* __restrict__ acc) {
    const int ntile = tile[G];
    for (int t = blockIdx.x; t < ntile; t += gridDim.x) {
        int g, e0, e1;
        if (!locate_tile(G, seg, tile, t, g, e0, e1)) break;
        const BJob& jb = J[gjob[g]];
        const int e = e0 + threadIdx.x;
        if (e < e1) {
            const uint32_t a = sorted[e];
            forward_one<NT, false>(jb.A, jb.A.progs[g - jb.g0], g - jb.g0, e, a, jb.A.a0 + (uint64_t)(a - jb.att0), mom, nmax, acc,
                                   PointRec{}, nullptr);
        }
    }
}

// The kT-jet cut of the jobs that use it (persistent, over the forward
// kernel's tiles): an attempt that passed the forward kernel and fails the
// clustering loses its flag, as in the single-context path.
template <int NT>
__global__ __launch_bounds__(kTile) void k_bktcut(int G, const BJob* __restrict__ J, const int* __restrict__ gjob,
                                                  const int* __restrict__ seg, const int* __restrict__ tile,
                                                  const uint32_t* __restrict__ sorted, const double* __restrict__ mom, int64_t nmax,
                                                  uint8_t* __restrict__ acc) {
    const int ntile = tile[G];
    for (int t = blockIdx.x; t < ntile; t += gridDim.x) {
        int g, e0, e1;
        if (!locate_tile(G, seg, tile, t, g, e0, e1)) break;
        const BJob& jb = J[gjob[g]];
        const int e = e0 + threadIdx.x;
        if (e < e1 && jb.kt) {
            const uint32_t a = sorted[e];
            if (acc[a] && !kt_cut_pass<NT>(jb.A.f, mom, nmax, e)) acc[a] = 0;
        }
    }
}

// Per job: the attempt at which `need` accepted attempts are reached (local
// index + 1, or n) and the kept count (MultiChannel.hh:214-248 on the device).
__global__ void k_bcutoff(const BJob* __restrict__ J, const int* __restrict__ blkacc, const uint8_t* __restrict__ acc,
                          int64_t* __restrict__ cut, int64_t* __restrict__ kept) {
    if (threadIdx.x != 0) return;
    const BJob& jb = J[blockIdx.x];
    int64_t run = 0, c = jb.n;
    for (int lb = 0; lb < jb.nblk; ++lb) {
        const int bc = blkacc[jb.blk0 + lb];
        if (run + bc >= jb.need) {
            for (int i = 0; i < kTile; ++i) {
                const int64_t local = (int64_t)lb * kTile + i;
                if (local >= jb.n) break;
                run += acc[jb.att0 + local];
                if (run == jb.need) { c = local + 1; break; }
            }
            break;
        }
        run += bc;
    }
    cut[blockIdx.x] = c;
    kept[blockIdx.x] = run < jb.need ? run : jb.need;
}

__device__ __forceinline__ bool bkeep(const BJob* __restrict__ J, const int* __restrict__ gjob, const int64_t* __restrict__ cut,
                                      const uint8_t* __restrict__ acc, int g, uint32_t a) {
    const int j = gjob[g];
    return (int64_t)a - J[j].att0 < cut[j] && acc[a];
}

__global__ __launch_bounds__(kTile) void k_bkeep_count(int G, const BJob* __restrict__ J, const int* __restrict__ gjob,
                                                       const int* __restrict__ seg, const int* __restrict__ tile,
                                                       const uint32_t* __restrict__ sorted, const uint8_t* __restrict__ acc,
                                                       const int64_t* __restrict__ cut, int* __restrict__ kept) {
    __shared__ int s[kWaves];
    int g, e0, e1;
    if (!locate_tile(G, seg, tile, blockIdx.x, g, e0, e1)) {
        if (threadIdx.x == 0) kept[blockIdx.x] = 0;
        return;
    }
    const int e = e0 + threadIdx.x;
    const bool k = e < e1 && bkeep(J, gjob, cut, acc, g, sorted[e]);
    const uint64_t m = __ballot(k);
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = __popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int w = 0; w < kWaves; ++w) t += s[w];
        kept[blockIdx.x] = t;
    }
}

__global__ __launch_bounds__(kTile) void k_bkeep_scatter(int G, const BJob* __restrict__ J, const int* __restrict__ gjob,
                                                         const int* __restrict__ seg, const int* __restrict__ tile,
                                                         const uint32_t* __restrict__ sorted, const uint8_t* __restrict__ acc,
                                                         const int64_t* __restrict__ cut, const int* __restrict__ kept_off,
                                                         uint32_t* __restrict__ comp) {
    __shared__ int s[kWaves];
    int g, e0, e1;
    if (!locate_tile(G, seg, tile, blockIdx.x, g, e0, e1)) return;
    const int e = e0 + threadIdx.x;
    const bool k = e < e1 && bkeep(J, gjob, cut, acc, g, sorted[e]);
    const uint64_t m = __ballot(k);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) s[w] = __popcll(m);
    __syncthreads();
    if (k) {
        int before = 0;
        for (int ww = 0; ww < w; ++ww) before += s[ww];
        const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
        comp[kept_off[blockIdx.x] + before + __popcll(m & lt)] = (uint32_t)e;
    }
}

// Kept segments (tiles of tsize) from the per-forward-tile kept offsets; the
// kept total is the scan's device total.
__global__ void k_bkeep_segments(int G, const int* __restrict__ tile, const int* __restrict__ kept_off,
                                 const int64_t* __restrict__ kept_total, int* seg2, int* tile2, int tsize) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const int kt = (int)*kept_total;
    int t = 0;
    for (int g = 0; g < G; ++g) {
        const int s0 = (tile[g] >= tile[G]) ? kt : kept_off[tile[g]];
        const int e = (g + 1 < G && tile[g + 1] < tile[G]) ? kept_off[tile[g + 1]] : kt;
        seg2[g] = s0;
        tile2[g] = t;
        t += (e - s0 + tsize - 1) / tsize;
    }
    seg2[G] = kt;
    tile2[G] = t;
}

template <int NT, int RND>
__global__ __launch_bounds__(kTile) void k_bmain(int G, const BJob* __restrict__ J, const int* __restrict__ gjob,
                                                 const int* __restrict__ seg2, const int* __restrict__ tile2,
                                                 const uint32_t* __restrict__ comp, const double* __restrict__ mom, int64_t nmax,
                                                 double* __restrict__ psbuf, int64_t psstride, int train, double* __restrict__ part,
                                                 int64_t rec) {
    extern __shared__ double s_mem[];
    const int ntile = tile2[G];
    for (int t = blockIdx.x; t < ntile; t += gridDim.x) {
        int g, e0, e1;
        if (!locate_tile(G, seg2, tile2, t, g, e0, e1, kTile * RND)) break;
        const BJob& jb = J[gjob[g]];
        main_tile<NT, RND, GenericLoop>(jb.A, g - jb.g0, e0, e1, s_mem, blockIdx.x, psbuf, psstride, train, comp, mom, nmax,
                           part + (int64_t)t * rec, nullptr, 0, 0, nullptr, PointRec{});
    }
}

// StatsData start values of every job's totals (max starts at the smallest
// positive normal, Statistics.hh:141)
__global__ void k_btot_init(const BJob* __restrict__ J, int njobs, int DB) {
    const int j = blockIdx.x;
    if (j >= njobs) return;
    const BJob& jb = J[j];
    const int n = kStatW + jb.A.C + jb.A.C * DB;
    for (int i = threadIdx.x; i < n; i += blockDim.x)
        jb.tot[i] = (i == 1) ? std::numeric_limits<double>::max() : (i == 2) ? std::numeric_limits<double>::min() : 0.0;
}

// Per job (blockIdx.x) the fixed-order sums of its tiles' partial records
// into its totals: stats, T_c, and each channel's histogram over that
// channel's tiles.
__global__ __launch_bounds__(256) void k_breduce(const BJob* __restrict__ J, const int* __restrict__ tile2, const double* __restrict__ part,
                                                 int64_t rec, int DB, int train) {
    const BJob& jb = J[blockIdx.x];
    const int C = jb.A.C;
    const int t0 = tile2[jb.g0], t1 = tile2[jb.g0 + C];
    const int64_t nout = train ? 1 + (int64_t)C + (int64_t)C * DB : 1;
    for (int64_t i = (int64_t)blockIdx.y * blockDim.x + threadIdx.x; i < nout; i += (int64_t)gridDim.y * blockDim.x) {
        if (i == 0) {
            double a[7] = {jb.tot[0], jb.tot[1], jb.tot[2], jb.tot[3], jb.tot[4], jb.tot[5], jb.tot[6]};
            for (int t = t0; t < t1; ++t) stat_merge(a, part + (int64_t)t * rec);
            for (int k = 0; k < kStatW; ++k) jb.tot[k] = a[k];
        } else if (i <= C) {
            double s = jb.tot[kStatW + i - 1];
            for (int t = t0; t < t1; ++t) s += part[(int64_t)t * rec + kStatW + i - 1];
            jb.tot[kStatW + i - 1] = s;
        } else {
            const int64_t h = i - 1 - C;
            const int c = (int)(h / DB), k = (int)(h % DB);
            double s = jb.tot[kStatW + C + h];
            for (int t = tile2[jb.g0 + c]; t < tile2[jb.g0 + c + 1]; ++t) s += part[(int64_t)t * rec + kStatW + C + k];
            jb.tot[kStatW + C + h] = s;
        }
    }
}

}  // namespace

namespace {

// ===========================================================================
// Batched jobs: host side
// ===========================================================================
struct BatchGroup {
    int NT = 0, D = 0, G = 0, Cmax = 0;
    bool any_kt = false;                   // some job of the group uses the kT-jet cut
    std::vector<int> jobs;                 // indices into chili_batch::jobs
    std::vector<int64_t> n;                // planned attempts per job (this group's order)
    int64_t natt = 0, ncnt = 0;
    int nblk = 0, tgrid = 0;
    std::vector<int64_t> cnt0, att0;
    std::vector<int> blk0, g0;
    std::vector<int64_t> tot_off;
    DevBuf<BJob> d_jobs;
    DevBuf<int> d_blkjob, d_gjob, d_counts, d_offs, d_seg, d_tile, d_blkacc, d_kept, d_keptoff, d_seg2, d_tile2;
    DevBuf<uint16_t> d_pick;
    DevBuf<uint32_t> d_sorted, d_comp;
    DevBuf<uint8_t> d_acc;
    DevBuf<double> d_mom, d_ps, d_part, d_tot;
    DevBuf<int64_t> d_total, d_total2, d_bsum, d_cut, d_keptj;
    std::vector<BJob> h_jobs;
    std::pair<hipEvent_t, hipEvent_t> ev_fwd{nullptr, nullptr}, ev_main{nullptr, nullptr};
};

}  // namespace

struct chili_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    int ncu = 256;
    std::vector<chili_mc*> jobs;
    std::vector<int> job_group;
    std::vector<BatchGroup> groups;
    std::vector<double> acc_est;
    int B = 0;
    // launch shape (autotuned): workgroups per CU of the persistent kernels,
    // kTile sub-tiles per weight-kernel tile
    int fwd_wpc = 4, main_wpc = 2, main_rnd = 1;
    // hipGraph of one pass of every group
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool graph_valid = false;
    int graph_train = -1;
    int64_t captures = 0, replays = 0, topups = 0;
    ~chili_batch() {
        if (exec) hipGraphExecDestroy(exec);
        if (graph) hipGraphDestroy(graph);
        for (auto& g : groups) {
            if (g.ev_fwd.first) { hipEventDestroy(g.ev_fwd.first); hipEventDestroy(g.ev_fwd.second); }
            if (g.ev_main.first) { hipEventDestroy(g.ev_main.first); hipEventDestroy(g.ev_main.second); }
        }
    }
};

namespace {

void bscan(hipStream_t st, const int* in, int* out, int64_t len, DevBuf<int64_t>& bsum, int64_t* total) {
    const int64_t nb = (len + kScanBlock - 1) / kScanBlock;
    bsum.resize(nb);
    hipLaunchKernelGGL(k_scan_sums, dim3((unsigned)nb), dim3(kScanBlock), 0, st, in, len, bsum.ptr);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kScanBlock), 0, st, bsum.ptr, nb, total);
    hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nb), dim3(kScanBlock), 0, st, in, out, len, bsum.ptr);
}

int batch_fgrid(const chili_batch* b, const BatchGroup& g) {
    return std::max(1, std::min(g.tgrid, b->fwd_wpc * b->ncu));
}
int batch_mgrid(const chili_batch* b, const BatchGroup& g) {
    const int mtiles = (int)(g.natt / (kTile * b->main_rnd)) + g.G + 1;
    return std::max(1, std::min(mtiles, b->main_wpc * b->ncu));
}

// Attempt plan of a group: n_j from the acceptance estimate; buffers sized to
// it.  Returns true when the plan (and so the captured graph) changed.
bool batch_plan(chili_batch* b, BatchGroup& g, const int64_t* ncalls) {
    std::vector<int64_t> n(g.jobs.size());
    bool changed = g.n.size() != g.jobs.size();
    for (size_t q = 0; q < g.jobs.size(); ++q) {
        const int j = g.jobs[q];
        const double acc = std::max(b->acc_est[j], 0.01);
        int64_t want = (int64_t)((double)ncalls[j] / acc * 1.2) + 1024;
        want = ((want + kTile - 1) / kTile) * kTile;
        if (want > (int64_t(1) << 26)) want = int64_t(1) << 26;
        // keep the previous plan (and graph) while it covers the need without
        // more than 3x waste
        if (!changed && g.n[q] >= (int64_t)((double)ncalls[j] / acc * 1.05) && g.n[q] <= 3 * want) n[q] = g.n[q];
        else { n[q] = want; changed = true; }
    }
    if (!changed) return false;
    g.n = n;
    const int J = (int)g.jobs.size();
    g.cnt0.assign(J, 0); g.att0.assign(J, 0); g.blk0.assign(J, 0); g.g0.assign(J, 0); g.tot_off.assign(J, 0);
    int64_t att = 0, cnt = 0, toff = 0;
    int blk = 0, gg = 0;
    std::vector<int> blkjob, gjob;
    for (int q = 0; q < J; ++q) {
        chili_mc* m = b->jobs[g.jobs[q]];
        const int nb = (int)(g.n[q] / kTile);
        g.att0[q] = att; g.blk0[q] = blk; g.cnt0[q] = cnt; g.g0[q] = gg; g.tot_off[q] = toff;
        for (int k = 0; k < nb; ++k) blkjob.push_back(q);
        for (int c = 0; c < m->C; ++c) gjob.push_back(q);
        att += g.n[q];
        blk += nb;
        cnt += (int64_t)m->C * nb;
        gg += m->C;
        toff += kStatW + m->C + (int64_t)m->C * m->d * m->B;
    }
    g.natt = att; g.nblk = blk; g.ncnt = cnt; g.G = gg;
    g.tgrid = blk + gg;
    if (att >= (int64_t(1) << 31)) throw chili_error(CHILI_ERANGE, "batched attempts exceed 2^31 per multiplicity class");
    g.d_blkjob.upload(blkjob, b->stream);
    g.d_gjob.upload(gjob, b->stream);
    g.d_pick.resize(att);
    g.d_sorted.resize(att);
    g.d_comp.resize(att);
    g.d_acc.resize(att);
    g.d_mom.resize((size_t)att * g.NT * 4);
    g.d_counts.resize(cnt);
    g.d_offs.resize(cnt);
    g.d_seg.resize(gg + 1);
    g.d_tile.resize(gg + 1);
    g.d_seg2.resize(gg + 1);
    g.d_tile2.resize(gg + 1);
    g.d_blkacc.resize(blk);
    g.d_kept.resize(g.tgrid);
    g.d_keptoff.resize(g.tgrid);
    g.d_tot.resize(toff);
    g.d_total.resize(1);
    g.d_total2.resize(1);
    g.d_cut.resize(J);
    g.d_keptj.resize(J);
    // both scans of a pass share the block-sum buffer: size it once (no
    // allocation may happen while a pass is being captured)
    g.d_bsum.resize((std::max<int64_t>(cnt, g.tgrid) + kScanBlock - 1) / kScanBlock + 1);
    return true;
}

// Buffers that depend on the launch shape (persistent grids, rounds).
void batch_shape_buffers(chili_batch* b, BatchGroup& g) {
    const int mgrid = batch_mgrid(b, g);
    const int64_t rec = kStatW + g.Cmax + (int64_t)g.D * b->B;
    const int64_t mtiles = g.natt / kTile + g.G + 1;   // at rnd = 1 (the most tiles)
    g.d_part.resize((size_t)mtiles * rec);
    g.d_ps.resize((size_t)g.Cmax * mgrid * kTile);
}

template <int NT, int RND>
void launch_bmain(chili_batch* b, BatchGroup& g, int train) {
    constexpr int D = 3 * (NT - 2) - 2;
    const int mgrid = batch_mgrid(b, g);
    const int64_t rec = kStatW + g.Cmax + (int64_t)D * b->B;
    const size_t lds = sizeof(double) * (D * b->B + kWaves * (8 + g.Cmax)) + sizeof(short) * D * kTile;
    if (lds > 160 * 1024) throw chili_error(CHILI_ERANGE, "channel count x bins exceed the LDS of one workgroup");
    hipLaunchKernelGGL((k_bmain<NT, RND>), dim3(mgrid), dim3(kTile), lds, b->stream, g.G, g.d_jobs.ptr, g.d_gjob.ptr, g.d_seg2.ptr,
                       g.d_tile2.ptr, g.d_comp.ptr, g.d_mom.ptr, g.natt, g.d_ps.ptr, (int64_t)mgrid * kTile, train, g.d_part.ptr, rec);
}

template <int NT>
void enqueue_group(chili_batch* b, BatchGroup& g, int train, bool timed) {
    hipStream_t st = b->stream;
    const int J = (int)g.jobs.size();
    const int DB = g.D * b->B;
    hipLaunchKernelGGL(k_btot_init, dim3(J), dim3(256), 0, st, g.d_jobs.ptr, J, DB);
    hipLaunchKernelGGL(chili_nf::k_fill<uint8_t>, dim3((unsigned)((g.natt + 255) / 256)), dim3(256), 0, st, g.natt, (uint8_t)0, g.d_acc.ptr);
    hipLaunchKernelGGL(k_bpick, dim3(g.nblk), dim3(kTile), sizeof(int) * g.Cmax, st, g.d_jobs.ptr, g.d_blkjob.ptr, g.d_pick.ptr,
                       g.d_counts.ptr);
    bscan(st, g.d_counts.ptr, g.d_offs.ptr, g.ncnt, g.d_bsum, g.d_total.ptr);
    hipLaunchKernelGGL(k_bscatter, dim3(g.nblk), dim3(kTile), sizeof(int) * kWaves * g.Cmax, st, g.d_jobs.ptr, g.d_blkjob.ptr,
                       g.d_pick.ptr, g.d_offs.ptr, g.d_sorted.ptr);
    hipLaunchKernelGGL(k_bsegments, dim3(1), dim3(1), 0, st, g.G, g.d_jobs.ptr, g.d_gjob.ptr, g.d_offs.ptr, g.d_total.ptr, g.d_seg.ptr,
                       g.d_tile.ptr);
    if (timed) HIP_CHECK(hipEventRecord(g.ev_fwd.first, st));
    hipLaunchKernelGGL(k_bforward<NT>, dim3(batch_fgrid(b, g)), dim3(kTile), 0, st, g.G, g.d_jobs.ptr, g.d_gjob.ptr, g.d_seg.ptr,
                       g.d_tile.ptr, g.d_sorted.ptr, g.d_mom.ptr, g.natt, g.d_acc.ptr);
    if (g.any_kt)
        hipLaunchKernelGGL(k_bktcut<NT>, dim3(batch_fgrid(b, g)), dim3(kTile), 0, st, g.G, g.d_jobs.ptr, g.d_gjob.ptr, g.d_seg.ptr,
                           g.d_tile.ptr, g.d_sorted.ptr, g.d_mom.ptr, g.natt, g.d_acc.ptr);
    if (timed) HIP_CHECK(hipEventRecord(g.ev_fwd.second, st));
    hipLaunchKernelGGL(k_acc_count, dim3(g.nblk), dim3(kTile), 0, st, g.d_acc.ptr, g.natt, g.d_blkacc.ptr);
    hipLaunchKernelGGL(k_bcutoff, dim3(J), dim3(64), 0, st, g.d_jobs.ptr, g.d_blkacc.ptr, g.d_acc.ptr, g.d_cut.ptr, g.d_keptj.ptr);
    hipLaunchKernelGGL(k_bkeep_count, dim3(g.tgrid), dim3(kTile), 0, st, g.G, g.d_jobs.ptr, g.d_gjob.ptr, g.d_seg.ptr, g.d_tile.ptr,
                       g.d_sorted.ptr, g.d_acc.ptr, g.d_cut.ptr, g.d_kept.ptr);
    bscan(st, g.d_kept.ptr, g.d_keptoff.ptr, g.tgrid, g.d_bsum, g.d_total2.ptr);
    hipLaunchKernelGGL(k_bkeep_scatter, dim3(g.tgrid), dim3(kTile), 0, st, g.G, g.d_jobs.ptr, g.d_gjob.ptr, g.d_seg.ptr, g.d_tile.ptr,
                       g.d_sorted.ptr, g.d_acc.ptr, g.d_cut.ptr, g.d_keptoff.ptr, g.d_comp.ptr);
    hipLaunchKernelGGL(k_bkeep_segments, dim3(1), dim3(1), 0, st, g.G, g.d_tile.ptr, g.d_keptoff.ptr, g.d_total2.ptr, g.d_seg2.ptr,
                       g.d_tile2.ptr, kTile * b->main_rnd);
    if (timed) HIP_CHECK(hipEventRecord(g.ev_main.first, st));
    switch (b->main_rnd) {
        case 1: launch_bmain<NT, 1>(b, g, train); break;
        case 2: launch_bmain<NT, 2>(b, g, train); break;
        case 4: launch_bmain<NT, 4>(b, g, train); break;
        default: throw chili_error(CHILI_EINVAL, "points per thread must be 1, 2 or 4");
    }
    if (timed) HIP_CHECK(hipEventRecord(g.ev_main.second, st));
    const int ny = train ? std::max(1, (int)std::min<int64_t>(64, (1 + g.Cmax + (int64_t)g.Cmax * DB + 255) / 256)) : 1;
    hipLaunchKernelGGL(k_breduce, dim3(J, ny), dim3(256), 0, st, g.d_jobs.ptr, g.d_tile2.ptr, g.d_part.ptr,
                       kStatW + g.Cmax + (int64_t)DB, DB, train);
}

void enqueue_all(chili_batch* b, int train, bool timed) {
    for (auto& g : b->groups) CHILI_DISPATCH_NT(g.NT, enqueue_group<NT>(b, g, train, timed));
}

// Per-job device tables of this pass (Philox stream and seed per job, a0 = 0).
void batch_upload_jobs(chili_batch* b, const uint64_t* seed, const uint32_t* streams, const int64_t* ncalls) {
    for (auto& g : b->groups) {
        const int J = (int)g.jobs.size();
        g.h_jobs.assign(J, BJob{});
        for (int q = 0; q < J; ++q) {
            const int j = g.jobs[q];
            chili_mc* m = b->jobs[j];
            BJob& jb = g.h_jobs[q];
            jb.A = make_args(m, nullptr, 0, seed[j], streams[j], 0);
            jb.att0 = g.att0[q];
            jb.n = g.n[q];
            jb.need = ncalls[j];
            jb.cnt0 = g.cnt0[q];
            jb.blk0 = g.blk0[q];
            jb.nblk = (int)(g.n[q] / kTile);
            jb.g0 = g.g0[q];
            jb.kt = (m->f.cuts == CHILI_CUTS_KT_JETS) ? 1 : 0;
            jb.tot = g.d_tot.ptr + g.tot_off[q];
        }
        g.d_jobs.upload(g.h_jobs, b->stream);
    }
}

}  // namespace

extern "C" {

int chili_construct_channels(const chili_process* proc, chili_channels** out) {
    return guarded([&] {
        if (!proc || !out) throw chili_error(CHILI_EINVAL, "null argument");
        auto* ch = new chili_channels();
        std::unique_ptr<chili_channels> hold(ch);
        auto& S = ch->set;
        S.model = chili_host::model_from_abi(*proc, S.cuts);
        S.topo = chili_host::construct_topologies(S.model, S.cuts);
        S.ntot = (int)S.model.flavs.size();
        S.ndims = 3 * (S.ntot - 2) - 2;
        S.sexternal.assign(S.ntot, 0.0);
        for (int i = 2; i < S.ntot; ++i) S.sexternal[i] = S.cuts.sexternal[i - 2];
        for (const auto& t : S.topo) S.progs.push_back(chili_host::compile_channel(t, S.cuts, S.model.sqrts, S.ntot));
        *out = hold.release();
    });
}
int chili_channels_count(const chili_channels* ch) { return ch ? (int)ch->set.topo.size() : CHILI_EINVAL; }
int chili_channels_ndims(const chili_channels* ch) { return ch ? ch->set.ndims : CHILI_EINVAL; }
int chili_channels_get(const chili_channels* ch, int i, chili_channel_desc* out) {
    return guarded([&] {
        if (!ch || !out || i < 0 || i >= (int)ch->set.topo.size()) throw chili_error(CHILI_EINVAL, "bad channel index");
        std::memset(out, 0, sizeof(*out));
        const auto& t = ch->set.topo[i];
        auto cv = [](const chili_host::Particle& p) { chili_particle_info q; q.idx = p.idx; q.pid = p.pid; q.mass = p.mass; q.width = p.width; return q; };
        out->ninfo = (int)t.tchan.size();
        for (size_t k = 0; k < t.tchan.size(); ++k) out->info[k] = cv(t.tchan[k]);
        out->ndecay = (int)t.decays.size();
        int k = 0;
        for (const auto& dk : t.decays) {
            out->decay[k][0] = cv(dk.first);
            out->decay[k][1] = cv(dk.second.first);
            out->decay[k][2] = cv(dk.second.second);
            ++k;
        }
    });
}
void chili_channels_free(chili_channels* ch) { delete ch; }

int chili_mc_create(const chili_channels* ch, int nbins, const chili_integrand* f, int device, chili_mc** out) {
    return guarded([&] {
        if (!ch || !f || !out) throw chili_error(CHILI_EINVAL, "null argument");
        if (nbins < 2 || nbins > 256) throw chili_error(CHILI_ERANGE, "nbins must be in [2, 256]");
        if (ch->set.topo.empty()) throw chili_error(CHILI_EINVAL, "process has no channels");
        if (ch->set.ntot < 4 || ch->set.ntot > CHILI_MAX_FLAV) throw chili_error(CHILI_ERANGE, "device path supports 2->2 .. 2->8");
        if (ch->set.topo.size() > 4096) throw chili_error(CHILI_ERANGE, "at most 4096 channels");
        if (f->cuts < CHILI_CUTS_NONE || f->cuts > CHILI_CUTS_KT_JETS) throw chili_error(CHILI_EINVAL, "unknown cut kind");
        if (f->cuts == CHILI_CUTS_KT_JETS && !(f->drmin > 0)) throw chili_error(CHILI_EINVAL, "kT jets need a radius drmin > 0");
        select_device(device);
        auto m = std::make_unique<chili_mc>();
        m->device = device;
        HIP_CHECK(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
        m->set = ch->set;
        m->f = *f;
        m->C = (int)ch->set.topo.size();
        m->d = ch->set.ndims;
        m->NT = ch->set.ntot;
        m->B = nbins;
        m->P2 = next_pow2(nbins);
        m->d_progs.upload(m->set.progs, m->stream);
        m->d_sext.upload(m->set.sexternal, m->stream);
        std::vector<chili_integrand> fv(1, *f);
        m->d_f.upload(fv, m->stream);
        m->d_seg.resize(m->C + 1);
        m->d_tiles.resize(m->C + 1);
        m->d_seg2.resize(m->C + 1);
        m->d_tiles2.resize(m->C + 1);
        m->d_grp2.resize(m->C + 1);
        m->d_total.resize(1);
        m->d_tot.resize(kStatW + m->C + (size_t)m->C * m->d * m->B);
        // uniform grids and alpha = 1/C (AdaptiveMap.hh:29-40, MultiChannel.cc:62-67)
        m->grids.resize((size_t)m->C * m->d * (m->B + 1));
        for (int c = 0; c < m->C; ++c)
            for (int k = 0; k < m->d; ++k)
                for (int b = 0; b <= m->B; ++b) m->grids[((size_t)c * m->d + k) * (m->B + 1) + b] = (double)b / (double)m->B;
        m->alpha.assign(m->C, 1.0 / (double)m->C);
        upload_state(m.get());
        HIP_CHECK(hipStreamSynchronize(m->stream));
        *out = m.release();
    });
}

int chili_mc_destroy(chili_mc* m) {
    return guarded([&] {
        if (!m) return;
        select_device(m->device);
        // teardown is best effort: a sticky error of earlier work must not
        // keep the context from being freed
        if (m->stream) (void)hipStreamSynchronize(m->stream);
        hipStream_t s = m->stream;
        delete m;
        if (s) (void)hipStreamDestroy(s);
    });
}

int chili_mc_info(const chili_mc* m, int* nchannels, int* ndims, int* nbins, int* ntot) {
    return guarded([&] {
        if (!m) throw chili_error(CHILI_EINVAL, "null context");
        if (nchannels) *nchannels = m->C;
        if (ndims) *ndims = m->d;
        if (nbins) *nbins = m->B;
        if (ntot) *ntot = m->NT;
    });
}

int chili_mc_set_grids(chili_mc* m, const double* grids, int nbins) {
    return guarded([&] {
        if (!m || !grids) throw chili_error(CHILI_EINVAL, "null argument");
        if (nbins < 2 || nbins > 256) throw chili_error(CHILI_ERANGE, "nbins must be in [2, 256]");
        // validate before any state changes: a rejected grid leaves the
        // context (bin count, buffers sized by it) as it was
        check_edges(grids, (size_t)m->C * m->d, nbins);
        select_device(m->device);
        if (nbins != m->B) {
            m->B = nbins;
            m->P2 = next_pow2(nbins);
            m->d_tot.resize(kStatW + m->C + (size_t)m->C * m->d * m->B);
            m->nmax = 0;   // record sizes changed
        }
        m->grids.assign(grids, grids + (size_t)m->C * m->d * (m->B + 1));
        upload_state(m);
    });
}
int chili_mc_get_grids(const chili_mc* m, double* grids) {
    return guarded([&] {
        if (!m || !grids) throw chili_error(CHILI_EINVAL, "null argument");
        std::copy(m->grids.begin(), m->grids.end(), grids);
    });
}
int chili_mc_set_alpha(chili_mc* m, const double* alpha) {
    return guarded([&] {
        if (!m || !alpha) throw chili_error(CHILI_EINVAL, "null argument");
        select_device(m->device);
        m->alpha.assign(alpha, alpha + m->C);
        upload_state(m);
    });
}
int chili_mc_get_alpha(const chili_mc* m, double* alpha) {
    return guarded([&] {
        if (!m || !alpha) throw chili_error(CHILI_EINVAL, "null argument");
        std::copy(m->alpha.begin(), m->alpha.end(), alpha);
    });
}

// One iteration: attempts from a0 until ncalls pass the cut with a finite
// value (MultiChannel.hh:214-248).  stats/T/hist are host outputs (T, hist
// only when train != 0).  vals_dev (device pointer, optional) receives the
// value of every kept point (channel-sorted order), at most vals_cap of them.
int chili_mc_iterate(chili_mc* m, const chili_uniforms* uni, uint64_t a0, int64_t ncalls, int train, chili_stats* stats,
                     double* T, double* hist, int64_t* n_attempts, double* vals_dev, int64_t vals_cap) {
    return guarded([&] {
        if (!m || !uni || !stats) throw chili_error(CHILI_EINVAL, "null argument");
        if (ncalls <= 0) throw chili_error(CHILI_EINVAL, "ncalls must be positive");
        select_device(m->device);
        std::vector<double> tot;
        int64_t att = 0;
        // the host callback path moves every attempt's momenta to the host:
        // smaller sub-batches
        const int64_t cap = m->host_fn ? (int64_t(1) << 16) : kSubMax;
        iterate_loop(m, uni, a0, ncalls, cap, tot, att, stats,
                     [&](const double* ud, int64_t ustride, uint64_t ab, int64_t n, int64_t need, int64_t got, int64_t& kept,
                         int64_t& used) {
                         run_subbatch(m, ud, ustride, a0, uni->seed, uni->stream, ab, n, need, train, vals_dev, got, vals_cap, kept,
                                      used);
                     });
        if (T && train) std::copy(tot.begin() + kStatW, tot.begin() + kStatW + m->C, T);
        if (hist && train) std::copy(tot.begin() + kStatW + m->C, tot.end(), hist);
        if (n_attempts) *n_attempts = att;
    });
}

// MultiChannel::operator() (MultiChannel.hh:196-251) without training and
// with an unweighting PostProcess (:240-245): a point of value v passing the
// cut is kept with probability min(1, |v| / wmax) (the attempt's Philox
// uniform number d+1), else it adds a zero and the loop retries, until
// nevents are kept.  Device outputs: ev [4 ntot][cap] momenta of the kept
// events in attempt order, ew [cap] their weights sign(v) max(1, |v| / wmax).
int chili_mc_generate(chili_mc* m, const chili_uniforms* uni, uint64_t a0, int64_t nevents, double wmax, double* ev_dev,
                      double* ew_dev, int64_t cap, chili_stats* stats, int64_t* n_attempts) {
    return guarded([&] {
        if (!m || !uni || !stats || !ev_dev || !ew_dev) throw chili_error(CHILI_EINVAL, "null argument");
        if (nevents <= 0) throw chili_error(CHILI_EINVAL, "nevents must be positive");
        if (!(wmax > 0) || !std::isfinite(wmax)) throw chili_error(CHILI_EINVAL, "wmax must be positive and finite");
        if (cap < nevents) throw chili_error(CHILI_ERANGE, "event buffer too small");
        if (uni->u) throw chili_error(CHILI_EINVAL, "the unweighting stage needs device-generated (Philox) uniforms");
        no_host_callback(m, "chili_mc_generate");
        select_device(m->device);
        std::vector<double> tot;
        int64_t att = 0;
        iterate_loop(m, uni, a0, nevents, kSubMax, tot, att, stats,
                     [&](const double*, int64_t, uint64_t ab, int64_t n, int64_t need, int64_t got, int64_t& kept, int64_t& used) {
                         run_subbatch_gen(m, uni->seed, uni->stream, ab, n, need, wmax, ev_dev, ew_dev, cap, got, kept, used);
                     },
                     &m->gen_rate, 1e-4);
        if (n_attempts) *n_attempts = att;
    });
}

// Batch evaluation on device buffers (the reference's ChiliCPU.evaluate /
// CallChili, nanobind.hh:23-50, Tensorflow.cc:22-25): rans [n][ld] row-major.
// channel in [0, C): that channel's mapping on the randoms as given (ld >= d),
// out = f * phase-space weight (0 when the cut fails), the reference's
// Evaluate.  channel = -1 (multichannel): column d picks the channel, the
// VEGAS map of that channel, every channel's inverse map and the
// multichannel weight as in chili_mc_iterate (ld >= d + 1, randoms in
// [0, 1)); out = f * W, 0 when the cut fails.  No retries: one value per row.
int chili_mc_evaluate(chili_mc* m, int channel, const double* rans_dev, int64_t n, int64_t ld, double* out_dev,
                      void* producer_stream) {
    return guarded([&] {
        if (!m || !rans_dev || !out_dev) throw chili_error(CHILI_EINVAL, "null argument");
        if (n <= 0 || n > (int64_t(1) << 26)) throw chili_error(CHILI_ERANGE, "n must be in [1, 2^26]");
        if (channel < -1 || channel >= m->C) throw chili_error(CHILI_EINVAL, "channel out of range");
        if (ld < m->d + (channel < 0 ? 1 : 0)) throw chili_error(CHILI_EINVAL, "row stride shorter than the randoms needed");
        no_host_callback(m, "chili_mc_evaluate");
        select_device(m->device);
        // the randoms (and the output's allocation) are ordered on the
        // caller's stream: the context's non-blocking stream waits for it
        if (!m->producer_ev) HIP_CHECK(hipEventCreateWithFlags(&m->producer_ev, hipEventDisableTiming));
        HIP_CHECK(hipEventRecord(m->producer_ev, (hipStream_t)producer_stream));
        HIP_CHECK(hipStreamWaitEvent(m->stream, m->producer_ev, 0));
        if (channel >= 0) {
            MCArgs A = make_args(m, nullptr, 0, 0, 0, 0);
            const unsigned g = (unsigned)((n + kTile - 1) / kTile);
            CHILI_DISPATCH_NT(m->NT, hipLaunchKernelGGL((k_eval_single<NT>), dim3(g), dim3(kTile), 0, m->stream, A, channel, rans_dev,
                                                        n, ld, out_dev));
            HIP_CHECK(hipGetLastError());
            HIP_CHECK(hipStreamSynchronize(m->stream));
            return;
        }
        const int nk = m->d + 1;
        m->d_u.resize((size_t)nk * n);
        DevBuf<int> bad;
        bad.resize(1);
        const unsigned g = (unsigned)((n + kTile - 1) / kTile);
        hipLaunchKernelGGL(chili_nf::k_fill<int>, dim3(1), dim3(kTile), 0, m->stream, (int64_t)1, 0, bad.ptr);
        hipLaunchKernelGGL(k_rows_to_soa, dim3(g), dim3(kTile), 0, m->stream, rans_dev, n, ld, nk, m->d_u.ptr, bad.ptr);
        int nbad = 0;
        bad.download(&nbad, 1, m->stream);
        HIP_CHECK(hipStreamSynchronize(m->stream));
        if (nbad) throw chili_error(CHILI_EINVAL, "randoms must lie in [0, 1)");
        hipLaunchKernelGGL(chili_nf::k_fill<double>, dim3(g), dim3(kTile), 0, m->stream, n, 0.0, out_dev);
        ensure_capacity(m, n);
        const int nblk = (int)((n + kTile - 1) / kTile);
        MCArgs A = make_args(m, m->d_u.ptr, n, 0, 0, 0);
        const int tgrid = pick_sort_forward(m, A, n, PointRec{});
        int64_t passed = 0, all = 0;
        cutoff_and_compact(m, n, nblk, tgrid, std::numeric_limits<int64_t>::max(), passed, all);
        if (passed > 0) {
            m->d_gval.resize(passed);
            const int mgrid = (int)((passed + kMainTile - 1) / kMainTile) + m->C;
            CHILI_DISPATCH_NT(m->NT, launch_main<NT>(m, A, mgrid, 0, m->d_gval.ptr, 0, passed, PointRec{}, passed));
            hipLaunchKernelGGL(k_vals_to_rows, dim3((unsigned)((passed + kTile - 1) / kTile)), dim3(kTile), 0, m->stream, passed,
                               m->d_comp.ptr, m->d_sorted.ptr, m->d_gval.ptr, out_dev);
        }
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipStreamSynchronize(m->stream));
    });
}

// Host integrand / PreProcess (the reference's std::function slow path,
// include/Channel/Integrand.hh:41-47): chili_mc_iterate and chili_mc_points
// hand the momenta of the attempts passing the device cut to fn; fn = NULL
// returns to the device integrand.
int chili_mc_set_host_integrand(chili_mc* m, chili_host_integrand_fn fn, void* user) {
    return guarded([&] {
        if (!m) throw chili_error(CHILI_EINVAL, "null context");
        select_device(m->device);
        HIP_CHECK(hipStreamSynchronize(m->stream));
        m->host_fn = fn;
        m->host_user = user;
    });
}

// Kernel timing on the context's stream (HIP events around every k_main and
// k_forward launch).  enable: 1 on, 0 off; reset clears the totals.  Outputs
// (each may be NULL): summed k_main milliseconds, its launch count and the
// points it processed, summed k_forward milliseconds.
int chili_mc_timing(chili_mc* m, int enable, int reset, double* main_ms, int64_t* launches, int64_t* points, double* fwd_ms) {
    return guarded([&] {
        if (!m) throw chili_error(CHILI_EINVAL, "null context");
        select_device(m->device);
        HIP_CHECK(hipStreamSynchronize(m->stream));
        harvest_events(m);
        if (main_ms) *main_ms = m->main_ms;
        if (launches) *launches = m->main_launches;
        if (points) *points = m->main_points;
        if (fwd_ms) *fwd_ms = m->fwd_ms;
        if (reset) { m->main_ms = 0; m->fwd_ms = 0; m->main_launches = 0; m->main_points = 0; m->fwd_launches = 0; m->fwd_points = 0; }
        m->timing = enable != 0;
    });
}

// Launches and attempts of the forward kernel timed since the last reset of
// chili_mc_timing (fwd_ms covers them).
int chili_mc_timing_forward(chili_mc* m, int64_t* launches, int64_t* attempts) {
    return guarded([&] {
        if (!m) throw chili_error(CHILI_EINVAL, "null context");
        select_device(m->device);
        HIP_CHECK(hipStreamSynchronize(m->stream));
        harvest_events(m);
        if (launches) *launches = m->fwd_launches;
        if (attempts) *attempts = m->fwd_points;
    });
}

// Per-attempt records for attempts [a0, a0+n) without the retry loop (every
// attempt evaluated once): the batch FFI's view of the path
// (include/Interfaces/nanobind.hh:23-50) extended to the multichannel weight.
// Host outputs (each may be NULL): status[n] (0 cut-rejected, 1 accepted,
// 2 NaN value), pick[n], w[n], f[n], val[n], ps[C][n], x[d][n] (inverse point
// of the picked channel), bins[d][n] (its FindBin), mom[ntot*4][n].
int chili_mc_points(chili_mc* m, const chili_uniforms* uni, uint64_t a0, int64_t n, int32_t* status, int32_t* pick, double* w,
                    double* f, double* val, double* ps, double* x, int32_t* bins, double* mom) {
    return guarded([&] {
        if (!m || !uni) throw chili_error(CHILI_EINVAL, "null argument");
        if (n <= 0 || n > (int64_t(1) << 24)) throw chili_error(CHILI_ERANGE, "n must be in [1, 2^24]");
        if (m->host_fn && n > (int64_t(1) << 22)) throw chili_error(CHILI_ERANGE, "with a host callback n must be <= 2^22");
        select_device(m->device);
        const double* ud = nullptr;
        int64_t ustride = 0;
        if (uni->u) {
            if (uni->stride < n) throw chili_error(CHILI_ERANGE, "host uniforms shorter than n");
            const size_t cnt = (size_t)(m->d + 1) * uni->stride;
            check_uniforms(uni->u, cnt);
            m->d_u.resize(cnt);
            HIP_CHECK(hipMemcpyAsync(m->d_u.ptr, uni->u, cnt * sizeof(double), hipMemcpyHostToDevice, m->stream));
            ud = m->d_u.ptr;
            ustride = uni->stride;
        }
        DevBuf<int32_t> st, pk, bn;
        DevBuf<double> wv, fv, vv, pv, xv, mv;
        st.resize(n); pk.resize(n); bn.resize((size_t)m->d * n);
        wv.resize(n); fv.resize(n); vv.resize(n); pv.resize((size_t)m->C * n); xv.resize((size_t)m->d * n);
        mv.resize((size_t)m->NT * 4 * n);
        HIP_CHECK(hipMemsetAsync(wv.ptr, 0, n * sizeof(double), m->stream));
        HIP_CHECK(hipMemsetAsync(fv.ptr, 0, n * sizeof(double), m->stream));
        HIP_CHECK(hipMemsetAsync(vv.ptr, 0, n * sizeof(double), m->stream));
        HIP_CHECK(hipMemsetAsync(pv.ptr, 0, (size_t)m->C * n * sizeof(double), m->stream));
        HIP_CHECK(hipMemsetAsync(xv.ptr, 0, (size_t)m->d * n * sizeof(double), m->stream));
        HIP_CHECK(hipMemsetAsync(bn.ptr, 0xff, (size_t)m->d * n * sizeof(int32_t), m->stream));
        PointRec R{st.ptr, pk.ptr, wv.ptr, fv.ptr, vv.ptr, pv.ptr, xv.ptr, bn.ptr, mv.ptr, n};
        const size_t ntot = kStatW + m->C + (size_t)m->C * m->d * m->B;
        std::vector<double> tot(ntot, 0.0);
        m->d_tot.upload(tot, m->stream);
        int64_t kept = 0, used = 0;
        run_subbatch(m, ud, ustride, a0, uni->seed, uni->stream, a0, n, n, 0, nullptr, 0, 0, kept, used, R);
        HIP_CHECK(hipGetLastError());
        if (status) st.download(status, n, m->stream);
        if (pick) pk.download(pick, n, m->stream);
        if (w) wv.download(w, n, m->stream);
        if (f) fv.download(f, n, m->stream);
        if (val) vv.download(val, n, m->stream);
        if (ps) pv.download(ps, (size_t)m->C * n, m->stream);
        if (x) xv.download(x, (size_t)m->d * n, m->stream);
        if (bins) bn.download(bins, (size_t)m->d * n, m->stream);
        if (mom) mv.download(mom, (size_t)m->NT * 4 * n, m->stream);
        HIP_CHECK(hipStreamSynchronize(m->stream));
    });
}

// ---------------------------------------------------------------------------
// Learned proposal (no reference counterpart).  fl (chili_flow_create with
// nflows = channels, ndims = dimensions) may be NULL: every q_j = 1, which is
// the plain multichannel weight evaluated through the same kernels.
// ---------------------------------------------------------------------------
// One production iteration: attempts from a0 until ncalls points pass (the
// retry semantics of chili_mc_iterate); grids, alpha and flows stay fixed.
int chili_mc_flow_iterate(chili_mc* m, chili_flow* fl, const chili_uniforms* uni, uint64_t a0, int64_t ncalls, chili_stats* stats,
                          int64_t* n_attempts, double* vals_dev, int64_t vals_cap) {
    return guarded([&] {
        if (!m || !uni || !stats) throw chili_error(CHILI_EINVAL, "null argument");
        if (ncalls <= 0) throw chili_error(CHILI_EINVAL, "ncalls must be positive");
        check_flow(m, fl);
        no_host_callback(m, "the flow path");
        select_device(m->device);
        std::vector<double> tot;
        int64_t att = 0;
        iterate_loop(m, uni, a0, ncalls, flow_sub_max(m->C, m->d), tot, att, stats,
                     [&](const double* ud, int64_t ustride, uint64_t ab, int64_t n, int64_t need, int64_t got, int64_t& kept,
                         int64_t& used) {
                         run_subbatch_flow(m, fl, ud, ustride, a0, uni->seed, uni->stream, ab, n, need, vals_dev, got, vals_cap,
                                           FlowData{nullptr, nullptr, nullptr, 0, 0}, kept, used);
                     });
        if (n_attempts) *n_attempts = att;
    });
}

// The unweighting stage on the learned proposal (chili_mc_generate's loop
// and outputs, the weights being the flow path's).  With the early
// rejection of chili_mc_specialize bit 3 and no mixture, the other
// channels' maps and densities are evaluated only for the points whose
// picked channel's term can pass the accept test.
int chili_mc_flow_generate(chili_mc* m, chili_flow* fl, const chili_uniforms* uni, uint64_t a0, int64_t nevents, double wmax,
                           double* ev_dev, double* ew_dev, int64_t cap, chili_stats* stats, int64_t* n_attempts) {
    return guarded([&] {
        if (!m || !uni || !stats || !ev_dev || !ew_dev) throw chili_error(CHILI_EINVAL, "null argument");
        if (nevents <= 0) throw chili_error(CHILI_EINVAL, "nevents must be positive");
        if (!(wmax > 0) || !std::isfinite(wmax)) throw chili_error(CHILI_EINVAL, "wmax must be positive and finite");
        if (cap < nevents) throw chili_error(CHILI_ERANGE, "event buffer too small");
        if (uni->u) throw chili_error(CHILI_EINVAL, "the unweighting stage needs device-generated (Philox) uniforms");
        check_flow(m, fl);
        no_host_callback(m, "chili_mc_flow_generate");
        select_device(m->device);
        std::vector<double> tot;
        int64_t att = 0;
        iterate_loop(m, uni, a0, nevents, flow_sub_max(m->C, m->d), tot, att, stats,
                     [&](const double*, int64_t, uint64_t ab, int64_t n, int64_t need, int64_t got, int64_t& kept, int64_t& used) {
                         const GenOut g{wmax, ev_dev, ew_dev, cap, got};
                         run_subbatch_flow(m, fl, nullptr, 0, ab, uni->seed, uni->stream, ab, n, need, nullptr, 0, 0,
                                           FlowData{nullptr, nullptr, nullptr, 0, 0}, kept, used, PointRec{}, &g);
                     },
                     &m->flow_gen_rate, 1e-4);
        if (n_attempts) *n_attempts = att;
    });
}

// Training records of one iteration drawn from the current proposal: for the
// first min(ncalls, cap) kept points r_j (device [C][cap][d]), beta_j = alpha_j
// h_j / g (device [C][cap]) and value^2 (device [cap]); n_kept receives the
// number written.
int chili_mc_flow_data(chili_mc* m, chili_flow* fl, const chili_uniforms* uni, uint64_t a0, int64_t ncalls, double* r_dev,
                       double* beta_dev, double* w2_dev, int64_t cap, chili_stats* stats, int64_t* n_kept, int64_t* n_attempts) {
    return guarded([&] {
        if (!m || !uni || !stats || !r_dev || !beta_dev || !w2_dev) throw chili_error(CHILI_EINVAL, "null argument");
        if (ncalls <= 0 || cap <= 0) throw chili_error(CHILI_EINVAL, "ncalls and cap must be positive");
        check_flow(m, fl);
        no_host_callback(m, "the flow path");
        select_device(m->device);
        std::vector<double> tot;
        int64_t att = 0, total_kept = 0;
        iterate_loop(m, uni, a0, ncalls, flow_sub_max(m->C, m->d), tot, att, stats,
                     [&](const double* ud, int64_t ustride, uint64_t ab, int64_t n, int64_t need, int64_t got, int64_t& kept,
                         int64_t& used) {
                         run_subbatch_flow(m, fl, ud, ustride, a0, uni->seed, uni->stream, ab, n, need, nullptr, 0, 0,
                                           FlowData{r_dev, beta_dev, w2_dev, got, cap}, kept, used);
                         total_kept = got + kept;
                     });
        if (n_kept) *n_kept = std::min(total_kept, cap);
        if (n_attempts) *n_attempts = att;
    });
}

// Per-attempt records of attempts [a0, a0+n) (no retry loop): status, pick,
// w, f, val as chili_mc_points; lq[C][n] = log q_j at the point (0 without
// flows); r[d][n] = the hypercube point the picked channel's flow drew.
int chili_mc_flow_points(chili_mc* m, chili_flow* fl, const chili_uniforms* uni, uint64_t a0, int64_t n, int32_t* status,
                         int32_t* pick, double* w, double* f, double* val, double* lq, double* r) {
    return guarded([&] {
        if (!m || !uni) throw chili_error(CHILI_EINVAL, "null argument");
        if (n <= 0 || n > (int64_t(1) << 22)) throw chili_error(CHILI_ERANGE, "n must be in [1, 2^22]");
        check_flow(m, fl);
        no_host_callback(m, "the flow path");
        select_device(m->device);
        int64_t ustride = 0;
        const double* ud = stage_uniforms(m, uni, ustride);
        if (ud && ustride < n) throw chili_error(CHILI_ERANGE, "host uniforms shorter than n");
        DevBuf<int32_t> st, pk;
        DevBuf<double> wv, fv, vv, lv, rv, mv;
        st.resize(n); pk.resize(n); wv.resize(n); fv.resize(n); vv.resize(n);
        lv.resize((size_t)m->C * n); rv.resize((size_t)m->d * n); mv.resize((size_t)m->NT * 4 * n);
        HIP_CHECK(hipMemsetAsync(wv.ptr, 0, n * sizeof(double), m->stream));
        HIP_CHECK(hipMemsetAsync(fv.ptr, 0, n * sizeof(double), m->stream));
        HIP_CHECK(hipMemsetAsync(vv.ptr, 0, n * sizeof(double), m->stream));
        HIP_CHECK(hipMemsetAsync(lv.ptr, 0, (size_t)m->C * n * sizeof(double), m->stream));
        HIP_CHECK(hipMemsetAsync(rv.ptr, 0, (size_t)m->d * n * sizeof(double), m->stream));
        PointRec R{st.ptr, pk.ptr, wv.ptr, fv.ptr, vv.ptr, lv.ptr, fl ? rv.ptr : nullptr, nullptr, mv.ptr, n};
        const size_t ntot = kStatW + m->C + (size_t)m->C * m->d * m->B;
        std::vector<double> tot(ntot, 0.0);
        m->d_tot.upload(tot, m->stream);
        int64_t kept = 0, used = 0;
        run_subbatch_flow(m, fl, ud, ustride, a0, uni->seed, uni->stream, a0, n, n, nullptr, 0, 0, FlowData{nullptr, nullptr, nullptr, 0, 0},
                          kept, used, R);
        HIP_CHECK(hipGetLastError());
        if (status) st.download(status, n, m->stream);
        if (pick) pk.download(pick, n, m->stream);
        if (w) wv.download(w, n, m->stream);
        if (f) fv.download(f, n, m->stream);
        if (val) vv.download(val, n, m->stream);
        if (lq) lv.download(lq, (size_t)m->C * n, m->stream);
        if (r) rv.download(r, (size_t)m->d * n, m->stream);
        HIP_CHECK(hipStreamSynchronize(m->stream));
    });
}

// ---------------------------------------------------------------------------
// Batched jobs (no single reference counterpart: many MultiChannel::operator()
// iterations, MultiChannel.hh:196-251, of independent integrands at once).
// ---------------------------------------------------------------------------
int chili_batch_create(int njobs, chili_mc* const* jobs, chili_batch** out) {
    return guarded([&] {
        if (!jobs || !out || njobs < 1) throw chili_error(CHILI_EINVAL, "need at least one job");
        auto b = std::make_unique<chili_batch>();
        b->device = jobs[0]->device;
        b->B = jobs[0]->B;
        for (int j = 0; j < njobs; ++j) {
            if (!jobs[j]) throw chili_error(CHILI_EINVAL, "null job");
            if (jobs[j]->device != b->device) throw chili_error(CHILI_EINVAL, "jobs of a batch share one device");
            if (jobs[j]->B != b->B) throw chili_error(CHILI_EINVAL, "jobs of a batch share one bin count");
            no_host_callback(jobs[j], "the batched path");
            b->jobs.push_back(jobs[j]);
        }
        select_device(b->device);
        HIP_CHECK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
        HIP_CHECK(hipDeviceGetAttribute(&b->ncu, hipDeviceAttributeMultiprocessorCount, b->device));
        b->acc_est.assign(njobs, 0.3);
        b->job_group.assign(njobs, -1);
        for (int nt = 4; nt <= CHILI_MAX_FLAV; ++nt) {
            BatchGroup g;
            g.NT = nt;
            g.D = 3 * (nt - 2) - 2;
            for (int j = 0; j < njobs; ++j)
                if (b->jobs[j]->NT == nt) {
                    b->job_group[j] = (int)b->groups.size();
                    g.jobs.push_back(j);
                    g.Cmax = std::max(g.Cmax, b->jobs[j]->C);
                    g.any_kt = g.any_kt || b->jobs[j]->f.cuts == CHILI_CUTS_KT_JETS;
                }
            if (g.jobs.empty()) continue;
            HIP_CHECK(hipEventCreate(&g.ev_fwd.first));
            HIP_CHECK(hipEventCreate(&g.ev_fwd.second));
            HIP_CHECK(hipEventCreate(&g.ev_main.first));
            HIP_CHECK(hipEventCreate(&g.ev_main.second));
            b->groups.push_back(std::move(g));
        }
        *out = b.release();
    });
}

int chili_batch_destroy(chili_batch* b) {
    return guarded([&] {
        if (!b) return;
        select_device(b->device);
        hipStream_t s = b->stream;
        if (s) hipStreamSynchronize(s);
        delete b;
        if (s) hipStreamDestroy(s);
    });
}

// Launch shape: workgroups per CU of the persistent forward and weight
// kernels, and kTile sub-tiles (points per thread) of a weight-kernel tile.
int chili_batch_set_shape(chili_batch* b, int fwd_wpc, int main_wpc, int main_rnd) {
    return guarded([&] {
        if (!b) throw chili_error(CHILI_EINVAL, "null batch");
        if (fwd_wpc < 1 || fwd_wpc > 16 || main_wpc < 1 || main_wpc > 8) throw chili_error(CHILI_ERANGE, "workgroups per CU out of range");
        if (main_rnd != 1 && main_rnd != 2 && main_rnd != 4) throw chili_error(CHILI_ERANGE, "points per thread must be 1, 2 or 4");
        b->fwd_wpc = fwd_wpc;
        b->main_wpc = main_wpc;
        b->main_rnd = main_rnd;
        b->graph_valid = false;
    });
}

int chili_batch_get_shape(const chili_batch* b, int* fwd_wpc, int* main_wpc, int* main_rnd, int64_t* captures, int64_t* replays,
                          int64_t* topups) {
    return guarded([&] {
        if (!b) throw chili_error(CHILI_EINVAL, "null batch");
        if (fwd_wpc) *fwd_wpc = b->fwd_wpc;
        if (main_wpc) *main_wpc = b->main_wpc;
        if (main_rnd) *main_rnd = b->main_rnd;
        if (captures) *captures = b->captures;
        if (replays) *replays = b->replays;
        if (topups) *topups = b->topups;
    });
}

// One iteration of every job: seed[j] / stream[j] select job j's Philox
// stream (attempts from 0), ncalls[j] accepted points.  One device pass runs
// every job (replayed from a hipGraph when use_graph != 0); a job whose pass
// falls short (acceptance below the plan's estimate, or NaN values) continues
// from its cutoff through the single-job path, so every job gets exactly the
// reference's retry semantics.  Outputs per job: stats[j], attempts[j]; with
// train != 0, T[j] (host [C_j]) and hist[j] (host [C_j][d_j][B]) when the
// pointer arrays are given.
int chili_batch_iterate(chili_batch* b, const uint64_t* seed, const uint32_t* streams, const int64_t* ncalls, int train, int use_graph,
                        chili_stats* stats, int64_t* attempts, double* const* T, double* const* hist) {
    return guarded([&] {
        if (!b || !seed || !streams || !ncalls || !stats) throw chili_error(CHILI_EINVAL, "null argument");
        const int NJ = (int)b->jobs.size();
        for (int j = 0; j < NJ; ++j)
            if (ncalls[j] <= 0) throw chili_error(CHILI_EINVAL, "ncalls must be positive");
        for (chili_mc* jm : b->jobs) no_host_callback(jm, "the batched path");
        select_device(b->device);
        for (chili_mc* m : b->jobs) HIP_CHECK(hipStreamSynchronize(m->stream));   // state uploads
        train = train ? 1 : 0;
        for (auto& g : b->groups) {
            if (batch_plan(b, g, ncalls)) b->graph_valid = false;
            batch_shape_buffers(b, g);
        }
        batch_upload_jobs(b, seed, streams, ncalls);
        if (use_graph) {
            if (!b->graph_valid || b->graph_train != train) {
                if (b->exec) { HIP_CHECK(hipGraphExecDestroy(b->exec)); b->exec = nullptr; }
                if (b->graph) { HIP_CHECK(hipGraphDestroy(b->graph)); b->graph = nullptr; }
                HIP_CHECK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
                enqueue_all(b, train, false);
                HIP_CHECK(hipStreamEndCapture(b->stream, &b->graph));
                HIP_CHECK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
                b->graph_valid = true;
                b->graph_train = train;
                b->captures += 1;
            }
            HIP_CHECK(hipGraphLaunch(b->exec, b->stream));
            b->replays += 1;
        } else {
            enqueue_all(b, train, false);
        }
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipStreamSynchronize(b->stream));
        for (auto& g : b->groups) {
            const int J = (int)g.jobs.size();
            std::vector<int64_t> cut(J), kept(J);
            std::vector<double> tot(g.d_tot.n);
            g.d_cut.download(cut.data(), J, b->stream);
            g.d_keptj.download(kept.data(), J, b->stream);
            g.d_tot.download(tot.data(), tot.size(), b->stream);
            HIP_CHECK(hipStreamSynchronize(b->stream));
            for (int q = 0; q < J; ++q) {
                const int j = g.jobs[q];
                chili_mc* m = b->jobs[j];
                const double* t = tot.data() + g.tot_off[q];
                const size_t ntot = kStatW + m->C + (size_t)m->C * m->d * m->B;
                std::vector<double> sum(t, t + ntot);
                const int64_t rej = cut[q] - kept[q];
                double n = t[0] + (double)rej, mn = t[1], mx = t[2];
                if (rej > 0) { mn = std::min(mn, 0.0); mx = std::max(mx, 0.0); }
                double s1 = t[3], s2 = t[4], nf = t[5];
                int64_t att = cut[q];
                const int64_t deficit = ncalls[j] - (kept[q] - (int64_t)t[6]);
                if (deficit > 0) {
                    // continue the job's attempt sequence from its cutoff
                    chili_uniforms uni{nullptr, 0, seed[j], streams[j], 0};
                    std::vector<double> tt;
                    int64_t att2 = 0;
                    chili_stats st2{};
                    iterate_loop(m, &uni, (uint64_t)cut[q], deficit, int64_t(1) << 23, tt, att2, &st2,
                                 [&](const double* ud, int64_t ustride, uint64_t ab, int64_t nn, int64_t need, int64_t got,
                                     int64_t& kk, int64_t& used) {
                                     run_subbatch(m, ud, ustride, (uint64_t)cut[q], seed[j], streams[j], ab, nn, need, train, nullptr,
                                                  0, 0, kk, used);
                                 });
                    n += st2.n; mn = std::min(mn, st2.min); mx = std::max(mx, st2.max);
                    s1 += st2.sum; s2 += st2.sum2; nf += st2.n_finite;
                    for (size_t k = kStatW; k < ntot; ++k) sum[k] += tt[k];
                    att += att2;
                    b->topups += 1;
                }
                stats[j].n = n; stats[j].min = mn; stats[j].max = mx; stats[j].sum = s1; stats[j].sum2 = s2; stats[j].n_finite = nf;
                if (attempts) attempts[j] = att;
                if (train && T && T[j]) std::copy(sum.begin() + kStatW, sum.begin() + kStatW + m->C, T[j]);
                if (train && hist && hist[j]) std::copy(sum.begin() + kStatW + m->C, sum.end(), hist[j]);
                if (cut[q] > 0) b->acc_est[j] = std::max(0.01, (double)kept[q] / (double)cut[q]);
            }
        }
    });
}

// Autotuner: times the persistent forward kernel at 1, 2, 4, 8 workgroups per
// CU and the weight kernel at 1, 2 workgroups per CU x 1, 2, 4 points per
// thread on one pass of this batch (no training), and keeps the fastest of
// each.  best_ms (optional): [0] forward, [1] weight kernel time of the
// chosen shapes.
int chili_batch_tune(chili_batch* b, const uint64_t* seed, const uint32_t* streams, const int64_t* ncalls, double* best_ms) {
    return guarded([&] {
        if (!b || !seed || !streams || !ncalls) throw chili_error(CHILI_EINVAL, "null argument");
        select_device(b->device);
        for (chili_mc* m : b->jobs) HIP_CHECK(hipStreamSynchronize(m->stream));
        auto pass_ms = [&](double& fwd, double& mainms) {
            for (auto& g : b->groups) {
                if (batch_plan(b, g, ncalls)) b->graph_valid = false;
                batch_shape_buffers(b, g);
            }
            batch_upload_jobs(b, seed, streams, ncalls);
            enqueue_all(b, 0, true);
            HIP_CHECK(hipStreamSynchronize(b->stream));
            fwd = 0;
            mainms = 0;
            for (auto& g : b->groups) {
                float a = 0, c = 0;
                HIP_CHECK(hipEventElapsedTime(&a, g.ev_fwd.first, g.ev_fwd.second));
                HIP_CHECK(hipEventElapsedTime(&c, g.ev_main.first, g.ev_main.second));
                fwd += a;
                mainms += c;
            }
        };
        double f0, m0;
        pass_ms(f0, m0);                 // warm-up (plans and buffers)
        int best_f = b->fwd_wpc;
        double tf = 1e300;
        for (int w : {1, 2, 4, 8}) {
            b->fwd_wpc = w;
            double f, mm;
            pass_ms(f, mm);
            if (f < tf) { tf = f; best_f = w; }
        }
        b->fwd_wpc = best_f;
        int best_w = b->main_wpc, best_r = b->main_rnd;
        double tm = 1e300;
        for (int w : {1, 2})
            for (int r : {1, 2, 4}) {
                b->main_wpc = w;
                b->main_rnd = r;
                double f, mm;
                pass_ms(f, mm);
                if (mm < tm) { tm = mm; best_w = w; best_r = r; }
            }
        b->main_wpc = best_w;
        b->main_rnd = best_r;
        b->graph_valid = false;
        if (best_ms) { best_ms[0] = tf; best_ms[1] = tm; }
    });
}

// ---------------------------------------------------------------------------
// Run-time specialisation (no reference counterpart): compile the forward and
// weight kernels of this context with its channel programs as compile-time
// constants (hipRTC, gfx950) and use them from now on: enable bit 0 selects
// the specialised forward kernel (uniform attempts), bit 1 the specialised fused
// weight kernel, bit 2 the split weight pass and on the learned-proposal
// path the per-channel geometry kernels, bit 3 the unweighting stage's exact
// early rejection (0: the generic kernels, no early rejection).
// compile_s (optional) receives the compile time of the first enable.
int chili_mc_specialize(chili_mc* m, int enable, double* compile_s) {
    return guarded([&] {
        if (!m) throw chili_error(CHILI_EINVAL, "null context");
        select_device(m->device);
        if (enable && !m->jit) {
            auto mod = std::make_unique<chili_jit::Module>();
            chili_jit::compile(chili_jit::kernel_source(m->set.progs, m->NT), *mod, &m->jit_compile_s, true, m->C);
            m->jit = std::move(mod);
        }
        m->jit_fwd = (enable & 1) != 0;
        m->jit_split = (enable & 4) != 0;
        m->jit_early = (enable & 8) != 0;
        m->jit_main = (enable & 2) != 0 && !m->jit_split;
        if (compile_s) *compile_s = m->jit_compile_s;
    });
}

// Run-time compilation of a channel set's specialised kernels without
// loading them (a build check that needs no GPU).
int chili_channels_compile_check(const chili_channels* ch, double* compile_s) {
    return guarded([&] {
        if (!ch) throw chili_error(CHILI_EINVAL, "null channels");
        chili_jit::Module mod;
        chili_jit::compile(chili_jit::kernel_source(ch->set.progs, ch->set.ntot), mod, compile_s, false);
    });
}

// The generated specialised source of a channel set (diagnostics): writes
// at most cap bytes (NUL-terminated) and the full length to *len.
int chili_channels_kernel_source(const chili_channels* ch, char* buf, int64_t cap, int64_t* len) {
    return guarded([&] {
        if (!ch) throw chili_error(CHILI_EINVAL, "null channels");
        const std::string src = chili_jit::kernel_source(ch->set.progs, ch->set.ntot);
        if (len) *len = (int64_t)src.size();
        if (buf && cap > 0) {
            const size_t n = std::min<size_t>(src.size(), (size_t)cap - 1);
            std::memcpy(buf, src.data(), n);
            buf[n] = '\0';
        }
    });
}

}  // extern "C"
