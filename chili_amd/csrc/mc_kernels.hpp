// mc_kernels.hpp -- gfx950 kernels of the multichannel VEGAS iteration.
//
// One iteration of MultiChannel::operator() (include/Channel/MultiChannel.hh:
// 196-251) over a batch of attempts runs as:
//   k_pick        channel of every attempt from its last uniform
//                 (std::discrete_distribution rule, lower_bound on cumulative alpha)
//   k_scan        exclusive scan of the per-(channel, block) counts
//   k_scatter     stable counting sort of the attempts by channel
//   k_forward     VEGAS forward map + FSMapper::GeneratePoint + cut, one
//                 channel per workgroup (no divergence between op lists)
//   k_compact     stable compaction of the accepted attempts below the cutoff
//                 (the reference retries rejected attempts until ncalls pass)
//   k_main        integrand, every channel's inverse map and VEGAS density,
//                 alpha-weighted total, StatsData / alpha-training / f^2
//                 histogram partials per workgroup in a fixed order
//   k_reduce      fixed-order sum of the workgroup partials
#pragma once
#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#endif
#include <stdint.h>

#include "../../include/chili_gpu.h"
#include "chanprog.hpp"
#include "phys.hpp"

namespace chili_dev {

using chili_host::ChanProg;

constexpr double kPi3 = 0x1.f019b59389d7bp+4;   // pow(M_PI, 3), correctly rounded
constexpr int kTile = 256;     // attempts per workgroup
constexpr int kWaves = kTile / 64;

// Optional per-attempt records (device arrays indexed by attempt, SoA).
struct PointRec {
    int32_t* status;   // 0 cut-rejected, 1 accepted, 2 NaN value
    int32_t* pick;
    double* w;
    double* f;
    double* val;
    double* ps;        // [C][n]
    double* x;         // [d][n] inverse point of the picked channel
    int32_t* bins;     // [d][n]
    double* mom;       // [ntot*4][n]
    int64_t n;
};

struct MCArgs {
    const ChanProg* progs;
    const double* edges;     // [C][d][B+1]
    const double* search;    // [C][d][P2] interior edges e_1..e_{B-1} padded with +inf
    const uint16_t* cells;   // [C][d][kCells + 1]: interior edges below j / kCells (null: plain search)
    const double* alpha;     // [C]
    const double* cum;       // [C] cumulative alpha (last entry 1.0)
    const double* sext;      // [ntot] external masses^2 (0 for incoming)
    const chili_integrand* f;
    const double* u;         // host-fed uniforms on the device ([k][ustride]) or null
    int64_t ustride;
    uint64_t ua0;            // attempt index of u[.][0]
    uint64_t seed;
    uint64_t a0;             // attempt index of the batch start
    uint32_t stream;
    int C, d, B, P2;
    double sqrts, ptmax;
    const double* fhost;     // integrand values from the host callback, by sorted position (null: device integrand)
    // split weight pass: the per-channel inverse kernels' results by kept
    // index e: ps [C][pre_stride], VEGAS weight [C][pre_stride] after them,
    // the picked channel's bins [d][pre_stride]
    const double* pre_pv;
    const short* pre_bins;
    int64_t pre_stride;
};

// --------------------------------------------------------------------------
// Momentum slots of one point, indexed by wave-uniform slot numbers: four
// scalar arrays touched only through get/set so that the compiler can keep
// them in VGPRs (indexed register moves) instead of private memory.
// --------------------------------------------------------------------------
template <int N>
struct MomRF {
    double v[4 * N];
    CHILI_DEV V4 get(int i) const { return mk(v[4 * i], v[4 * i + 1], v[4 * i + 2], v[4 * i + 3]); }
    CHILI_DEV void set(int i, const V4& q) { v[4 * i] = q.t; v[4 * i + 1] = q.x; v[4 * i + 2] = q.y; v[4 * i + 3] = q.z; }
    CHILI_DEV double mass2(int i) const { return vmass2(get(i)); }
};

// --------------------------------------------------------------------------
// uniforms
// --------------------------------------------------------------------------
template <int D>
CHILI_DEV void load_uniforms(const MCArgs& A, uint64_t att, double (&r)[D + 1]) {
    if (A.u) {
#pragma unroll
        for (int k = 0; k <= D; ++k) r[k] = A.u[(int64_t)k * A.ustride + (int64_t)(att - A.ua0)];
    } else {
#pragma unroll
        for (int k = 0; k <= D; k += 2) {
            double u0, u1;
            philox_pair(A.seed, A.stream, att, (uint32_t)k, u0, u1);
            r[k] = u0;
            if (k + 1 <= D) r[k + 1] = u1;
        }
    }
}
CHILI_DEV double channel_uniform(const MCArgs& A, uint64_t att) {
    if (A.u) return A.u[(int64_t)A.d * A.ustride + (int64_t)(att - A.ua0)];
    double u0, u1;
    philox_pair(A.seed, A.stream, att, (uint32_t)(A.d & ~1), u0, u1);
    return (A.d & 1) ? u1 : u0;
}
// Uniform d+1 of an attempt's Philox stream (device-generated uniforms only):
// the component choice of a mixture proposal, the accept / reject decision of
// the unweighting stage.
CHILI_DEV double aux_uniform(const MCArgs& A, uint64_t att) {
    double u0, u1;
    philox_pair(A.seed, A.stream, att, (uint32_t)((A.d + 1) & ~1), u0, u1);
    return ((A.d + 1) & 1) ? u1 : u0;
}
CHILI_DEV int pick_channel(const MCArgs& A, double u) {      // lower_bound(cum, u)
    int lo = 0, n = A.C;
    if (n < 2) return 0;
    while (n > 0) {
        int half = n >> 1;
        if (A.cum[lo + half] < u) { lo = lo + half + 1; n = n - half - 1; }
        else n = half;
    }
    return lo;
}

// FindBin (AdaptiveMap.cc:34-38): lower_bound over e_0..e_{B-1} minus one,
// with x <= e_0 or NaN clamped to bin 0.  For non-decreasing edges that is the
// number of interior edges e_1..e_{B-1} below x, found by a branch-free
// search over the +inf-padded power-of-two table (P2 >= B).
CHILI_DEV int find_bin(const double* __restrict__ s, int P2, double x) {
    int lo = 0;
    for (int step = P2 >> 1; step > 0; step >>= 1)
        lo += (s[lo + step - 1] < x) ? step : 0;
    return lo;
}

// The same count through a uniform cell index: for x in [j/K, (j+1)/K) it
// lies between t[j] and t[j+1] (the counts at the cell's ends), so a short
// search over that range replaces most of the dependent loads.  x outside
// [0, 1) (or NaN) takes the full search.
constexpr int kCells = 256;
CHILI_DEV int find_bin_cells(const double* __restrict__ s, const uint16_t* __restrict__ t, int P2, double x) {
    if (!(x >= 0.0 && x < 1.0)) return find_bin(s, P2, x);
    const int j = (int)(x * (double)kCells);   // exact: kCells is a power of two
    int lo = t[j], n = (int)t[j + 1] - lo;
    while (n > 0) {
        const int h = n >> 1;
        if (s[lo + h] < x) { lo += h + 1; n -= h + 1; } else { n = h; }
    }
    return lo;
}

// --------------------------------------------------------------------------
// FSMapper::GeneratePoint (Channel.cc:22-38) for channel program `p`.
// P: momentum slots, m2: masses^2 per slot, r: the VEGAS-mapped point.
// --------------------------------------------------------------------------
template <int NT>
CHILI_DEV void generate_point(const MCArgs& A, const ChanProg& p, const double* r, MomRF<2 * NT>& P, double (&m2)[2 * NT]) {
    constexpr int ND = NT - 3;       // max decays
    constexpr int NTT = NT - 2;      // max t-channel objects
#pragma unroll
    for (int i = 0; i < NT; ++i) m2[i] = A.sext[i];
    const int nd = p.nd, nt = p.nt;
    // GenDecays, Channel.cc:63-81
#pragma unroll
    for (int k = 0; k < ND; ++k) {
        if (k < nd) {
            double a = sqrt(m2[p.dec_c1[k]]) + sqrt(m2[p.dec_c2[k]]);
            double smin1 = a * a;
            double smin = dmax(dmax(smin1, p.dec_smin2[k]), p.dec_smincut[k]);
            double s = p.dec_massive[k] ? massive_prop_momenta(p.dec_mass[k], p.dec_width[k], smin, p.dec_smax[k], r[k])
                                        : massless_prop_momenta(smin, p.dec_smax[k], r[k]);
            m2[p.dec_par[k]] = s;
        }
    }
    // GenTChan, Channel.cc:115-158
    int ir = nd;
    V4 psum = mk(0., 0., 0., 0.);
    const double sqrts = A.sqrts, ptmax = A.ptmax;
#pragma unroll
    for (int i = 0; i < NTT - 1; ++i) {
        if (i < nt - 1) {
            const double ran = r[ir++];
            const double ptmin = p.t_ptmin[i];
            double pt = 2 * ptmin * ptmax * ran / (2 * ptmin + ptmax * (1 - ran));
            double etamax = sqrts / 2 / pt;
            etamax = log(etamax + sqrt(etamax * etamax - 1));
            etamax = dmin(etamax, p.t_etacut[i]);
            double y = etamax * (2 * r[ir++] - 1);
            double sinhy = sinh(y);
            double coshy = sqrt(1 + sinhy * sinhy);
            double phi = 2 * kPi * r[ir++];
            double mt = sqrt(pt * pt + m2[p.t_slot[i]]);
            double sp, cp;
            sincos(phi, &sp, &cp);
            V4 q = mk(mt * coshy, pt * cp, pt * sp, mt * sinhy);
            P.set(p.t_slot[i], q);
            psum = vadd(psum, q);
        }
    }
    {
        const int last = p.t_slot[nt - 1];
        double mjets2 = vmass2(psum);
        double yjets = vrap(psum);
        double ptj2 = vpt2(psum);
        double m2l = m2[last];
        double qt = sqrt(m2l + ptj2);
        double mt = sqrt(mjets2 + ptj2);
        double ymin = -log(sqrts / qt * (1 - mt / sqrts * exp(-yjets)));
        double ymax = log(sqrts / qt * (1 - mt / sqrts * exp(yjets)));
        double dely = ymax - ymin;
        double yv = ymin + r[ir++] * dely;
        double sinhy = sinh(yv);
        double coshy = sqrt(1 + sinhy * sinhy);
        V4 q = mk(qt * coshy, -psum.x, -psum.y, qt * sinhy);
        P.set(last, q);
        V4 tot = vadd(q, psum);
        double pp = tot.t + tot.z;
        double pm = tot.t - tot.z;
        P.set(0, mk(-pp / 2, 0., 0., -pp / 2));
        P.set(1, mk(-pm / 2, 0., 0., pm / 2));
    }
    // GenSChan / DecayParts, Channel.cc:204-245
#pragma unroll
    for (int k = 0; k < ND; ++k) {
        if (k < nd) {
            double r1 = r[ir++];
            double r2 = r[ir++];
            V4 pa, pb;
            schannel_momenta(P.get(p.s_par[k]), m2[p.s_a[k]], m2[p.s_b[k]], pa, pb, r1, r2);
            P.set(p.s_a[k], pa);
            P.set(p.s_b[k], pb);
        }
    }
}

// --------------------------------------------------------------------------
// FSMapper::GenerateWeight (Channel.cc:40-61) fused with the VEGAS density of
// the channel's grid (AdaptiveMap::GenerateWeight, AdaptiveMap.cc:58-65).
// Returns the phase-space weight; *vw receives the VEGAS weight; for the
// picked channel the bins of its inverse point go to `bins_out` (LDS).
// --------------------------------------------------------------------------
template <int NT, bool kOneLds = false>
CHILI_DEV double generate_weight(const MCArgs& A, const ChanProg& p, int ch, MomRF<2 * NT>& P, double* vw_out,
                                 short* bins_out, int bstride, double* x_out = nullptr, int32_t* bin_rec = nullptr,
                                 int64_t xstride = 0, double* r_out = nullptr, const double* search = nullptr,
                                 bool search_one = false) {
    constexpr int ND = NT - 3;
    constexpr int NTT = NT - 2;
    const int nd = p.nd, nt = p.nt;
    const double Bd = (double)A.B;
    const double* __restrict__ edges = A.edges + (int64_t)ch * A.d * (A.B + 1);
    // search tables: device memory, a workgroup's LDS copy of every channel's,
    // or (search_one) an LDS copy of this channel's alone; kOneLds: the last,
    // fixed at compile time, so that the pointer is known to address LDS and
    // the search compiles to LDS reads rather than flat loads
    const double* __restrict__ srch =
        kOneLds ? search : (search ? search : A.search) + (search_one ? 0 : (int64_t)ch * A.d * A.P2);
    double vw = 1.0;
    int slot = 0;
    const uint16_t* __restrict__ cells = (kOneLds || search || !A.cells) ? nullptr : A.cells + (int64_t)ch * A.d * (kCells + 1);
    auto take = [&](double x) {
        int b = cells ? find_bin_cells(srch + slot * A.P2, cells + slot * (kCells + 1), A.P2, x) : find_bin(srch + slot * A.P2, A.P2, x);
        const double* e = edges + slot * (A.B + 1);
        vw *= (e[b + 1] - e[b]) * Bd;
        if (bins_out) bins_out[slot * bstride] = (short)b;
        if (x_out) { x_out[slot * xstride] = x; bin_rec[slot * xstride] = b; }
        if (r_out) {
            // VEGAS inverse: the hypercube point the forward map sends to x
            const double wb = e[b + 1] - e[b];
            double r = (wb > 0) ? ((double)b + (x - e[b]) / wb) / Bd : (double)b / Bd;
            r_out[slot] = fmin(fmax(r, 0.0), 1.0);
        }
        ++slot;
    };
    // composite momenta, ascending decays (Channel.cc:47-50)
#pragma unroll
    for (int k = 0; k < ND; ++k)
        if (k < nd) P.set(p.dec_par[k], vadd(P.get(p.dec_c1[k]), P.get(p.dec_c2[k])));
    // WgtDecays, Channel.cc:83-113
    double wd = 1;
#pragma unroll
    for (int k = 0; k < ND; ++k) {
        if (k < nd) {
            double a = sqrt(fabs(P.mass2(p.dec_c1[k]))) + sqrt(fabs(P.mass2(p.dec_c2[k])));
            double smin1 = a * a;
            double smin = dmax(dmax(smin1, p.dec_smin2[k]), p.dec_smincut[k]);
            double s = P.mass2(p.dec_par[k]);
            double xr;
            if (p.dec_massive[k]) wd *= massive_prop_weight(p.dec_mass[k], p.dec_width[k], smin, p.dec_smax[k], s, xr);
            else wd *= massless_prop_weight(smin, p.dec_smax[k], s, xr);
            wd /= (2 * kPi);
            take(xr);
        }
    }
    // WgtTChan, Channel.cc:160-202
    const double sqrts = A.sqrts, ptmax = A.ptmax;
    double wt = 2 * kPi;
    V4 psum = mk(0., 0., 0., 0.);
#pragma unroll
    for (int i = 0; i < NTT - 1; ++i) {
        if (i < nt - 1) {
            wt *= 1.0 / (16 * kPi3);
            const V4 m = P.get(p.t_slot[i]);
            const double ptmin = p.t_ptmin[i];
            double pt = vpt(m);
            double tp = 2 * ptmin + pt;
            wt *= pt * ptmax / 2 / ptmin / (2 * ptmin + ptmax) * (tp * tp);
            take(pt * (ptmax + 2 * ptmin) / (ptmax * (pt + 2 * ptmin)));
            double etamax = sqrts / 2 / pt;
            etamax = log(etamax + sqrt(etamax * etamax - 1));
            etamax = isnan(etamax) ? 99 : etamax;
            etamax = dmin(etamax, p.t_etacut[i]);
            wt *= 2 * etamax * 2 * kPi;
            double y = vrap(m);
            double phi = vphi(m);
            take((y / etamax + 1) / 2);
            take(phi / 2 / kPi);
            psum = vadd(psum, m);
        }
    }
    {
        const V4 ml = P.get(p.t_slot[nt - 1]);
        double mjets2 = vmass2(psum);
        double yjets = vrap(psum);
        double ptj2 = vpt2(psum);
        double m2l = vmass2(ml);
        double qt = sqrt(m2l + ptj2);
        double mt = sqrt(mjets2 + ptj2);
        double yv = vrap(ml);
        double ymin = -log(sqrts / qt * (1 - mt / sqrts * exp(-yjets)));
        double ymax = log(sqrts / qt * (1 - mt / sqrts * exp(yjets)));
        double dely = ymax - ymin;
        if (isnan(ymax) || isnan(ymin)) {
            // The reference returns 0 here without writing this slot; the
            // remaining slots only feed a VEGAS weight that a zero
            // phase-space weight makes irrelevant.
            wt = 0;
            take(0.5);
        } else {
            wt *= dely / sqrts / sqrts;
            take((yv - ymin) / dely);
        }
    }
    // WgtSChan / WgtDecayParts, Channel.cc:222-256
    double ws = 1;
#pragma unroll
    for (int k = 0; k < ND; ++k) {
        if (k < nd) {
            double r1, r2;
            ws *= schannel_weight(P.get(p.s_a[k]), P.get(p.s_b[k]), r1, r2);
            take(r1);
            take(r2);
        }
    }
    *vw_out = vw;
    return wd * wt * ws;
}

// --------------------------------------------------------------------------
// device integrands and cuts
// --------------------------------------------------------------------------
CHILI_DEV double pair_s(const V4& a, const V4& b) { return vmass2(vadd(a, b)); }

// JetCluster (src/Tools/JetCluster.cc:8-78; Q2i, Q2ij JetCluster.hh; R2
// JetCluster.cc:80-84) for the question the kT-jet cut asks: do all N
// partons end as jets, i.e. does the inclusive kT sequence merge none of
// them?  Until a first merge no momentum changes, so every distance is
// computed once; each step scans the remaining partons in their order as the
// reference's loop does (beam distance of i, then pairs (i, j < i), strict <,
// the first remaining parton's beam distance as the start except in the first
// scan, which starts from DBL_MAX) and a pair at the minimum is that merge.
template <int N>
CHILI_DEV bool kt_all_jets(const V4* p, double R) {
    double pt2[N], y[N], ph[N], d[N][N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
        pt2[i] = vpt2(p[i]);
        y[i] = vrap(p[i]);
        ph[i] = vphi(p[i]);
    }
#pragma unroll
    for (int i = 0; i < N; ++i) {
        d[i][i] = pt2[i];
#pragma unroll
        for (int j = 0; j < i; ++j) {
            const double dy = y[i] - y[j], dphi = ph[i] - ph[j];   // phi in [0, 2 pi), not wrapped (JetCluster.cc:82)
            const double mn = (pt2[j] < pt2[i]) ? pt2[j] : pt2[i];
            d[i][j] = mn * ((dy * dy + dphi * dphi) / R / R);
        }
    }
    unsigned rem = (1u << N) - 1;
#pragma unroll
    for (int step = 0; step < N; ++step) {
        int ii = 0, jj = 0;
        double dmin = 0;
        bool seen = false;
#pragma unroll
        for (int i = 0; i < N; ++i) {
            if (!((rem >> i) & 1u)) continue;
            if (!seen) {
                seen = true;
                ii = jj = i;
                dmin = (step == 0) ? __DBL_MAX__ : d[i][i];
            }
            if (d[i][i] < dmin) { dmin = d[i][i]; ii = i; jj = i; }
#pragma unroll
            for (int j = 0; j < i; ++j) {
                if (!((rem >> j) & 1u)) continue;
                if (d[i][j] < dmin) { dmin = d[i][j]; ii = i; jj = j; }
            }
        }
        if (ii != jj) return false;
        rem &= ~(1u << ii);
    }
    return true;
}

template <int NT>
CHILI_DEV bool cut_pass(const chili_integrand* __restrict__ f, const MomRF<2 * NT>& PR, double sqrts) {
    V4 P[NT];
#pragma unroll
    for (int i = 0; i < NT; ++i) P[i] = PR.get(i);
    if (f->cuts == CHILI_CUTS_NONE) return true;
    if (isnan(P[0].t)) return false;
    // physical incoming momenta: x1, x2 <= 1 (incoming legs carry E = -x sqrt(s)/2)
    if (!(-2 * P[0].t <= sqrts) || !(-2 * P[1].t <= sqrts)) return false;
    // kT jets: the clustering is a pass of its own over the kept momenta (k_ktcut)
    if (f->cuts == CHILI_CUTS_KT_JETS) return true;
    const double ptmin = f->ptmin, etamax = f->etamax, drmin = f->drmin;
    double y[NT], ph[NT];
    bool ok = true;
#pragma unroll
    for (int i = 2; i < NT; ++i) {
        if (!(vpt(P[i]) > ptmin)) ok = false;
        y[i] = vrap(P[i]);
        ph[i] = vphi(P[i]);
        if (!(fabs(y[i]) < etamax)) ok = false;
    }
#pragma unroll
    for (int i = 2; i < NT; ++i)
#pragma unroll
        for (int j = i + 1; j < NT; ++j) {
            double dy = y[i] - y[j];
            double dphi = fabs(ph[i] - ph[j]);
            if (dphi > kPi) dphi = 2 * kPi - dphi;
            double dr = sqrt(dy * dy + dphi * dphi);
            if (!(dr > drmin)) ok = false;
        }
    return ok;
}

template <int NT>
CHILI_DEV double integrand(const chili_integrand* __restrict__ f, const MomRF<2 * NT>& PR) {
    V4 P[NT];
#pragma unroll
    for (int i = 0; i < NT; ++i) P[i] = PR.get(i);
    const int kind = f->kind;
    if (kind == CHILI_F_UNIT) return 1.0;
    constexpr int NP = (NT - 2) * (NT - 3) / 2;
    double s[NP > 0 ? NP : 1];
    {
        int q = 0;
#pragma unroll
        for (int i = 2; i < NT; ++i)
#pragma unroll
            for (int j = i + 1; j < NT; ++j) s[q++] = pair_s(P[i], P[j]);
    }
    if (kind == CHILI_F_PROPPROD) {
        const double s0 = f->params[0];
        if (NP == 1) return s0 / s[0];
        double tot = 0;
#pragma unroll
        for (int a = 0; a < NP; ++a)
#pragma unroll
            for (int b = a + 1; b < NP; ++b) tot += s0 * s0 / (s[a] * s[b]);
        return tot;
    }
    if (kind == CHILI_F_RESONANT) {
        const int nbw = (int)f->params[0];
        double bw = 1.0;
        for (int k = 0; k < nbw; ++k) {
            unsigned mask = (unsigned)f->params[1 + 3 * k];
            double M = f->params[2 + 3 * k], G = f->params[3 + 3 * k];
            V4 qs = mk(0., 0., 0., 0.);
#pragma unroll
            for (int i = 0; i < NT; ++i)
                if (mask & (1u << i)) qs = vadd(qs, P[i]);
            double sv = vmass2(qs);
            double dd = sv - M * M;
            double mg = M * G;
            bw *= (mg * mg) / (dd * dd + mg * mg);
        }
        const double S = f->params[31];
        double x[NP > 0 ? NP : 1];
#pragma unroll
        for (int a = 0; a < NP; ++a) x[a] = s[a] / S;
        double qf = 0;
        const double* Q = f->params + 33;
#pragma unroll
        for (int a = 0; a < NP; ++a) {
            double row = 0;
#pragma unroll
            for (int b = 0; b < NP; ++b) row += Q[a * NP + b] * x[b];
            qf += x[a] * row;
        }
        return bw * qf;
    }
    return 0.0;
}

}  // namespace chili_dev
