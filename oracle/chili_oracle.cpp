// chili_oracle.cpp -- CPU oracle for the Chili hot path.
//
// TEST INFRASTRUCTURE ONLY.  This file restates, statement by statement and
// in the same floating-point operation order, the reference algorithm of
// jxi24/Chili for VEGAS grid sampling/adaptation and the multichannel
// s-/t-channel phase-space maps.  Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg load it, and only as the checker (or the timed
// CPU baseline); the product library never links or calls it.
//
// Build (oracle/Makefile): g++ -std=c++17 -O2 -ffp-contract=off.  With
// -DORACLE_COUNT every FP64 operation is counted (N_arith: + - * / sqrt;
// N_trans: log exp pow sin cos tan atan atan2 sinh acosh) to give the
// algorithmic flop-equivalents per point F = N_arith + 20 N_trans.
//
// Parity pins (which check covers which function) are listed in the header
// of tests/test_oracle.py.  Functions whose result no closed form, invariant
// or brute force fixes are marked "parity unpinned" below.
//
// Reference citations are file:line in the jxi24/Chili tree.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <set>
#include <stack>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "../include/chili_gpu.h"

// ---------------------------------------------------------------------------
// Real type: plain double, or a counting wrapper for the flop tally.
// ---------------------------------------------------------------------------
#ifdef ORACLE_COUNT
static unsigned long long g_arith = 0, g_trans = 0;
struct Real {
    double v;
    Real() : v(0) {}
    Real(double x) : v(x) {}
};
static inline Real operator+(Real a, Real b) { ++g_arith; return Real(a.v + b.v); }
static inline Real operator-(Real a, Real b) { ++g_arith; return Real(a.v - b.v); }
static inline Real operator*(Real a, Real b) { ++g_arith; return Real(a.v * b.v); }
static inline Real operator/(Real a, Real b) { ++g_arith; return Real(a.v / b.v); }
static inline Real operator-(Real a) { return Real(-a.v); }
static inline Real& operator+=(Real& a, Real b) { a = a + b; return a; }
static inline Real& operator-=(Real& a, Real b) { a = a - b; return a; }
static inline Real& operator*=(Real& a, Real b) { a = a * b; return a; }
static inline Real& operator/=(Real& a, Real b) { a = a / b; return a; }
static inline bool operator<(Real a, Real b) { return a.v < b.v; }
static inline bool operator>(Real a, Real b) { return a.v > b.v; }
static inline bool operator<=(Real a, Real b) { return a.v <= b.v; }
static inline bool operator>=(Real a, Real b) { return a.v >= b.v; }
static inline bool operator==(Real a, Real b) { return a.v == b.v; }
static inline bool operator!=(Real a, Real b) { return a.v != b.v; }
static inline double D(Real a) { return a.v; }
static inline Real r_sqrt(Real a) { ++g_arith; return Real(std::sqrt(a.v)); }
static inline Real r_log(Real a) { ++g_trans; return Real(std::log(a.v)); }
static inline Real r_exp(Real a) { ++g_trans; return Real(std::exp(a.v)); }
static inline Real r_pow(Real a, Real b) { ++g_trans; return Real(std::pow(a.v, b.v)); }
static inline Real r_sin(Real a) { ++g_trans; return Real(std::sin(a.v)); }
static inline Real r_cos(Real a) { ++g_trans; return Real(std::cos(a.v)); }
static inline Real r_tan(Real a) { ++g_trans; return Real(std::tan(a.v)); }
static inline Real r_atan(Real a) { ++g_trans; return Real(std::atan(a.v)); }
static inline Real r_atan2(Real a, Real b) { ++g_trans; return Real(std::atan2(a.v, b.v)); }
static inline Real r_sinh(Real a) { ++g_trans; return Real(std::sinh(a.v)); }
static inline bool r_isnan(Real a) { return std::isnan(a.v); }
static inline Real r_abs(Real a) { return Real(std::fabs(a.v)); }
#else
typedef double Real;
static inline double D(double a) { return a; }
static inline double r_sqrt(double a) { return std::sqrt(a); }
static inline double r_log(double a) { return std::log(a); }
static inline double r_exp(double a) { return std::exp(a); }
static inline double r_pow(double a, double b) { return std::pow(a, b); }
static inline double r_sin(double a) { return std::sin(a); }
static inline double r_cos(double a) { return std::cos(a); }
static inline double r_tan(double a) { return std::tan(a); }
static inline double r_atan(double a) { return std::atan(a); }
static inline double r_atan2(double a, double b) { return std::atan2(a, b); }
static inline double r_sinh(double a) { return std::sinh(a); }
static inline bool r_isnan(double a) { return std::isnan(a); }
static inline double r_abs(double a) { return std::fabs(a); }
#endif

// std::max / std::min semantics (NaN in the first argument propagates).
static inline Real r_max(Real a, Real b) { return (a < b) ? b : a; }
static inline Real r_min(Real a, Real b) { return (b < a) ? b : a; }

static const double kPi = M_PI;

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).  Uniform double from two
// 32-bit words: 32 high bits + 21 high bits of the low word, times 2^-53.
// ---------------------------------------------------------------------------
static inline void philox4x32_10(uint32_t ctr[4], uint32_t key0, uint32_t key1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)M0 * ctr[0];
        uint64_t p1 = (uint64_t)M1 * ctr[2];
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ ctr[1] ^ key0;
        uint32_t n1 = lo1;
        uint32_t n2 = hi0 ^ ctr[3] ^ key1;
        uint32_t n3 = lo0;
        ctr[0] = n0; ctr[1] = n1; ctr[2] = n2; ctr[3] = n3;
        key0 += W0; key1 += W1;
    }
}

static inline double philox_uniform(uint64_t seed, uint32_t stream, uint64_t attempt, uint32_t k) {
    uint32_t c[4] = {(uint32_t)attempt, (uint32_t)(attempt >> 32), k >> 1, stream};
    philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    uint32_t hi = (k & 1u) ? c[2] : c[0];
    uint32_t lo = (k & 1u) ? c[3] : c[1];
    uint64_t m = ((uint64_t)hi << 21) | (uint64_t)(lo >> 11);
    return (double)m * 0x1p-53;
}

// ---------------------------------------------------------------------------
// Four-vectors, include/Tools/FourVector.hh + src/Tools/FourVector.cc.
// Storage (E, px, py, pz), metric (+,-,-,-).  pow(x, 2) is x*x.
// ---------------------------------------------------------------------------
struct V4 { Real c[4]; };
static inline V4 mk(Real e, Real x, Real y, Real z) { V4 v; v.c[0] = e; v.c[1] = x; v.c[2] = y; v.c[3] = z; return v; }
static inline V4 vadd(const V4& a, const V4& b) { return mk(a.c[0] + b.c[0], a.c[1] + b.c[1], a.c[2] + b.c[2], a.c[3] + b.c[3]); }   // FourVector.cc:130-137
static inline V4 vsub(const V4& a, const V4& b) { return mk(a.c[0] - b.c[0], a.c[1] - b.c[1], a.c[2] - b.c[2], a.c[3] - b.c[3]); }
static inline V4 vscale(const V4& a, Real s) { return mk(a.c[0] * s, a.c[1] * s, a.c[2] * s, a.c[3] * s); }        // FourVector.cc:148-155
static inline V4 vneg(const V4& a) { return mk(-a.c[0], -a.c[1], -a.c[2], -a.c[3]); }
static inline Real vdot(const V4& a, const V4& b) {                                                                 // FourVector.cc:166-169
    return a.c[0] * b.c[0] - (a.c[1] * b.c[1] + a.c[2] * b.c[2] + a.c[3] * b.c[3]);
}
static inline Real vpt2(const V4& a) { return a.c[1] * a.c[1] + a.c[2] * a.c[2]; }                                  // FourVector.hh:141
static inline Real vpt(const V4& a) { return r_sqrt(vpt2(a)); }
static inline Real vp2(const V4& a) { return vpt2(a) + a.c[3] * a.c[3]; }                                          // FourVector.hh:149
static inline Real vp(const V4& a) { return r_sqrt(vp2(a)); }
static inline Real vmass2(const V4& a) { return vdot(a, a); }
static inline Real vrap(const V4& a) {                                                                              // FourVector.cc:26-29
    if (a.c[0] == 0) return 99;
    return r_log((a.c[0] + a.c[3]) / (a.c[0] - a.c[3])) / 2;
}
static inline Real vphi(const V4& a) {                                                                              // FourVector.cc:20-24
    Real phi = r_atan2(a.c[2], a.c[1]);
    if (phi < 0) return phi + 2 * kPi;
    return phi;
}
// three-vector helpers (src/Tools/ThreeVector.cc)
static inline Real dot3(const V4& a, const V4& b) { return a.c[1] * b.c[1] + a.c[2] * b.c[2] + a.c[3] * b.c[3]; }   // ThreeVector.cc:96-98
static inline V4 cross3(const V4& a, const V4& b) {                                                                 // ThreeVector.cc:17-23
    return mk(Real(0), a.c[2] * b.c[3] - a.c[3] * b.c[2], a.c[3] * b.c[1] - a.c[1] * b.c[3], a.c[1] * b.c[2] - a.c[2] * b.c[1]);
}
static inline Real p3mag(const V4& a) { return r_sqrt(dot3(a, a)); }                                                // ThreeVector.hh:156-160

static Real small_omct(const V4& a, const V4& v) {                                                                  // FourVector.cc:57-65
    Real mag = r_sqrt(vp2(a) * vp2(v));
    Real pq = a.c[1] * v.c[1] + a.c[2] * v.c[2] + a.c[3] * v.c[3];
    Real ct = r_min(r_max(pq / mag, Real(-1.)), Real(1.));
    if (ct < 0.) return 1. - ct;
    Real st = p3mag(cross3(a, v)) / mag;
    Real st2 = st / (2. * r_sqrt((ct + 1) / 2.));
    return 2. * st2 * st2;
}
static inline Real small_mldp(const V4& a, const V4& v) { return a.c[0] * v.c[0] * small_omct(a, v); }              // FourVector.cc:67-69

// ---------------------------------------------------------------------------
// Poincare, src/Tools/Poincare.cc.  Only the boost (type 1) and the transverse
// axis PT() of the rotation (type 2) feed the path.
// ---------------------------------------------------------------------------
struct Boost { V4 l; Real rsq; };
static inline Boost make_boost(const V4& v) { Boost b; b.l = v; b.rsq = r_sqrt(vmass2(v)); return b; }              // Poincare.cc:7-8 (rsq=-1 -> Mass())
static inline void boost_fwd(const Boost& b, V4& v) {                                                                // Poincare.cc:36-42
    Real lv = b.l.c[1] * v.c[1] + b.l.c[2] * v.c[2] + b.l.c[3] * v.c[3];
    Real v0 = (b.l.c[0] * v.c[0] - lv) / b.rsq;
    Real c1 = (v.c[0] + v0) / (b.rsq + b.l.c[0]);
    v = mk(v0, v.c[1] - b.l.c[1] * c1, v.c[2] - b.l.c[2] * c1, v.c[3] - b.l.c[3] * c1);
}
static inline void boost_back(const Boost& b, V4& v) {                                                               // Poincare.cc:44-50
    Real lv = b.l.c[1] * v.c[1] + b.l.c[2] * v.c[2] + b.l.c[3] * v.c[3];
    Real v0 = (b.l.c[0] * v.c[0] + lv) / b.rsq;
    Real c1 = (v.c[0] + v0) / (b.rsq + b.l.c[0]);
    v = mk(v0, v.c[1] + b.l.c[1] * c1, v.c[2] + b.l.c[2] * c1, v.c[3] + b.l.c[3] * c1);
}
// Poincare(v1, v2) rotation: the transverse unit axis m_t (Poincare.cc:10-34).
static V4 rotation_pt(const V4& v1, const V4& v2) {
    Real n2 = vp(v2);
    V4 b = mk(Real(0.), v2.c[1] / n2, v2.c[2] / n2, v2.c[3] / n2);
    Real n1 = vp(v1);
    V4 l = mk(Real(0.), v1.c[1] / n1, v1.c[2] / n1, v1.c[3] / n1);
    Real lb = vdot(l, b);
    V4 t = vadd(b, vscale(l, lb));
    Real mt = vp2(t);
    if (mt != 0.) t = vscale(t, 1. / r_sqrt(mt));
    int ix[3] = {1, 2, 3};
    Real ml[4] = {Real(0.), r_abs(l.c[1]), r_abs(l.c[2]), r_abs(l.c[3])};
    if (ml[ix[2]] > ml[ix[1]]) std::swap(ix[1], ix[2]);
    if (ml[ix[1]] > ml[ix[0]]) std::swap(ix[0], ix[1]);
    if (ml[ix[2]] > ml[ix[1]]) std::swap(ix[1], ix[2]);
    Real tdp = t.c[ix[1]] * l.c[ix[1]] + t.c[ix[2]] * l.c[ix[2]];
    if (tdp != 0.) t.c[ix[0]] = -tdp / l.c[ix[0]];
    if (vp2(t) == 0.) t.c[ix[1]] = 1.;
    return t;
}

// ---------------------------------------------------------------------------
// Channel elements, src/Tools/ChannelElements.cc
// ---------------------------------------------------------------------------
static V4 lt4(const V4& a, const V4& b, const V4& c) {                                                              // ChannelElements.cc:9-19
    Real t = a.c[1] * b.c[2] * c.c[3] + a.c[2] * b.c[3] * c.c[1] + a.c[3] * b.c[1] * c.c[2]
           - a.c[1] * b.c[3] * c.c[2] - a.c[3] * b.c[2] * c.c[1] - a.c[2] * b.c[1] * c.c[3];
    Real x = -a.c[0] * b.c[2] * c.c[3] - a.c[2] * b.c[3] * c.c[0] - a.c[3] * b.c[0] * c.c[2]
           + a.c[0] * b.c[3] * c.c[2] + a.c[3] * b.c[2] * c.c[0] + a.c[2] * b.c[0] * c.c[3];
    Real y = -a.c[1] * b.c[0] * c.c[3] - a.c[0] * b.c[3] * c.c[1] - a.c[3] * b.c[1] * c.c[0]
           + a.c[1] * b.c[3] * c.c[0] + a.c[3] * b.c[0] * c.c[1] + a.c[0] * b.c[1] * c.c[3];
    Real z = -a.c[1] * b.c[2] * c.c[0] - a.c[2] * b.c[0] * c.c[1] - a.c[0] * b.c[1] * c.c[2]
           + a.c[1] * b.c[0] * c.c[2] + a.c[0] * b.c[2] * c.c[1] + a.c[2] * b.c[1] * c.c[0];
    return mk(t, -x, -y, -z);
}
static Real sqlam(Real s, Real s1, Real s2) {                                                                       // ChannelElements.cc:21-25
    Real d = s - s1 - s2;
    Real arg = d * d - 4. * s1 * s2;
    if (arg > 0.) return r_sqrt(arg) / s;
    return 0.;
}
static Real peaked_dist(Real a, Real cn, Real cxm, Real cxp, int k, Real ran) {                                     // ChannelElements.cc:27-32
    Real ce = 1. - cn;
    Real kk = (double)k;
    if (ce != 0.) return kk * (r_pow(ran * r_pow(a + kk * cxp, ce) + (1. - ran) * r_pow(a + kk * cxm, ce), 1 / ce) - a);
    return kk * ((a + kk * cxm) * r_pow((a + kk * cxp) / (a + kk * cxm), ran) - a);
}
static Real peaked_weight(Real a, Real cn, Real cxm, Real cxp, Real res, int k, Real& ran) {                         // ChannelElements.cc:34-50
    Real ce = 1. - cn, w;
    Real kk = (double)k;
    if (ce != 0.) {
        Real amin = r_pow(a + kk * cxm, ce);
        w = r_pow(a + kk * cxp, ce) - amin;
        ran = (r_pow(a + kk * res, ce) - amin) / w;
        w /= kk * ce;
    } else {
        Real amin = a + kk * cxm;
        w = r_log((a + kk * cxp) / amin);
        ran = r_log((a + kk * res) / amin) / w;
        w /= kk;
    }
    return w;
}
static Real massless_prop_weight(Real sexp, Real smin, Real smax, Real s, Real& ran) {                              // ChannelElements.cc:52-56
    return peaked_weight(Real(0.), sexp, smin, smax, s, 1, ran) / r_pow(s, -sexp);
}
static Real massless_prop_momenta(Real sexp, Real smin, Real smax, Real ran) {                                      // ChannelElements.cc:58-62
    return peaked_dist(Real(0.), sexp, smin, smax, 1, ran);
}
static Real massive_prop_weight(Real m, Real g, Real smin, Real smax, Real s, Real& ran) {                           // ChannelElements.cc:64-75
    Real m2 = m * m, mw = m * g;
    Real ymax = r_atan((smin - m2) / mw), ymin = r_atan((smax - m2) / mw);
    Real y = r_atan((s - m2) / mw);
    ran = (y - ymin) / (ymax - ymin);
    Real w = mw / ((s - m2) * (s - m2) + mw * mw);
    w = (ymin - ymax) / w;
    return w;
}
static Real massive_prop_momenta(Real m, Real g, Real smin, Real smax, Real ran) {                                   // ChannelElements.cc:77-85
    Real m2 = m * m, mw = m * g;
    Real ymax = r_atan((smin - m2) / mw), ymin = r_atan((smax - m2) / mw);
    return m2 + mw * r_tan(ymin + ran * (ymax - ymin));
}
// Isotropic two-body decay, ChannelElements.cc:87-104 (ctmin=-1, ctmax=1, xref=(1,1,0,0)).
static void schannel_momenta(V4 p, Real s1, Real s2, V4& p1, V4& p2, Real ran1, Real ran2) {
    const Real ctmin = -1.0, ctmax = 1.0;
    Real s = vmass2(p), rs = r_sqrt(r_abs(s));
    Real e1 = (s + s1 - s2) / rs / 2., m1 = r_sqrt(e1 * e1 - s1);
    Real ct = ctmin + (ctmax - ctmin) * ran1, st = r_sqrt(1. - ct * ct);
    Real phi = 2. * kPi * ran2;
    V4 xref = mk(Real(1.), Real(1.), Real(0.), Real(0.));
    V4 pl = (vp2(p) != 0.) ? p : mk(Real(1.), Real(0.), Real(0.), Real(1.));
    Boost cms = make_boost(p);
    V4 n_perp = rotation_pt(pl, xref);
    V4 l_perp = lt4(pl, xref, n_perp);
    l_perp = vscale(l_perp, 1.0 / r_sqrt(r_abs(vmass2(l_perp))));
    Real mct = m1 * ct, plp = vp(pl);
    p1 = mk(e1, pl.c[1] * mct / plp, pl.c[2] * mct / plp, pl.c[3] * mct / plp);
    V4 dir = vadd(vscale(l_perp, r_cos(phi)), vscale(n_perp, r_sin(phi)));
    p1 = vadd(p1, vscale(dir, m1 * st));
    boost_back(cms, p1);
    p2 = vsub(p, p1);
}
static Real schannel_weight(const V4& p1, const V4& p2, Real& ran1, Real& ran2) {                                   // ChannelElements.cc:106-135
    const Real ctmin = -1.0, ctmax = 1.0;
    V4 p = vadd(p1, p2), p1h = p1;
    V4 xref = mk(Real(1.), Real(1.), Real(0.), Real(0.));
    V4 pl = (vp2(p) != 0.) ? p : mk(Real(1.), Real(0.), Real(0.), Real(1.));
    Boost cms = make_boost(p);
    V4 n_perp = rotation_pt(pl, xref);
    V4 l_perp = lt4(pl, xref, n_perp);
    l_perp = vscale(l_perp, 1.0 / r_sqrt(r_abs(vmass2(l_perp))));
    boost_fwd(cms, p1h);
    Real ct = dot3(p1h, pl) / r_sqrt(vp2(p1h) * vp2(pl));
    Real cp = vdot(vneg(l_perp), p1), sp = vdot(vneg(n_perp), p1), norm = r_sqrt(cp * cp + sp * sp);
    cp /= norm;
    sp /= norm;
    ran1 = (ct - ctmin) / (ctmax - ctmin);
    ran2 = r_atan2(sp, cp) / (2. * kPi);
    if (ran2 < 0.) ran2 += 1.;
    Real w = (ctmax - ctmin) / 2.;
    Real s1 = r_max(vmass2(p1), Real(0.));
    Real s2 = r_max(vmass2(p2), Real(0.));
    Real s = vmass2(p);
    if (s1 == 0. && s2 == 0.) s = small_mldp(p1, p2);
    w *= sqlam(s, s1, s2) / (8. * kPi);
    return w;
}

// ---------------------------------------------------------------------------
// AdaptiveMap, src/Integrator/AdaptiveMap.cc
// ---------------------------------------------------------------------------
struct Grid {
    size_t dims = 0, bins = 0;
    std::vector<double> h;   // dims*(bins+1) edges
    double lower(size_t d, size_t b) const { return h[d * (bins + 1) + b]; }
    double upper(size_t d, size_t b) const { return h[d * (bins + 1) + b + 1]; }
    double width(size_t d, size_t b) const { return upper(d, b) - lower(d, b); }
};
static Grid grid_uniform(size_t dims, size_t bins) {                                                                 // AdaptiveMap.hh:29-40
    Grid g; g.dims = dims; g.bins = bins; g.h.resize(dims * (bins + 1));
    for (size_t i = 0; i < bins + 1; ++i) g.h[i] = (double)i / (double)bins;
    for (size_t d = 1; d < dims; ++d)
        std::copy(g.h.begin(), g.h.begin() + (long)(bins + 1), g.h.begin() + (long)(d * (bins + 1)));
    return g;
}
// lower_bound over e_0..e_{B-1}, minus one (AdaptiveMap.cc:34-38).  The
// reference underflows to SIZE_MAX for x <= e_0 or NaN and then indexes out
// of bounds; this restatement (and the device code) clamps that case to bin 0.
static size_t grid_find_bin(const Grid& g, size_t d, double x) {
    const double* e = g.h.data() + d * (g.bins + 1);
    size_t k = (size_t)(std::lower_bound(e, e + g.bins, x) - e);
    return k == 0 ? 0 : k - 1;
}
static Real grid_forward(const Grid& g, Real* r) {                                                                   // AdaptiveMap.cc:40-56
    Real jac = 1.0;
    for (size_t i = 0; i < g.dims; ++i) {
        Real pos = r[i] * (double)g.bins;
        size_t idx = std::min((size_t)D(pos), g.bins - 1);   // r = 1 (only from a learned proposal) stays in the last bin
        Real loc = pos - (double)idx;
        double left = g.lower(i, idx);
        double size = g.width(i, idx);
        r[i] = left + loc * size;
        jac *= size * (double)g.bins;
    }
    return jac;
}
static Real grid_weight(const Grid& g, const Real* x) {                                                              // AdaptiveMap.cc:58-65
    Real jac = 1.0;
    for (size_t i = 0; i < g.dims; ++i) {
        size_t idx = grid_find_bin(g, i, D(x[i]));
        jac *= g.width(i, idx) * (double)g.bins;
    }
    return jac;
}
static void grid_adapt(Grid& g, double alpha, const double* data) {                                                  // AdaptiveMap.cc:67-125
    const size_t B = g.bins;
    std::vector<double> tmp(B);
    std::vector<double> nh(g.h.size());
    for (size_t i = 0; i < g.dims; ++i) {
        tmp.assign(data + i * B, data + (i + 1) * B);
        double previous = tmp[0], current = tmp[1];
        tmp[0] = (previous + current) / 2;
        double norm = tmp[0];
        for (size_t bin = 1; bin < B - 1; ++bin) {
            const double sum = previous + current;
            previous = current;
            current = tmp[bin + 1];
            tmp[bin] = (sum + current) / 3;
            norm += tmp[bin];
        }
        tmp.back() = (previous + current) / 2;
        norm += tmp.back();
        if (norm == 0) continue;
        double avg = 0;
        for (size_t bin = 0; bin < B; ++bin) {
            if (tmp[bin] != 0) {
                const double r = tmp[bin] / norm;
                const double fac = std::pow((r - 1.0) / std::log(r), alpha);
                avg += fac;
                tmp[bin] = fac;
            }
        }
        avg /= (double)B;
        double cbin = 0;
        size_t ibin = 0;
        for (size_t nbin = 1; nbin < B; ++nbin) {
            // The reference scan (AdaptiveMap.cc:112) has no bound; roundoff can
            // carry it past B.  Here it stops at B so no read leaves the row.
            for (; cbin < avg && ibin < B; ++ibin) cbin += tmp[ibin];
            const double prev = g.lower(i, ibin - 1);
            const double curr = g.lower(i, ibin);
            cbin -= avg;
            const double delta = (curr - prev) * cbin;
            nh[i * (B + 1) + nbin] = curr - delta / tmp[ibin - 1];
        }
        nh[i * (B + 1) + B] = 1.0;
    }
    g.h = nh;
}
static void grid_split(Grid& g, int nsplit) {                                                                        // AdaptiveMap.cc:127-160
    std::vector<double> h(g.dims * (nsplit * g.bins + 1));
    for (size_t d = 0; d < g.dims; ++d) {
        for (size_t bin = 0; bin < g.bins; ++bin) {
            const double curr = g.lower(d, bin);
            const double size = g.width(d, bin);
            for (int i = 0; i < nsplit; ++i) {
                const size_t idx = d * (nsplit * g.bins + 1) + bin * nsplit + i;
                h[idx] = curr + (double)i * size / (double)nsplit;
            }
        }
        h[(d + 1) * (nsplit * g.bins) + d] = 1.0;
    }
    g.bins = nsplit * g.bins;
    g.h.swap(h);
}

// ---------------------------------------------------------------------------
// StatsData, include/Integrator/Statistics.hh:62-142
// ---------------------------------------------------------------------------
struct Stats {
    double n = 0, mn = std::numeric_limits<double>::max(), mx = std::numeric_limits<double>::min(), sum = 0, sum2 = 0, nf = 0;
    void add(double x) { n++; sum += x; sum2 += x * x; mn = std::min(mn, x); mx = std::max(mx, x); if (x != 0) nf++; }
    void merge(const Stats& o) { n += o.n; nf += o.nf; sum += o.sum; sum2 += o.sum2; mn = std::min(mn, o.mn); mx = std::max(mx, o.mx); }
    double mean() const { return sum / n; }
    double var() const { return (sum2 / n - mean() * mean()) / (n - 1); }
    double err() const { return std::sqrt(var()); }
    void out(chili_stats* s) const { s->n = n; s->min = mn; s->max = mx; s->sum = sum; s->sum2 = sum2; s->n_finite = nf; }
};

// ---------------------------------------------------------------------------
// ParticleInfo / ChannelDescription ordering, src/Channel/ChannelNode.cc:293-316
// ---------------------------------------------------------------------------
struct PInfo { unsigned idx = 0; int pid = 0; double mass = 0, width = 0; };
static bool operator<(const PInfo& a, const PInfo& b) {
    if (a.idx == b.idx) {
        if (a.mass == b.mass) return a.width < b.width;
        return a.mass < b.mass;
    }
    return a.idx < b.idx;
}
static bool operator==(const PInfo& a, const PInfo& b) { return a.mass == b.mass && a.idx == b.idx && a.width == b.width; }
static bool operator!=(const PInfo& a, const PInfo& b) { return !(a == b); }
static bool operator>(const PInfo& a, const PInfo& b) { return b < a && b != a; }
struct PGreater { bool operator()(const PInfo& a, const PInfo& b) const { return a > b; } };

struct ChanDesc {
    std::vector<PInfo> info;
    std::map<PInfo, std::pair<PInfo, PInfo>> decays;
};
static bool operator<(const ChanDesc& a, const ChanDesc& b) {
    if (a.info.size() == b.info.size()) {
        for (size_t i = 0; i < a.info.size(); ++i) {
            if (a.info[i] < b.info[i]) return true;
            else if (a.info[i] > b.info[i]) return false;
        }
        return false;
    }
    return a.info.size() < b.info.size();
}

// Cuts (include/Tools/Cuts.hh:13-29)
struct CutsR {
    std::vector<double> sexternal;
    std::map<std::pair<int, int>, double> deltaR;
    std::unordered_map<unsigned, double> smin, ptmin, etamax;
};
static double smax_of(double sqrts, const CutsR& c, unsigned idx) {                                                  // Cuts.hh:21-29
    double r = sqrts;
    for (unsigned i = 0; i < c.sexternal.size(); ++i)
        if (!(idx & (1u << (i + 2)))) r -= std::sqrt(c.sexternal[i]);
    return r * r;
}

// Model (include/Model/Model.hh) backed by the chili_process tables.
struct ModelR {
    std::unordered_map<unsigned, double> mass, width;
    std::vector<std::array<int, 3>> vtx;
    double Mass(int f) const { return mass.at((unsigned)std::abs(f)); }
    double Width(int f) const { return width.at((unsigned)std::abs(f)); }
    std::vector<int> Combinable(int a, int b) const {
        std::vector<int> r;
        for (auto& v : vtx) if (v[0] == a && v[1] == b) r.push_back(v[2]);
        return r;
    }
};

// next bit permutation with the same population count (lexicographic order)
static inline unsigned next_perm(unsigned v) {
    unsigned t = v | (v - 1);
    return (t + 1) | (((~t & (0u - ~t)) - 1) >> (__builtin_ctz(v) + 1));
}
static inline unsigned popcnt(unsigned v) { return (unsigned)__builtin_popcount(v); }

// ConstructChannels, src/Channel/ChannelNode.cc:98-291 (restated).
static std::vector<ChanDesc> construct_channels(const std::vector<int>& flavs, const ModelR& model, CutsR& cuts, size_t smax) {
    std::unordered_map<unsigned, std::set<PInfo>> comps;
    std::map<PInfo, std::set<std::pair<PInfo, PInfo>>, PGreater> chain;
    const unsigned nf = (unsigned)flavs.size();
    for (unsigned i = 0; i < nf; ++i) {
        PInfo p; p.idx = 1u << i; p.pid = flavs[i]; p.mass = model.Mass(p.pid); p.width = model.Width(p.pid);
        comps[p.idx].insert(p);
    }
    for (unsigned nset = 2; nset < nf; ++nset) {
        unsigned cur = (1u << nset) - 1;
        while (cur < (1u << nf) && popcnt(cur) - 1 <= smax) {
            std::vector<unsigned> bits;
            for (unsigned i = 0; i < nf; ++i) if (cur & (1u << i)) bits.push_back(cur & (1u << i));
            for (unsigned iset = 1; iset < nset; ++iset) {
                unsigned sel = (1u << iset) - 1;
                while (sel < (1u << (nset - 1))) {
                    unsigned s1 = 0;
                    for (unsigned i = 0; i < bits.size(); ++i) s1 += bits[i] * ((sel >> i) & 1u);
                    unsigned s2 = cur ^ s1;
                    if ((s1 & 1u) || (s2 & 1u)) break;
                    if ((s1 & 2u) || (s2 & 2u)) break;
                    if (comps.find(s1) == comps.end() || comps.find(s2) == comps.end()) { sel = next_perm(sel); continue; }
                    std::set<int> combined;
                    for (const auto& c1 : comps[s1])
                        for (const auto& c2 : comps[s2]) {
                            auto t = model.Combinable(c1.pid, c2.pid);
                            combined.insert(t.begin(), t.end());
                        }
                    if (combined.empty()) { sel = next_perm(sel); continue; }
                    for (int e : combined) {
                        PInfo p; p.pid = e; p.idx = cur; p.mass = model.Mass(e); p.width = model.Width(e);
                        comps[cur].insert(p);
                    }
                    cuts.ptmin[cur] = 0;
                    cuts.etamax[cur] = 99;
                    auto lp = comps[s1], rp = comps[s2];
                    for (const auto& li : lp)
                        for (const auto& ri : rp) {
                            auto allowed = model.Combinable(li.pid, ri.pid);
                            for (int e : allowed) {
                                PInfo p; p.pid = e; p.idx = cur; p.mass = model.Mass(e); p.width = model.Width(e);
                                if (li < ri) chain[p].insert({li, ri});
                                else chain[p].insert({ri, li});
                            }
                        }
                    sel = next_perm(sel);
                }
            }
            cur = next_perm(cur);
        }
    }
    std::vector<unsigned> avail;
    for (const auto& kv : comps) avail.push_back(kv.first);
    struct Frame { std::set<int> pid; std::vector<unsigned> avail, currents; unsigned sum = 0; size_t schans = 0; };
    std::stack<Frame> st;
    const unsigned max_id = (1u << nf) - 1;
    Frame f0; f0.pid.insert(comps[1].begin()->pid); f0.avail = avail;
    st.push(f0);
    std::set<ChanDesc> chans;
    while (!st.empty()) {
        Frame top = st.top(); st.pop();
        if (top.sum > max_id) continue;
        if (top.sum == max_id) {
            std::vector<ChanDesc> old(1);
            for (unsigned c : top.currents) {
                if (c == 1 || c == 2) continue;
                std::vector<PInfo> infos;
                for (const auto& ci : comps[c]) { PInfo p{c, ci.pid, model.Mass(ci.pid), model.Width(ci.pid)}; infos.push_back(p); }
                std::vector<ChanDesc> nw;
                for (const auto& ch : old)
                    for (const auto& p : infos) {
                        ChanDesc c2 = ch;
                        c2.info.push_back(p);
                        std::sort(c2.info.begin(), c2.info.end(), PGreater());
                        nw.push_back(c2);
                    }
                old = nw;
            }
            chans.insert(old.begin(), old.end());
            continue;
        }
        if (top.avail.empty()) continue;
        for (size_t i = 0; i < top.avail.size(); ++i) {
            Frame d = top;
            d.schans += popcnt(top.avail[i]) - 1;
            if (d.schans > smax) continue;
            if (d.sum & top.avail[i]) continue;
            std::set<int> combined;
            for (int p1 : d.pid)
                for (const auto& ci : comps[top.avail[i]]) {
                    auto t = model.Combinable(p1, ci.pid);
                    combined.insert(t.begin(), t.end());
                }
            if (combined.empty()) continue;
            d.pid = combined;
            d.sum += top.avail[i];
            d.currents.push_back(top.avail[i]);
            d.avail.erase(d.avail.begin() + (long)i);
            st.push(d);
        }
    }
    for (const auto& dk : chain) {
        std::set<ChanDesc> nw;
        for (const auto& ch : chans) {
            bool add = std::binary_search(ch.info.begin(), ch.info.end(), dk.first, PGreater());
            for (const auto& pr : ch.decays) add |= (pr.second.first == dk.first) || (pr.second.second == dk.first);
            if (add) {
                for (const auto& prods : dk.second) {
                    ChanDesc c2 = ch;
                    c2.decays[dk.first] = prods;
                    nw.insert(c2);
                }
            } else {
                nw.insert(ch);
            }
        }
        chans = nw;
    }
    return std::vector<ChanDesc>(chans.begin(), chans.end());
}

// ---------------------------------------------------------------------------
// FSMapper, include/Channel/Channel.hh + src/Channel/Channel.cc (restated with
// the same sparse-momentum maps keyed by current bitmask).
// ---------------------------------------------------------------------------
struct Mapper {
    double sqrts = 0, ptmax = 0;
    size_t ntot = 0, nout = 0;
    ChanDesc ch;
    CutsR cuts;
    std::vector<double> ptmin;        // m_ptmin per info entry
    size_t iran = 0;
    std::map<unsigned, V4> tmom;      // SparseMom
    std::map<unsigned, Real> tmass;   // SparseMass
    size_t ndims() const { return 3 * nout - 4 + 2; }                                                               // Channel.hh:33

    void init(double s, size_t nt, const ChanDesc& c, const CutsR& cu) {                                            // Channel.hh:18-30
        sqrts = s; ptmax = s / 2.0; ntot = nt; nout = nt - 2; ch = c; cuts = cu;
        for (size_t i = 0; i < ch.info.size(); ++i) {
            double v = 1;
            if (cuts.ptmin.at(ch.info[i].idx) != 0) v = cuts.ptmin.at(ch.info[i].idx);
            ptmin.push_back(v);
        }
    }
    double dR(int a, int b) const {
        auto it = cuts.deltaR.find({a, b});
        return it == cuts.deltaR.end() ? 0 : it->second;
    }
    void gen_decays(const Real* r) {                                                                                // Channel.cc:63-81
        for (const auto& dk : ch.decays) {
            const auto& p1 = dk.second.first;
            const auto& p2 = dk.second.second;
            Real a = r_sqrt(tmass[p1.idx]) + r_sqrt(tmass[p2.idx]);
            Real smin1 = a * a;
            double dr = dR(p1.pid, p2.pid);
            double smin2 = cuts.ptmin.at(p1.idx) * cuts.ptmin.at(p2.idx) * 2 / kPi / kPi * (dr * dr);
            Real smin = r_max(r_max(smin1, Real(smin2)), Real(cuts.smin[p1.idx | p2.idx]));
            double smax = smax_of(sqrts, cuts, dk.first.idx);
            if (std::abs(dk.first.mass) < 1e-16 || dk.first.width < 1e-16)
                tmass[dk.first.idx] = massless_prop_momenta(0.99, smin, smax, r[iran++]);
            else
                tmass[dk.first.idx] = massive_prop_momenta(dk.first.mass, dk.first.width, smin, smax, r[iran++]);
        }
    }
    Real wgt_decays(Real* r) {                                                                                      // Channel.cc:83-113
        Real w = 1;
        for (const auto& dk : ch.decays) {
            const auto& p1 = dk.second.first;
            const auto& p2 = dk.second.second;
            Real a = r_sqrt(r_abs(vmass2(tmom.at(p1.idx)))) + r_sqrt(r_abs(vmass2(tmom.at(p2.idx))));
            Real smin1 = a * a;
            double dr = dR(p1.pid, p2.pid);
            double smin2 = cuts.ptmin.at(p1.idx) * cuts.ptmin.at(p2.idx) * 2 / kPi / kPi * (dr * dr);
            Real smin = r_max(r_max(smin1, Real(smin2)), Real(cuts.smin[p1.idx | p2.idx]));
            double smax = smax_of(sqrts, cuts, dk.first.idx);
            if (std::abs(dk.first.mass) < 1e-16 || dk.first.width < 1e-16)
                w *= massless_prop_weight(0.99, smin, smax, vmass2(tmom.at(dk.first.idx)), r[iran++]);
            else
                w *= massive_prop_weight(dk.first.mass, dk.first.width, smin, smax, vmass2(tmom.at(dk.first.idx)), r[iran++]);
            w /= (2 * kPi);
        }
        return w;
    }
    void gen_tchan(const Real* r) {                                                                                 // Channel.cc:115-158
        V4 psum = mk(0., 0., 0., 0.);
        for (size_t i = 0; i + 1 < ch.info.size(); ++i) {
            const Real ran = r[iran++];
            Real pt = 2 * ptmin[i] * ptmax * ran / (2 * ptmin[i] + ptmax * (1 - ran));
            Real etamax = sqrts / 2 / pt;
            etamax = r_log(etamax + r_sqrt(etamax * etamax - 1));
            etamax = r_min(etamax, Real(cuts.etamax.at(ch.info[i].idx)));
            Real y = etamax * (2 * r[iran++] - 1);
            Real sinhy = r_sinh(y);
            Real coshy = r_sqrt(1 + sinhy * sinhy);
            Real phi = 2 * kPi * r[iran++];
            Real mt = r_sqrt(pt * pt + tmass.at(ch.info[i].idx));
            tmom[ch.info[i].idx] = mk(mt * coshy, pt * r_cos(phi), pt * r_sin(phi), mt * sinhy);
            psum = vadd(psum, tmom[ch.info[i].idx]);
        }
        Real mjets2 = vmass2(psum);
        Real yjets = vrap(psum);
        Real ptj2 = vpt2(psum);
        Real m2 = tmass.at(ch.info.back().idx);
        Real qt = r_sqrt(m2 + ptj2);
        Real mt = r_sqrt(mjets2 + ptj2);
        Real ymin = -r_log(sqrts / qt * (1 - mt / sqrts * r_exp(-yjets)));
        Real ymax = r_log(sqrts / qt * (1 - mt / sqrts * r_exp(yjets)));
        Real dely = ymax - ymin;
        Real yv = ymin + r[iran++] * dely;
        Real sinhy = r_sinh(yv);
        Real coshy = r_sqrt(1 + sinhy * sinhy);
        tmom[ch.info.back().idx] = mk(qt * coshy, -psum.c[1], -psum.c[2], qt * sinhy);
        V4 tot = vadd(tmom[ch.info.back().idx], psum);
        Real pp = tot.c[0] + tot.c[3];
        Real pm = tot.c[0] - tot.c[3];
        tmom[1] = mk(-pp / 2, 0., 0., -pp / 2);
        tmom[2] = mk(-pm / 2, 0., 0., pm / 2);
    }
    Real wgt_tchan(Real* r) {                                                                                       // Channel.cc:160-202
        Real w = 2 * kPi;
        V4 psum = mk(0., 0., 0., 0.);
        for (size_t i = 0; i + 1 < ch.info.size(); ++i) {
            w *= 1.0 / (16 * std::pow(kPi, 3));
            const V4& m = tmom.at(ch.info[i].idx);
            Real pt = vpt(m);
            Real tp = 2 * ptmin[i] + pt;
            w *= pt * ptmax / 2 / ptmin[i] / (2 * ptmin[i] + ptmax) * (tp * tp);
            r[iran++] = pt * (ptmax + 2 * ptmin[i]) / (ptmax * (pt + 2 * ptmin[i]));
            Real etamax = sqrts / 2 / pt;
            etamax = r_log(etamax + r_sqrt(etamax * etamax - 1));
            etamax = r_isnan(etamax) ? Real(99) : etamax;
            etamax = r_min(etamax, Real(cuts.etamax.at(ch.info[i].idx)));
            w *= 2 * etamax * 2 * kPi;
            Real y = vrap(m);
            Real phi = vphi(m);
            r[iran++] = (y / etamax + 1) / 2;
            r[iran++] = phi / 2 / kPi;
            psum = vadd(psum, m);
        }
        Real mjets2 = vmass2(psum);
        Real yjets = vrap(psum);
        Real ptj2 = vpt2(psum);
        Real m2 = vmass2(tmom.at(ch.info.back().idx));
        Real qt = r_sqrt(m2 + ptj2);
        Real mt = r_sqrt(mjets2 + ptj2);
        Real yv = vrap(tmom.at(ch.info.back().idx));
        Real ymin = -r_log(sqrts / qt * (1 - mt / sqrts * r_exp(-yjets)));
        Real ymax = r_log(sqrts / qt * (1 - mt / sqrts * r_exp(yjets)));
        Real dely = ymax - ymin;
        if (r_isnan(ymax) || r_isnan(ymin)) return 0;
        w *= dely / sqrts / sqrts;
        r[iran++] = (yv - ymin) / dely;
        return w;
    }
    void decay_parts(const PInfo& ab, const PInfo& a, const PInfo& b, const Real* r) {                             // Channel.cc:204-220
        Real r1 = r[iran++];
        Real r2 = r[iran++];
        V4 pa = mk(0., 0., 0., 0.), pb = mk(0., 0., 0., 0.);
        schannel_momenta(tmom[ab.idx], tmass.at(a.idx), tmass.at(b.idx), pa, pb, r1, r2);
        tmom[a.idx] = pa;
        tmom[b.idx] = pb;
        if (popcnt(a.idx) != 1) { auto d = ch.decays.at(a); decay_parts(a, d.first, d.second, r); }
        if (popcnt(b.idx) != 1) { auto d = ch.decays.at(b); decay_parts(b, d.first, d.second, r); }
    }
    Real wgt_decay_parts(const PInfo& a, const PInfo& b, Real* r) {                                                // Channel.cc:222-236
        Real r1, r2;
        Real w = schannel_weight(tmom.at(a.idx), tmom.at(b.idx), r1, r2);
        r[iran++] = r1;
        r[iran++] = r2;
        if (popcnt(a.idx) != 1) { auto d = ch.decays.at(a); w *= wgt_decay_parts(d.first, d.second, r); }
        if (popcnt(b.idx) != 1) { auto d = ch.decays.at(b); w *= wgt_decay_parts(d.first, d.second, r); }
        return w;
    }
    void gen_schan(const Real* r) {                                                                                 // Channel.cc:238-245
        for (size_t i = 0; i < ch.info.size(); ++i)
            if (popcnt(ch.info[i].idx) != 1) { auto d = ch.decays.at(ch.info[i]); decay_parts(ch.info[i], d.first, d.second, r); }
    }
    Real wgt_schan(Real* r) {                                                                                       // Channel.cc:247-256
        Real w = 1;
        for (size_t i = 0; i < ch.info.size(); ++i)
            if (popcnt(ch.info[i].idx) != 1) { auto d = ch.decays.at(ch.info[i]); w *= wgt_decay_parts(d.first, d.second, r); }
        return w;
    }
    void generate_point(std::vector<V4>& mom, const Real* r) {                                                      // Channel.cc:22-38
        mom.resize(ntot);
        iran = 0;
        for (size_t i = 0; i < nout; ++i) tmass[1u << (i + 2)] = cuts.sexternal[i];
        gen_decays(r);
        gen_tchan(r);
        gen_schan(r);
        for (size_t i = 0; i < mom.size(); ++i) mom[i] = tmom[1u << i];
    }
    Real generate_weight(const std::vector<V4>& mom, std::vector<Real>& r) {                                        // Channel.cc:40-61
        r.resize(ndims());
        iran = 0;
        for (size_t i = 0; i < mom.size(); ++i) tmom[1u << i] = mom[i];
        for (const auto& dk : ch.decays) tmom[dk.first.idx] = vadd(tmom[dk.second.first.idx], tmom[dk.second.second.idx]);
        tmom[3] = vadd(tmom[1], tmom[2]);
        Real wd = wgt_decays(r.data());
        Real wt = wgt_tchan(r.data());
        Real ws = wgt_schan(r.data());
        return wd * wt * ws;
    }
};

// ---------------------------------------------------------------------------
// Device-style integrands and the PreProcess cut (this project's definitions;
// the reference takes them as user callbacks, Integrand.hh:41-47).
// ---------------------------------------------------------------------------
// JetCluster::operator() (src/Tools/JetCluster.cc:8-78), the reference's loop
// with its position list and distance table: inclusive kT with
// Q2i = pT^2, Q2ij = min(pT_i^2, pT_j^2) R2_ij, R2 = (dy^2 + dphi^2) / R / R
// (phi in [0, 2 pi), unwrapped; JetCluster.hh, JetCluster.cc:80-84).  The
// number of jets it returns for the final state p[2..].
static size_t jet_cluster_count(const std::vector<V4>& parts, double R) {
    std::vector<V4> tmp(parts.begin() + 2, parts.end());
    size_t njets = 0;
    size_t n = tmp.size();
    auto q2i = [&](const V4& a) { return D(vpt2(a)); };
    auto q2ij = [&](const V4& a, const V4& b) {
        const double dy = D(vrap(a)) - D(vrap(b)), dphi = D(vphi(a)) - D(vphi(b));
        return std::min(q2i(a), q2i(b)) * ((dy * dy + dphi * dphi) / R / R);
    };
    std::vector<size_t> pos(n);                       // remaining partons, in order
    for (size_t i = 0; i < n; ++i) pos[i] = i;
    std::vector<std::vector<double>> dist(n, std::vector<double>(n));
    size_t ii = 0, jj = 0;
    double dmin = std::numeric_limits<double>::max();
    for (size_t i = 0; i < n; ++i) {
        dist[i][i] = q2i(tmp[i]);
        if (dist[i][i] < dmin) { dmin = dist[i][i]; ii = jj = i; }
        for (size_t j = 0; j < i; ++j) {
            dist[i][j] = q2ij(tmp[i], tmp[j]);
            if (dist[i][j] < dmin) { dmin = dist[i][j]; ii = i; jj = j; }
        }
    }
    while (true) {
        if (ii != jj) tmp[pos[jj]] = vadd(tmp[pos[jj]], tmp[pos[ii]]);       // merge i into j
        if (n == 0) break;
        --n;
        if (ii == jj) ++njets;                                               // beam distance: a jet
        pos.erase(pos.begin() + (long)ii);
        // distances of the parton now at position jj (when jj is past the
        // end the reference recomputes entries of the removed parton, which
        // no later scan reads)
        if (jj < n) {
            const size_t jx = pos[jj];
            dist[jx][jx] = q2i(tmp[jx]);
            for (size_t j = 0; j < n; ++j) {
                const size_t o = pos[j];
                if (o < jx) dist[jx][o] = q2ij(tmp[jx], tmp[o]);
                else if (o > jx) dist[o][jx] = q2ij(tmp[jx], tmp[o]);
            }
        }
        ii = jj = 0;
        if (n == 0) continue;
        dmin = dist[pos[0]][pos[0]];
        for (size_t i = 0; i < n; ++i) {
            const size_t ix = pos[i];
            if (dist[ix][ix] < dmin) { dmin = dist[ix][ix]; ii = jj = i; }
            for (size_t j = 0; j < i; ++j) {
                const double dij = dist[ix][pos[j]];
                if (dij < dmin) { dmin = dij; ii = i; jj = j; }
            }
        }
    }
    return njets;
}

static bool cut_pass(const chili_integrand& f, const std::vector<V4>& p, double sqrts) {
    if (f.cuts == CHILI_CUTS_NONE) return true;
    if (r_isnan(p[0].c[0])) return false;
    // physical incoming momenta: x1, x2 <= 1 (incoming legs carry E = -x sqrt(s)/2)
    if (!(D(-2 * p[0].c[0]) <= sqrts) || !(D(-2 * p[1].c[0]) <= sqrts)) return false;
    if (f.cuts == CHILI_CUTS_KT_JETS) {
        // the reference example's PreProcess, src/example.cc:24-37
        if (f.ptmin > 0 && D(vmass2(vadd(p[0], p[1]))) > f.ptmin * f.ptmin) return false;
        return jet_cluster_count(p, f.drmin) >= p.size() - 2;
    }
    const size_t n = p.size();
    for (size_t i = 2; i < n; ++i) {
        if (!(D(vpt(p[i])) > f.ptmin)) return false;
        if (!(std::fabs(D(vrap(p[i]))) < f.etamax)) return false;
    }
    for (size_t i = 2; i < n; ++i)
        for (size_t j = i + 1; j < n; ++j) {
            double dy = D(vrap(p[i])) - D(vrap(p[j]));
            double dphi = std::fabs(D(vphi(p[i])) - D(vphi(p[j])));
            if (dphi > kPi) dphi = 2 * kPi - dphi;
            double dr = std::sqrt(dy * dy + dphi * dphi);
            if (!(dr > f.drmin)) return false;
        }
    return true;
}
static Real pair_s(const std::vector<V4>& p, size_t i, size_t j) { return vmass2(vadd(p[i], p[j])); }
static Real mask_s(const std::vector<V4>& p, unsigned mask) {
    V4 q = mk(0., 0., 0., 0.);
    for (size_t i = 0; i < p.size(); ++i) if (mask & (1u << i)) q = vadd(q, p[i]);
    return vmass2(q);
}
// f for phase-space integrands (momenta in, incoming legs first).
static Real integrand_ps(const chili_integrand& f, const std::vector<V4>& p) {
    const size_t n = p.size();
    switch (f.kind) {
    case CHILI_F_UNIT: return 1.0;
    case CHILI_F_PROPPROD: {
        // s0^2 / (s_ij s_kl) summed over unordered pairs of distinct final-state
        // pairs (ij) < (kl); with a single pair, s0 / s_ij.
        const double s0 = f.params[0];
        std::vector<Real> s;
        for (size_t i = 2; i < n; ++i)
            for (size_t j = i + 1; j < n; ++j) s.push_back(pair_s(p, i, j));
        if (s.size() == 1) return s0 / s[0];
        Real tot = 0;
        for (size_t a = 0; a < s.size(); ++a)
            for (size_t b = a + 1; b < s.size(); ++b) tot += s0 * s0 / (s[a] * s[b]);
        return tot;
    }
    case CHILI_F_RESONANT: {
        // params: [0] nbw, then nbw x (mask, M, Gamma), [31] S (scale), [32] npair,
        // [33..] Q (npair x npair, row major) over x_k = s_ij / S in pair order.
        const int nbw = (int)f.params[0];
        Real bw = 1.0;
        for (int k = 0; k < nbw; ++k) {
            unsigned mask = (unsigned)f.params[1 + 3 * k];
            double M = f.params[2 + 3 * k], G = f.params[3 + 3 * k];
            Real s = mask_s(p, mask);
            Real d = s - M * M;
            double mg = M * G;
            bw *= (mg * mg) / (d * d + mg * mg);
        }
        const double S = f.params[31];
        const int np = (int)f.params[32];
        std::vector<Real> x;
        for (size_t i = 2; i < n; ++i)
            for (size_t j = i + 1; j < n; ++j) x.push_back(pair_s(p, i, j) / S);
        Real q = 0;
        for (int a = 0; a < np; ++a) {
            Real row = 0;
            for (int b = 0; b < np; ++b) row += f.params[33 + a * np + b] * x[b];
            q += x[a] * row;
        }
        return bw * q;
    }
    default: return 0.0;
    }
}
// e+e- -> mu+mu- through gamma/Z at tree level, on [0,1]^2 (plain VEGAS).
// params: [0] MZ [1] GZ [2] sin^2 thetaW [3] alpha [4] smin [5] smax.
// x0 -> s-hat by the Breit-Wigner map (ChannelElements.cc:64-85), x1 -> cos theta.
static Real integrand_eemumu(const chili_integrand& f, const Real* x) {
    const double MZ = f.params[0], GZ = f.params[1], sw2 = f.params[2], al = f.params[3], smin = f.params[4], smax = f.params[5];
    Real s = massive_prop_momenta(MZ, GZ, smin, smax, x[0]);
    Real dummy;
    Real wbw = massive_prop_weight(MZ, GZ, smin, smax, s, dummy);
    Real ct = 2. * x[1] - 1.;
    const double cw2 = 1. - sw2;
    const double kap = 1. / (4. * sw2 * cw2);
    const double ve = -0.5 + 2. * sw2, ae = -0.5;
    Real dm = s - MZ * MZ;
    Real den = dm * dm + MZ * MZ * GZ * GZ;
    Real rechi = kap * s * dm / den;
    Real abschi2 = kap * kap * s * s / den;
    Real A0 = 1. + 2. * ve * ve * rechi + (ae * ae + ve * ve) * (ae * ae + ve * ve) * abschi2;
    Real A1 = 4. * ae * ae * rechi + 8. * ve * ae * ve * ae * abschi2;
    const double gev2pb = 0.3893793721e9;
    Real dsig = kPi * al * al / (2. * s) * (A0 * (1. + ct * ct) + A1 * ct) * gev2pb;
    return wbw * 2. * dsig;
}

// ---------------------------------------------------------------------------
// Oracle process: channels + grids + alpha.
// ---------------------------------------------------------------------------
struct OProc {
    double sqrts = 0;
    size_t ntot = 0;
    std::vector<Mapper> maps;
    std::vector<Grid> grids;
    std::vector<double> alpha;
    chili_integrand f{};
    std::vector<Real> xr;            // shared inverse-slot vector (Integrand.hh:115)
};

static void build_cuts(const chili_process& P, const ModelR& model, CutsR& cuts) {
    std::vector<int> fl(P.flavs, P.flavs + P.nflav);
    if (P.use_default_cuts) {                                                                                         // ChannelNode.cc:69-95
        for (int a : fl)
            for (int b : fl) cuts.deltaR[{a, b}] = (std::abs(a) == 6 || std::abs(b) == 6) ? 0. : 0.4;
        for (size_t i = 0; i < fl.size(); ++i) {
            unsigned idx = 1u << i;
            double m = model.Mass(fl[i]);
            cuts.smin[idx] = m * m;
            cuts.ptmin[idx] = 30;
            cuts.etamax[idx] = 99;
            if (i > 1) cuts.sexternal.push_back(m * m);
        }
    } else {
        for (int k = 0; k < P.ndeltaR; ++k) cuts.deltaR[{P.deltaR_pid[k][0], P.deltaR_pid[k][1]}] = P.deltaR[k];
        for (size_t i = 0; i < fl.size(); ++i) {
            unsigned idx = 1u << i;
            double m = model.Mass(fl[i]);
            cuts.smin[idx] = P.smin[i];
            cuts.ptmin[idx] = P.ptmin[i];
            cuts.etamax[idx] = P.etamax[i];
            if (i > 1) cuts.sexternal.push_back(m * m);
        }
    }
}

static double pick_uniform(const chili_uniforms* U, uint64_t a, uint32_t k) {
    if (U->u) return U->u[(int64_t)k * U->stride + (int64_t)a];
    return philox_uniform(U->seed, U->stream, a, k);
}

// Channel pick: std::discrete_distribution on alpha (random.tcc:2656-2677),
// lower_bound on the normalised cumulative sum with the last entry set to 1.
static size_t pick_channel(const std::vector<double>& w, double u) {
    if (w.size() < 2) return 0;
    double sum = 0.0;
    for (double x : w) sum += x;
    std::vector<double> cp(w.size());
    double acc = 0;
    for (size_t i = 0; i < w.size(); ++i) { double p = w[i] / sum; acc = (i == 0) ? p : acc + p; cp[i] = acc; }
    cp.back() = 1.0;
    return (size_t)(std::lower_bound(cp.begin(), cp.end(), u) - cp.begin());
}

extern "C" {

int oracle_pick_channel(int n, const double* alpha, double u) {
    return (int)pick_channel(std::vector<double>(alpha, alpha + n), u);
}
double oracle_stats(int n, const double* x, chili_stats* out) {
    Stats s;
    for (int i = 0; i < n; ++i) s.add(x[i]);
    s.out(out);
    return s.err();
}

double oracle_philox_uniform(uint64_t seed, uint32_t stream, uint64_t attempt, uint32_t k) {
    return philox_uniform(seed, stream, attempt, k);
}
void oracle_philox_block(uint32_t* ctr4, uint32_t k0, uint32_t k1) { philox4x32_10(ctr4, k0, k1); }

void oracle_count_reset(void) {
#ifdef ORACLE_COUNT
    g_arith = 0; g_trans = 0;
#endif
}
void oracle_count_get(unsigned long long* arith, unsigned long long* trans) {
#ifdef ORACLE_COUNT
    *arith = g_arith; *trans = g_trans;
#else
    *arith = 0; *trans = 0;
#endif
}

void* oracle_create(const chili_process* P, int nbins, const chili_integrand* f, int* nchannels, int* ndims) {
    try {
        ModelR model;
        for (int i = 0; i < P->nmodel; ++i) {
            model.mass[(unsigned)std::abs(P->model_pid[i])] = P->model_mass[i];
            model.width[(unsigned)std::abs(P->model_pid[i])] = P->model_width[i];
        }
        for (int i = 0; i < P->nvertex; ++i) model.vtx.push_back({P->vertex[i][0], P->vertex[i][1], P->vertex[i][2]});
        CutsR cuts;
        build_cuts(*P, model, cuts);
        std::vector<int> fl(P->flavs, P->flavs + P->nflav);
        auto descs = construct_channels(fl, model, cuts, (size_t)P->smax);
        OProc* o = new OProc();
        o->sqrts = P->sqrts;
        o->ntot = fl.size();
        o->maps.resize(descs.size());
        for (size_t i = 0; i < descs.size(); ++i) o->maps[i].init(P->sqrts, fl.size(), descs[i], cuts);
        size_t d = descs.empty() ? 0 : o->maps[0].ndims();
        for (size_t i = 0; i < descs.size(); ++i) o->grids.push_back(grid_uniform(d, (size_t)nbins));
        o->alpha.assign(descs.size(), descs.empty() ? 0.0 : 1.0 / (double)descs.size());
        if (f) o->f = *f;
        *nchannels = (int)descs.size();
        *ndims = (int)d;
        return o;
    } catch (...) {
        return nullptr;
    }
}
void oracle_destroy(void* h) { delete (OProc*)h; }

int oracle_channel_desc(void* h, int i, chili_channel_desc* out) {
    OProc* o = (OProc*)h;
    if (i < 0 || (size_t)i >= o->maps.size()) return -1;
    std::memset(out, 0, sizeof(*out));
    const ChanDesc& c = o->maps[i].ch;
    out->ninfo = (int)c.info.size();
    for (size_t k = 0; k < c.info.size(); ++k) out->info[k] = {c.info[k].idx, c.info[k].pid, c.info[k].mass, c.info[k].width};
    out->ndecay = (int)c.decays.size();
    int k = 0;
    for (const auto& dk : c.decays) {
        out->decay[k][0] = {dk.first.idx, dk.first.pid, dk.first.mass, dk.first.width};
        out->decay[k][1] = {dk.second.first.idx, dk.second.first.pid, dk.second.first.mass, dk.second.first.width};
        out->decay[k][2] = {dk.second.second.idx, dk.second.second.pid, dk.second.second.mass, dk.second.second.width};
        ++k;
    }
    return 0;
}

int oracle_set_grids(void* h, const double* grids, int nbins) {
    OProc* o = (OProc*)h;
    for (size_t c = 0; c < o->grids.size(); ++c) {
        Grid& g = o->grids[c];
        g.bins = (size_t)nbins;
        g.h.assign(grids + c * g.dims * (nbins + 1), grids + (c + 1) * g.dims * (nbins + 1));
    }
    return 0;
}
int oracle_get_nbins(void* h) { return (int)((OProc*)h)->grids[0].bins; }
int oracle_get_grids(void* h, double* grids) {
    OProc* o = (OProc*)h;
    size_t off = 0;
    for (auto& g : o->grids) { std::copy(g.h.begin(), g.h.end(), grids + off); off += g.h.size(); }
    return 0;
}
int oracle_set_alpha(void* h, const double* a) { OProc* o = (OProc*)h; o->alpha.assign(a, a + o->alpha.size()); return 0; }
int oracle_get_alpha(void* h, double* a) { OProc* o = (OProc*)h; std::copy(o->alpha.begin(), o->alpha.end(), a); return 0; }

// Mapper forward / inverse for one channel (Channel.cc:22-61).
int oracle_generate_point(void* h, int ch, const double* x, double* mom) {
    OProc* o = (OProc*)h;
    Mapper& m = o->maps[ch];
    std::vector<Real> xr(x, x + m.ndims());
    std::vector<V4> p;
    m.generate_point(p, xr.data());
    for (size_t i = 0; i < p.size(); ++i) for (int k = 0; k < 4; ++k) mom[4 * i + k] = D(p[i].c[k]);
    return 0;
}
double oracle_generate_weight(void* h, int ch, const double* mom, double* x) {
    OProc* o = (OProc*)h;
    Mapper& m = o->maps[ch];
    std::vector<V4> p(o->ntot);
    for (size_t i = 0; i < o->ntot; ++i) p[i] = mk(mom[4 * i], mom[4 * i + 1], mom[4 * i + 2], mom[4 * i + 3]);
    std::vector<Real> xr(m.ndims(), Real(0.));
    Real w = m.generate_weight(p, xr);
    for (size_t i = 0; i < xr.size(); ++i) x[i] = D(xr[i]);
    return D(w);
}

// One attempt of MultiChannel::operator() (MultiChannel.hh:214-247) without
// the loop control: everything the per-point records hold.
struct Attempt {
    size_t pick = 0;
    int status = 0;          // 0 cut-rejected, 1 accepted, 2 NaN value
    double w = 0, f = 0, val = 0;
    std::vector<V4> mom;
    std::vector<Real> xpick;
    std::vector<double> ps;
};
// mom_in (optional, ntot x 4): use these momenta instead of the forward map's,
// so that the weight stage can be compared on identical inputs.
static void run_attempt(OProc* o, const chili_uniforms* U, uint64_t a, Attempt& A, size_t d, const double* mom_in = nullptr) {
    const size_t C = o->maps.size();
    std::vector<Real> r(d);
    for (size_t k = 0; k < d; ++k) r[k] = pick_uniform(U, a, (uint32_t)k);
    double uch = pick_uniform(U, a, (uint32_t)d);
    A.pick = pick_channel(o->alpha, uch);                                                                          // MultiChannel.hh:147
    if (mom_in) {
        A.mom.resize(o->ntot);
        for (size_t i = 0; i < o->ntot; ++i) A.mom[i] = mk(mom_in[4 * i], mom_in[4 * i + 1], mom_in[4 * i + 2], mom_in[4 * i + 3]);
    } else {
        grid_forward(o->grids[A.pick], r.data());                                                                  // Integrand.hh:108 (jacobian dropped)
        o->maps[A.pick].generate_point(A.mom, r.data());
    }
    if (!cut_pass(o->f, A.mom, o->sqrts)) { A.status = 0; return; }                                                // MultiChannel.hh:219-223
    // Integrand::GenerateWeight, Integrand.hh:112-123 (shared _rans vector)
    std::vector<Real> xr(o->ntot, Real(0.));
    A.ps.assign(C, 0.0);
    Real weight = 0;
    for (size_t i = 0; i < C; ++i) {
        Real psi = o->maps[i].generate_weight(A.mom, xr);
        A.ps[i] = D(psi);
        if (i == A.pick) A.xpick = xr;
        Real vw = grid_weight(o->grids[i], xr.data());
        weight += o->alpha[i] / psi / vw;
    }
    Real W = 1.0 / weight;
    Real fv = integrand_ps(o->f, A.mom);
    Real val = (W == 0) ? Real(0) : fv * W;
    A.w = D(W); A.f = D(fv); A.val = D(val);
    A.status = r_isnan(val) ? 2 : 1;
    if (A.status == 2) { A.val = 0; A.w = 0; }                                                                      // MultiChannel.hh:230-236
}

// Per-attempt records for attempts [a0, a0+n): pick, status, W, f, val,
// momenta (ntot*4), x of the picked channel's inverse (d), its FindBin bins
// (d) and the phase-space weights of every channel (C).
// mom_in (optional, n x ntot x 4) replaces the forward map's momenta.
int oracle_mc_points2(void* h, const chili_uniforms* U, uint64_t a0, int64_t n, const double* mom_in,
                      int32_t* pick, int32_t* status, double* w, double* f, double* val,
                      double* mom, double* xpick, int32_t* bins, double* ps) {
    OProc* o = (OProc*)h;
    const size_t C = o->maps.size(), d = o->maps[0].ndims(), nt = o->ntot;
    for (int64_t i = 0; i < n; ++i) {
        Attempt A;
        run_attempt(o, U, a0 + (uint64_t)i, A, d, mom_in ? mom_in + (size_t)i * nt * 4 : nullptr);
        if (pick) pick[i] = (int32_t)A.pick;
        if (status) status[i] = A.status;
        if (w) w[i] = A.w;
        if (f) f[i] = A.f;
        if (val) val[i] = A.val;
        if (mom) for (size_t k = 0; k < nt; ++k) for (int c = 0; c < 4; ++c) mom[(i * nt + k) * 4 + c] = D(A.mom[k].c[c]);
        if (A.status != 0) {
            if (xpick) for (size_t k = 0; k < d; ++k) xpick[i * d + k] = D(A.xpick[k]);
            if (bins) for (size_t k = 0; k < d; ++k) bins[i * d + k] = (int32_t)grid_find_bin(o->grids[A.pick], k, D(A.xpick[k]));
            if (ps) for (size_t k = 0; k < C; ++k) ps[i * C + k] = A.ps[k];
        } else {
            if (xpick) for (size_t k = 0; k < d; ++k) xpick[i * d + k] = 0;
            if (bins) for (size_t k = 0; k < d; ++k) bins[i * d + k] = -1;
            if (ps) for (size_t k = 0; k < C; ++k) ps[i * C + k] = 0;
        }
    }
    return 0;
}

// Flop tally of attempts [a0, a0+n) with the current state (meaningful in
// the ORACLE_COUNT build): out = {arith, trans} of the forward phase (uniforms
// -> VEGAS map -> GeneratePoint -> cut, every attempt), {arith, trans} of the
// weight phase (integrand + every channel's inverse map and VEGAS density +
// alpha-weighted total, accepted attempts), and the accepted count.
int oracle_count_points(void* h, const chili_uniforms* U, uint64_t a0, int64_t n, double* out) {
    OProc* o = (OProc*)h;
    const size_t C = o->maps.size(), d = o->maps[0].ndims();
    double fa = 0, ft = 0, wa = 0, wt = 0, nacc = 0;
    for (int64_t i = 0; i < n; ++i) {
        unsigned long long a0c = 0, t0c = 0, a1c = 0, t1c = 0, a2c = 0, t2c = 0;
        oracle_count_get(&a0c, &t0c);
        std::vector<Real> r(d);
        for (size_t k = 0; k < d; ++k) r[k] = pick_uniform(U, a0 + (uint64_t)i, (uint32_t)k);
        size_t pk = pick_channel(o->alpha, pick_uniform(U, a0 + (uint64_t)i, (uint32_t)d));
        grid_forward(o->grids[pk], r.data());
        std::vector<V4> mom;
        o->maps[pk].generate_point(mom, r.data());
        bool ok = cut_pass(o->f, mom, o->sqrts);
        oracle_count_get(&a1c, &t1c);
        fa += (double)(a1c - a0c); ft += (double)(t1c - t0c);
        if (!ok) continue;
        std::vector<Real> xr(o->ntot, Real(0.));
        Real weight = 0;
        for (size_t j = 0; j < C; ++j) {
            Real psj = o->maps[j].generate_weight(mom, xr);
            Real vw = grid_weight(o->grids[j], xr.data());
            weight += o->alpha[j] / psj / vw;
        }
        Real W = 1.0 / weight;
        Real fv = integrand_ps(o->f, mom);
        Real val = fv * W;
        (void)val;
        oracle_count_get(&a2c, &t2c);
        wa += (double)(a2c - a1c); wt += (double)(t2c - t1c);
        nacc += 1;
    }
    out[0] = fa; out[1] = ft; out[2] = wa; out[3] = wt; out[4] = nacc;
    return 0;
}

int oracle_mc_points(void* h, const chili_uniforms* U, uint64_t a0, int64_t n,
                     int32_t* pick, int32_t* status, double* w, double* f, double* val,
                     double* mom, double* xpick, int32_t* bins, double* ps) {
    return oracle_mc_points2(h, U, a0, n, nullptr, pick, status, w, f, val, mom, xpick, bins, ps);
}

// One iteration of MultiChannel::operator() (MultiChannel.hh:196-251) with
// should_optimize == train: accumulates StatsData, the alpha training sums
// T[C] (MultiChannel.hh:167-180) and every channel's VEGAS f^2 histogram
// hist[c][dim][bin] (Integrand.hh:89-96).  Attempts start at a0; the number
// used is returned in *n_attempts.
int oracle_mc_iterate2(void* h, const chili_uniforms* U, uint64_t a0, int64_t ncalls, int train,
                       const double* mom_in, int64_t n_in,
                       chili_stats* stats, double* T, double* hist, int64_t* n_attempts) {
    OProc* o = (OProc*)h;
    const size_t C = o->maps.size(), d = o->maps[0].ndims(), B = o->grids[0].bins;
    Stats res;
    std::vector<double> t(C, 0.0), hh(C * d * B, 0.0);
    uint64_t a = a0;
    for (int64_t i = 0; i < ncalls; ++i) {
        if (U->u && (int64_t)(a - a0) >= U->stride) return -2;
        if (mom_in && (int64_t)(a - a0) >= n_in) return -2;
        Attempt A;
        run_attempt(o, U, a, A, d, mom_in ? mom_in + (size_t)(a - a0) * o->ntot * 4 : nullptr);
        ++a;
        if (A.status == 0) { --i; res.add(0); continue; }
        if (A.status == 2) --i;
        if (train) {
            double v2 = A.val * A.val;
            for (size_t k = 0; k < d; ++k)
                hh[(A.pick * d + k) * B + grid_find_bin(o->grids[A.pick], k, D(A.xpick[k]))] += v2;
            for (size_t j = 0; j < C; ++j) t[j] += v2 * A.w / A.ps[j];
        }
        res.add(A.val);
    }
    res.out(stats);
    if (T) std::copy(t.begin(), t.end(), T);
    if (hist) std::copy(hh.begin(), hh.end(), hist);
    *n_attempts = (int64_t)(a - a0);
    return 0;
}

// MultiChannel::operator() without training and with an unweighting
// PostProcess (MultiChannel.hh:240-245): a point of value v passing the cut
// is kept when the attempt's uniform number d+1 is below |v| / wmax, else it
// adds a zero and is retried.  This build's definition for a NaN value: 0,
// so one rejection (the reference would decrement its counter twice).
// events [nevents][ntot][4], ew [nevents] = sign(v) max(1, |v| / wmax).
int oracle_mc_generate2(void* h, const chili_uniforms* U, uint64_t a0, int64_t nevents, double wmax, const double* mom_in,
                        int64_t n_in, chili_stats* stats, int64_t* n_attempts, double* events, double* ew) {
    OProc* o = (OProc*)h;
    const size_t d = o->maps[0].ndims();
    if (U->u) return -3;
    Stats res;
    uint64_t a = a0;
    for (int64_t i = 0; i < nevents; ++i) {
        if (mom_in && (int64_t)(a - a0) >= n_in) return -2;
        Attempt A;
        run_attempt(o, U, a, A, d, mom_in ? mom_in + (size_t)(a - a0) * o->ntot * 4 : nullptr);
        const double u = pick_uniform(U, a, (uint32_t)(d + 1));
        ++a;
        if (A.status == 0) { --i; res.add(0); continue; }
        const double v = A.val;
        if (!(u < std::fabs(v) / wmax)) { --i; res.add(0); continue; }
        res.add(v);
        if (events)
            for (size_t p = 0; p < o->ntot; ++p)
                for (int q = 0; q < 4; ++q) events[((size_t)i * o->ntot + p) * 4 + q] = D(A.mom[p].c[q]);
        if (ew) ew[i] = (std::fabs(v) > wmax) ? v / wmax : std::copysign(1.0, v);
    }
    res.out(stats);
    *n_attempts = (int64_t)(a - a0);
    return 0;
}

// Jets JetCluster returns for momenta [n][4] (incoming first).
int oracle_jet_count(const double* mom, int n, double R) {
    std::vector<V4> p(n);
    for (int i = 0; i < n; ++i) p[i] = mk(mom[4 * i], mom[4 * i + 1], mom[4 * i + 2], mom[4 * i + 3]);
    return (int)jet_cluster_count(p, R);
}

int oracle_mc_iterate(void* h, const chili_uniforms* U, uint64_t a0, int64_t ncalls, int train,
                      chili_stats* stats, double* T, double* hist, int64_t* n_attempts) {
    return oracle_mc_iterate2(h, U, a0, ncalls, train, nullptr, 0, stats, T, hist, n_attempts);
}

// MultiChannel::Adapt, src/Channel/MultiChannel.cc:109-136
void oracle_alpha_adapt(int n, double* alpha, const double* train, double beta, double min_alpha) {
    std::vector<double> nw(n);
    double sum = 0;
    for (int i = 0; i < n; ++i) { nw[i] = alpha[i] * std::pow(train[i], beta); sum += nw[i]; }
    double ns = 0;
    for (auto& w : nw) { if (w == 0) continue; w /= sum; w = std::max(w, min_alpha); ns += w; }
    for (auto& w : nw) w /= ns;
    std::copy(nw.begin(), nw.end(), alpha);
}
// MultiChannel::MaxDifference, src/Channel/MultiChannel.cc:138-154 (the max part)
double oracle_max_difference(int n, const double* train) {
    double mx = 0;
    for (int i = 0; i + 1 < n; ++i)
        for (int j = i + 1; j < n; ++j) mx = std::max(mx, std::abs(train[i] - train[j]));
    return mx;
}
// AdaptiveMap::Adapt / Split on a single grid (dims x (bins+1) edges).
void oracle_grid_adapt(int dims, int bins, double* edges, double alpha, const double* data) {
    Grid g; g.dims = (size_t)dims; g.bins = (size_t)bins; g.h.assign(edges, edges + dims * (bins + 1));
    grid_adapt(g, alpha, data);
    std::copy(g.h.begin(), g.h.end(), edges);
}
void oracle_grid_split(int dims, int bins, const double* edges, int nsplit, double* out) {
    Grid g; g.dims = (size_t)dims; g.bins = (size_t)bins; g.h.assign(edges, edges + dims * (bins + 1));
    grid_split(g, nsplit);
    std::copy(g.h.begin(), g.h.end(), out);
}
int oracle_grid_find_bin(int dims, int bins, const double* edges, int dim, double x) {
    Grid g; g.dims = (size_t)dims; g.bins = (size_t)bins; g.h.assign(edges, edges + dims * (bins + 1));
    return (int)grid_find_bin(g, (size_t)dim, x);
}
double oracle_grid_forward(int dims, int bins, const double* edges, double* r) {
    Grid g; g.dims = (size_t)dims; g.bins = (size_t)bins; g.h.assign(edges, edges + dims * (bins + 1));
    std::vector<Real> rr(r, r + dims);
    Real j = grid_forward(g, rr.data());
    for (int i = 0; i < dims; ++i) r[i] = D(rr[i]);
    return D(j);
}

// AdaptiveMap::GenerateWeight (AdaptiveMap.cc:58-65) of x, and the
// hypercube point r that the forward map (AdaptiveMap.cc:40-56) sends to x:
// r_i = (k + (x_i - e_k) / w_k) / B with k = FindBin(x_i).
double oracle_grid_inverse(int dims, int bins, const double* edges, const double* x, double* r) {
    Grid g; g.dims = (size_t)dims; g.bins = (size_t)bins; g.h.assign(edges, edges + dims * (bins + 1));
    double jac = 1.0;
    for (size_t i = 0; i < g.dims; ++i) {
        const size_t k = grid_find_bin(g, i, x[i]);
        const double w = g.width(i, k);
        jac *= w * (double)g.bins;
        double ri = (w > 0) ? ((double)k + (x[i] - g.lower(i, k)) / w) / (double)g.bins : (double)k / (double)g.bins;
        r[i] = std::min(std::max(ri, 0.0), 1.0);
    }
    return jac;
}

// Plain VEGAS iteration, Vegas::operator() (src/Integrator/Vegas.cc:18-52),
// without the final Adapt: returns the StatsData and the f^2 histogram.
int oracle_vegas_iterate(int dims, int bins, const double* edges, const chili_integrand* f,
                         const chili_uniforms* U, uint64_t a0, int64_t ncalls,
                         chili_stats* stats, double* hist, double* vals) {
    Grid g; g.dims = (size_t)dims; g.bins = (size_t)bins; g.h.assign(edges, edges + dims * (bins + 1));
    Stats res;
    std::vector<double> hh((size_t)dims * bins, 0.0);
    std::vector<Real> r((size_t)dims);
    for (int64_t i = 0; i < ncalls; ++i) {
        for (int k = 0; k < dims; ++k) r[k] = pick_uniform(U, a0 + (uint64_t)i, (uint32_t)k);
        Real wgt = grid_forward(g, r.data());
        Real val = integrand_eemumu(*f, r.data()) * wgt;
        double v2 = D(val) * D(val);
        res.add(D(val));
        if (vals) vals[i] = D(val);
        for (int k = 0; k < dims; ++k) hh[(size_t)k * bins + grid_find_bin(g, (size_t)k, D(r[k]))] += v2;
    }
    res.out(stats);
    if (hist) std::copy(hh.begin(), hh.end(), hist);
    return 0;
}

// ---------------------------------------------------------------------------
// Optimisation loops.  MultiChannel::Optimize (MultiChannel.hh:253-268) and
// its parts: operator() with training (:196-251), Train (:182-194 ->
// MultiChannel::Adapt MultiChannel.cc:109-136, Integrand::Train
// Integrand.hh:97-104, MaxDifference MultiChannel.cc:138-154), UpdateResults
// (MultiChannel.cc:69-77), RefineChannels (MultiChannel.hh:122-130 ->
// Vegas::Refine Vegas.cc:79-82 -> AdaptiveMap::Split).  Iteration k draws
// the uniform stream U->stream + k from attempt 0 (the reference draws one
// unseeded stream); mom_in replaces the forward map of that iteration.
// ---------------------------------------------------------------------------
static Stats stats_in(const chili_stats& c) {
    Stats s; s.n = c.n; s.mn = c.min; s.mx = c.max; s.sum = c.sum; s.sum2 = c.sum2; s.nf = c.n_finite; return s;
}
int oracle_mc_optimize_step(void* h, const chili_uniforms* U, const double* mom_in, int64_t n_in, chili_mc_params* p,
                            chili_mc_summary* s, double* best, chili_stats* st_out) {
    OProc* o = (OProc*)h;
    const size_t C = o->maps.size(), d = o->maps[0].ndims();
    chili_uniforms u = *U;
    u.stream = U->stream + (uint32_t)s->nresults;
    std::vector<double> T(C), hist;
    const size_t B = o->grids[0].bins;
    hist.assign(C * d * B, 0.0);
    chili_stats st{};
    int64_t na = 0;
    const bool optimize = p->should_optimize != 0;
    int rc = oracle_mc_iterate2(h, &u, 0, p->ncalls, optimize ? 1 : 0, mom_in, n_in, &st, T.data(), hist.data(), &na);
    if (rc != 0) return rc;
    if (optimize) {                                                                                    // Train
        oracle_alpha_adapt((int)C, o->alpha.data(), T.data(), p->beta, p->min_alpha);
        for (size_t c = 0; c < C; ++c) {
            const double* hc = hist.data() + c * d * o->grids[c].bins;
            bool zero = true;
            for (size_t i = 0; i < d * o->grids[c].bins; ++i) zero = zero && hc[i] == 0;
            if (zero) continue;
            grid_adapt(o->grids[c], p->vegas_alpha, hc);
        }
        double mx = 0;                                                                                 // MaxDifference
        for (size_t i = 0; i + 1 < C; ++i)
            for (size_t j = i + 1; j < C; ++j) mx = std::max(mx, std::abs(T[i] - T[j]));
        if (mx < s->min_diff) { std::copy(o->alpha.begin(), o->alpha.end(), best); s->min_diff = mx; }
    }
    Stats sum = stats_in(s->sum);                                                                      // UpdateResults
    sum.merge(stats_in(st));
    sum.out(&s->sum);
    s->nresults += 1;
    if (st_out) *st_out = st;
    if (++p->iteration == p->nrefine) {                                                                // RefineChannels
        p->iteration = 0;
        p->ncalls = (int64_t)((double)p->ncalls * p->refine_size);
        for (auto& g : o->grids)
            if ((int64_t)g.bins < p->max_bins) grid_split(g, 2);
    }
    return 0;
}
int oracle_mc_optimize(void* h, const chili_uniforms* U, chili_mc_params* p, chili_mc_summary* s, double* best,
                       chili_stats* results, int64_t cap, int64_t max_steps) {
    double rel = std::numeric_limits<double>::max();
    int64_t steps = 0;
    while (rel > p->rtol && s->nresults < p->niterations && (max_steps <= 0 || steps < max_steps)) {
        chili_stats st{};
        const int64_t k = s->nresults;
        int rc = oracle_mc_optimize_step(h, U, nullptr, 0, p, s, best, &st);
        if (rc != 0) return rc;
        if (results && k < cap) results[k] = st;
        Stats sum = stats_in(s->sum);
        rel = sum.err() / std::abs(sum.mean());
        ++steps;
    }
    return 0;
}
// Vegas::Optimize (Vegas.cc:54-77) with operator()'s Adapt (:49) and Refine
// (:79-82).  edges holds dims x (bins+1) and has room for edges_cap doubles.
int oracle_vegas_optimize(int dims, int* bins, double* edges, int64_t edges_cap, const chili_integrand* f, const chili_uniforms* U,
                          chili_vegas_params* p, chili_stats* sum, int64_t* nres, chili_stats* results, int64_t cap,
                          int64_t max_steps) {
    double abs_err = std::numeric_limits<double>::max(), rel_err = std::numeric_limits<double>::max();
    int64_t steps = 0;
    while (((abs_err > p->atol && rel_err > p->rtol) || *nres < p->niterations) && steps < max_steps) {
        Grid g; g.dims = (size_t)dims; g.bins = (size_t)*bins; g.h.assign(edges, edges + dims * (*bins + 1));
        chili_uniforms u = *U;
        u.stream = U->stream + (uint32_t)*nres;
        chili_stats st{};
        std::vector<double> hist((size_t)dims * *bins);
        oracle_vegas_iterate(dims, *bins, edges, f, &u, 0, p->ncalls, &st, hist.data(), nullptr);
        grid_adapt(g, p->alpha, hist.data());
        Stats acc = stats_in(*sum);
        acc.merge(stats_in(st));
        acc.out(sum);
        if (results && *nres < cap) results[*nres] = st;
        *nres += 1;
        abs_err = acc.err();
        rel_err = abs_err / std::abs(acc.mean());
        ++steps;
        if (++p->irefine == p->nrefine) {
            grid_split(g, 2);
            p->ncalls *= 2;
            p->irefine = 0;
        }
        if ((int64_t)g.h.size() > edges_cap) return -2;
        std::copy(g.h.begin(), g.h.end(), edges);
        *bins = (int)g.bins;
    }
    return 0;
}

// Phase-space integrand and cut on explicit momenta (for tests).
double oracle_integrand(const chili_integrand* f, int ntot, const double* mom, double sqrts, int* pass) {
    std::vector<V4> p((size_t)ntot);
    for (int i = 0; i < ntot; ++i) p[i] = mk(mom[4 * i], mom[4 * i + 1], mom[4 * i + 2], mom[4 * i + 3]);
    if (pass) *pass = cut_pass(*f, p, sqrts) ? 1 : 0;
    return D(integrand_ps(*f, p));
}
double oracle_eemumu(const chili_integrand* f, const double* x) {
    Real xr[2] = {x[0], x[1]};
    return D(integrand_eemumu(*f, xr));
}

// Channel-element kernels exposed for closed-form tests.
double oracle_sqlam(double s, double s1, double s2) { return D(sqlam(s, s1, s2)); }
double oracle_massless_prop_momenta(double sexp, double smin, double smax, double ran) { return D(massless_prop_momenta(sexp, smin, smax, ran)); }
double oracle_massless_prop_weight(double sexp, double smin, double smax, double s, double* ran) { Real r; Real w = massless_prop_weight(sexp, smin, smax, s, r); *ran = D(r); return D(w); }
double oracle_massive_prop_momenta(double m, double g, double smin, double smax, double ran) { return D(massive_prop_momenta(m, g, smin, smax, ran)); }
double oracle_massive_prop_weight(double m, double g, double smin, double smax, double s, double* ran) { Real r; Real w = massive_prop_weight(m, g, smin, smax, s, r); *ran = D(r); return D(w); }
void oracle_schannel_momenta(const double* p, double s1, double s2, double r1, double r2, double* p1, double* p2) {
    V4 a, b;
    schannel_momenta(mk(p[0], p[1], p[2], p[3]), s1, s2, a, b, r1, r2);
    for (int k = 0; k < 4; ++k) { p1[k] = D(a.c[k]); p2[k] = D(b.c[k]); }
}
double oracle_schannel_weight(const double* p1, const double* p2, double* r1, double* r2) {
    Real a, b;
    Real w = schannel_weight(mk(p1[0], p1[1], p1[2], p1[3]), mk(p2[0], p2[1], p2[2], p2[3]), a, b);
    *r1 = D(a); *r2 = D(b);
    return D(w);
}

}  // extern "C"
