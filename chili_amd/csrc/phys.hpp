// phys.hpp -- FP64 kinematics and phase-space elements for the gfx950 kernels.
//
// Every function keeps the floating-point operation order of the reference
// so that the device results agree with it to rounding of the libm calls
// (the kernels are compiled with -ffp-contract=off).  Citations are file:line
// in the jxi24/Chili tree.
#pragma once
#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#endif
#include <stdint.h>

#define CHILI_DEV __device__ __forceinline__

namespace chili_dev {

constexpr double kPi = 3.14159265358979323846;

struct V4 { double t, x, y, z; };

CHILI_DEV V4 mk(double t, double x, double y, double z) { V4 v; v.t = t; v.x = x; v.y = y; v.z = z; return v; }
CHILI_DEV V4 vadd(const V4& a, const V4& b) { return mk(a.t + b.t, a.x + b.x, a.y + b.y, a.z + b.z); }   // FourVector.cc:130-137
CHILI_DEV V4 vsub(const V4& a, const V4& b) { return mk(a.t - b.t, a.x - b.x, a.y - b.y, a.z - b.z); }
CHILI_DEV V4 vscale(const V4& a, double s) { return mk(a.t * s, a.x * s, a.y * s, a.z * s); }            // FourVector.cc:148-155
CHILI_DEV V4 vneg(const V4& a) { return mk(-a.t, -a.x, -a.y, -a.z); }
CHILI_DEV double vdot(const V4& a, const V4& b) { return a.t * b.t - (a.x * b.x + a.y * b.y + a.z * b.z); }  // FourVector.cc:166-169
CHILI_DEV double vpt2(const V4& a) { return a.x * a.x + a.y * a.y; }                                          // FourVector.hh:141
CHILI_DEV double vpt(const V4& a) { return sqrt(vpt2(a)); }
CHILI_DEV double vp2(const V4& a) { return vpt2(a) + a.z * a.z; }                                             // FourVector.hh:149
CHILI_DEV double vp(const V4& a) { return sqrt(vp2(a)); }
CHILI_DEV double vmass2(const V4& a) { return vdot(a, a); }
CHILI_DEV double vrap(const V4& a) {                                                                           // FourVector.cc:26-29
    if (a.t == 0) return 99;
    return log((a.t + a.z) / (a.t - a.z)) / 2;
}
CHILI_DEV double vphi(const V4& a) {                                                                           // FourVector.cc:20-24
    double phi = atan2(a.y, a.x);
    if (phi < 0) return phi + 2 * kPi;
    return phi;
}
CHILI_DEV double dmax(double a, double b) { return (a < b) ? b : a; }   // std::max
CHILI_DEV double dmin(double a, double b) { return (b < a) ? b : a; }   // std::min
// Spatial component i in {1, 2, 3} of a four-vector, by value selects only (no
// address arithmetic, so the vector stays in registers).
CHILI_DEV double comp(const V4 v, int i) { return i == 1 ? v.x : (i == 2 ? v.y : v.z); }
CHILI_DEV V4 with_comp(const V4 v, int i, double s) { return mk(v.t, i == 1 ? s : v.x, i == 2 ? s : v.y, i == 3 ? s : v.z); }

CHILI_DEV double small_omct(const V4& a, const V4& v) {                                                        // FourVector.cc:57-65
    double mag = sqrt(vp2(a) * vp2(v));
    double pq = a.x * v.x + a.y * v.y + a.z * v.z;
    double ct = dmin(dmax(pq / mag, -1.), 1.);
    if (ct < 0.) return 1. - ct;
    double cx = a.y * v.z - a.z * v.y, cy = a.z * v.x - a.x * v.z, cz = a.x * v.y - a.y * v.x;
    double st = sqrt(cx * cx + cy * cy + cz * cz) / mag;
    double st2 = st / (2. * sqrt((ct + 1) / 2.));
    return 2. * st2 * st2;
}

// Poincare boost (Poincare.cc:7-8, 36-50) with rsq = Mass().
CHILI_DEV void boost_fwd(const V4& l, double rsq, V4& v) {
    double lv = l.x * v.x + l.y * v.y + l.z * v.z;
    double v0 = (l.t * v.t - lv) / rsq;
    double c1 = (v.t + v0) / (rsq + l.t);
    v = mk(v0, v.x - l.x * c1, v.y - l.y * c1, v.z - l.z * c1);
}
CHILI_DEV void boost_back(const V4& l, double rsq, V4& v) {
    double lv = l.x * v.x + l.y * v.y + l.z * v.z;
    double v0 = (l.t * v.t + lv) / rsq;
    double c1 = (v.t + v0) / (rsq + l.t);
    v = mk(v0, v.x + l.x * c1, v.y + l.y * c1, v.z + l.z * c1);
}
// Transverse axis PT() of the rotation Poincare(v1, v2) (Poincare.cc:10-34).
CHILI_DEV V4 rotation_pt(const V4& v1, const V4& v2) {
    double n2 = vp(v2);
    V4 b = mk(0., v2.x / n2, v2.y / n2, v2.z / n2);
    double n1 = vp(v1);
    V4 l = mk(0., v1.x / n1, v1.y / n1, v1.z / n1);
    double lb = vdot(l, b);
    V4 t = vadd(b, vscale(l, lb));
    double mt = vp2(t);
    if (mt != 0.) t = vscale(t, 1. / sqrt(mt));
    // the reference's three-compare sorting network over |l_i|, carrying the
    // index together with its value
    int i0 = 1, i1 = 2, i2 = 3;
    double m0 = fabs(l.x), m1 = fabs(l.y), m2 = fabs(l.z);
    if (m2 > m1) { int s = i1; i1 = i2; i2 = s; double q = m1; m1 = m2; m2 = q; }
    if (m1 > m0) { int s = i0; i0 = i1; i1 = s; double q = m0; m0 = m1; m1 = q; }
    if (m2 > m1) { int s = i1; i1 = i2; i2 = s; double q = m1; m1 = m2; m2 = q; }
    double tdp = comp(t, i1) * comp(l, i1) + comp(t, i2) * comp(l, i2);
    if (tdp != 0.) t = with_comp(t, i0, -tdp / comp(l, i0));
    if (vp2(t) == 0.) t = with_comp(t, i1, 1.);
    return t;
}

// 4-D Levi-Civita contraction, ChannelElements.cc:9-19.
CHILI_DEV V4 lt4(const V4& a, const V4& b, const V4& c) {
    double t = a.x * b.y * c.z + a.y * b.z * c.x + a.z * b.x * c.y
             - a.x * b.z * c.y - a.z * b.y * c.x - a.y * b.x * c.z;
    double x = -a.t * b.y * c.z - a.y * b.z * c.t - a.z * b.t * c.y
             + a.t * b.z * c.y + a.z * b.y * c.t + a.y * b.t * c.z;
    double y = -a.x * b.t * c.z - a.t * b.z * c.x - a.z * b.x * c.t
             + a.x * b.z * c.t + a.z * b.t * c.x + a.t * b.x * c.z;
    double z = -a.x * b.y * c.t - a.y * b.t * c.x - a.t * b.x * c.y
             + a.x * b.t * c.y + a.t * b.y * c.x + a.y * b.x * c.t;
    return mk(t, -x, -y, -z);
}
CHILI_DEV double sqlam(double s, double s1, double s2) {                                                       // ChannelElements.cc:21-25
    double d = s - s1 - s2;
    double arg = d * d - 4. * s1 * s2;
    if (arg > 0.) return sqrt(arg) / s;
    return 0.;
}
// Power-law s sampling with a = 0, k = 1, cn = 0.99 (ChannelElements.cc:27-62).
CHILI_DEV double massless_prop_momenta(double smin, double smax, double ran) {
    const double a = 0., kk = 1.0, ce = 1. - 0.99;
    return kk * (pow(ran * pow(a + kk * smax, ce) + (1. - ran) * pow(a + kk * smin, ce), 1 / ce) - a);
}
CHILI_DEV double massless_prop_weight(double smin, double smax, double s, double& ran) {
    const double a = 0., kk = 1.0, ce = 1. - 0.99;
    double amin = pow(a + kk * smin, ce);
    double w = pow(a + kk * smax, ce) - amin;
    ran = (pow(a + kk * s, ce) - amin) / w;
    w /= kk * ce;
    return w / pow(s, -0.99);
}
// Breit-Wigner s sampling, ChannelElements.cc:64-85.
CHILI_DEV double massive_prop_momenta(double m, double g, double smin, double smax, double ran) {
    double m2 = m * m, mw = m * g;
    double ymax = atan((smin - m2) / mw), ymin = atan((smax - m2) / mw);
    return m2 + mw * tan(ymin + ran * (ymax - ymin));
}
CHILI_DEV double massive_prop_weight(double m, double g, double smin, double smax, double s, double& ran) {
    double m2 = m * m, mw = m * g;
    double ymax = atan((smin - m2) / mw), ymin = atan((smax - m2) / mw);
    double y = atan((s - m2) / mw);
    ran = (y - ymin) / (ymax - ymin);
    double w = mw / ((s - m2) * (s - m2) + mw * mw);
    w = (ymin - ymax) / w;
    return w;
}
// Isotropic two-body decay p -> (s1, s2), ChannelElements.cc:87-104.
CHILI_DEV void schannel_momenta(const V4& p, double s1, double s2, V4& p1, V4& p2, double ran1, double ran2) {
    const double ctmin = -1.0, ctmax = 1.0;
    double s = vmass2(p), rs = sqrt(fabs(s));
    double e1 = (s + s1 - s2) / rs / 2., m1 = sqrt(e1 * e1 - s1);
    double ct = ctmin + (ctmax - ctmin) * ran1, st = sqrt(1. - ct * ct);
    double phi = 2. * kPi * ran2;
    const V4 xref = mk(1., 1., 0., 0.);
    V4 pl = (vp2(p) != 0.) ? p : mk(1., 0., 0., 1.);
    double rsq = sqrt(vmass2(p));
    V4 n_perp = rotation_pt(pl, xref);
    V4 l_perp = lt4(pl, xref, n_perp);
    l_perp = vscale(l_perp, 1.0 / sqrt(fabs(vmass2(l_perp))));
    double mct = m1 * ct, plp = vp(pl);
    V4 q = mk(e1, pl.x * mct / plp, pl.y * mct / plp, pl.z * mct / plp);
    double sphi, cphi;
    sincos(phi, &sphi, &cphi);
    V4 dir = vadd(vscale(l_perp, cphi), vscale(n_perp, sphi));
    q = vadd(q, vscale(dir, m1 * st));
    boost_back(p, rsq, q);
    p1 = q;
    p2 = vsub(p, q);
}
// Inverse two-body map and weight, ChannelElements.cc:106-135.
CHILI_DEV double schannel_weight(const V4& p1, const V4& p2, double& ran1, double& ran2) {
    const double ctmin = -1.0, ctmax = 1.0;
    V4 p = vadd(p1, p2), p1h = p1;
    const V4 xref = mk(1., 1., 0., 0.);
    V4 pl = (vp2(p) != 0.) ? p : mk(1., 0., 0., 1.);
    double rsq = sqrt(vmass2(p));
    V4 n_perp = rotation_pt(pl, xref);
    V4 l_perp = lt4(pl, xref, n_perp);
    l_perp = vscale(l_perp, 1.0 / sqrt(fabs(vmass2(l_perp))));
    boost_fwd(p, rsq, p1h);
    double ct = (p1h.x * pl.x + p1h.y * pl.y + p1h.z * pl.z) / sqrt(vp2(p1h) * vp2(pl));
    double cp = vdot(vneg(l_perp), p1), sp = vdot(vneg(n_perp), p1), norm = sqrt(cp * cp + sp * sp);
    cp /= norm;
    sp /= norm;
    ran1 = (ct - ctmin) / (ctmax - ctmin);
    ran2 = atan2(sp, cp) / (2. * kPi);
    if (ran2 < 0.) ran2 += 1.;
    double w = (ctmax - ctmin) / 2.;
    double s1 = dmax(vmass2(p1), 0.);
    double s2 = dmax(vmass2(p2), 0.);
    double s = vmass2(p);
    if (s1 == 0. && s2 == 0.) s = p1.t * p2.t * small_omct(p1, p2);                                           // FourVector.cc:67-69
    w *= sqlam(s, s1, s2) / (8. * kPi);
    return w;
}

// ---------------------------------------------------------------------------
// Philox4x32-10 counter-based uniforms: counter (attempt lo, attempt hi,
// dim pair, stream), key = seed; double = (32 high bits, 21 bits) * 2^-53.
// ---------------------------------------------------------------------------
CHILI_DEV void philox4x32_10(uint32_t& c0, uint32_t& c1, uint32_t& c2, uint32_t& c3, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        // each 32 x 32 -> 64-bit product as one 64-bit multiply (one
        // v_mad_u64_u32 on the device instead of a mul_hi / mul_lo pair)
        const uint64_t p0 = (uint64_t)0xD2511F53u * (uint64_t)c0, p1 = (uint64_t)0xCD9E8D57u * (uint64_t)c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
}
CHILI_DEV double u53(uint32_t hi, uint32_t lo) {
    uint64_t m = ((uint64_t)hi << 21) | (uint64_t)(lo >> 11);
    return (double)m * 0x1p-53;
}
// Uniforms k0, k0+1 of one attempt (k0 even).
CHILI_DEV void philox_pair(uint64_t seed, uint32_t stream, uint64_t attempt, uint32_t k0, double& u0, double& u1) {
    uint32_t c0 = (uint32_t)attempt, c1 = (uint32_t)(attempt >> 32), c2 = k0 >> 1, c3 = stream;
    philox4x32_10(c0, c1, c2, c3, (uint32_t)seed, (uint32_t)(seed >> 32));
    u0 = u53(c0, c1);
    u1 = u53(c2, c3);
}

}  // namespace chili_dev
