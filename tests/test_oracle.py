"""Pins of the CPU oracle (oracle/chili_oracle.cpp).

The reference ships no tests, golden vectors or fixtures for this path and
cannot be built here (its build fetches fmt/spdlog/yaml-cpp), so every oracle
function is pinned by a known-answer vector, a closed form, an invariant, a
library routine or brute force -- chosen so that a dropped term, a wrong
sign or a transposed index fails:

  Philox4x32-10 ............ Random123 known-answer vectors
  channel pick ............. libstdc++ std::discrete_distribution itself (compiled probe)
  ConstructChannels ........ set-partition counts (Bell numbers, brute force), reference example
  FindBin .................. bisect_left (textbook lower_bound), edge cases
  grid map / weight ........ identity on the uniform grid, numpy.interp, product of widths
  grid Adapt ............... fixed point on uniform data, all-zero row, equal-mass invariant
  grid Split ............... closed form
  MultiChannel::Adapt ...... hand-computed closed form
  StatsData ................ numpy mean / unbiased variance
  s sampling (both) ........ weight == |ds/dr| (finite differences), round trip
  two-body decay ........... momentum conservation, on-shell masses, round trip,
                             weight == sqrt(lambda)/(8 pi s)
  FSMapper ................. round trip x -> p -> x, sum p = 0, masses, channel
                             consistency of the phase-space integral
  integrands / cut ......... explicit numpy restatement of the definitions
  plain VEGAS loop ......... integral vs scipy quadrature of the closed-form sigma(s)
  retry loop ............... iterate() against the per-attempt records
  JetCluster (kT cut) ...... independent textbook inclusive kT (400 random
                             final states, two radii), closed forms (dR vs R,
                             unwrapped phi)
  Optimize loops ........... composition of the pinned parts (iterate, Adapt,
                             grid Adapt, MaxDifference, merge, Split) and the
                             stop / refinement rules
  unweighting loop ......... its definition on the per-attempt records and
                             the Philox uniforms (keep rule, first n kept,
                             event weights, StatsData)
  ConstructChannels order .. std::set / std::map semantics (tests/test_channel_order.py)
  W conditioning ........... one-ulp momentum changes on the oracle alone
"""
import math
import os
import subprocess

import numpy as np
import pytest

from chili_amd import abi
from chili_amd import processes as P
from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))


# ---------------------------------------------------------------------------
def test_philox_known_answers():
    # Random123 kat_vectors for philox4x32_10
    assert O.philox_block([0, 0, 0, 0], [0, 0]) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert O.philox_block([0xffffffff] * 4, [0xffffffff] * 2) == [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert O.philox_block([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0]) == \
        [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


def test_philox_uniform_layout():
    seed, stream, att = 0x1234567890ABCDEF, 7, (5 << 32) + 11
    for k in range(6):
        c = O.philox_block([att & 0xffffffff, att >> 32, k >> 1, stream], [seed & 0xffffffff, seed >> 32])
        hi, lo = (c[2], c[3]) if k & 1 else (c[0], c[1])
        want = ((hi << 21) | (lo >> 11)) * 2.0 ** -53
        assert O.philox_uniform(seed, stream, att, k) == want
    u = np.array([O.philox_uniform(1, 0, a, 0) for a in range(20000)])
    assert 0 <= u.min() and u.max() < 1 and abs(u.mean() - 0.5) < 0.01


# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def discrete_probe(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("probe") / "discrete_ref")
    subprocess.run(["g++", "-O2", "-std=c++17", os.path.join(HERE, "golden", "discrete_ref.cpp"), "-o", exe], check=True)
    return exe


def test_channel_pick_matches_libstdcxx(discrete_probe):
    rng = np.random.default_rng(3)
    for n in (2, 3, 15, 52):
        w = rng.uniform(0.0, 1.0, n)
        w[rng.integers(0, n)] = 0.0          # a zero-weight channel is never picked
        words = rng.integers(0, 2 ** 32, (400, 2), dtype=np.uint64)
        words[0] = (0, 0)                     # u = 0
        words[1] = (2 ** 32 - 1, 2 ** 32 - 1)  # u just below 1
        inp = f"{n}\n" + " ".join(repr(float(x)) for x in w) + f"\n{len(words)}\n" + \
            "\n".join(f"{a} {b}" for a, b in words) + "\n"
        out = subprocess.run([discrete_probe], input=inp, capture_output=True, text=True, check=True).stdout.split("\n")
        for line in out:
            if not line.strip():
                continue
            us, ks = line.split()
            assert O.pick_channel(w, float.fromhex(us)) == int(ks)


# ---------------------------------------------------------------------------
def partitions_with_cost(n, smax):
    """Brute-force count of set partitions of n labelled items with
    sum(|block| - 1) <= smax."""
    def rec(items, cost):
        if not items:
            return 1
        first, rest = items[0], items[1:]
        tot = 0
        for r in range(len(rest) + 1):
            for combo in __import__("itertools").combinations(rest, r):
                c = cost + r
                if c > smax:
                    continue
                tot += rec([x for x in rest if x not in combo], c)
        return tot
    return rec(list(range(n)), 0)


@pytest.mark.parametrize("n,smax", [(2, 0), (2, 1), (3, 1), (3, 2), (4, 1), (4, 2), (4, 3), (5, 2), (5, 4)])
def test_construct_channels_counts(n, smax):
    o = O.Oracle(P.gluon_process(n, smax=smax), 8, P.integrand_unit())
    assert o.nchannels == partitions_with_cost(n, smax)
    assert o.ndims == 3 * n - 2
    if smax == n - 1:
        bell = [1, 1, 2, 5, 15, 52, 203]
        assert o.nchannels == bell[n]


def test_construct_channels_reference_example():
    # src/example.cc:98 -- g g -> g g at smax = 0: one channel, t-objects 8 > 4
    o = O.Oracle(P.example_process(), 2, P.integrand_unit())
    assert o.channels() == [(((8, 21, 0.0, 0.0), (4, 21, 0.0, 0.0)), ())]


def test_construct_channels_structure():
    o = O.Oracle(P.gluon_process(5, smax=4), 8, P.integrand_unit())
    full = sum(1 << i for i in range(2, 7))
    for info, dec in o.channels():
        idx = [p[0] for p in info]
        assert idx == sorted(idx, reverse=True)
        assert sum(idx) == full and all((a & b) == 0 for i, a in enumerate(idx) for b in idx[i + 1:])
        parents = [d[0][0] for d in dec]
        assert parents == sorted(parents)
        for par, c1, c2 in dec:
            assert c1[0] | c2[0] == par[0] and (c1[0] & c2[0]) == 0
        comp = [x for x in idx if bin(x).count("1") > 1]
        assert len(dec) == sum(bin(x).count("1") - 1 for x in comp)


def test_tt_channels():
    o = O.Oracle(P.tt_process(smax=5), 64, P.integrand_resonant())
    ch = o.channels()
    assert len(ch) == 2 and o.ndims == 16
    tchan = [tuple(p[0] for p in c[0]) for c in ch]
    assert (252,) in tchan and (224, 28) in tchan


# ---------------------------------------------------------------------------
def random_edges(rng, dims, B):
    return np.array([np.concatenate([[0.0], np.sort(rng.uniform(0, 1, B - 1)), [1.0]]) for _ in range(dims)])


def test_find_bin_lower_bound():
    import bisect
    rng = np.random.default_rng(1)
    e = random_edges(rng, 3, 17)
    for dim in range(3):
        row = list(e[dim, :17])           # lower edges e_0..e_{B-1}
        for x in list(rng.uniform(0, 1, 400)) + list(e[dim, 1:17]) + [1.0, 0.5]:
            k = bisect.bisect_left(row, x) - 1
            assert O.grid_find_bin(e, dim, x) == max(k, 0)
        # x on an interior edge lands in the bin to its left: bins are (e_k, e_{k+1}]
        assert O.grid_find_bin(e, dim, e[dim, 5]) == 4
        assert O.grid_find_bin(e, dim, 0.0) == 0        # reference: size_t(-1), clamped
        assert O.grid_find_bin(e, dim, float("nan")) == 0


def test_grid_forward_map():
    rng = np.random.default_rng(2)
    u = O.uniform_grid(4, 8)
    r = rng.uniform(0, 1, 4)
    x, jac = O.grid_forward(u, r)
    assert np.allclose(x, r, rtol=0, atol=1e-15) and jac == 1.0
    e = random_edges(rng, 4, 8)
    r = rng.uniform(0, 1, 4)
    x, jac = O.grid_forward(e, r)
    for k in range(4):
        assert abs(x[k] - np.interp(r[k], np.linspace(0, 1, 9), e[k])) < 1e-15
    idx = np.floor(r * 8).astype(int)
    assert math.isclose(jac, np.prod([(e[k, idx[k] + 1] - e[k, idx[k]]) * 8 for k in range(4)]), rel_tol=1e-14)


def test_grid_adapt_invariants():
    rng = np.random.default_rng(4)
    B = 16
    # uniform data: fixed point (up to rounding)
    e = O.uniform_grid(2, B)
    assert np.allclose(O.grid_adapt(e, np.ones((2, B))), e, atol=1e-14)
    # an all-zero row is left with every edge at zero (AdaptiveMap.cc:69,93)
    d = rng.uniform(0.1, 2.0, (2, B))
    d[1] = 0
    out = O.grid_adapt(e, d)
    assert np.all(out[1] == 0) and out[0, 0] == 0 and out[0, -1] == 1
    # equal-mass invariant: with smoothed, damped weights w_i the new edges
    # split the piecewise-linear cumulative weight into B equal parts
    e = random_edges(rng, 1, B)
    d = rng.uniform(0.0, 3.0, (1, B))
    out = O.grid_adapt(e, d)
    t = d[0].copy()
    sm = np.empty(B)
    sm[0] = (t[0] + t[1]) / 2
    sm[1:-1] = (t[:-2] + t[1:-1] + t[2:]) / 3
    sm[-1] = (t[-2] + t[-1]) / 2
    rr = sm / sm.sum()
    w = (rr - 1) / np.log(rr)
    cum = np.concatenate([[0], np.cumsum(w)])
    target = np.interp(out[0], e[0], cum)
    assert np.allclose(target, np.arange(B + 1) / B * cum[-1], rtol=0, atol=1e-9 * cum[-1])
    assert np.all(np.diff(out[0]) > 0)


def test_grid_split():
    rng = np.random.default_rng(5)
    e = random_edges(rng, 2, 5)
    for ns in (2, 3, 4):
        out = O.grid_split(e, ns)
        assert out.shape == (2, 5 * ns + 1)
        for k in range(2):
            want = np.concatenate([e[k, b] + np.arange(ns) * (e[k, b + 1] - e[k, b]) / ns for b in range(5)] + [[1.0]])
            assert np.allclose(out[k], want, rtol=0, atol=1e-16)


def test_alpha_adapt_closed_form():
    a = np.array([0.5, 0.25, 0.25])
    t = np.array([16.0, 1.0, 0.0])
    out = O.alpha_adapt(a, t, beta=0.25, min_alpha=1e-5)
    # new = [0.5*2, 0.25*1, 0] -> normalised [0.8, 0.2, 0] -> zeros stay zero
    assert np.allclose(out, [0.8, 0.2, 0.0], rtol=1e-15)
    out = O.alpha_adapt([0.5, 0.5], [1.0, 1e-40], beta=0.25, min_alpha=0.01)
    # [0.5, 0.5e-10] -> [1, 1e-10] -> clamp [1, 0.01] -> renormalise
    assert np.allclose(out, [1 / 1.01, 0.01 / 1.01], rtol=1e-12)
    assert O.max_difference([3.0, -1.0, 2.5]) == 4.0


def test_stats_textbook():
    x = np.random.default_rng(6).normal(2.0, 3.0, 1000)
    x[::7] = 0.0
    st, err = O.stats(x)
    assert st.n == 1000 and st.n_finite == np.count_nonzero(x)
    assert math.isclose(st.sum / st.n, x.mean(), rel_tol=1e-13)
    # Error() is the error of the mean: (sum2/n - mean^2)/(n-1) = s^2/n (Statistics.hh:90,125)
    assert math.isclose(err, np.sqrt(np.var(x, ddof=1) / len(x)), rel_tol=1e-10)
    assert st.min == x.min() and st.max == x.max()


# ---------------------------------------------------------------------------
def test_massless_prop_jacobian():
    smin, smax = 12.97, 1.0e6
    for r in (1e-3, 0.1, 0.5, 0.9, 0.999):
        s, w, rback = O.massless_prop(smin, smax, r)
        assert smin <= s <= smax and abs(rback - r) < 1e-12
        h = 1e-6
        ds = (O.massless_prop(smin, smax, r + h)[0] - O.massless_prop(smin, smax, r - h)[0]) / (2 * h)
        assert math.isclose(w, abs(ds), rel_tol=1e-6)


def test_massive_prop_jacobian():
    m, g, smin, smax = 80.4, 2.08, 100.0, 1.0e5
    for r in (1e-3, 0.2, 0.5, 0.8, 0.999):
        s, w, rback = O.massive_prop(m, g, smin, smax, r)
        assert smin <= s <= smax and abs(rback - r) < 1e-12
        h = 1e-7
        ds = (O.massive_prop(m, g, smin, smax, r + h)[0] - O.massive_prop(m, g, smin, smax, r - h)[0]) / (2 * h)
        assert math.isclose(w, abs(ds), rel_tol=1e-5)
    # r = 0 gives smax: the reference's ymin/ymax names are swapped
    assert math.isclose(O.massive_prop(m, g, smin, smax, 0.0)[0], smax, rel_tol=1e-12)


def kallen(a, b, c):
    return a * a + b * b + c * c - 2 * a * b - 2 * a * c - 2 * b * c


def test_sqlam():
    for s, s1, s2 in [(100.0, 1.0, 4.0), (1e4, 0.0, 0.0), (100.0, 20.0, 20.0)]:
        assert math.isclose(O.sqlam(s, s1, s2), math.sqrt(kallen(s, s1, s2)) / s, rel_tol=1e-14)
    assert O.sqlam(10.0, 9.0, 9.0) == 0.0


@pytest.mark.parametrize("s1,s2", [(0.0, 0.0), (25.0, 0.0), (100.0, 400.0)])
def test_schannel_decay(s1, s2):
    rng = np.random.default_rng(8)
    for _ in range(20):
        m = 60.0 + 100 * rng.uniform()
        pv = rng.normal(0, 80, 3)
        p = np.array([math.sqrt(m * m + pv @ pv), *pv])
        r1, r2 = rng.uniform(0, 1, 2)
        p1, p2 = O.schannel_momenta(p, s1, s2, r1, r2)
        mink = lambda v: v[0] ** 2 - v[1:] @ v[1:]
        assert np.allclose(p1 + p2, p, rtol=0, atol=1e-12 * p[0])
        assert abs(mink(p1) - s1) < 1e-9 * p[0] ** 2 and abs(mink(p2) - s2) < 1e-9 * p[0] ** 2
        w, q1, q2 = O.schannel_weight(p1, p2)
        assert abs(q1 - r1) < 1e-9 and abs(q2 - r2) < 1e-9
        want = math.sqrt(kallen(m * m, s1, s2)) / (m * m) / (8 * math.pi)
        assert math.isclose(w, want, rel_tol=1e-8)


@pytest.mark.parametrize("n,smax", [(2, 1), (3, 2), (4, 3), (5, 4)])
def test_mapper_round_trip(n, smax):
    o = O.Oracle(P.gluon_process(n, smax=smax), 8, P.integrand_unit())
    rng = np.random.default_rng(9 + n)
    for c in range(o.nchannels):
        done = 0
        for _ in range(200):
            x = rng.uniform(0.02, 0.98, o.ndims)
            mom = o.generate_point(c, x)
            if not np.all(np.isfinite(mom)):
                continue
            tot = mom.sum(axis=0)
            E = np.abs(mom[:, 0]).max()
            xin = -2 * mom[:2, 0] / 1000.0
            if np.any(xin > 1):          # outside the beam: not a physical point
                continue
            assert np.allclose(tot, 0, atol=1e-10 * E)
            m2 = mom[2:, 0] ** 2 - (mom[2:, 1:] ** 2).sum(axis=1)
            assert np.all(np.abs(m2) < 1e-8 * E * E)
            w, xb = o.generate_weight(c, mom)
            assert w > 0
            # WgtDecays takes sqrt(|m^2|) of massless children (Channel.cc:167-168):
            # the rounding noise of m^2 ~ 1e-16 E^2 moves s_min by ~1e-8 relative
            assert np.allclose(xb, x, rtol=0, atol=1e-6), (c, x, xb)
            done += 1
            if done >= 20:
                break
        assert done >= 5


def test_channel_consistency_of_phase_space():
    """Every channel's weight is a Jacobian of the same measure: the
    integral of one cut-bounded test function through each channel alone
    must agree within errors."""
    proc = P.gluon_process(3, smax=2)
    f = P.integrand_propprod()
    o = O.Oracle(proc, 8, f)
    rng = np.random.default_rng(12)
    est = []
    for c in range(o.nchannels):
        vals = []
        for _ in range(30000):
            x = rng.uniform(0, 1, o.ndims)
            mom = o.generate_point(c, x)
            if not np.all(np.isfinite(mom)):
                vals.append(0.0)
                continue
            fv, ok = O.integrand(f, mom, 1000.0)
            if not ok:
                vals.append(0.0)
                continue
            w, _ = o.generate_weight(c, mom)
            vals.append(fv * w)
        v = np.array(vals)
        est.append((v.mean(), v.std() / math.sqrt(len(v))))
    m = np.array([e[0] for e in est])
    s = np.array([e[1] for e in est])
    mean = np.sum(m / s ** 2) / np.sum(1 / s ** 2)
    chi2 = np.sum((m - mean) ** 2 / s ** 2)
    assert chi2 < 4 * (len(m) - 1) + 10, (m, s)


# ---------------------------------------------------------------------------
def test_integrand_definitions():
    rng = np.random.default_rng(13)
    mom = np.zeros((6, 4))
    for i in range(2, 6):
        v = rng.normal(0, 100, 3)
        mom[i] = [np.linalg.norm(v), *v]
    tot = mom[2:].sum(axis=0)
    mom[0] = [-(tot[0] + tot[3]) / 2, 0, 0, -(tot[0] + tot[3]) / 2]
    mom[1] = [-(tot[0] - tot[3]) / 2, 0, 0, (tot[0] - tot[3]) / 2]
    mink = lambda v: v[0] ** 2 - v[1:] @ v[1:]
    s = [mink(mom[i] + mom[j]) for i in range(2, 6) for j in range(i + 1, 6)]
    want = sum(1e8 / (s[a] * s[b]) for a in range(6) for b in range(a + 1, 6))
    f = P.integrand_propprod(s0=1e4, cuts=None)
    v, ok = O.integrand(f, mom, 1e6)
    assert ok and math.isclose(v, want, rel_tol=1e-12)
    # resonant: Breit-Wigners of the four currents times x^T Q x
    mom8 = np.zeros((8, 4))
    for i in range(2, 8):
        vv = rng.normal(0, 100, 3)
        mom8[i] = [np.linalg.norm(vv), *vv]
    q = P.resonant_q()
    fr = P.integrand_resonant(q=q, cuts=None)
    bw = 1.0
    for mask, M, G in [(0b11000, P.MW, P.GW), (0b11100, P.MT, P.GT), (0b11000000, P.MW, P.GW), (0b11100000, P.MT, P.GT)]:
        sv = mink(sum(mom8[i] for i in range(8) if mask >> i & 1))
        bw *= (M * G) ** 2 / ((sv - M * M) ** 2 + (M * G) ** 2)
    x = np.array([mink(mom8[i] + mom8[j]) for i in range(2, 8) for j in range(i + 1, 8)]) / 1e5
    v, ok = O.integrand(fr, mom8, 1e6)
    assert ok and math.isclose(v, bw * (x @ q @ x), rel_tol=1e-10)


def test_cut_definition():
    f = P.integrand_unit((20.0, 2.5, 0.4))

    def event(pts, ys, phis):
        mom = np.zeros((2 + len(pts), 4))
        for i, (pt, y, ph) in enumerate(zip(pts, ys, phis)):
            mom[2 + i] = [pt * math.cosh(y), pt * math.cos(ph), pt * math.sin(ph), pt * math.sinh(y)]
        tot = mom[2:].sum(axis=0)
        mom[0] = [-(tot[0] + tot[3]) / 2, 0, 0, -(tot[0] + tot[3]) / 2]
        mom[1] = [-(tot[0] - tot[3]) / 2, 0, 0, (tot[0] - tot[3]) / 2]
        return mom

    assert O.integrand(f, event([30, 40], [0.1, -0.3], [0.5, 3.5]), 1000.0)[1]
    assert not O.integrand(f, event([19.9, 40], [0.1, -0.3], [0.5, 3.5]), 1000.0)[1]
    assert not O.integrand(f, event([30, 40], [2.6, -0.3], [0.5, 3.5]), 1000.0)[1]
    # dR uses the wrapped azimuth difference: phi = 0.1 and 2 pi - 0.1 are 0.2 apart
    assert not O.integrand(f, event([30, 40], [0.1, 0.1], [0.1, 2 * math.pi - 0.1]), 1000.0)[1]
    assert O.integrand(f, event([30, 40], [0.1, 0.6], [0.1, 2 * math.pi - 0.1]), 1000.0)[1]
    # x1 > 1 (beyond the beam energy) fails
    assert not O.integrand(f, event([300, 300], [1.5, 1.4], [0.0, 3.1]), 1000.0)[1]


def test_vegas_integral_vs_quadrature():
    from scipy.integrate import quad
    f = P.integrand_eemumu()
    exact = quad(P.eemumu_sigma, P.EE_SMIN, P.EE_SMAX, points=[P.EE_MZ ** 2], limit=200, epsabs=0, epsrel=1e-10)[0]
    edges = O.uniform_grid(2, 50)
    uni = abi.philox_uniforms(0xC0FFEE00)
    tot = np.zeros(3)
    for it in range(6):
        st, hist, _ = O.vegas_iterate(edges, f, uni, 10000, a0=it * 10000)
        edges = O.grid_adapt(edges, hist)
        tot += [st.n, st.sum, st.sum2]
    mean = tot[1] / tot[0]
    err = math.sqrt((tot[2] / tot[0] - mean ** 2) / (tot[0] - 1))
    assert abs(mean - exact) < 4 * err, (mean, err, exact)
    assert err / exact < 0.01


def test_retry_loop_accounting():
    proc, f = P.gluon_process(3, smax=2), P.integrand_propprod()
    o = O.Oracle(proc, 16, f)
    uni = abi.philox_uniforms(77, 2)
    ncalls = 500
    st, T, hist, na = o.iterate(uni, ncalls, a0=3)
    rec = o.points(uni, 3, na)
    acc = rec["status"] == 1
    assert rec["status"][na - 1] == 1 and acc.sum() == ncalls     # stops at the ncalls-th accepted attempt
    assert st.n == na and st.n_finite == np.count_nonzero(rec["val"][acc])
    assert math.isclose(st.sum, rec["val"][acc].sum(), rel_tol=1e-12)
    assert st.min == min(0.0, rec["val"][acc].min())               # rejected attempts add zeros
    # alpha training and picked-channel histograms from the records
    v2 = rec["val"] ** 2
    Tw = (v2[acc, None] * rec["w"][acc, None] / rec["ps"][acc]).sum(axis=0)
    assert np.allclose(T, Tw, rtol=1e-12)
    H = np.zeros_like(hist)
    for i in np.nonzero(acc)[0]:
        for k in range(o.ndims):
            H[rec["pick"][i], k, rec["bins"][i, k]] += v2[i]
    assert np.allclose(hist, H, rtol=1e-12, atol=0)


@pytest.mark.parametrize("wmax_q", [0.5, 0.99])
def test_unweighting_loop_from_records(wmax_q):
    """The oracle's unweighting loop (MultiChannel::operator() with a
    PostProcess that keeps a point of value v with probability
    min(1, |v| / wmax), MultiChannel.hh:196-251) against its definition
    applied to the per-attempt records: attempt a is kept when it passes the
    cut and u_{d+1}(a) < |v| / wmax (u from the attempt's Philox stream),
    every other attempt adds a zero to StatsData, the loop ends at the n-th
    kept attempt, the events are the kept attempts' momenta in order and
    their weights sign(v) max(1, |v| / wmax)."""
    proc, f = P.gluon_process(3, smax=2), P.integrand_propprod()
    o = O.Oracle(proc, 16, f)
    seed, stream, a0, n = 91, 4, 5, 300
    uni = abi.philox_uniforms(seed, stream)
    probe = o.points(uni, a0, 4000)
    ok = probe["status"] == 1
    wmax = float(np.quantile(np.abs(probe["val"][ok]), wmax_q))
    st, na, ev, ew = o.generate(uni, n, wmax, a0=a0)
    rec = o.points(uni, a0, na)
    u = np.array([O.philox_uniform(seed, stream, a0 + i, o.ndims + 1) for i in range(na)])
    keep = (rec["status"] == 1) & (u < np.abs(rec["val"]) / wmax)
    assert keep.sum() == n and keep[-1], "the loop ends at the n-th kept attempt"
    kept = np.nonzero(keep)[0]
    assert np.array_equal(ev, rec["mom"][kept])
    v = rec["val"][kept]
    assert np.array_equal(ew, np.where(np.abs(v) > wmax, v / wmax, np.sign(v)))
    vals = np.where(keep, rec["val"], 0.0)
    assert st.n == na and st.n_finite == np.count_nonzero(vals)
    assert math.isclose(st.sum, vals.sum(), rel_tol=1e-12) and math.isclose(st.sum2, (vals ** 2).sum(), rel_tol=1e-12)
    assert st.min == vals.min() and st.max == vals.max()
    if wmax_q < 0.9:
        assert (np.abs(ew) > 1).any()      # overweight events keep |v| / wmax


def test_weight_conditioning_one_ulp():
    """The multichannel weight of the reference algorithm is ill-conditioned
    in the momenta: WgtDecays takes sqrt(|m^2|) of the massless children
    (Channel.cc:167-168), so a last-bit change of the momenta (what two libm
    implementations of the forward map's pow / sin / cos produce) moves W by
    far more than 1e-12 at some points.  Shown on the oracle alone: each
    momentum component of the accepted points moved by at most one ulp gives
    a W spread of the same order as the one measured between the GPU's and the
    oracle's independent forward maps (<= 1e-6), while identical momenta give
    identical weights."""
    for proc, f in ((P.gluon_process(4, smax=3), P.integrand_propprod()), (P.tt_process(smax=5), P.integrand_resonant())):
        o = O.Oracle(proc, 64, f)
        uni = abi.philox_uniforms(0xC0FFEE01, 3)
        n = 3000
        rec = o.points(uni, 17, n)
        same = o.points(uni, 17, n, mom_in=rec["mom"])
        ok = (rec["status"] == 1) & (rec["w"] != 0)
        assert np.array_equal(same["w"][ok], rec["w"][ok])
        rng = np.random.default_rng(5)
        m = rec["mom"]
        step = rng.integers(-1, 2, m.shape)
        mp = np.where(step > 0, np.nextafter(m, np.inf), np.where(step < 0, np.nextafter(m, -np.inf), m))
        pert = o.points(uni, 17, n, mom_in=mp)
        both = ok & (pert["status"] == 1)
        d = np.abs(pert["w"][both] - rec["w"][both]) / np.abs(rec["w"][both])
        assert d.max() > 1e-9, d.max()          # far above the 1e-12 of identical momenta
        assert d.max() < 1e-6, d.max()


def test_optimize_step_is_its_parts():
    """The oracle's MultiChannel::Optimize loop body (operator() with
    training, Train, UpdateResults, RefineChannels, MultiChannel.hh:122-130,
    182-268) equals its separately pinned parts composed by hand: the retry
    loop's iterate, MultiChannel::Adapt, AdaptiveMap::Adapt of every channel
    with a non-zero histogram, MaxDifference (best weights = the new alpha
    when it improves), StatsData merge, and the refinement (ncalls x 1.5
    truncated, grids below max_bins split in two)."""
    from chili_amd import lib as L
    proc, f = P.gluon_process(3, smax=2), P.integrand_propprod()
    a = O.Oracle(proc, 8, f)
    b = O.Oracle(proc, 8, f)
    params = L.mc_params(ncalls=3001, nrefine=2, max_bins=16)
    summ = L.mc_summary()
    best = np.full(a.nchannels, -1.0)
    uni = abi.philox_uniforms(123, 40)
    pooled = None
    min_diff, best_ref, ncalls, it, bins = np.inf, best.copy(), 3001, 0, 8
    for k in range(5):
        st = a.optimize_step(uni, params, summ, best)
        # by hand on b
        alpha, grids = b.alpha(), b.grids()
        sb, T, hist, _ = b.iterate(abi.philox_uniforms(123, 40 + k), ncalls)
        assert stats_eq(st, sb)
        alpha = O.alpha_adapt(alpha, T)
        for c in range(b.nchannels):
            if np.any(hist[c] != 0):
                grids[c] = O.grid_adapt(grids[c], hist[c], 1.0)
        md = O.max_difference(T)
        if md < min_diff:
            min_diff, best_ref = md, alpha.copy()
        it += 1
        if it == 2:
            it, ncalls = 0, int(ncalls * 1.5)
            if bins < 16:
                grids = np.stack([O.grid_split(g, 2) for g in grids])
                bins *= 2
        b.set_grids(grids)
        b.set_alpha(alpha)
        pooled = sb.as_tuple() if pooled is None else merge(pooled, sb.as_tuple())
        assert np.array_equal(a.alpha(), b.alpha()) and np.array_equal(a.grids(), b.grids())
        assert np.array_equal(best, best_ref) and summ.min_diff == min_diff
        assert (params.ncalls, params.iteration, a.nbins, summ.nresults) == (ncalls, it, bins, k + 1)
        assert summ.sum.as_tuple() == pooled
    assert bins == 16 and ncalls == int(int(3001 * 1.5) * 1.5)


def stats_eq(x, y):
    return x.as_tuple() == y.as_tuple()


def merge(a, b):
    """StatsData::operator+=(StatsData), Statistics.hh:106-115."""
    return (a[0] + b[0], min(a[1], b[1]), max(a[2], b[2]), a[3] + b[3], a[4] + b[4], a[5] + b[5])


def test_optimize_stop_rules():
    """The loop ends at niterations, or as soon as the pooled relative error
    is <= rtol (MultiChannel.hh:256); a should_optimize = false run leaves
    alpha and grids untouched; nrefine = 0 never refines."""
    from chili_amd import lib as L
    proc, f = P.gluon_process(3, smax=2), P.integrand_propprod()
    o = O.Oracle(proc, 8, f)
    p = L.mc_params(ncalls=2000, niterations=4, rtol=0.0, nrefine=0)
    s = L.mc_summary()
    best = np.zeros(o.nchannels)
    res = o.optimize(abi.philox_uniforms(5, 0), p, s, best)
    assert len(res) == 4 and s.nresults == 4 and p.ncalls == 2000 and o.nbins == 8
    pooled = res[0].as_tuple()
    for r in res[1:]:
        pooled = merge(pooled, r.as_tuple())
    assert s.sum.as_tuple() == pooled
    rel = abi.Stats(*pooled).error() / abs(abi.Stats(*pooled).mean())
    # rtol just above the error after two iterations: stops after two
    o2 = O.Oracle(proc, 8, f)
    p2 = L.mc_params(ncalls=2000, niterations=10, nrefine=0)
    s2 = L.mc_summary()
    two = o2.optimize(abi.philox_uniforms(5, 0), p2, s2, np.zeros(o.nchannels), max_steps=2)
    r2 = s2.sum.error() / abs(s2.sum.mean())
    o3 = O.Oracle(proc, 8, f)
    p3 = L.mc_params(ncalls=2000, niterations=10, nrefine=0, rtol=r2 * (1 + 1e-9))
    s3 = L.mc_summary()
    o3.optimize(abi.philox_uniforms(5, 0), p3, s3, np.zeros(o.nchannels))
    assert s3.nresults == 2 and rel > 0
    # frozen: no training
    o4 = O.Oracle(proc, 8, f)
    g0, a0 = o4.grids(), o4.alpha()
    p4 = L.mc_params(ncalls=2000, niterations=2, should_optimize=0, nrefine=0)
    o4.optimize(abi.philox_uniforms(5, 0), p4, L.mc_summary(), np.zeros(o.nchannels))
    assert np.array_equal(o4.grids(), g0) and np.array_equal(o4.alpha(), a0)


def test_vegas_optimize_schedule():
    """Vegas::Optimize (Vegas.cc:54-82): at least niterations passes, the grid
    adapted after every pass (AdaptiveMap::Adapt, pinned above), split in two
    and ncalls doubled every nrefine passes; each pass equals the oracle's
    single Vegas pass on the adapted grid."""
    from chili_amd import lib as L
    f = P.integrand_eemumu()
    edges = O.uniform_grid(2, 50)
    p = L.vegas_params(ncalls=5000, nrefine=2, niterations=5, rtol=0.0, atol=0.0)
    tot = abi.Stats(0, np.finfo(float).max, np.finfo(float).tiny, 0, 0, 0)
    e, res, n = O.vegas_optimize(edges, f, abi.philox_uniforms(9, 0), p, tot, 0, 5)
    assert n == 5 and len(res) == 5 and e.shape == (2, 201) and p.ncalls == 20000 and p.irefine == 1
    g, nc = O.uniform_grid(2, 50), 5000
    for k in range(5):
        st, hist, _ = O.vegas_iterate(g, f, abi.philox_uniforms(9, k), nc)
        assert stats_eq(st, res[k])
        g = O.grid_adapt(g, hist, 1.0)
        if k % 2 == 1:
            g, nc = O.grid_split(g, 2), nc * 2
    assert np.array_equal(g, e)


# ---------------------------------------------------------------------------
# JetCluster restatement (kT-jet PreProcess) against an independent textbook
# inclusive kT: repeatedly the smallest of d_iB = pT_i^2 and
# d_ij = min(pT_i^2, pT_j^2) (dy^2 + dphi^2) / R^2 (phi in [0, 2 pi), not
# wrapped, as JetCluster.cc:80-84); a beam minimum makes a jet, a pair merges
# (four-momentum sum).
# ---------------------------------------------------------------------------
def _textbook_kt_jets(mom, R):
    import math

    def pt2(p):
        return p[1] * p[1] + p[2] * p[2]

    def rap(p):
        return 99.0 if p[0] == 0 else math.log((p[0] + p[3]) / (p[0] - p[3])) / 2

    def phi(p):
        a = math.atan2(p[2], p[1])
        return a + 2 * math.pi if a < 0 else a

    ps = [list(map(float, p)) for p in mom[2:]]
    jets = 0
    while ps:
        best = (float("inf"), None, None)
        for i in range(len(ps)):
            if pt2(ps[i]) < best[0]:
                best = (pt2(ps[i]), i, i)
            for j in range(i):
                dy, dp = rap(ps[i]) - rap(ps[j]), phi(ps[i]) - phi(ps[j])
                d = min(pt2(ps[i]), pt2(ps[j])) * ((dy * dy + dp * dp) / R / R)
                if d < best[0]:
                    best = (d, i, j)
        _, i, j = best
        if i == j:
            jets += 1
        else:
            ps[j] = [a + b for a, b in zip(ps[j], ps[i])]
        del ps[i]
    return jets


def _random_final_state(rng, n, collinear=False):
    ps = [[0.0, 0, 0, 0], [0.0, 0, 0, 0]]
    for k in range(n):
        pt, y, ph = rng.uniform(5, 200), rng.uniform(-3, 3), rng.uniform(0, 2 * np.pi)
        if collinear and k > 0 and rng.random() < 0.5:
            y0, ph0 = ps[-1][4], ps[-1][5]
            y, ph = y0 + rng.normal(0, 0.2), (ph0 + rng.normal(0, 0.2)) % (2 * np.pi)
        ps.append([pt * np.cosh(y), pt * np.cos(ph), pt * np.sin(ph), pt * np.sinh(y), y, ph])
    return np.array([p[:4] for p in ps])


def test_jet_cluster_vs_textbook_kt():
    rng = np.random.default_rng(41)
    counts = []
    for trial in range(400):
        n = 2 + trial % 5
        mom = _random_final_state(rng, n, collinear=trial % 2 == 1)
        for R in (0.4, 1.0):
            got = O.jet_count(mom, R)
            assert got == _textbook_kt_jets(mom, R), (trial, R)
            counts.append(got < n)
    # both outcomes occur
    assert 0.1 < np.mean(counts) < 0.9


def test_jet_cluster_closed_forms():
    # two partons: one jet iff dR^2 < R^2 (then d_12 < min pT^2)
    for dphi, want in ((0.3, 1), (0.5, 2)):
        mom = np.array([[0.0] * 4, [0.0] * 4,
                        [50.0, 50.0, 0.0, 0.0],
                        [40.0, 40.0 * np.cos(dphi), 40.0 * np.sin(dphi), 0.0]])
        assert O.jet_count(mom, 0.4) == want
    # phi is not wrapped: partons at phi = 0.1 and 2 pi - 0.1 are 2 pi - 0.2 apart
    mom = np.array([[0.0] * 4, [0.0] * 4,
                    [50.0, 50.0 * np.cos(0.1), 50.0 * np.sin(0.1), 0.0],
                    [40.0, 40.0 * np.cos(-0.1), 40.0 * np.sin(-0.1), 0.0]])
    assert O.jet_count(mom, 0.4) == 2
