"""Process, cut and integrand definitions of the benchmark workloads.

Workloads (dims follow the reference map, NDims = 3 n_out - 2,
include/Channel/Channel.hh:33):

  ee_mumu       e+e- -> mu+mu-, plain VEGAS on [0,1]^2 (Breit-Wigner in s-hat,
                flat cos theta), 2 dims x 50 bins, 10 x 1e4 then 1e5 points.
  gg_4g         2 -> 4 massless at sqrt(s) = 1 TeV, every s-/t-channel
                topology (15 channels, smax = 3), 1/s propagator maps,
                cuts pT > 20, |y| < 2.5, dR > 0.4, 10 dims x 64 bins.
  batched       2 -> 2 .. 2 -> 5 massless jobs (4/7/10/13 dims).
  tt_6j         2 -> 6 with two t(173, 1.42) -> W(80.4, 2.08) j, W -> j j
                chains at sqrt(s) = 1 TeV, 16 dims x 64 bins.
"""
import numpy as np

from . import abi

GEV2PB = 0.3893793721e9


def make_process(sqrts, flavs, masses, widths, vertices, smax, cuts=None):
    """cuts: None for the ConstructChannels defaults, else a dict with
    ptmin, etamax (per outgoing particle, scalars) and deltaR (scalar for
    every pid pair)."""
    p = abi.Process()
    p.sqrts = sqrts
    p.nflav = len(flavs)
    for i, f in enumerate(flavs):
        p.flavs[i] = f
    pids = sorted(masses)
    p.nmodel = len(pids)
    for i, pid in enumerate(pids):
        p.model_pid[i] = pid
        p.model_mass[i] = masses[pid]
        p.model_width[i] = widths[pid]
    p.nvertex = len(vertices)
    for i, v in enumerate(vertices):
        for k in range(3):
            p.vertex[i][k] = v[k]
    p.smax = smax
    if cuts is None:
        p.use_default_cuts = 1
    else:
        p.use_default_cuts = 0
        for i, f in enumerate(flavs):
            m = masses[abs(f)]
            p.ptmin[i] = cuts["ptmin"]
            p.etamax[i] = cuts["etamax"]
            p.smin[i] = m * m
        allp = sorted(set(flavs))
        k = 0
        for a in allp:
            for b in allp:
                p.deltaR_pid[k][0] = a
                p.deltaR_pid[k][1] = b
                p.deltaR[k] = cuts["deltaR"]
                k += 1
        p.ndeltaR = k
    return p


def gluon_process(n_out, sqrts=1000.0, smax=None, cuts=None):
    """All-gluon 2 -> n_out with g g -> g as the only vertex: every set
    partition of the final state is a t-channel topology."""
    if smax is None:
        smax = n_out - 1
    if cuts is None:
        cuts = dict(ptmin=20.0, etamax=2.5, deltaR=0.4)
    return make_process(sqrts, [21] * (n_out + 2), {21: 0.0}, {21: 0.0},
                        [(21, 21, 21)], smax, cuts)


def example_process():
    """The reference's own example (src/example.cc:56-98): g g -> g g at
    13 TeV, smax = 0, default cuts, with its Model::combine rule."""
    flav_all = [1, -1, 2, -2, 3, -3, 5, -5, 21, 23]
    masses = {1: 0.0, 2: 0.0, 3: 0.0, 5: 172.0, 21: 0.0, 23: 100.0}
    widths = {1: 0.0, 2: 0.0, 3: 0.0, 5: 5.0, 21: 0.0, 23: 2.5}

    def combine(i, j):
        if (i == 21 and j == 23) or (i == 23 and j == 21):
            return []
        if i == 21 and j == 21:
            return [21]
        if i == -j:
            return [21, 23]
        if i != j:
            if i in (21, 23):
                return [j]
            if j in (21, 23):
                return [i]
        return []

    verts = [(i, j, k) for i in flav_all for j in flav_all for k in combine(i, j)]
    return make_process(13000.0, [21, 21, 21, 21], masses, widths, verts, 0, None)


# ---- top pair to six jets ------------------------------------------------
MT, GT = 173.0, 1.42
MW, GW = 80.4, 2.08
TT_FLAVS = [21, 21, 5, 2, -1, -5, -2, 1]


def tt_process(sqrts=1000.0, smax=4, cuts=None):
    """g g -> t t~ -> (b u d~)(b~ u~ d): W+ = (u, d~) -> t = (W+, b),
    W- = (u~, d) -> t~ = (W-, b~); t-channel top exchange g t -> t,
    s-channel g g -> g, g t t~."""
    masses = {21: 0.0, 1: 0.0, 2: 0.0, 5: 0.0, 6: MT, 24: MW}
    widths = {21: 0.0, 1: 0.0, 2: 0.0, 5: 0.0, 6: GT, 24: GW}
    verts = []

    def sym(a, b, c):
        verts.append((a, b, c))
        if a != b:
            verts.append((b, a, c))

    sym(2, -1, 24)
    sym(-2, 1, -24)
    sym(24, 5, 6)
    sym(-24, -5, -6)
    sym(21, 6, 6)       # t-channel top line: g t* -> t
    sym(21, -6, -6)
    sym(6, -6, 21)      # g* -> t t~
    sym(21, 21, 21)
    if cuts is None:
        cuts = dict(ptmin=20.0, etamax=2.5, deltaR=0.4)
    return make_process(sqrts, TT_FLAVS, masses, widths, verts, smax, cuts)


# ---- integrands ------------------------------------------------------------
def integrand_unit(cuts=None):
    f = abi.Integrand()
    f.kind = abi.F_UNIT
    _set_cuts(f, cuts)
    return f


def _set_cuts(f, cuts):
    """cuts: None, (ptmin, etamax, drmin), or ("kt", R[, sqrt_s_max]) for the
    reference example's kT-jet PreProcess (every parton a jet of radius R)."""
    if cuts is None:
        f.cuts = abi.CUTS_NONE
    elif cuts[0] == "kt":
        f.cuts = abi.CUTS_KT_JETS
        f.drmin = cuts[1]
        f.ptmin = cuts[2] if len(cuts) > 2 else 0.0
        f.etamax = 0.0
    else:
        f.cuts = abi.CUTS_PT_ETA_DR
        f.ptmin, f.etamax, f.drmin = cuts


def integrand_propprod(s0=1.0e4, cuts=(20.0, 2.5, 0.4)):
    f = abi.Integrand()
    f.kind = abi.F_PROPPROD
    f.params[0] = s0
    _set_cuts(f, cuts)
    return f


def resonant_q(seed=0xC0FFEE04, npair=15):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((npair, npair))
    return a.T @ a + 1.0e-3 * np.eye(npair)


def integrand_resonant(q=None, scale=1.0e5, cuts=(20.0, 2.5, 0.4)):
    """Breit-Wigners of W+ (u d~), t (b u d~), W- (u~ d), t~ (b~ u~ d)
    times x^T Q x with x_k = s_ij / scale over the 15 final-state pairs."""
    f = abi.Integrand()
    f.kind = abi.F_RESONANT
    _set_cuts(f, cuts)
    if q is None:
        q = resonant_q()
    bws = [((1 << 3) | (1 << 4), MW, GW), ((1 << 2) | (1 << 3) | (1 << 4), MT, GT),
           ((1 << 6) | (1 << 7), MW, GW), ((1 << 5) | (1 << 6) | (1 << 7), MT, GT)]
    f.params[0] = len(bws)
    for k, (mask, m, g) in enumerate(bws):
        f.params[1 + 3 * k] = mask
        f.params[2 + 3 * k] = m
        f.params[3 + 3 * k] = g
    f.params[31] = scale
    n = q.shape[0]
    f.params[32] = n
    for a in range(n):
        for b in range(n):
            f.params[33 + a * n + b] = q[a, b]
    return f


EE_MZ, EE_GZ, EE_SW2, EE_ALPHA = 91.1876, 2.4952, 0.2312, 1.0 / 128.0
EE_SMIN, EE_SMAX = 60.0 ** 2, 120.0 ** 2


def integrand_eemumu():
    f = abi.Integrand()
    f.kind = abi.F_EEMUMU
    f.cuts = abi.CUTS_NONE
    for i, v in enumerate([EE_MZ, EE_GZ, EE_SW2, EE_ALPHA, EE_SMIN, EE_SMAX]):
        f.params[i] = v
    return f


def eemumu_sigma(s):
    """Total tree-level sigma(s) [pb] of the e+e- -> mu+mu- integrand
    (closed form of the cos theta integral: int (1 + c^2) = 8/3, int c = 0)."""
    cw2 = 1.0 - EE_SW2
    kap = 1.0 / (4.0 * EE_SW2 * cw2)
    ve, ae = -0.5 + 2.0 * EE_SW2, -0.5
    dm = s - EE_MZ ** 2
    den = dm * dm + EE_MZ ** 2 * EE_GZ ** 2
    rechi = kap * s * dm / den
    abschi2 = kap * kap * s * s / den
    a0 = 1.0 + 2.0 * ve * ve * rechi + (ae * ae + ve * ve) ** 2 * abschi2
    return np.pi * EE_ALPHA ** 2 / (2.0 * s) * a0 * (8.0 / 3.0) * GEV2PB


# ---- batched workload: 32 independent jobs ----------------------------------
BATCH_SEED = 0xC0FFEE02


def batched_jobs(seed=BATCH_SEED, njobs=32):
    """The batched workload's jobs: job j is 2 -> 2 + j % 4 all-gluon
    (every s-/t-channel topology, smax = n_out - 1: 2/5/15/52 channels,
    4/7/10/13 VEGAS dimensions), propagator-product integrand with pT / |y| /
    dR cuts, accepted points per iteration floor(10^(4 + 2 u)) with u from
    the seed (log-uniform in [1e4, 1e6]).  Returns [(process, integrand,
    ncalls)]."""
    rng = np.random.default_rng(seed)
    jobs = []
    for j in range(njobs):
        n_out = 2 + j % 4
        jobs.append((gluon_process(n_out, smax=n_out - 1), integrand_propprod(), int(10 ** (4 + 2 * rng.random()))))
    return jobs
