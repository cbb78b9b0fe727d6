// channels.cpp -- ConstructChannels for the device path and the compiler that
// turns each FSMapper topology into the fixed-size op tables the kernels read.
//
// Semantics follow src/Channel/ChannelNode.cc:69-291 (currents, decay chains,
// t-channel sets, decay extension).  The t-channel search is organised
// differently from the reference's explicit stack over ordered current
// sequences: the final state is partitioned into available currents and a
// partition is kept when some ordering of its currents satisfies the
// Model::Combinable chain -- the same set of channels, found without
// enumerating every permutation.
#include "channels.hpp"

#include <algorithm>
#include <cmath>
#include <functional>
#include <stdexcept>

namespace chili_host {

bool TopologyLess::operator()(const Topology& a, const Topology& b) const {     // ChannelNode.cc:308-316
    if (a.tchan.size() != b.tchan.size()) return a.tchan.size() < b.tchan.size();
    ParticleLess lt;
    for (size_t i = 0; i < a.tchan.size(); ++i) {
        if (lt(a.tchan[i], b.tchan[i])) return true;
        if (greater(a.tchan[i], b.tchan[i])) return false;
    }
    return false;
}

double ProcessModel::Mass(int f) const {
    auto it = mass.find((unsigned)std::abs(f));
    if (it == mass.end()) throw std::invalid_argument("model has no mass for pid " + std::to_string(f));
    return it->second;
}
double ProcessModel::Width(int f) const {
    auto it = width.find((unsigned)std::abs(f));
    if (it == width.end()) throw std::invalid_argument("model has no width for pid " + std::to_string(f));
    return it->second;
}
std::set<int> ProcessModel::Combinable(int a, int b) const {
    std::set<int> r;
    for (const auto& v : vertices) if (v[0] == a && v[1] == b) r.insert(v[2]);
    return r;
}
static std::vector<int> combinable_list(const ProcessModel& m, int a, int b) {
    std::vector<int> r;
    for (const auto& v : m.vertices) if (v[0] == a && v[1] == b) r.push_back(v[2]);
    return r;
}

ProcessModel model_from_abi(const chili_process& p, CutTable& cuts) {
    if (p.nflav < 4 || p.nflav > CHILI_MAX_FLAV) throw std::invalid_argument("nflav must be in [4, CHILI_MAX_FLAV]");
    if (p.nmodel < 0 || p.nmodel > CHILI_MAX_MODEL || p.nvertex < 0 || p.nvertex > CHILI_MAX_VERTEX || p.ndeltaR < 0 || p.ndeltaR > CHILI_MAX_DR)
        throw std::invalid_argument("model table sizes out of range");
    ProcessModel m;
    m.sqrts = p.sqrts;
    m.flavs.assign(p.flavs, p.flavs + p.nflav);
    for (int i = 0; i < p.nmodel; ++i) {
        m.mass[(unsigned)std::abs(p.model_pid[i])] = p.model_mass[i];
        m.width[(unsigned)std::abs(p.model_pid[i])] = p.model_width[i];
    }
    for (int i = 0; i < p.nvertex; ++i) m.vertices.push_back({p.vertex[i][0], p.vertex[i][1], p.vertex[i][2]});
    m.smax = p.smax;
    const auto& fl = m.flavs;
    if (p.use_default_cuts) {                                                     // ChannelNode.cc:69-95
        for (int a : fl)
            for (int b : fl) cuts.deltaR[{a, b}] = (std::abs(a) == 6 || std::abs(b) == 6) ? 0. : 0.4;
        for (size_t i = 0; i < fl.size(); ++i) {
            double ms = m.Mass(fl[i]);
            cuts.smin[1u << i] = ms * ms;
            cuts.ptmin[1u << i] = 30;
            cuts.etamax[1u << i] = 99;
            if (i > 1) cuts.sexternal.push_back(ms * ms);
        }
    } else {
        for (int k = 0; k < p.ndeltaR; ++k) cuts.deltaR[{p.deltaR_pid[k][0], p.deltaR_pid[k][1]}] = p.deltaR[k];
        for (size_t i = 0; i < fl.size(); ++i) {
            double ms = m.Mass(fl[i]);
            cuts.smin[1u << i] = p.smin[i];
            cuts.ptmin[1u << i] = p.ptmin[i];
            cuts.etamax[1u << i] = p.etamax[i];
            if (i > 1) cuts.sexternal.push_back(ms * ms);
        }
    }
    return m;
}

static unsigned next_subset_same_size(unsigned v) {   // next larger integer with the same popcount
    unsigned low = v & (0u - v);
    unsigned ripple = v + low;
    unsigned ones = ((v ^ ripple) >> 2) / low;
    return ripple | ones;
}

std::vector<Topology> construct_topologies(const ProcessModel& m, CutTable& cuts) {
    const unsigned nf = (unsigned)m.flavs.size();
    const unsigned full = (1u << nf) - 1;
    const unsigned smax = (unsigned)m.smax;
    std::map<unsigned, std::set<Particle, ParticleLess>> comps;
    std::map<Particle, std::set<Decay, DecayLess>, std::function<bool(const Particle&, const Particle&)>> chain(
        [](const Particle& a, const Particle& b) { return greater(a, b); });
    auto make = [&](unsigned idx, int pid) { Particle q; q.idx = idx; q.pid = pid; q.mass = m.Mass(pid); q.width = m.Width(pid); return q; };

    for (unsigned i = 0; i < nf; ++i) comps[1u << i].insert(make(1u << i, m.flavs[i]));

    // Currents of nset particles from two smaller currents (ChannelNode.cc:114-186).
    for (unsigned nset = 2; nset < nf; ++nset) {
        for (unsigned cur = (1u << nset) - 1; cur < (1u << nf) && (unsigned)__builtin_popcount(cur) - 1 <= smax;
             cur = next_subset_same_size(cur)) {
            std::vector<unsigned> bits;
            for (unsigned i = 0; i < nf; ++i) if (cur & (1u << i)) bits.push_back(1u << i);
            for (unsigned iset = 1; iset < nset; ++iset) {
                // subsets of the first nset-1 members of size iset (the last member
                // always stays in the second half, so each split is seen once)
                for (unsigned sel = (1u << iset) - 1; sel < (1u << (nset - 1)); sel = next_subset_same_size(sel)) {
                    unsigned s1 = 0;
                    for (unsigned i = 0; i < bits.size(); ++i) if ((sel >> i) & 1u) s1 |= bits[i];
                    unsigned s2 = cur ^ s1;
                    if (((s1 | s2) & 3u) != 0) break;            // currents never contain an incoming leg
                    auto i1 = comps.find(s1), i2 = comps.find(s2);
                    if (i1 == comps.end() || i2 == comps.end()) continue;
                    std::set<int> combined;
                    for (const auto& a : i1->second)
                        for (const auto& b : i2->second) {
                            auto t = m.Combinable(a.pid, b.pid);
                            combined.insert(t.begin(), t.end());
                        }
                    if (combined.empty()) continue;
                    for (int e : combined) comps[cur].insert(make(cur, e));
                    cuts.ptmin[cur] = 0;
                    cuts.etamax[cur] = 99;
                    const auto lp = comps[s1], rp = comps[s2];
                    for (const auto& li : lp)
                        for (const auto& ri : rp)
                            for (int e : combinable_list(m, li.pid, ri.pid)) {
                                Particle par = make(cur, e);
                                if (ParticleLess()(li, ri)) chain[par].insert({li, ri});
                                else chain[par].insert({ri, li});
                            }
                }
            }
        }
    }

    // t-channel sets: partitions of the full mask into available currents with
    // s-channel cost sum(|c| - 1) <= smax, kept if some order of the currents
    // passes the Combinable chain started from the first incoming flavour.
    std::vector<unsigned> avail;
    for (const auto& kv : comps) avail.push_back(kv.first);
    const int pid0 = comps[1].begin()->pid;
    std::set<Topology, TopologyLess> chans;
    std::vector<unsigned> blocks;

    std::function<bool(unsigned, const std::set<int>&, std::map<std::pair<unsigned, std::set<int>>, bool>&)> orderable =
        [&](unsigned used, const std::set<int>& pids, std::map<std::pair<unsigned, std::set<int>>, bool>& memo) -> bool {
        if (used == (1u << blocks.size()) - 1) return true;
        auto key = std::make_pair(used, pids);
        auto it = memo.find(key);
        if (it != memo.end()) return it->second;
        bool ok = false;
        for (unsigned b = 0; b < blocks.size() && !ok; ++b) {
            if (used & (1u << b)) continue;
            std::set<int> nxt;
            for (int p : pids)
                for (const auto& ci : comps[blocks[b]]) {
                    auto t = m.Combinable(p, ci.pid);
                    nxt.insert(t.begin(), t.end());
                }
            if (!nxt.empty()) ok = orderable(used | (1u << b), nxt, memo);
        }
        memo[key] = ok;
        return ok;
    };
    std::function<void(unsigned, unsigned)> partition = [&](unsigned rest, unsigned cost) {
        if (rest == 0) {
            std::map<std::pair<unsigned, std::set<int>>, bool> memo;
            if (!orderable(0, std::set<int>{pid0}, memo)) return;
            std::vector<Topology> cur(1);
            for (unsigned c : blocks) {
                if (c == 1 || c == 2) continue;
                std::vector<Topology> nw;
                for (const auto& t : cur)
                    for (const auto& ci : comps[c]) {
                        Topology t2 = t;
                        t2.tchan.push_back(make(c, ci.pid));
                        nw.push_back(t2);
                    }
                cur.swap(nw);
            }
            for (auto& t : cur) {
                std::sort(t.tchan.begin(), t.tchan.end(), [](const Particle& a, const Particle& b) { return greater(a, b); });
                chans.insert(t);
            }
            return;
        }
        unsigned low = rest & (0u - rest);
        for (unsigned c : avail) {
            if (!(c & low) || (c & ~rest)) continue;
            unsigned cc = cost + (unsigned)__builtin_popcount(c) - 1;
            if (cc > smax) continue;
            blocks.push_back(c);
            partition(rest & ~c, cc);
            blocks.pop_back();
        }
    };
    partition(full, 0);

    // Add the decays of every decaying current, parents first (ChannelNode.cc:261-281).
    for (const auto& dk : chain) {
        std::set<Topology, TopologyLess> nw;
        for (const auto& t : chans) {
            bool add = false;
            for (const auto& q : t.tchan) add |= same(q, dk.first);
            for (const auto& pr : t.decays) add |= same(pr.second.first, dk.first) || same(pr.second.second, dk.first);
            if (add) {
                for (const auto& prods : dk.second) {
                    Topology t2 = t;
                    t2.decays[dk.first] = prods;
                    nw.insert(t2);
                }
            } else {
                nw.insert(t);
            }
        }
        chans.swap(nw);
    }
    return std::vector<Topology>(chans.begin(), chans.end());
}

static double smax_of(double sqrts, const CutTable& c, unsigned idx) {            // Cuts.hh:21-29
    double r = sqrts;
    for (unsigned i = 0; i < c.sexternal.size(); ++i)
        if (!(idx & (1u << (i + 2)))) r -= std::sqrt(c.sexternal[i]);
    return r * r;
}

ChanProg compile_channel(const Topology& t, const CutTable& cuts, double sqrts, int ntot) {
    ChanProg p{};
    std::map<unsigned, int> slot;
    for (int i = 0; i < ntot; ++i) slot[1u << i] = i;
    int k = 0;
    for (const auto& dk : t.decays) slot[dk.first.idx] = ntot + k++;
    p.nd = (int)t.decays.size();
    p.nt = (int)t.tchan.size();
    if (p.nd > kMaxDec || p.nt > kMaxT || p.nt < 1) throw std::invalid_argument("channel too large for the device tables");
    auto at = [](const std::unordered_map<unsigned, double>& mp, unsigned idx) {
        auto it = mp.find(idx);
        if (it == mp.end()) throw std::out_of_range("cut table has no entry for current " + std::to_string(idx));
        return it->second;
    };
    k = 0;
    for (const auto& dk : t.decays) {                                           // GenDecays / WgtDecays, Channel.cc:63-113
        const Particle &a = dk.second.first, &b = dk.second.second;
        p.dec_par[k] = slot.at(dk.first.idx);
        p.dec_c1[k] = slot.at(a.idx);
        p.dec_c2[k] = slot.at(b.idx);
        p.dec_massive[k] = !(std::abs(dk.first.mass) < 1e-16 || dk.first.width < 1e-16);
        p.dec_mass[k] = dk.first.mass;
        p.dec_width[k] = dk.first.width;
        auto dr_it = cuts.deltaR.find({a.pid, b.pid});
        double dr = dr_it == cuts.deltaR.end() ? 0 : dr_it->second;
        p.dec_smin2[k] = at(cuts.ptmin, a.idx) * at(cuts.ptmin, b.idx) * 2 / M_PI / M_PI * (dr * dr);
        auto sm = cuts.smin.find(a.idx | b.idx);
        p.dec_smincut[k] = sm == cuts.smin.end() ? 0.0 : sm->second;
        p.dec_smax[k] = smax_of(sqrts, cuts, dk.first.idx);
        ++k;
    }
    for (int i = 0; i < p.nt; ++i) {                                            // FSMapper ctor, Channel.hh:18-30
        const Particle& q = t.tchan[i];
        p.t_slot[i] = slot.at(q.idx);
        double v = at(cuts.ptmin, q.idx);
        p.t_ptmin[i] = v != 0 ? v : 1.0;
        p.t_etacut[i] = at(cuts.etamax, q.idx);
    }
    // s-channel decays in the recursion order of GenSChan/DecayParts (Channel.cc:204-256)
    int ns = 0;
    std::function<void(const Particle&)> rec = [&](const Particle& ab) {
        auto it = t.decays.find(ab);
        if (it == t.decays.end()) throw std::invalid_argument("topology lacks a decay for a composite current");
        const Particle &a = it->second.first, &b = it->second.second;
        p.s_par[ns] = slot.at(ab.idx);
        p.s_a[ns] = slot.at(a.idx);
        p.s_b[ns] = slot.at(b.idx);
        ++ns;
        if (__builtin_popcount(a.idx) != 1) rec(a);
        if (__builtin_popcount(b.idx) != 1) rec(b);
    };
    for (const auto& q : t.tchan)
        if (__builtin_popcount(q.idx) != 1) rec(q);
    if (ns != p.nd) throw std::invalid_argument("decays not reachable from the t-channel objects");
    return p;
}

}  // namespace chili_host
