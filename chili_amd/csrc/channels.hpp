// channels.hpp -- host-side channel topologies and their compiled device form.
#pragma once
#include <array>
#include <map>
#include <set>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "../../include/chili_gpu.h"
#include "chanprog.hpp"

namespace chili_host {

// ParticleInfo and its ordering (idx, mass, width; pid does not take part),
// include/Channel/ChannelNode.hh:26-32 and src/Channel/ChannelNode.cc:293-306.
struct Particle {
    unsigned idx = 0;
    int pid = 0;
    double mass = 0, width = 0;
};
struct ParticleLess {
    bool operator()(const Particle& a, const Particle& b) const {
        if (a.idx != b.idx) return a.idx < b.idx;
        if (a.mass != b.mass) return a.mass < b.mass;
        return a.width < b.width;
    }
};
inline bool same(const Particle& a, const Particle& b) { return a.mass == b.mass && a.idx == b.idx && a.width == b.width; }
inline bool greater(const Particle& a, const Particle& b) { return ParticleLess()(b, a) && !same(a, b); }

using Decay = std::pair<Particle, Particle>;
struct DecayLess {
    bool operator()(const Decay& a, const Decay& b) const {
        ParticleLess lt;
        if (lt(a.first, b.first)) return true;
        if (lt(b.first, a.first)) return false;
        return lt(a.second, b.second);
    }
};

struct Topology {            // ChannelDescription
    std::vector<Particle> tchan;                          // descending
    std::map<Particle, Decay, ParticleLess> decays;       // ascending parent
};
struct TopologyLess {
    bool operator()(const Topology& a, const Topology& b) const;
};

// Cuts as the mappers see them after ConstructChannels has filled in the
// composite currents (include/Tools/Cuts.hh:13-29).
struct CutTable {
    std::vector<double> sexternal;
    std::map<std::pair<int, int>, double> deltaR;
    std::unordered_map<unsigned, double> smin, ptmin, etamax;
};

struct ProcessModel {
    double sqrts = 0;
    std::vector<int> flavs;
    std::unordered_map<unsigned, double> mass, width;
    std::vector<std::array<int, 3>> vertices;
    int smax = 2;
    double Mass(int f) const;
    double Width(int f) const;
    std::set<int> Combinable(int a, int b) const;
};

ProcessModel model_from_abi(const chili_process& p, CutTable& cuts);
std::vector<Topology> construct_topologies(const ProcessModel& m, CutTable& cuts);

struct ChannelSet {
    ProcessModel model;
    CutTable cuts;
    std::vector<Topology> topo;
    int ntot = 0, ndims = 0;
    std::vector<ChanProg> progs;
    std::vector<double> sexternal;   // per external slot (0 for incoming)
};

ChanProg compile_channel(const Topology& t, const CutTable& cuts, double sqrts, int ntot);

}  // namespace chili_host

struct chili_channels {
    chili_host::ChannelSet set;
};
