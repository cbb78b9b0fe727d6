// checkpoint.cpp -- results.chili, the reference's checkpoint file, written
// and read byte for byte in its layout (x86-64: size_t is 8 bytes, little
// endian, no padding between fields):
//
//   Integrand::Serialize (include/Channel/Integrand.hh:48-55)
//     size_t C
//     C x Channel::Serialize (Integrand.hh:18-22)
//       AdaptiveMap::Serialize (src/Integrator/AdaptiveMap.cc:23-32):
//         size_t bins, size_t dims, double edges[dims][bins + 1]
//       Mapper::Serialize (include/Channel/Mapper.hh:49-55):
//         size_t len, char name[len] ("FSMapper")
//       FSMapper::_Serialize (src/Channel/Channel.cc:7-20):
//         ChannelDescription::Serialize (src/Channel/ChannelNode.cc:35-48):
//           size_t ninfo, ninfo x ParticleInfo
//           size_t ndecay, ndecay x (parent, child 1, child 2) ParticleInfo
//           ParticleInfo (ChannelNode.cc:19-25): uint32 idx, int32 pid,
//                                                double mass, double width
//         double sqrts, double ptmax, size_t ntot, size_t nout,
//         size_t n, double ptmin[n]
//   MultiChannel::Serialize (src/Channel/MultiChannel.cc:94-107)
//     size_t ndims
//     MultiChannelParams::Serialize (MultiChannel.cc:48-59): size_t ncalls,
//       size_t niterations, double rtol, size_t nrefine, double beta,
//       double min_alpha, double refine_size, size_t max_bins
//     size_t C, double best_weights[C], double min_diff
//     MultiChannelSummary::Serialize (MultiChannel.cc:21-33):
//       size_t n, n x StatsData (6 doubles: n, min, max, sum, sum2, n_finite,
//       include/Integrator/Statistics.hh:80-88), size_t C, double best_weights[C]
//
// The reference stores no cuts with a mapper (Channel.hh:49, Channel.cc:18),
// so a mapper it loads has none; here the cuts always come from the process
// the reader constructs its channels from, and the stored pT minima of the
// t-channel objects are checked against them.
#include <array>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "channels.hpp"
#include "common.hpp"

namespace {

using chili_host::ChannelSet;
using chili_host::Particle;
using chili_host::Topology;

static_assert(sizeof(size_t) == 8, "the checkpoint layout assumes a 64-bit size_t");

class Writer {
  public:
    explicit Writer(std::ostream& out) : out_(out) {}
    template <class T>
    void put(const T& v) { out_.write(reinterpret_cast<const char*>(&v), sizeof(T)); }
    void put_size(size_t v) { put<size_t>(v); }
    void put_particle(const Particle& p) {
        put<uint32_t>(p.idx);
        put<int32_t>(p.pid);
        put<double>(p.mass);
        put<double>(p.width);
    }
    void finish() {
        out_.flush();
        if (!out_) throw chili_error(CHILI_EINVAL, "write failed");
    }

  private:
    std::ostream& out_;
};

class Reader {
  public:
    explicit Reader(std::istream& in) : in_(in) {}
    template <class T>
    T get() {
        T v{};
        in_.read(reinterpret_cast<char*>(&v), sizeof(T));
        if (!in_) throw chili_error(CHILI_EINVAL, "checkpoint ends early");
        return v;
    }
    size_t get_size(size_t limit) {
        const size_t v = get<size_t>();
        if (v > limit) throw chili_error(CHILI_EINVAL, "checkpoint holds an implausible count");
        return v;
    }
    Particle get_particle() {
        Particle p;
        p.idx = get<uint32_t>();
        p.pid = get<int32_t>();
        p.mass = get<double>();
        p.width = get<double>();
        return p;
    }
    std::string get_name() {
        const size_t n = get_size(256);
        std::string s(n, '\0');
        if (n) in_.read(&s[0], (std::streamsize)n);
        if (!in_) throw chili_error(CHILI_EINVAL, "checkpoint ends early");
        return s;
    }

  private:
    std::istream& in_;
};

// One channel as the file holds it.
struct FileChannel {
    size_t bins = 0, dims = 0;
    std::vector<double> edges;
    std::vector<Particle> info;
    std::vector<std::array<Particle, 3>> decays;
    double sqrts = 0, ptmax = 0;
    size_t ntot = 0, nout = 0;
    std::vector<double> ptmin;
};

struct FileState {
    std::vector<FileChannel> channels;
    size_t ndims = 0;
    chili_checkpoint_state st{};
    std::vector<double> best;
    std::vector<chili_stats> results;
};

// FSMapper ctor (include/Channel/Channel.hh:18-30): per t-channel object the
// cut's pT minimum, 1 where there is none.
std::vector<double> mapper_ptmin(const ChannelSet& set, const Topology& t) {
    std::vector<double> v;
    for (const Particle& q : t.tchan) {
        auto it = set.cuts.ptmin.find(q.idx);
        const double c = it == set.cuts.ptmin.end() ? 0.0 : it->second;
        v.push_back(c != 0 ? c : 1.0);
    }
    return v;
}

bool same_particle(const Particle& a, const Particle& b) {
    return a.idx == b.idx && a.pid == b.pid && a.mass == b.mass && a.width == b.width;
}

bool same_channel(const FileChannel& f, const Topology& t) {
    if (f.info.size() != t.tchan.size() || f.decays.size() != t.decays.size()) return false;
    for (size_t i = 0; i < f.info.size(); ++i)
        if (!same_particle(f.info[i], t.tchan[i])) return false;
    size_t k = 0;
    for (const auto& dk : t.decays) {
        const auto& fd = f.decays[k++];
        if (!same_particle(fd[0], dk.first) || !same_particle(fd[1], dk.second.first) || !same_particle(fd[2], dk.second.second))
            return false;
    }
    return true;
}

// Integrand::Serialize (Integrand.hh:48-55): the channels
void read_channels(Reader& r, FileState& fs) {
    const size_t C = r.get_size(1 << 20);
    fs.channels.resize(C);
    for (auto& ch : fs.channels) {
        ch.bins = r.get_size(1 << 16);
        ch.dims = r.get_size(1 << 10);
        ch.edges.resize((ch.bins + 1) * ch.dims);
        for (auto& e : ch.edges) e = r.get<double>();
        const std::string name = r.get_name();
        if (name != "FSMapper") throw chili_error(CHILI_EINVAL, "checkpoint mapper '" + name + "' is not an FSMapper");
        const size_t ninfo = r.get_size(CHILI_MAX_FLAV);
        for (size_t i = 0; i < ninfo; ++i) ch.info.push_back(r.get_particle());
        const size_t ndec = r.get_size(CHILI_MAX_FLAV);
        for (size_t i = 0; i < ndec; ++i) {
            std::array<Particle, 3> d;
            for (auto& p : d) p = r.get_particle();
            ch.decays.push_back(d);
        }
        ch.sqrts = r.get<double>();
        ch.ptmax = r.get<double>();
        ch.ntot = r.get_size(CHILI_MAX_FLAV);
        ch.nout = r.get_size(CHILI_MAX_FLAV);
        const size_t np = r.get_size(CHILI_MAX_FLAV);
        for (size_t i = 0; i < np; ++i) ch.ptmin.push_back(r.get<double>());
    }
}

// MultiChannel::Serialize (MultiChannel.cc:94-107)
void read_multichannel(Reader& r, FileState& fs) {
    fs.ndims = r.get_size(1 << 10);
    chili_checkpoint_state& st = fs.st;
    st.ncalls = r.get<uint64_t>();
    st.niterations = r.get<uint64_t>();
    st.rtol = r.get<double>();
    st.nrefine = r.get<uint64_t>();
    st.beta = r.get<double>();
    st.min_alpha = r.get<double>();
    st.refine_size = r.get<double>();
    st.max_bins = r.get<uint64_t>();
    const size_t nb = r.get_size(1 << 20);
    fs.best.resize(nb);
    for (auto& w : fs.best) w = r.get<double>();
    st.min_diff = r.get<double>();
    const size_t nres = r.get_size(size_t(1) << 32);
    fs.results.resize(nres);
    for (auto& s : fs.results) {
        s.n = r.get<double>();
        s.min = r.get<double>();
        s.max = r.get<double>();
        s.sum = r.get<double>();
        s.sum2 = r.get<double>();
        s.n_finite = r.get<double>();
    }
    st.nresults = (int64_t)nres;
    const size_t nb2 = r.get_size(1 << 20);
    for (size_t i = 0; i < nb2; ++i) (void)r.get<double>();   // the summary's copy of best_weights
}

FileState read_file(const char* path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw chili_error(CHILI_EINVAL, std::string("cannot open ") + path);
    Reader r(in);
    FileState fs;
    read_channels(r, fs);
    read_multichannel(r, fs);
    return fs;
}

// MultiChannel::Serialize's bytes for a state.
void write_multichannel(Writer& w, size_t ndims, const chili_checkpoint_state* st, size_t nbest, const double* best,
                        const chili_stats* results) {
    w.put_size(ndims);
    w.put<uint64_t>(st->ncalls);
    w.put<uint64_t>(st->niterations);
    w.put<double>(st->rtol);
    w.put<uint64_t>(st->nrefine);
    w.put<double>(st->beta);
    w.put<double>(st->min_alpha);
    w.put<double>(st->refine_size);
    w.put<uint64_t>(st->max_bins);
    w.put_size(nbest);
    for (size_t c = 0; c < nbest; ++c) w.put<double>(best[c]);
    w.put<double>(st->min_diff);
    w.put_size((size_t)st->nresults);
    for (int64_t i = 0; i < st->nresults; ++i) {
        const chili_stats& s = results[i];
        for (double v : {s.n, s.min, s.max, s.sum, s.sum2, s.n_finite}) w.put<double>(v);
    }
    w.put_size(nbest);
    for (size_t c = 0; c < nbest; ++c) w.put<double>(best[c]);
}

// Index in the channel set of each file channel (matched by description);
// checks the process, the grids' shape and the pT minima.
std::vector<int> match_channels(const ChannelSet& set, const FileState& fs) {
    const size_t C = set.topo.size(), d = (size_t)set.ndims;
    if (fs.channels.size() != C) throw chili_error(CHILI_EINVAL, "a different number of channels");
    if (C == 0) return {};
    const size_t B = fs.channels[0].bins;
    std::vector<int> used(C, 0), perm(C, -1);
    for (size_t k = 0; k < C; ++k) {
        const FileChannel& f = fs.channels[k];
        if (f.bins != B || f.dims != d) throw chili_error(CHILI_EINVAL, "grids differ in shape");
        if (f.ntot != (size_t)set.ntot || f.nout != (size_t)set.ntot - 2 || f.sqrts != set.model.sqrts)
            throw chili_error(CHILI_EINVAL, "the channels belong to a different process");
        size_t c = C;
        for (size_t j = 0; j < C; ++j)
            if (!used[j] && same_channel(f, set.topo[j])) { c = j; break; }
        if (c == C) throw chili_error(CHILI_EINVAL, "channel " + std::to_string(k) + " matches none of the process's");
        used[c] = 1;
        if (f.ptmin != mapper_ptmin(set, set.topo[c])) throw chili_error(CHILI_EINVAL, "pT minima differ from the cuts");
        perm[k] = (int)c;
    }
    return perm;
}

}  // namespace

extern "C" {

int chili_checkpoint_write(const chili_channels* ch, int nbins, const double* grids, const double* best_weights,
                           const chili_checkpoint_state* st, const chili_stats* results, const char* path) {
    return guarded([&] {
        if (!ch || !grids || !best_weights || !st || !path || (st->nresults > 0 && !results))
            throw chili_error(CHILI_EINVAL, "null argument");
        if (nbins < 1) throw chili_error(CHILI_EINVAL, "nbins must be positive");
        const ChannelSet& set = ch->set;
        const size_t C = set.topo.size(), d = (size_t)set.ndims, B = (size_t)nbins;
        std::ofstream out(path, std::ios::binary);
        if (!out) throw chili_error(CHILI_EINVAL, std::string("cannot open ") + path + " for writing");
        Writer w(out);
        w.put_size(C);
        for (size_t c = 0; c < C; ++c) {
            const Topology& t = set.topo[c];
            w.put_size(B);
            w.put_size(d);
            for (size_t i = 0; i < d * (B + 1); ++i) w.put<double>(grids[c * d * (B + 1) + i]);
            const char* name = "FSMapper";
            w.put_size(std::strlen(name));
            for (const char* p = name; *p; ++p) w.put<char>(*p);
            w.put_size(t.tchan.size());
            for (const Particle& p : t.tchan) w.put_particle(p);
            w.put_size(t.decays.size());
            for (const auto& dk : t.decays) {
                w.put_particle(dk.first);
                w.put_particle(dk.second.first);
                w.put_particle(dk.second.second);
            }
            w.put<double>(set.model.sqrts);
            w.put<double>(set.model.sqrts / 2.0);
            w.put_size((size_t)set.ntot);
            w.put_size((size_t)set.ntot - 2);
            const std::vector<double> ptmin = mapper_ptmin(set, t);
            w.put_size(ptmin.size());
            for (double v : ptmin) w.put<double>(v);
        }
        write_multichannel(w, d, st, C, best_weights, results);
        w.finish();
    });
}

int chili_checkpoint_info(const char* path, int* nchannels, int* ndims, int* nbins, int64_t* nresults) {
    return guarded([&] {
        if (!path) throw chili_error(CHILI_EINVAL, "null path");
        const FileState fs = read_file(path);
        size_t bins = fs.channels.empty() ? 0 : fs.channels[0].bins;
        for (const auto& c : fs.channels)
            if (c.bins != bins) throw chili_error(CHILI_EINVAL, "channels of the checkpoint differ in their bin counts");
        if (nchannels) *nchannels = (int)fs.channels.size();
        if (ndims) *ndims = (int)fs.ndims;
        if (nbins) *nbins = (int)bins;
        if (nresults) *nresults = fs.st.nresults;
    });
}

int chili_checkpoint_read(const chili_channels* ch, const char* path, double* grids, double* best_weights,
                          chili_checkpoint_state* st, chili_stats* results, int64_t results_cap) {
    return guarded([&] {
        if (!ch || !path || !grids || !best_weights || !st) throw chili_error(CHILI_EINVAL, "null argument");
        const ChannelSet& set = ch->set;
        const FileState fs = read_file(path);
        const size_t C = set.topo.size(), d = (size_t)set.ndims;
        if (fs.channels.size() != C) throw chili_error(CHILI_EINVAL, "checkpoint has a different number of channels");
        if (fs.ndims != d) throw chili_error(CHILI_EINVAL, "checkpoint has a different number of dimensions");
        if (fs.best.size() != C) throw chili_error(CHILI_EINVAL, "checkpoint weights do not match its channels");
        if (fs.st.nresults > results_cap) throw chili_error(CHILI_ERANGE, "results buffer too small");
        const size_t B = fs.channels[0].bins;
        // the reference's channel order need not be this build's: match by description
        const std::vector<int> perm = match_channels(set, fs);
        for (size_t k = 0; k < C; ++k) {
            const size_t c = (size_t)perm[k];
            std::copy(fs.channels[k].edges.begin(), fs.channels[k].edges.end(), grids + c * d * (B + 1));
            best_weights[c] = fs.best[k];
        }
        *st = fs.st;
        for (int64_t i = 0; i < fs.st.nresults; ++i) results[i] = fs.results[(size_t)i];
    });
}

// The channels of an Integrand::Serialize byte stream (include/Channel/
// Integrand.hh:48-55) matched to a channel set by description: perm[k] is
// the set's index of the stream's channel k; grids (optional, [C][d][B+1])
// receive the stream's grids in the set's order, nbins their bin count.
int chili_channels_match(const chili_channels* ch, const void* bytes, int64_t len, int32_t* perm, double* grids, int* nbins) {
    return guarded([&] {
        if (!ch || !bytes || len < 0 || !perm) throw chili_error(CHILI_EINVAL, "null argument");
        std::istringstream in(std::string((const char*)bytes, (size_t)len), std::ios::binary);
        Reader r(in);
        FileState fs;
        read_channels(r, fs);
        const std::vector<int> p = match_channels(ch->set, fs);
        const size_t C = p.size(), d = (size_t)ch->set.ndims;
        const size_t B = C ? fs.channels[0].bins : 0;
        for (size_t k = 0; k < C; ++k) {
            perm[k] = p[k];
            if (grids) std::copy(fs.channels[k].edges.begin(), fs.channels[k].edges.end(), grids + (size_t)p[k] * d * (B + 1));
        }
        if (nbins) *nbins = (int)B;
    });
}

// MultiChannel::Deserialize's input (src/Channel/MultiChannel.cc:79-92), read
// from a MultiChannel::Serialize byte stream: ndims, the parameters and
// min_diff in st (st->nresults = the summary's entry count), best weights
// (*nbest of them, at most best_cap) and the summary's results (at most
// results_cap).
int chili_mc_state_read(const void* bytes, int64_t len, int* ndims, chili_checkpoint_state* st, double* best, int best_cap, int* nbest,
                        chili_stats* results, int64_t results_cap) {
    return guarded([&] {
        if (!bytes || len < 0 || !st || !nbest) throw chili_error(CHILI_EINVAL, "null argument");
        std::istringstream in(std::string((const char*)bytes, (size_t)len), std::ios::binary);
        Reader r(in);
        FileState fs;
        read_multichannel(r, fs);
        // a NULL buffer only asks for the counts
        if ((best && (int64_t)fs.best.size() > best_cap) || (results && fs.st.nresults > results_cap))
            throw chili_error(CHILI_ERANGE, "output buffers too small");
        if (ndims) *ndims = (int)fs.ndims;
        *st = fs.st;
        *nbest = (int)fs.best.size();
        if (best) std::copy(fs.best.begin(), fs.best.end(), best);
        if (results) std::copy(fs.results.begin(), fs.results.end(), results);
    });
}

// MultiChannel::Serialize's bytes (src/Channel/MultiChannel.cc:94-107) of a
// state: written to buf (cap bytes); *len receives the size (call with
// buf = NULL to query it).
int chili_mc_state_write(int ndims, const chili_checkpoint_state* st, int nbest, const double* best, const chili_stats* results, void* buf,
                         int64_t cap, int64_t* len) {
    return guarded([&] {
        if (!st || !len || (nbest > 0 && !best) || (st->nresults > 0 && !results) || ndims < 0 || nbest < 0)
            throw chili_error(CHILI_EINVAL, "null argument");
        std::ostringstream out(std::ios::binary);
        Writer w(out);
        write_multichannel(w, (size_t)ndims, st, (size_t)nbest, best, results);
        w.finish();
        const std::string s = out.str();
        *len = (int64_t)s.size();
        if (buf) {
            if (cap < (int64_t)s.size()) throw chili_error(CHILI_ERANGE, "buffer too small");
            std::memcpy(buf, s.data(), s.size());
        }
    });
}

}  // extern "C"
