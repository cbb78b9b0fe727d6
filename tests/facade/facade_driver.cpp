// facade_driver.cpp -- test program for include/chili_gpu_facade.hh.
//
// The facade is written against the reference's MultiChannel, Integrand,
// Channel, Vegas, AdaptiveMap and FourVector.  The reference cannot be built
// here, so this program instantiates the facade with small test doubles that
// have exactly the members the facade uses (cited in the facade's header)
// and the reference's byte layouts for them.  Usage:
//
//   facade_driver <process.bin> <integrand.bin> <results.chili> <mode> <ncalls> <niterations> <nrefine> <seed> <stream>
//
// process.bin / integrand.bin hold a chili_process / chili_integrand struct
// as raw bytes; results.chili supplies the channel descriptions (written by
// chili_checkpoint_write); the test Integrand takes its channels in reversed
// order (so the facade's channel matching is exercised).  mode "device" uses
// the device integrand, "host" the facade's host callback with f = 1 and a
// PreProcess of x1, x2 <= 1 and pT > 20 on every outgoing particle.  Prints
// one JSON object: the reference-order grids, channel weights, summary and
// parameters after chili_gpu::Optimize.
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <functional>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/chili_gpu_facade.hh"

namespace testdouble {

struct FourVector {
    std::array<double, 4> v;
    FourVector(double e, double x, double y, double z) : v{e, x, y, z} {}
    double operator[](size_t i) const { return v[i]; }
};

struct AdaptiveMap {        // AdaptiveMap::Serialize / Deserialize layout: bins, dims, edges
    size_t bins = 0, dims = 0;
    std::vector<double> hist;
    bool Deserialize(std::istream& in) {
        in.read(reinterpret_cast<char*>(&bins), sizeof bins);
        in.read(reinterpret_cast<char*>(&dims), sizeof dims);
        hist.resize((bins + 1) * dims);
        in.read(reinterpret_cast<char*>(hist.data()), (std::streamsize)(sizeof(double) * hist.size()));
        return bool(in);
    }
    bool Serialize(std::ostream& out) const {
        out.write(reinterpret_cast<const char*>(&bins), sizeof bins);
        out.write(reinterpret_cast<const char*>(&dims), sizeof dims);
        out.write(reinterpret_cast<const char*>(hist.data()), (std::streamsize)(sizeof(double) * hist.size()));
        return true;
    }
};

struct Vegas {
    AdaptiveMap grid;
    AdaptiveMap& Grid() { return grid; }
};

template <class T>
struct Channel {
    Vegas integrator;
    std::string mapping;    // Mapper::Serialize's bytes: name, FSMapper description
};

template <class T>
class Integrand {
  public:
    std::vector<Channel<T>> channels;
    std::function<double(const std::vector<T>&)> func;
    std::function<bool(const std::vector<T>&)> pre;
    bool Serialize(std::ostream& out) const {
        const size_t n = channels.size();
        out.write(reinterpret_cast<const char*>(&n), sizeof n);
        for (const auto& c : channels) {
            c.integrator.grid.Serialize(out);
            out.write(c.mapping.data(), (std::streamsize)c.mapping.size());
        }
        return true;
    }
    std::vector<Channel<T>>& Channels() { return channels; }
    double operator()(const std::vector<T>& p) const { return func(p); }
    std::function<bool(const std::vector<T>&)>& PreProcess() { return pre; }
};

struct MultiChannelParams {
    size_t ncalls{10000}, niterations{10};
    double rtol{1e-5};
    size_t nrefine{1};
    double beta{0.25}, min_alpha{1e-5};
    double refine_size{1.5};
    size_t max_bins{200};
    size_t iteration{};
    bool should_optimize{true};
};

// Holds what MultiChannel::Serialize writes, through the library's own
// writer / reader of that layout.
class MultiChannel {
  public:
    MultiChannelParams params;
    size_t ndims = 0;
    std::vector<double> best;
    double min_diff = std::numeric_limits<double>::infinity();
    std::vector<chili_stats> results;
    MultiChannelParams& Parameters() { return params; }
    const MultiChannelParams& Parameters() const { return params; }
    bool Serialize(std::ostream& out) const {
        chili_checkpoint_state st{};
        st.ncalls = params.ncalls;
        st.niterations = params.niterations;
        st.rtol = params.rtol;
        st.nrefine = params.nrefine;
        st.beta = params.beta;
        st.min_alpha = params.min_alpha;
        st.refine_size = params.refine_size;
        st.max_bins = params.max_bins;
        st.min_diff = min_diff;
        st.nresults = (int64_t)results.size();
        int64_t len = 0;
        chili_gpu::check(chili_mc_state_write((int)ndims, &st, (int)best.size(), best.data(), results.data(), nullptr, 0, &len));
        std::string b((size_t)len, '\0');
        chili_gpu::check(chili_mc_state_write((int)ndims, &st, (int)best.size(), best.data(), results.data(), &b[0], len, &len));
        out.write(b.data(), len);
        return true;
    }
    bool Deserialize(std::istream& in) {
        std::string b((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
        chili_checkpoint_state st{};
        int nd = 0, nb = 0;
        best.assign(64, 0.0);
        chili_gpu::check(chili_mc_state_read(b.data(), (int64_t)b.size(), &nd, &st, best.data(), 64, &nb, nullptr, 0));
        results.resize((size_t)st.nresults);
        chili_gpu::check(chili_mc_state_read(b.data(), (int64_t)b.size(), &nd, &st, best.data(), 64, &nb, results.data(), st.nresults));
        best.resize((size_t)nb);
        ndims = (size_t)nd;
        params.ncalls = st.ncalls;
        params.niterations = st.niterations;
        params.rtol = st.rtol;
        params.nrefine = st.nrefine;
        params.beta = st.beta;
        params.min_alpha = st.min_alpha;
        params.refine_size = st.refine_size;
        params.max_bins = st.max_bins;
        min_diff = st.min_diff;
        return true;
    }
};

}  // namespace testdouble

static std::string slurp(const char* path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) { std::fprintf(stderr, "cannot read %s\n", path); std::exit(2); }
    return std::string((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
}

// The channels of a results.chili (its Integrand part: per channel the grid,
// then the mapper's name and description).
template <class T>
static std::vector<testdouble::Channel<T>> channels_of(const std::string& file) {
    size_t off = 0;
    auto take = [&](size_t n) { size_t o = off; off += n; if (off > file.size()) { std::fprintf(stderr, "short file\n"); std::exit(2); } return o; };
    auto size_at = [&](size_t o) { size_t v; std::memcpy(&v, file.data() + o, sizeof v); return v; };
    const size_t C = size_at(take(8));
    std::vector<testdouble::Channel<T>> out(C);
    for (auto& c : out) {
        const size_t bins = size_at(take(8)), dims = size_at(take(8));
        c.integrator.grid.bins = bins;
        c.integrator.grid.dims = dims;
        c.integrator.grid.hist.resize((bins + 1) * dims);
        std::memcpy(c.integrator.grid.hist.data(), file.data() + take(8 * (bins + 1) * dims), 8 * (bins + 1) * dims);
        const size_t m0 = off;
        const size_t nlen = size_at(take(8));
        take(nlen);                                         // "FSMapper"
        const size_t ninfo = size_at(take(8));
        take(24 * ninfo);                                   // ParticleInfo: u32, i32, f64, f64
        const size_t ndec = size_at(take(8));
        take(72 * ndec);
        take(8 + 8 + 8 + 8);                                // sqrts, ptmax, ntot, nout
        const size_t np = size_at(take(8));
        take(8 * np);
        c.mapping = file.substr(m0, off - m0);
    }
    return out;
}

int main(int argc, char** argv) {
    if (argc != 10) {
        std::fprintf(stderr, "usage: %s process.bin integrand.bin results.chili device|host ncalls niterations nrefine seed stream\n", argv[0]);
        return 2;
    }
    using P = testdouble::FourVector;
    chili_gpu::Options opt;
    const std::string pb = slurp(argv[1]), ib = slurp(argv[2]);
    if (pb.size() != sizeof(chili_process) || ib.size() != sizeof(chili_integrand)) { std::fprintf(stderr, "struct size mismatch\n"); return 2; }
    std::memcpy(&opt.process, pb.data(), sizeof(chili_process));
    std::memcpy(&opt.integrand, ib.data(), sizeof(chili_integrand));
    opt.host_callback = std::string(argv[4]) == "host";
    opt.seed = std::strtoull(argv[8], nullptr, 0);
    opt.stream = (uint32_t)std::strtoul(argv[9], nullptr, 0);
    testdouble::Integrand<P> in;
    auto chans = channels_of<P>(slurp(argv[3]));
    in.channels.assign(chans.rbegin(), chans.rend());
    const double sqrts = opt.process.sqrts;
    in.func = [](const std::vector<P>&) { return 1.0; };
    in.pre = [sqrts](const std::vector<P>& p) {
        if (std::isnan(p[0][0]) || !(-2 * p[0][0] <= sqrts) || !(-2 * p[1][0] <= sqrts)) return false;
        for (size_t i = 2; i < p.size(); ++i)
            if (!(std::sqrt(p[i][1] * p[i][1] + p[i][2] * p[i][2]) > 20.0)) return false;
        return true;
    };
    testdouble::MultiChannel mc;
    mc.ndims = (size_t)in.channels[0].integrator.grid.dims;
    mc.params.ncalls = std::strtoull(argv[5], nullptr, 0);
    mc.params.niterations = std::strtoull(argv[6], nullptr, 0);
    mc.params.nrefine = std::strtoull(argv[7], nullptr, 0);
    std::ostringstream log;
    try {
        chili_gpu::Optimize<P>(mc, in, opt, log);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "facade error: %s\n", e.what());
        return 1;
    }
    std::cerr << log.str();
    std::printf("{\"grids\": [");
    for (size_t k = 0; k < in.channels.size(); ++k) {
        const auto& g = in.channels[k].integrator.grid;
        std::printf("%s{\"bins\": %zu, \"dims\": %zu, \"edges\": [", k ? ", " : "", g.bins, g.dims);
        for (size_t i = 0; i < g.hist.size(); ++i) std::printf("%s%.17g", i ? ", " : "", g.hist[i]);
        std::printf("]}");
    }
    std::printf("], \"best\": [");
    for (size_t k = 0; k < mc.best.size(); ++k) std::printf("%s%.17g", k ? ", " : "", mc.best[k]);
    std::printf("], \"min_diff\": %.17g, \"results\": [", mc.min_diff);
    for (size_t k = 0; k < mc.results.size(); ++k) {
        const auto& s = mc.results[k];
        std::printf("%s[%.17g, %.17g, %.17g, %.17g, %.17g, %.17g]", k ? ", " : "", s.n, s.min, s.max, s.sum, s.sum2, s.n_finite);
    }
    std::printf("], \"ncalls\": %zu, \"iteration\": %zu, \"should_optimize\": %d}\n", mc.params.ncalls, mc.params.iteration,
                (int)mc.params.should_optimize);
    return 0;
}
