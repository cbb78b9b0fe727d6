// chili_gpu_facade.hh -- runs the reference's MultiChannel::Optimize /
// operator() over an Integrand<FourVector> of FSMapper channels on the GPU
// (libchili_gpu.so), with the reference's own objects staying authoritative:
// the channels, grids, channel weights, parameters and summary are read from
// them before a run and written back after it, so MultiChannel::Summary,
// SaveAs / LoadFrom and the reference's own event generation keep working.
//
// Header-only C++17; include it after the reference's headers.  A call site
//     integrator.Optimize(integrand);
// becomes
//     chili_gpu::Optimize(integrator, integrand, options);
// with options holding the process (what ConstructChannels was given) and the
// device integrand and cut, or host_callback = true to evaluate the
// Integrand's own std::function integrand and PreProcess on the host.
//
// Reference members used (all public; paths relative to the jxi24/Chili tree):
//   Integrand<T>::Serialize(std::ostream&)             include/Channel/Integrand.hh:48-55
//   Integrand<T>::Channels() -> std::vector<Channel<T>>&  Integrand.hh:76
//   Channel<T>::integrator (Vegas)                     Integrand.hh:10
//   Vegas::Grid() -> AdaptiveMap&                      include/Integrator/Vegas.hh:68
//   AdaptiveMap::Deserialize(std::istream&)            include/Integrator/AdaptiveMap.hh:44,
//                                                      src/Integrator/AdaptiveMap.cc:11-21
//   Integrand<T>::operator()(point), PreProcess()      Integrand.hh:35, 40-41
//   MultiChannel::Serialize / Deserialize              include/Channel/MultiChannel.hh:62-63,
//                                                      src/Channel/MultiChannel.cc:79-107
//   MultiChannel::Parameters() -> MultiChannelParams&  MultiChannel.hh:61 (all fields public,
//                                                      MultiChannel.hh:29-49)
//   MultiChannel() and move assignment                 MultiChannel.hh:55
//   T(double E, double px, double py, double pz)       include/Tools/FourVector.hh:35
//
// State exchange: the channels come from Integrand::Serialize (matched by
// description to the channels chili_construct_channels builds for the
// process, so the reference's channel order is kept); the grids are written
// back with AdaptiveMap::Deserialize; the MultiChannel is rebuilt from the
// bytes MultiChannel::Deserialize reads (ndims, parameters, best weights,
// min_diff, the summary), then the parameters the stream does not hold
// (iteration, should_optimize) are set.  MultiChannel::Deserialize sets the
// channel weights to the best weights (MultiChannel.cc:88), as a run that the
// reference saves and loads again would; the device context keeps the
// current weights.  PostProcess is not called: the unweighting PostProcess
// of the reference's examples is chili_mc_generate.
#ifndef CHILI_GPU_FACADE_HH
#define CHILI_GPU_FACADE_HH

#include <cmath>
#include <cstdint>
#include <iostream>
#include <limits>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "chili_gpu.h"

namespace chili_gpu {

struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

inline void check(int rc) {
    if (rc != CHILI_OK) throw Error(chili_last_error());
}

struct Options {
    chili_process process{};        // what ConstructChannels was given: sqrts, flavours, model, cuts, smax
    chili_integrand integrand{};    // device integrand and cut (unused with host_callback)
    bool host_callback = false;     // evaluate Integrand::operator() and PreProcess() on the host
    int device = 0;
    uint64_t seed = 0xC0FFEE00;     // Philox key; iteration k draws stream + k
    uint32_t stream = 0;
};

// One GPU context bound to a reference MultiChannel and Integrand.
template <class MultiChannelT, class IntegrandT, class PointT>
class Driver {
  public:
    Driver(MultiChannelT& mc, IntegrandT& integrand, const Options& opt) : mc_(mc), in_(integrand), opt_(opt) {
        try {
            pull();
        } catch (...) {
            release();
            throw;
        }
    }
    ~Driver() { release(); }
    Driver(const Driver&) = delete;
    Driver& operator=(const Driver&) = delete;

    // MultiChannel::Optimize (MultiChannel.hh:253-268)
    void Optimize(std::ostream& out = std::cout) {
        while (true) {
            const int64_t before = summary_.nresults;
            run(1);
            if (summary_.nresults == before) break;
            print_iteration(out);
            if (!(rel_error() > params_.rtol)) break;
        }
        push();
    }

    // MultiChannel::operator() (MultiChannel.hh:196-251): one iteration,
    // training as should_optimize says; no refinement (that is Optimize's)
    void operator()() {
        const chili_mc_params saved = params_;
        params_.niterations = summary_.nresults + 1;
        params_.rtol = -1.0;
        params_.nrefine = 0;
        run(1);
        params_.niterations = saved.niterations;
        params_.rtol = saved.rtol;
        params_.nrefine = saved.nrefine;
        params_.iteration = saved.iteration;
        push();
    }

    chili_mc* context() { return ctx_; }

  private:
    // reference -> device
    void pull() {
        check(chili_construct_channels(&opt_.process, &ch_));
        C_ = chili_channels_count(ch_);
        d_ = chili_channels_ndims(ch_);
        std::ostringstream is(std::ios::binary);
        in_.Serialize(is);
        const std::string ib = is.str();
        perm_.assign((size_t)C_, -1);
        int nb = 0;
        check(chili_channels_match(ch_, ib.data(), (int64_t)ib.size(), perm_.data(), nullptr, &nb));
        std::vector<double> grids((size_t)C_ * d_ * (nb + 1));
        check(chili_channels_match(ch_, ib.data(), (int64_t)ib.size(), perm_.data(), grids.data(), &nb));
        check(chili_mc_create(ch_, nb, &opt_.integrand, opt_.device, &ctx_));
        check(chili_mc_set_grids(ctx_, grids.data(), nb));
        if (opt_.host_callback) check(chili_mc_set_host_integrand(ctx_, &Driver::host_cb, this));
        // MultiChannel state: parameters from Parameters(), the rest from its stream
        std::ostringstream ms(std::ios::binary);
        mc_.Serialize(ms);
        const std::string mb = ms.str();
        int ndims = 0, nbest = 0;
        chili_checkpoint_state st{};
        std::vector<double> best((size_t)C_);
        check(chili_mc_state_read(mb.data(), (int64_t)mb.size(), &ndims, &st, best.data(), C_, &nbest, nullptr, 0));
        results_.resize((size_t)st.nresults);
        check(chili_mc_state_read(mb.data(), (int64_t)mb.size(), &ndims, &st, best.data(), C_, &nbest, results_.data(),
                                  st.nresults));
        chili_mc_params_default(&params_);
        const auto& p = mc_.Parameters();
        params_.ncalls = (int64_t)p.ncalls;
        params_.niterations = (int64_t)p.niterations;
        params_.rtol = p.rtol;
        params_.nrefine = (int64_t)p.nrefine;
        params_.beta = p.beta;
        params_.min_alpha = p.min_alpha;
        params_.refine_size = p.refine_size;
        params_.max_bins = (int64_t)p.max_bins;
        params_.iteration = (int64_t)p.iteration;
        params_.should_optimize = p.should_optimize ? 1 : 0;
        chili_mc_summary_init(&summary_);
        for (const chili_stats& r : results_) merge(summary_.sum, r);
        summary_.nresults = (int64_t)results_.size();
        summary_.min_diff = st.min_diff;
        // channel weights: the best weights a loaded MultiChannel resumes
        // with (MultiChannel.cc:88), else 1/C as a new one (MultiChannel.cc:62-67)
        best_.assign((size_t)C_, 1.0 / C_);
        have_best_ = nbest == C_;
        if (have_best_)
            for (int k = 0; k < C_; ++k) best_[(size_t)perm_[(size_t)k]] = best[(size_t)k];
        check(chili_mc_set_alpha(ctx_, best_.data()));
    }

    // device -> reference
    void push() {
        int C = 0, d = 0, B = 0, ntot = 0;
        check(chili_mc_info(ctx_, &C, &d, &B, &ntot));
        std::vector<double> grids((size_t)C * d * (B + 1));
        check(chili_mc_get_grids(ctx_, grids.data()));
        auto& chans = in_.Channels();
        for (int k = 0; k < C; ++k) {
            // AdaptiveMap::Serialize's layout (AdaptiveMap.cc:23-32): bins, dims, edges
            std::ostringstream g(std::ios::binary);
            const size_t bins = (size_t)B, dims = (size_t)d;
            g.write(reinterpret_cast<const char*>(&bins), sizeof bins);
            g.write(reinterpret_cast<const char*>(&dims), sizeof dims);
            const double* e = grids.data() + (size_t)perm_[(size_t)k] * d * (B + 1);
            g.write(reinterpret_cast<const char*>(e), sizeof(double) * (size_t)d * (B + 1));
            std::istringstream gi(g.str(), std::ios::binary);
            chans[(size_t)k].integrator.Grid().Deserialize(gi);
        }
        // the weights MultiChannel::Deserialize installs: the best ones, or
        // (no training yet) the current ones
        std::vector<double> cur((size_t)C), ref((size_t)C);
        check(chili_mc_get_alpha(ctx_, cur.data()));
        const bool best_known = have_best_ || summary_.min_diff < std::numeric_limits<double>::infinity();
        for (int k = 0; k < C; ++k) ref[(size_t)k] = best_known ? best_[(size_t)perm_[(size_t)k]] : cur[(size_t)perm_[(size_t)k]];
        chili_checkpoint_state st{};
        st.ncalls = (uint64_t)params_.ncalls;
        st.niterations = (uint64_t)params_.niterations;
        st.rtol = params_.rtol;
        st.nrefine = (uint64_t)params_.nrefine;
        st.beta = params_.beta;
        st.min_alpha = params_.min_alpha;
        st.refine_size = params_.refine_size;
        st.max_bins = (uint64_t)params_.max_bins;
        st.min_diff = summary_.min_diff;
        st.nresults = (int64_t)results_.size();
        int64_t len = 0;
        check(chili_mc_state_write(d, &st, C, ref.data(), results_.data(), nullptr, 0, &len));
        std::string bytes((size_t)len, '\0');
        check(chili_mc_state_write(d, &st, C, ref.data(), results_.data(), &bytes[0], len, &len));
        // a fresh object: MultiChannelSummary::Deserialize adds to its sum
        mc_ = MultiChannelT();
        std::istringstream in(bytes, std::ios::binary);
        mc_.Deserialize(in);
        mc_.Parameters().iteration = (size_t)params_.iteration;
        mc_.Parameters().should_optimize = params_.should_optimize != 0;
    }

    void run(int64_t steps) {
        chili_uniforms u{};
        u.seed = opt_.seed;
        u.stream = opt_.stream;
        const size_t k0 = results_.size();
        std::vector<chili_stats> res((size_t)(summary_.nresults + steps));
        std::copy(results_.begin(), results_.end(), res.begin());
        check(chili_mc_optimize(ctx_, &u, &params_, &summary_, best_.data(), res.data(), (int64_t)res.size(), steps));
        for (size_t k = k0; k < (size_t)summary_.nresults; ++k) results_.push_back(res[k]);
    }

    double rel_error() const {
        const chili_stats& s = summary_.sum;
        const double mean = s.sum / s.n;
        return std::sqrt((s.sum2 / s.n - mean * mean) / (s.n - 1)) / std::abs(mean);
    }

    void print_iteration(std::ostream& out) const {
        const chili_stats& s = summary_.sum;
        const chili_stats& last = results_.back();
        const double mean = s.sum / s.n;
        out << "iteration " << results_.size() << ": " << mean << " +- " << rel_error() * std::abs(mean) << " ("
            << rel_error() * 100 << " %), points " << last.n_finite << " / " << last.n << "\n";
    }

    // Integrand::PreProcess then Integrand::operator() per point
    // (MultiChannel.hh:219-227) on the host
    static int host_cb(void* user, int64_t n, int ntot, const double* mom, double* f, int32_t* pass) {
        Driver* self = static_cast<Driver*>(user);
        try {
            std::vector<PointT> pt;
            pt.reserve((size_t)ntot);
            for (int64_t i = 0; i < n; ++i) {
                pt.clear();
                const double* m = mom + (size_t)i * ntot * 4;
                for (int j = 0; j < ntot; ++j) pt.emplace_back(m[4 * j], m[4 * j + 1], m[4 * j + 2], m[4 * j + 3]);
                const bool ok = self->in_.PreProcess() ? self->in_.PreProcess()(pt) : true;
                pass[i] = ok ? 1 : 0;
                f[i] = ok ? self->in_(pt) : 0.0;
            }
            return 0;
        } catch (...) {
            return 1;
        }
    }

    static void merge(chili_stats& a, const chili_stats& b) {      // StatsData::operator+=(StatsData)
        a.n += b.n;
        a.n_finite += b.n_finite;
        a.sum += b.sum;
        a.sum2 += b.sum2;
        a.min = b.min < a.min ? b.min : a.min;
        a.max = a.max < b.max ? b.max : a.max;
    }

    void release() {
        if (ctx_) chili_mc_destroy(ctx_);
        if (ch_) chili_channels_free(ch_);
        ctx_ = nullptr;
        ch_ = nullptr;
    }

    MultiChannelT& mc_;
    IntegrandT& in_;
    Options opt_;
    chili_channels* ch_ = nullptr;
    chili_mc* ctx_ = nullptr;
    int C_ = 0, d_ = 0;
    std::vector<int32_t> perm_;          // reference channel k -> device channel
    chili_mc_params params_{};
    chili_mc_summary summary_{};
    std::vector<chili_stats> results_;  // the summary's entries
    std::vector<double> best_;          // device channel order
    bool have_best_ = false;
};

// MultiChannel::Optimize(Integrand<T>&) on the GPU.
template <class PointT, class MultiChannelT, class IntegrandT>
void Optimize(MultiChannelT& mc, IntegrandT& integrand, const Options& opt, std::ostream& out = std::cout) {
    Driver<MultiChannelT, IntegrandT, PointT> drv(mc, integrand, opt);
    drv.Optimize(out);
}

}  // namespace chili_gpu

#endif  // CHILI_GPU_FACADE_HH
