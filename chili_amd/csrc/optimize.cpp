// optimize.cpp -- MultiChannel::Optimize over the device iterations: the
// reference's outer loop, Train and RefineChannels (include/Channel/
// MultiChannel.hh:122-130,182-194,253-268; src/Channel/MultiChannel.cc:69-77,
// 109-154) on the host, every iteration a chili_mc_iterate on the GPU.  The
// loop state is the caller's (params, summary, best weights), so a run can
// be stepped, checkpointed and resumed like the reference's MultiChannel.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

#include "common.hpp"

namespace {

// StatsData::operator+=(StatsData), include/Integrator/Statistics.hh:106-115
void stats_merge(chili_stats& a, const chili_stats& b) {
    a.n += b.n;
    a.n_finite += b.n_finite;
    a.sum += b.sum;
    a.sum2 += b.sum2;
    a.min = std::min(a.min, b.min);
    a.max = std::max(a.max, b.max);
}

// StatsData::Error() / |Mean()|, Statistics.hh:90,122,125
double stats_rel_error(const chili_stats& s) {
    const double mean = s.sum / s.n;
    const double var = (s.sum2 / s.n - mean * mean) / (s.n - 1);
    return std::sqrt(var) / std::abs(mean);
}

void check(int rc) {
    if (rc != CHILI_OK) throw chili_error(rc, chili_last_error());
}

}  // namespace

extern "C" {

void chili_mc_params_default(chili_mc_params* p) {
    if (!p) return;
    std::memset(p, 0, sizeof *p);
    p->ncalls = 10000;          // MultiChannelParams defaults, MultiChannel.hh:38-44
    p->niterations = 10;
    p->rtol = 1e-5;
    p->nrefine = 1;
    p->beta = 0.25;
    p->min_alpha = 1e-5;
    p->refine_size = 1.5;
    p->max_bins = 200;
    p->iteration = 0;
    p->should_optimize = 1;
    p->vegas_alpha = 1.0;       // VegasParams::alpha_default, Vegas.hh:36
}

void chili_mc_summary_init(chili_mc_summary* s) {
    if (!s) return;
    std::memset(s, 0, sizeof *s);
    s->sum.min = std::numeric_limits<double>::max();    // StatsData's initial min / max, Statistics.hh:141
    s->sum.max = std::numeric_limits<double>::min();
    s->nresults = 0;
    s->min_diff = std::numeric_limits<double>::infinity();
}

int chili_mc_optimize(chili_mc* m, const chili_uniforms* uni, chili_mc_params* p, chili_mc_summary* s, double* best_weights,
                      chili_stats* results, int64_t results_cap, int64_t max_steps) {
    return guarded([&] {
        if (!m || !uni || !p || !s || !best_weights) throw chili_error(CHILI_EINVAL, "null argument");
        if (uni->u) throw chili_error(CHILI_EINVAL, "chili_mc_optimize draws Philox uniforms (uni->u must be NULL)");
        if (p->ncalls <= 0) throw chili_error(CHILI_EINVAL, "ncalls must be positive");
        int C = 0, d = 0, B = 0, ntot = 0;
        check(chili_mc_info(m, &C, &d, &B, &ntot));
        std::vector<double> T(C), hist, grids, alpha(C);
        double rel = std::numeric_limits<double>::max();
        int64_t steps = 0;
        while (rel > p->rtol && s->nresults < p->niterations && (max_steps <= 0 || steps < max_steps)) {
            check(chili_mc_info(m, &C, &d, &B, &ntot));
            const bool refine_next = (p->iteration + 1 == p->nrefine) && B < p->max_bins;
            if (refine_next && 2 * B > 256)
                throw chili_error(CHILI_ERANGE, "refinement would split the grids beyond 256 bins (lower max_bins or nrefine)");
            // MultiChannel::operator() (MultiChannel.hh:196-251)
            const int64_t k = s->nresults;
            chili_uniforms u = *uni;
            u.stream = uni->stream + (uint32_t)k;
            chili_stats st{};
            int64_t attempts = 0;
            hist.assign((size_t)C * d * B, 0.0);
            const int train = p->should_optimize ? 1 : 0;
            check(chili_mc_iterate(m, &u, 0, p->ncalls, train, &st, T.data(), hist.data(), &attempts, nullptr, 0));
            if (train) {
                // MultiChannel::Train (MultiChannel.hh:182-194): Adapt, then
                // every channel's grid (Integrand::Train skips an all-zero
                // histogram), then MaxDifference (best weights after the update)
                check(chili_mc_get_alpha(m, alpha.data()));
                check(chili_alpha_adapt(C, alpha.data(), T.data(), p->beta, p->min_alpha));
                grids.resize((size_t)C * d * (B + 1));
                check(chili_mc_get_grids(m, grids.data()));
                for (int c = 0; c < C; ++c) {
                    const double* h = hist.data() + (size_t)c * d * B;
                    if (std::all_of(h, h + (size_t)d * B, [](double x) { return x == 0; })) continue;
                    check(chili_grid_adapt(d, B, grids.data() + (size_t)c * d * (B + 1), p->vegas_alpha, h));
                }
                check(chili_mc_set_grids(m, grids.data(), B));
                check(chili_mc_set_alpha(m, alpha.data()));
                const double md = chili_max_difference(C, T.data());
                if (md < s->min_diff) {
                    std::copy(alpha.begin(), alpha.end(), best_weights);
                    s->min_diff = md;
                }
            }
            // UpdateResults (MultiChannel.cc:69-77) and the loop's error
            stats_merge(s->sum, st);
            if (results && k < results_cap) results[k] = st;
            s->nresults = k + 1;
            rel = stats_rel_error(s->sum);
            ++steps;
            // RefineChannels (MultiChannel.hh:122-130)
            if (++p->iteration == p->nrefine) {
                p->iteration = 0;
                p->ncalls = (int64_t)((double)p->ncalls * p->refine_size);
                if (B < p->max_bins) {
                    grids.resize((size_t)C * d * (B + 1));
                    check(chili_mc_get_grids(m, grids.data()));
                    std::vector<double> split((size_t)C * d * (2 * B + 1));
                    for (int c = 0; c < C; ++c)
                        check(chili_grid_split(d, B, grids.data() + (size_t)c * d * (B + 1), 2, split.data() + (size_t)c * d * (2 * B + 1)));
                    check(chili_mc_set_grids(m, split.data(), 2 * B));
                }
            }
        }
    });
}

}  // extern "C"
