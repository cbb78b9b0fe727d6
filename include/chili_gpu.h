/*
 * chili_gpu.h -- C-ABI of the MI355X (gfx950) implementation of Chili's
 * batched hot path: VEGAS grid sampling and adaptation together with the
 * multichannel s-/t-channel phase-space mappings.
 *
 * Every entry point is `extern "C"`, takes plain pointers and sizes, returns
 * an int status (CHILI_OK == 0, negative on error; chili_last_error() gives
 * the message) and never lets a C++ exception cross the boundary.  Host
 * buffers are owned by the caller, device buffers by the library unless a
 * parameter says "device pointer".  Calls on one context are serialised and
 * stream ordered; use one context per host thread (the reference's FSMapper
 * is not re-entrant, include/Channel/Channel.hh:69-72).
 *
 * Reference interfaces replaced (paths relative to the jxi24/Chili tree):
 *   chili_construct_channels  <- chili::ConstructChannels      src/Channel/ChannelNode.cc:69-291
 *   chili_mc_*                <- chili::MultiChannel operator()/Optimize over an
 *                                Integrand<FourVector> of FSMapper channels
 *                                include/Channel/MultiChannel.hh:141-268,
 *                                include/Channel/Integrand.hh:82-123
 *   chili_mc_points           <- the per-point pair Mapper::GeneratePoint /
 *                                GenerateWeight as the batch FFI uses it
 *                                include/Interfaces/nanobind.hh:23-50,
 *                                src/Interfaces/Tensorflow.cc:22-25
 *   chili_vegas_*             <- chili::Vegas::operator()/Optimize  src/Integrator/Vegas.cc:18-69
 *   chili_grid_adapt/_split   <- chili::AdaptiveMap::Adapt/Split   src/Integrator/AdaptiveMap.cc:67-160
 *   chili_alpha_adapt         <- chili::MultiChannel::Adapt         src/Channel/MultiChannel.cc:109-136
 *   chili_mc_optimize         <- chili::MultiChannel::Optimize + Train + RefineChannels
 *                                include/Channel/MultiChannel.hh:122-130,182-194,253-268
 *   chili_vegas_optimize      <- chili::Vegas::Optimize + Refine   src/Integrator/Vegas.cc:54-82
 */
#ifndef CHILI_GPU_H
#define CHILI_GPU_H

#ifndef __HIPCC_RTC__
#include <stddef.h>
#endif
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CHILI_MAX_FLAV   10                        /* 2 incoming + up to 8 outgoing */
#define CHILI_MAX_MODEL  64
#define CHILI_MAX_VERTEX 256
#define CHILI_MAX_DR     256
#define CHILI_MAX_DIMS   (3 * (CHILI_MAX_FLAV - 2) - 2)
#define CHILI_NPARAMS    320

enum {
    CHILI_OK = 0,
    CHILI_EINVAL = -1,   /* bad argument */
    CHILI_ERANGE = -2,   /* size beyond a documented limit / uniforms exhausted */
    CHILI_EHIP = -3,     /* HIP runtime error */
    CHILI_ENOMEM = -4,
    CHILI_ESTATE = -5,   /* call out of order (e.g. no grids set) */
    CHILI_ENODEV = -6    /* no GPU */
};

/* A process: flavours, Model (include/Model/Model.hh:10-29) and Cuts
 * (include/Tools/Cuts.hh:13-29).  Model::Combinable(f1, f2) is the set
 * { v[2] : v in vertex, v[0] == f1, v[1] == f2 } (the reference only ever
 * inserts its result into std::sets, so the order is immaterial). */
typedef struct chili_process {
    double  sqrts;
    int32_t nflav;
    int32_t flavs[CHILI_MAX_FLAV];
    int32_t nmodel;                      /* masses / widths by |pid| */
    int32_t model_pid[CHILI_MAX_MODEL];
    double  model_mass[CHILI_MAX_MODEL];
    double  model_width[CHILI_MAX_MODEL];
    int32_t nvertex;
    int32_t vertex[CHILI_MAX_VERTEX][3];
    /* use_default_cuts != 0: the cuts ConstructChannels(sqrts, flavs, model,
     * smax) builds itself (src/Channel/ChannelNode.cc:69-95).  Otherwise
     * ptmin/etamax/smin per external particle i (current 1<<i), deltaR by
     * ordered pid pair, sexternal = mass^2 of the outgoing flavours. */
    int32_t use_default_cuts;
    double  ptmin[CHILI_MAX_FLAV];
    double  etamax[CHILI_MAX_FLAV];
    double  smin[CHILI_MAX_FLAV];
    int32_t ndeltaR;
    int32_t deltaR_pid[CHILI_MAX_DR][2];
    double  deltaR[CHILI_MAX_DR];
    int32_t smax;                        /* s-channel budget, ChannelNode.hh:87-88 */
} chili_process;

/* ParticleInfo, include/Channel/ChannelNode.hh:26-32 */
typedef struct chili_particle_info {
    uint32_t idx;
    int32_t  pid;
    double   mass;
    double   width;
} chili_particle_info;

/* ChannelDescription, include/Channel/ChannelNode.hh:46-51: t-channel
 * objects in descending ParticleInfo order, decays in ascending order of the
 * decaying current (std::map order). */
typedef struct chili_channel_desc {
    int32_t ninfo;
    chili_particle_info info[CHILI_MAX_FLAV];
    int32_t ndecay;
    chili_particle_info decay[CHILI_MAX_FLAV][3];   /* parent, child 1, child 2 */
} chili_channel_desc;

/* Device integrand + PreProcess cut, selected by id + parameter block (the
 * reference takes user std::functions, include/Channel/Integrand.hh:41-47). */
enum {
    CHILI_F_UNIT = 0,       /* f = 1 */
    CHILI_F_PROPPROD = 1,   /* sum over pairs of final-state pairs of s0^2/(s_ij s_kl) */
    CHILI_F_RESONANT = 2,   /* product of Breit-Wigners x positive quadratic form in s_ij */
    CHILI_F_EEMUMU = 3      /* plain VEGAS: e+e- -> mu+mu- line shape on [0,1]^2 */
};
enum {
    CHILI_CUTS_NONE = 0,
    CHILI_CUTS_PT_ETA_DR = 1, /* x1, x2 <= 1; pT > ptmin, |y| < etamax, dR > drmin on the final state */
    CHILI_CUTS_KT_JETS = 2    /* the reference example's PreProcess (src/example.cc:24-37): finite incoming
                                 momenta with x1, x2 <= 1, (p1 + p2)^2 <= ptmin^2 when ptmin > 0, and the kT
                                 clustering of the final state with radius R = drmin (JetCluster,
                                 src/Tools/JetCluster.cc:8-84) returns one jet per parton */
};
typedef struct chili_integrand {
    int32_t kind;
    int32_t cuts;
    double  ptmin, etamax, drmin;
    double  params[CHILI_NPARAMS];
} chili_integrand;

/* StatsData moments in the reference's field order
 * (include/Integrator/Statistics.hh:140). */
typedef struct chili_stats {
    double n, min, max, sum, sum2, n_finite;
} chili_stats;

/* Where the uniforms come from.  u == NULL: Philox4x32-10 keyed by seed,
 * counter (attempt, dim pair, stream).  Otherwise u is host memory in SoA
 * layout u[k * stride + attempt], k < ndims + 1 (the multichannel stream is
 * the d point coordinates then the channel number, MultiChannel.hh:145-147). */
typedef struct chili_uniforms {
    const double* u;
    int64_t  stride;
    uint64_t seed;
    uint32_t stream;
    uint32_t pad_;
} chili_uniforms;

typedef struct chili_channels chili_channels;
typedef struct chili_mc chili_mc;
typedef struct chili_vegas chili_vegas;
typedef struct chili_flow chili_flow;
typedef struct chili_batch chili_batch;

const char* chili_last_error(void);
int chili_device_count(void);

/* ---- Philox uniforms (shared by the kernels and the host) ---- */
double chili_philox_uniform(uint64_t seed, uint32_t stream, uint64_t attempt, uint32_t k);

/* ---- host setup: ConstructChannels, src/Channel/ChannelNode.cc:98-291 ---- */
int  chili_construct_channels(const chili_process* proc, chili_channels** out);
int  chili_channels_count(const chili_channels* ch);
int  chili_channels_ndims(const chili_channels* ch);
int  chili_channels_get(const chili_channels* ch, int i, chili_channel_desc* out);
void chili_channels_free(chili_channels* ch);

/* ---- multichannel context: MultiChannel + Integrand<FourVector> of FSMapper
 *      channels (include/Channel/MultiChannel.hh:51-139, Integrand.hh:9-133) ---- */
/* nbins: VEGAS bins per dimension, 2..256 (AdaptiveMap(dims, bins),
 * include/Integrator/AdaptiveMap.hh:29-40); grids start uniform, alpha = 1/C
 * (MultiChannel ctor, src/Channel/MultiChannel.cc:62-67). */
int chili_mc_create(const chili_channels* ch, int nbins, const chili_integrand* f, int device, chili_mc** out);
int chili_mc_destroy(chili_mc* m);
int chili_mc_info(const chili_mc* m, int* nchannels, int* ndims, int* nbins, int* ntot);
/* grids [C][d][nbins+1] (edges, non-decreasing; AdaptiveMap storage
 * AdaptiveMap.hh:71-72), alpha [C] (MultiChannel::channel_weights). */
int chili_mc_set_grids(chili_mc* m, const double* grids, int nbins);
int chili_mc_get_grids(const chili_mc* m, double* grids);
int chili_mc_set_alpha(chili_mc* m, const double* alpha);
int chili_mc_get_alpha(const chili_mc* m, double* alpha);
/* One iteration of MultiChannel::operator() (MultiChannel.hh:197-251): attempts
 * from a0 until ncalls pass the cut with a finite value; rejected attempts add
 * a zero to stats.  train != 0: T [C] (alpha training sums, MultiChannel.hh:
 * 168-180) and hist [C][d][nbins] (picked channel's f^2 histograms,
 * Integrand.hh:89-96) are returned for MultiChannel::Train (.hh:182-194).
 * vals_dev (device, optional): value of every kept point, at most vals_cap. */
int chili_mc_iterate(chili_mc* m, const chili_uniforms* uni, uint64_t a0, int64_t ncalls, int train, chili_stats* stats,
                     double* T, double* hist, int64_t* n_attempts, double* vals_dev, int64_t vals_cap);
/* MultiChannel::operator() (MultiChannel.hh:196-251) without training and
 * with an unweighting PostProcess (MultiChannel.hh:240-245): a point of value
 * v that passes the cut is kept with probability min(1, |v| / wmax), decided
 * by the attempt's Philox uniform number d+1 (device-generated uniforms
 * only); a rejected point adds a zero to stats and the loop retries until
 * nevents are kept (a NaN value counts as 0, i.e. one rejection).  Device
 * outputs: ev [4 ntot][cap] FP64 momenta of the kept events in attempt order
 * (external legs, incoming first), ew [cap] event weights
 * sign(v) max(1, |v| / wmax).  cap >= nevents. */
int chili_mc_generate(chili_mc* m, const chili_uniforms* uni, uint64_t a0, int64_t nevents, double wmax, double* ev_dev,
                      double* ew_dev, int64_t cap, chili_stats* stats, int64_t* n_attempts);
/* Batch evaluation on device buffers, the reference's batch FFI
 * (ChiliCPU.evaluate, include/Interfaces/nanobind.hh:23-50; CallChili,
 * src/Interfaces/Tensorflow.cc:22-25): rans [n][ld] row-major FP64, out [n].
 * channel in [0, C): that channel's FSMapper on the randoms as given
 * (ld >= d), out = f * GenerateWeight, 0 when PreProcess fails -- the
 * reference's Evaluate.  channel = -1 (multichannel): column d picks the
 * channel, then VEGAS map, mapper, cut and the multichannel weight as in
 * chili_mc_iterate (ld >= d + 1, randoms in [0, 1)); one value per row.
 * producer_stream: the hipStream_t (NULL: the null stream) on which rans and
 * out were produced / allocated; the library's stream waits for it before
 * reading, and the call returns when out is written. */
int chili_mc_evaluate(chili_mc* m, int channel, const double* rans_dev, int64_t n, int64_t ld, double* out_dev,
                      void* producer_stream);
/* Per-attempt records of attempts [a0, a0+n) (no retry loop): the batch FFI
 * view (include/Interfaces/nanobind.hh:23-50) of MultiChannel::GeneratePoint
 * / GenerateWeight (MultiChannel.hh:141-165).  Host outputs, each optional:
 * status[n] (0 cut, 1 accepted, 2 NaN), pick[n], w[n], f[n], val[n],
 * ps[C][n], x[d][n] (picked channel's inverse point), bins[d][n], mom[ntot*4][n]. */
int chili_mc_points(chili_mc* m, const chili_uniforms* uni, uint64_t a0, int64_t n, int32_t* status, int32_t* pick, double* w,
                    double* f, double* val, double* ps, double* x, int32_t* bins, double* mom);
/* Host integrand and PreProcess: the slow path for arbitrary user functions
 * (the reference's std::function integrand and PreProcess, include/Channel/
 * Integrand.hh:41-47, called in MultiChannel.hh:219-229).  With a callback
 * set, chili_mc_iterate and chili_mc_points hand the momenta of every
 * attempt of a sub-batch that passed the device cut (f->cuts of the context;
 * CHILI_CUTS_NONE passes every attempt) to fn in attempt order, attempts
 * beyond the iteration's last accepted one included: mom [n][ntot][4]
 * (E, px, py, pz per leg, incoming legs first with negative energies, as the
 * reference's FSMapper writes them); fn fills f[n] (the integrand) and
 * pass[n] (the PreProcess decision, 0 rejects) and returns 0, or nonzero to
 * abort the call (CHILI_EINVAL).  The multichannel weight, statistics and
 * training stay on the device.  Sub-batches are limited to 2^16 attempts.
 * The batched, flow, unweighting and batch-evaluation entry points refuse a
 * context with a callback.  fn = NULL restores the device integrand. */
typedef int (*chili_host_integrand_fn)(void* user, int64_t n, int ntot, const double* mom, double* f, int32_t* pass);
int chili_mc_set_host_integrand(chili_mc* m, chili_host_integrand_fn fn, void* user);
/* Run-time specialisation (no reference counterpart): compile the forward
 * and weight kernels with this context's channel programs as compile-time
 * constants (hipRTC, gfx950); enable bit 0 uses the specialised forward
 * kernel (uniform attempts), bit 1 the specialised fused
 * weight kernel, bit 2 the split weight pass (one specialised inverse kernel
 * per channel; on the learned-proposal path, one specialised kernel per
 * channel for the hypercube points and densities of the kept points; takes
 * precedence over bit 1), bit 3 exact early rejection in chili_mc_generate
 * (a passing point is weighted in full only when the bound f / (alpha_c /
 * ps_c / vegas_c) of its picked channel c can pass the accept test; the
 * multichannel weight never exceeds it, so the events are the same), 0 the
 * generic ones; compile_s receives the compile time.  Same results. */
int chili_mc_specialize(chili_mc* m, int enable, double* compile_s);
/* The generated specialised kernel source of a channel set (diagnostics),
 * and its run-time compilation without loading it (no GPU needed). */
int chili_channels_kernel_source(const chili_channels* ch, char* buf, int64_t cap, int64_t* len);
int chili_channels_compile_check(const chili_channels* ch, double* compile_s);
/* Kernel timing on the context's stream (HIP events around the weight and
 * forward kernels); no reference counterpart (measurement only). */
int chili_mc_timing(chili_mc* m, int enable, int reset, double* main_ms, int64_t* launches, int64_t* points, double* fwd_ms);
/* Launches and attempts of the forward kernel that fwd_ms covers (since the
 * last reset); no reference counterpart (measurement only). */
int chili_mc_timing_forward(chili_mc* m, int64_t* launches, int64_t* attempts);

/* ---- batched jobs: many independent MultiChannel::operator() iterations
 *      (include/Channel/MultiChannel.hh:196-251), each job a chili_mc context
 *      (same device and bin count), in one device pass per multiplicity
 *      class with persistent kernels, captured once in a hipGraph ---- */
int chili_batch_create(int njobs, chili_mc* const* jobs, chili_batch** out);
int chili_batch_destroy(chili_batch* b);
/* seed[j], stream[j]: job j's Philox stream (attempts from 0); ncalls[j]
 * accepted points; use_graph: replay the captured pass.  Outputs stats[j],
 * attempts[j] and, when train != 0, T[j] (host [C_j]), hist[j] (host
 * [C_j][d_j][nbins]); the pointer arrays may be NULL. */
int chili_batch_iterate(chili_batch* b, const uint64_t* seed, const uint32_t* stream, const int64_t* ncalls, int train, int use_graph,
                        chili_stats* stats, int64_t* attempts, double* const* T, double* const* hist);
/* launch shape: workgroups per CU of the persistent forward / weight kernels
 * and points per thread (1, 2, 4) of the weight kernel; the autotuner times
 * the candidates on one pass and keeps the fastest (best_ms: forward, weight). */
int chili_batch_set_shape(chili_batch* b, int fwd_wpc, int main_wpc, int main_rnd);
int chili_batch_get_shape(const chili_batch* b, int* fwd_wpc, int* main_wpc, int* main_rnd, int64_t* captures, int64_t* replays,
                          int64_t* topups);
int chili_batch_tune(chili_batch* b, const uint64_t* seed, const uint32_t* stream, const int64_t* ncalls, double* best_ms);

/* ---- plain VEGAS (chili::Vegas, include/Integrator/Vegas.hh:55-92,
 *      src/Integrator/Vegas.cc:18-69); device integrand CHILI_F_EEMUMU ---- */
int chili_vegas_create(int ndims, int nbins, const chili_integrand* f, int device, chili_vegas** out);
int chili_vegas_destroy(chili_vegas* v);
int chili_vegas_set_grid(chili_vegas* v, const double* edges, int nbins);
int chili_vegas_get_grid(const chili_vegas* v, double* edges, int* nbins);
/* One Vegas::operator() pass without its final Adapt (Vegas.cc:18-52):
 * stats and the f^2 histogram hist [d][nbins] for AdaptiveMap::Adapt. */
int chili_vegas_iterate(chili_vegas* v, const chili_uniforms* uni, uint64_t a0, int64_t ncalls, chili_stats* stats, double* hist,
                        double* vals_dev, int64_t vals_cap);
/* chili_vegas_iterate without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory during the call
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments and results (bit for bit); device-generated
 * uniforms only (uni->u must be NULL).  Its totals and per-workgroup
 * partials live in a static 4 MiB device array, one per GPU, held until the
 * process ends and shared by every call on that GPU: calls that overlap on
 * one GPU race.  Limit: 6 + ndims x nbins <= 1024 (always true for the
 * two-dimensional integrand); any ncalls up to 2^31 (processed in chunks). */
int chili_vegas_iterate_static(chili_vegas* v, const chili_uniforms* uni, uint64_t a0, int64_t ncalls, chili_stats* stats,
                               double* hist, double* vals_dev, int64_t vals_cap);

/* ---- optimisation drivers: the reference's Optimize loops over the device
 *      iterations (host control, device work) ---- */
/* MultiChannelParams (include/Channel/MultiChannel.hh:29-49) plus the VEGAS
 * damping of the channels' grids (VegasParams::alpha, Vegas.hh:30-38). */
typedef struct chili_mc_params {
    int64_t ncalls;          /* accepted points per iteration (10000) */
    int64_t niterations;     /* iteration cap of Optimize (10) */
    double  rtol;            /* stop when the pooled relative error <= rtol (1e-5) */
    int64_t nrefine;         /* refine when ++iteration == nrefine (1; 0 never) */
    double  beta;            /* alpha update exponent (0.25) */
    double  min_alpha;       /* alpha clamp (1e-5) */
    double  refine_size;     /* ncalls growth per refinement (1.5) */
    int64_t max_bins;        /* grids below this many bins are split in two (200) */
    int64_t iteration;       /* MultiChannelParams::iteration, the refinement counter (0) */
    int32_t should_optimize; /* train alpha and grids after each iteration (1) */
    int32_t pad_;
    double  vegas_alpha;     /* AdaptiveMap::Adapt damping of the channels' grids (1.0) */
} chili_mc_params;
/* MultiChannel's summary state (MultiChannel.hh:132-138, MultiChannel.cc:69-77,
 * 138-154): the pooled StatsData of every iteration, their count and the
 * smallest max|T_i - T_j| seen (the best weights' criterion). */
typedef struct chili_mc_summary {
    chili_stats sum;         /* starts as StatsData{}: min = DBL_MAX, max = DBL_MIN */
    int64_t nresults;
    double  min_diff;        /* starts at +inf */
} chili_mc_summary;
void chili_mc_params_default(chili_mc_params* p);
void chili_mc_summary_init(chili_mc_summary* s);
/* MultiChannel::Optimize (MultiChannel.hh:253-268): while the pooled relative
 * error > rtol and nresults < niterations: one iteration (chili_mc_iterate,
 * training when should_optimize), then Train -- MultiChannel::Adapt of alpha,
 * AdaptiveMap::Adapt of every channel grid whose histogram is not all zero
 * (Integrand::Train, Integrand.hh:97-104), MaxDifference keeping the new
 * alpha in best_weights when it improves min_diff -- then UpdateResults, and
 * RefineChannels when ++iteration == nrefine (iteration = 0, ncalls *=
 * refine_size truncated, grids with fewer than max_bins bins split in two,
 * MultiChannel.hh:122-130).  Iteration k (= summary->nresults) draws Philox
 * uniforms from stream uni->stream + k, attempts from 0 (device uniforms
 * only).  params and summary are updated in place; best_weights [C] (in/out)
 * holds the reference's best_weights; results [results_cap] (optional)
 * receives iteration k's StatsData at index k.  max_steps > 0 bounds the
 * iterations of this call.  A refinement beyond 256 bins is refused
 * (CHILI_ERANGE) before the iteration that would need it. */
int chili_mc_optimize(chili_mc* m, const chili_uniforms* uni, chili_mc_params* params, chili_mc_summary* summary,
                      double* best_weights, chili_stats* results, int64_t results_cap, int64_t max_steps);
/* VegasParams (include/Integrator/Vegas.hh:30-38) and Vegas::Optimize's
 * refinement counter. */
typedef struct chili_vegas_params {
    int64_t ncalls;          /* points per iteration (10000) */
    int64_t nrefine;         /* Refine when ++irefine == nrefine (5) */
    double  rtol, atol;      /* 1e-4, 1e-4 */
    double  alpha;           /* AdaptiveMap::Adapt damping (1.0) */
    int64_t niterations;     /* 10 */
    int64_t irefine;         /* refinement counter (0) */
} chili_vegas_params;
void chili_vegas_params_default(chili_vegas_params* p);
/* Vegas::Optimize (src/Integrator/Vegas.cc:54-77): while (abs err > atol and
 * rel err > rtol) or nresults < niterations: one Vegas::operator() pass
 * (chili_vegas_iterate) and AdaptiveMap::Adapt of the grid, the pass's
 * StatsData pooled into *sum and stored at results[k]; Refine (Split in two,
 * ncalls *= 2, Vegas.cc:79-82) when ++irefine == nrefine.  Iteration k draws
 * stream uni->stream + k, attempts from 0.  sum / nresults / params are
 * updated in place.  max_steps >= 1 bounds this call (the reference loop
 * need not end). */
int chili_vegas_optimize(chili_vegas* v, const chili_uniforms* uni, chili_vegas_params* params, chili_stats* sum, int64_t* nresults,
                         chili_stats* results, int64_t results_cap, int64_t max_steps);

/* ---- host-side adaptation (small, serial; the reference's own steps) ---- */
/* AdaptiveMap::Adapt (src/Integrator/AdaptiveMap.cc:67-125), in place on edges [dims][bins+1]. */
int    chili_grid_adapt(int dims, int bins, double* edges, double alpha, const double* data);
/* AdaptiveMap::Split (AdaptiveMap.cc:127-160): out [dims][nsplit*bins+1]. */
int    chili_grid_split(int dims, int bins, const double* edges, int nsplit, double* out);
/* MultiChannel::Adapt (src/Channel/MultiChannel.cc:109-136), in place on alpha [n]. */
int    chili_alpha_adapt(int n, double* alpha, const double* train, double beta, double min_alpha);
/* MultiChannel::MaxDifference (src/Channel/MultiChannel.cc:138-154). */
double chili_max_difference(int n, const double* train);

/* ---- checkpoints: results.chili, the reference's file (MultiChannel::SaveAs /
 *      LoadFrom, include/Channel/MultiChannel.hh:65-86), in its byte layout ---- */
/* MultiChannel's own part of the file: MultiChannelParams
 * (MultiChannel.hh:29-49, serialised MultiChannel.cc:48-59), min_diff and the
 * number of StatsData entries of its summary (MultiChannel.cc:21-33). */
typedef struct chili_checkpoint_state {
    uint64_t ncalls, niterations;
    double   rtol;
    uint64_t nrefine;
    double   beta, min_alpha, refine_size;
    uint64_t max_bins;
    double   min_diff;
    int64_t  nresults;
} chili_checkpoint_state;
/* Writes results.chili for the channels of ch: per channel its grid
 * (AdaptiveMap::Serialize, AdaptiveMap.cc:23-32; grids [C][d][nbins+1]) and
 * its FSMapper (Channel.cc:7-20, ChannelNode.cc:35-48), then the MultiChannel
 * (MultiChannel.cc:94-107) with best_weights [C] and results [nresults]. */
int chili_checkpoint_write(const chili_channels* ch, int nbins, const double* grids, const double* best_weights,
                           const chili_checkpoint_state* st, const chili_stats* results, const char* path);
/* Sizes of a results.chili: channels, dimensions, bins, summary results. */
int chili_checkpoint_info(const char* path, int* nchannels, int* ndims, int* nbins, int64_t* nresults);
/* Reads results.chili (MultiChannel::Load, MultiChannel.hh:82-86) for the
 * channels of ch: the file's channels are matched to ch's by description (the
 * reference's channel order may differ), grids [C][d][nbins+1] and
 * best_weights [C] come back in ch's order (best_weights are the alpha the
 * reference resumes with, MultiChannel.cc:88).  The cuts are ch's: the file
 * holds none (Channel.hh:49); its pT minima must equal ch's. */
int chili_checkpoint_read(const chili_channels* ch, const char* path, double* grids, double* best_weights,
                          chili_checkpoint_state* st, chili_stats* results, int64_t results_cap);

/* The state exchange of the C++ facade (chili_gpu_facade.hh) with the
 * reference's own objects, through their public byte streams. */
/* Channels of an Integrand::Serialize stream (include/Channel/Integrand.hh:
 * 48-55: per channel its AdaptiveMap and FSMapper) matched to ch by
 * description: perm[k] = ch's index of the stream's channel k; grids
 * (optional, [C][d][nbins+1]) receive the stream's grids in ch's order.  The
 * stream's process (sqrts, particle count, pT minima) must be ch's. */
int chili_channels_match(const chili_channels* ch, const void* bytes, int64_t len, int32_t* perm, double* grids, int* nbins);
/* A MultiChannel::Serialize stream (src/Channel/MultiChannel.cc:94-107)
 * read: ndims, MultiChannelParams and min_diff (st; st->nresults = the
 * summary's entries), best_weights (*nbest <= best_cap) and the summary's
 * StatsData (<= results_cap); best / results may be NULL to ask for the
 * counts first. */
int chili_mc_state_read(const void* bytes, int64_t len, int* ndims, chili_checkpoint_state* st, double* best, int best_cap, int* nbest,
                        chili_stats* results, int64_t results_cap);
/* The MultiChannel::Serialize bytes of a state (what MultiChannel::Deserialize,
 * MultiChannel.cc:79-92, reads): written to buf (cap bytes), *len = size;
 * buf = NULL queries the size. */
int chili_mc_state_write(int ndims, const chili_checkpoint_state* st, int nbest, const double* best, const chili_stats* results, void* buf,
                         int64_t cap, int64_t* len);

/* ---- learned proposal: per-channel coupling-layer spline flows on the
 *      VEGAS hypercube.  No counterpart in the reference; the sampler it
 *      feeds replaces MultiChannel::GeneratePoint's uniforms
 *      (MultiChannel.hh:141-150) and its density joins the channel
 *      densities of Integrand::GenerateWeight (Integrand.hh:112-123). ---- */
/* nflows flows of ndims dims, nlayers coupling layers, hidden == 64; the
 * last layer starts at zero (every flow is the identity). */
int chili_flow_create(int ndims, int nlayers, int hidden, int nflows, uint64_t seed, int device, chili_flow** out);
int chili_flow_destroy(chili_flow* f);
int chili_flow_info(const chili_flow* f, int64_t* nparams, int* ndims, int* nlayers, int* nflows);
int chili_flow_mask(const chili_flow* f, int layer, int32_t* B, int* nB);
int chili_flow_get_params(chili_flow* f, int flow, float* out);
int chili_flow_set_params(chili_flow* f, int flow, const float* in);
int chili_flow_get_grads(chili_flow* f, int flow, float* out);
/* Defensive mixture: the proposal of chili_mc_flow_* becomes
 * (1 - lambda) q + lambda (the uniform density on the hypercube; an attempt
 * takes its uniforms as the point when its uniform number d+1 is below
 * lambda), which bounds every weight by 1/lambda times the plain weight.
 * chili_flow_loss_grad / chili_flow_train then minimise the second moment of
 * the weights under that proposal.  0 <= lambda <= 1; default 0.  Needs
 * device-generated uniforms. */
int chili_flow_set_mixture(chili_flow* f, double lambda);
/* Training loss of chili_flow_loss_grad / chili_flow_train: the p-th moment
 * of the weights under the (mixture) proposal, estimated on the training
 * records, L_p = mean w^p / (sum_j beta_j q'_j)^(p-1); p = 2 (default) is
 * the variance, larger p weighs the largest weights more.  For p > 2 the
 * loss returned and differentiated is (1/p) log L_p (same minimiser, bounded
 * gradients).  2 <= p <= 64.  No counterpart in the reference (the learned
 * proposal is new). */
int chili_flow_set_moment(chili_flow* f, int p);
/* Timing of the coupling-layer kernel (no reference counterpart;
 * measurement only): with enable != 0 every k_couple launch of this flow set
 * is bracketed by HIP events on its stream and the rows of each pass are
 * counted on the device.  Outputs (each may be NULL): summed milliseconds,
 * launches, rows of the timed passes and their algorithmic BF16 flops
 * (2 (nA H + 2 H^2 + H nB (3K - 1)) per row and layer); reset != 0 clears the
 * totals after reading them. */
int chili_flow_timing(chili_flow* f, int enable, int reset, double* ms, int64_t* launches, int64_t* rows, double* flops);
/* device pointers: r, z [n][ndims], lq [n] = log q(r) */
int chili_flow_log_density(chili_flow* f, int flow, int64_t n, const double* r_dev, double* lq_dev, double* z_dev);
int chili_flow_sample(chili_flow* f, int flow, int64_t n, double* zr_dev, double* lq_dev);
/* weighted-variance loss mean(w2 / sum_j beta_j q_j(r_j)) and its gradients
 * (device: r [C][n][d], beta [C][n], w2 [n]) */
int chili_flow_loss_grad(chili_flow* f, int64_t n, const double* r_dev, const double* beta_dev, const double* w2_dev, double* loss);
int chili_flow_train(chili_flow* f, int64_t n, const double* r_dev, const double* beta_dev, const double* w2_dev, int64_t batch,
                     int epochs, double lr, uint64_t seed, double* loss_hist);
/* multichannel iteration with the flows' proposals (fl may be NULL: q = 1) */
int chili_mc_flow_iterate(chili_mc* m, chili_flow* fl, const chili_uniforms* uni, uint64_t a0, int64_t ncalls, chili_stats* stats,
                          int64_t* n_attempts, double* vals_dev, int64_t vals_cap);
/* The unweighting stage on the learned proposal: chili_mc_generate's loop,
 * accept rule and outputs (events ev_dev [4 ntot][cap], weights ew_dev
 * [cap], StatsData of the accepted events, attempts) with the weights of
 * chili_mc_flow_iterate.  No reference counterpart beyond
 * chili_mc_generate's (MultiChannel.hh:240-245 with the learned proposal).
 * With chili_mc_specialize bit 3 and no mixture the other channels' maps
 * and flow densities are evaluated only for the points whose picked
 * channel's term can pass the accept test; the events are the same. */
int chili_mc_flow_generate(chili_mc* m, chili_flow* fl, const chili_uniforms* uni, uint64_t a0, int64_t nevents, double wmax,
                           double* ev_dev, double* ew_dev, int64_t cap, chili_stats* stats, int64_t* n_attempts);
/* training records drawn from the current proposal (device outputs) */
int chili_mc_flow_data(chili_mc* m, chili_flow* fl, const chili_uniforms* uni, uint64_t a0, int64_t ncalls, double* r_dev,
                       double* beta_dev, double* w2_dev, int64_t cap, chili_stats* stats, int64_t* n_kept, int64_t* n_attempts);
/* per-attempt records; lq [C][n], r [d][n] host outputs */
int chili_mc_flow_points(chili_mc* m, chili_flow* fl, const chili_uniforms* uni, uint64_t a0, int64_t n, int32_t* status,
                         int32_t* pick, double* w, double* f, double* val, double* lq, double* r);
#ifdef __cplusplus
}
#endif

#endif /* CHILI_GPU_H */
