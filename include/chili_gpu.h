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
 */
#ifndef CHILI_GPU_H
#define CHILI_GPU_H

#include <stddef.h>
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
    CHILI_CUTS_PT_ETA_DR = 1  /* x1, x2 <= 1; pT > ptmin, |y| < etamax, dR > drmin on the final state */
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

#ifdef __cplusplus
}
#endif

#endif /* CHILI_GPU_H */
