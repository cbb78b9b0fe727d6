// chanprog.hpp -- compiled channel programs: the device form of one
// FSMapper (src/Channel/Channel.cc) as fixed-size tables.  Free of host
// library headers so that run-time compiled kernels can include it.
#pragma once
#include "../../include/chili_gpu.h"

namespace chili_host {

// ---------------------------------------------------------------------------
// Compiled channel: the FSMapper op lists as fixed-size tables (device form).
// Momentum slots: external particle i -> slot i; the k-th decaying current of
// the channel (ascending order) -> slot ntot + k.
// ---------------------------------------------------------------------------
constexpr int kMaxExt = CHILI_MAX_FLAV;
constexpr int kMaxDec = CHILI_MAX_FLAV - 3;
constexpr int kMaxT = CHILI_MAX_FLAV - 2;

struct ChanProg {
    int nd, nt, pad0, pad1;
    int dec_par[kMaxDec], dec_c1[kMaxDec], dec_c2[kMaxDec], dec_massive[kMaxDec];
    double dec_mass[kMaxDec], dec_width[kMaxDec], dec_smin2[kMaxDec], dec_smincut[kMaxDec], dec_smax[kMaxDec];
    int t_slot[kMaxT];
    double t_ptmin[kMaxT], t_etacut[kMaxT];
    int s_par[kMaxDec], s_a[kMaxDec], s_b[kMaxDec], pad2;
};

}  // namespace chili_host
