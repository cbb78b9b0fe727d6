// jit.hpp -- run-time specialisation of the multichannel kernels (jit.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "chanprog.hpp"

namespace chili_jit {

struct Module {
    hipModule_t mod = nullptr;
    hipFunction_t main = nullptr, main_lds = nullptr, forward = nullptr;
    hipFunction_t own = nullptr;         // unweighting stage: bound on the weight from the picked channel
    std::vector<hipFunction_t> inv;      // split weight pass: one inverse kernel per channel
    std::vector<hipFunction_t> geom;     // learned-proposal path: r_j and a_j, one kernel per channel
    Module() = default;
    Module(const Module&) = delete;
    Module& operator=(const Module&) = delete;
    ~Module();
};

// Kernel source with the channel programs as compile-time constants.
std::string kernel_source(const std::vector<chili_host::ChanProg>& progs, int NT);
// hipRTC compile for gfx950 and (load) load; seconds (optional) gets the
// compile time.
void compile(const std::string& src, Module& out, double* seconds, bool load = true, int nchannels = 0);

}  // namespace chili_jit
