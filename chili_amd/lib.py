"""ctypes binding of libchili_gpu.so (include/chili_gpu.h).

The product path is the HIP library: loading fails loudly when it is
missing, and nothing here falls back to a CPU implementation.
"""
import ctypes as C
import os

import numpy as np

from . import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libchili_gpu.so")
_lib = None

EXPORTS = [
    "chili_last_error", "chili_device_count", "chili_philox_uniform",
    "chili_construct_channels", "chili_channels_count", "chili_channels_ndims", "chili_channels_get", "chili_channels_free",
    "chili_mc_create", "chili_mc_destroy", "chili_mc_info", "chili_mc_set_grids", "chili_mc_get_grids",
    "chili_mc_set_alpha", "chili_mc_get_alpha", "chili_mc_iterate", "chili_mc_points", "chili_mc_timing", "chili_mc_timing_forward",
    "chili_vegas_create", "chili_vegas_destroy", "chili_vegas_set_grid", "chili_vegas_get_grid", "chili_vegas_iterate",
    "chili_vegas_iterate_static", "chili_mc_set_host_integrand", "chili_channels_match", "chili_mc_state_read",
    "chili_mc_state_write", "chili_mc_optimize", "chili_vegas_optimize", "chili_mc_params_default",
    "chili_mc_summary_init", "chili_vegas_params_default",
    "chili_grid_adapt", "chili_grid_split", "chili_alpha_adapt", "chili_max_difference",
    "chili_mc_flow_iterate", "chili_mc_flow_generate", "chili_mc_flow_data", "chili_mc_flow_points",
    "chili_flow_create", "chili_flow_destroy", "chili_flow_info", "chili_flow_mask", "chili_flow_get_params",
    "chili_flow_set_params", "chili_flow_get_grads", "chili_flow_log_density", "chili_flow_sample", "chili_flow_loss_grad",
    "chili_flow_train", "chili_flow_timing",
    "chili_batch_create", "chili_batch_destroy", "chili_batch_iterate", "chili_batch_set_shape", "chili_batch_get_shape",
    "chili_batch_tune", "chili_mc_specialize", "chili_flow_set_mixture", "chili_flow_set_moment", "chili_mc_generate", "chili_mc_evaluate", "chili_checkpoint_write", "chili_checkpoint_info", "chili_checkpoint_read", "chili_channels_kernel_source", "chili_channels_compile_check",
]


# chili_host_integrand_fn: int (void* user, int64_t n, int ntot, const double* mom, double* f, int32_t* pass)
HOST_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int64, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double),
                      C.POINTER(C.c_int32))


class ChiliError(RuntimeError):
    pass


def load():
    global _lib
    if _lib is not None:
        return _lib
    # PyTorch-ROCm ships its own HIP runtime under the same soname as the
    # system's; whichever loads first serves the whole process, and torch's
    # device code only initialises on its own.  Load torch (when present)
    # before this library so that both share torch's runtime.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not os.path.exists(LIB_PATH):
        raise ChiliError(f"{LIB_PATH} is not built (run `python -m chili_amd.build`); there is no CPU fallback")
    L = C.CDLL(LIB_PATH)
    P, D, I32, I64, U64, U32 = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int32), C.c_int64, C.c_uint64, C.c_uint32
    L.chili_last_error.restype = C.c_char_p
    L.chili_philox_uniform.restype = C.c_double
    L.chili_philox_uniform.argtypes = [U64, U32, U64, U32]
    L.chili_construct_channels.argtypes = [C.POINTER(abi.Process), C.POINTER(P)]
    L.chili_channels_count.argtypes = [P]
    L.chili_channels_ndims.argtypes = [P]
    L.chili_channels_get.argtypes = [P, C.c_int, C.POINTER(abi.ChannelDesc)]
    L.chili_channels_free.argtypes = [P]
    L.chili_channels_free.restype = None
    L.chili_mc_create.argtypes = [P, C.c_int, C.POINTER(abi.Integrand), C.c_int, C.POINTER(P)]
    L.chili_mc_destroy.argtypes = [P]
    L.chili_mc_info.argtypes = [P, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.chili_mc_set_grids.argtypes = [P, D, C.c_int]
    L.chili_mc_get_grids.argtypes = [P, D]
    L.chili_mc_set_alpha.argtypes = [P, D]
    L.chili_mc_get_alpha.argtypes = [P, D]
    L.chili_mc_iterate.argtypes = [P, C.POINTER(abi.Uniforms), U64, I64, C.c_int, C.POINTER(abi.Stats), D, D,
                                   C.POINTER(C.c_int64), C.c_void_p, I64]
    L.chili_mc_points.argtypes = [P, C.POINTER(abi.Uniforms), U64, I64, I32, I32, D, D, D, D, D, I32, D]
    L.chili_mc_timing.argtypes = [P, C.c_int, C.c_int, D, C.POINTER(C.c_int64), C.POINTER(C.c_int64), D]
    L.chili_mc_timing_forward.argtypes = [P, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    L.chili_vegas_create.argtypes = [C.c_int, C.c_int, C.POINTER(abi.Integrand), C.c_int, C.POINTER(P)]
    L.chili_vegas_destroy.argtypes = [P]
    L.chili_vegas_set_grid.argtypes = [P, D, C.c_int]
    L.chili_vegas_get_grid.argtypes = [P, D, C.POINTER(C.c_int)]
    L.chili_vegas_iterate.argtypes = [P, C.POINTER(abi.Uniforms), U64, I64, C.POINTER(abi.Stats), D, C.c_void_p, I64]
    L.chili_mc_set_host_integrand.argtypes = [P, HOST_FN, C.c_void_p]
    L.chili_mc_optimize.argtypes = [P, C.POINTER(abi.Uniforms), C.POINTER(abi.MCParams), C.POINTER(abi.MCSummary), D,
                                    C.POINTER(abi.Stats), I64, I64]
    L.chili_vegas_optimize.argtypes = [P, C.POINTER(abi.Uniforms), C.POINTER(abi.VegasParams), C.POINTER(abi.Stats),
                                       C.POINTER(C.c_int64), C.POINTER(abi.Stats), I64, I64]
    L.chili_mc_params_default.argtypes = [C.POINTER(abi.MCParams)]
    L.chili_mc_params_default.restype = None
    L.chili_mc_summary_init.argtypes = [C.POINTER(abi.MCSummary)]
    L.chili_mc_summary_init.restype = None
    L.chili_vegas_params_default.argtypes = [C.POINTER(abi.VegasParams)]
    L.chili_vegas_params_default.restype = None
    L.chili_vegas_iterate_static.argtypes = [P, C.POINTER(abi.Uniforms), U64, I64, C.POINTER(abi.Stats), D, C.c_void_p, I64]
    L.chili_grid_adapt.argtypes = [C.c_int, C.c_int, D, C.c_double, D]
    L.chili_grid_split.argtypes = [C.c_int, C.c_int, D, C.c_int, D]
    L.chili_alpha_adapt.argtypes = [C.c_int, D, D, C.c_double, C.c_double]
    L.chili_max_difference.restype = C.c_double
    L.chili_max_difference.argtypes = [C.c_int, D]
    VP, F = C.c_void_p, C.POINTER(C.c_float)
    L.chili_mc_flow_iterate.argtypes = [P, P, C.POINTER(abi.Uniforms), U64, I64, C.POINTER(abi.Stats), C.POINTER(C.c_int64), VP, I64]
    L.chili_mc_flow_generate.argtypes = [P, P, C.POINTER(abi.Uniforms), U64, I64, C.c_double, VP, VP, I64, C.POINTER(abi.Stats),
                                         C.POINTER(C.c_int64)]
    L.chili_mc_flow_data.argtypes = [P, P, C.POINTER(abi.Uniforms), U64, I64, VP, VP, VP, I64, C.POINTER(abi.Stats),
                                     C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    L.chili_mc_flow_points.argtypes = [P, P, C.POINTER(abi.Uniforms), U64, I64, I32, I32, D, D, D, D, D]
    L.chili_flow_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, U64, C.c_int, C.POINTER(P)]
    L.chili_flow_destroy.argtypes = [P]
    L.chili_flow_info.argtypes = [P, C.POINTER(C.c_int64), C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.chili_flow_mask.argtypes = [P, C.c_int, I32, C.POINTER(C.c_int)]
    L.chili_flow_get_params.argtypes = [P, C.c_int, F]
    L.chili_flow_set_params.argtypes = [P, C.c_int, F]
    L.chili_flow_set_mixture.argtypes = [P, C.c_double]
    L.chili_flow_set_moment.argtypes = [P, C.c_int]
    L.chili_mc_evaluate.argtypes = [P, C.c_int, VP, I64, I64, VP, VP]
    L.chili_checkpoint_write.argtypes = [P, C.c_int, D, D, C.POINTER(abi.CheckpointState), C.POINTER(abi.Stats), C.c_char_p]
    L.chili_checkpoint_info.argtypes = [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                        C.POINTER(C.c_int64)]
    L.chili_checkpoint_read.argtypes = [P, C.c_char_p, D, D, C.POINTER(abi.CheckpointState), C.POINTER(abi.Stats), I64]
    L.chili_mc_generate.argtypes = [P, C.POINTER(abi.Uniforms), U64, I64, C.c_double, VP, VP, I64, C.POINTER(abi.Stats),
                                    C.POINTER(C.c_int64)]
    L.chili_flow_get_grads.argtypes = [P, C.c_int, F]
    L.chili_flow_timing.argtypes = [P, C.c_int, C.c_int, D, C.POINTER(C.c_int64), C.POINTER(C.c_int64), D]
    L.chili_flow_log_density.argtypes = [P, C.c_int, I64, VP, VP, VP]
    L.chili_flow_sample.argtypes = [P, C.c_int, I64, VP, VP]
    L.chili_flow_loss_grad.argtypes = [P, I64, VP, VP, VP, D]
    L.chili_flow_train.argtypes = [P, I64, VP, VP, VP, I64, C.c_int, C.c_double, U64, D]
    L.chili_mc_specialize.argtypes = [P, C.c_int, D]
    L.chili_channels_kernel_source.argtypes = [P, C.c_char_p, C.c_int64, C.POINTER(C.c_int64)]
    L.chili_channels_compile_check.argtypes = [P, D]
    PP = C.POINTER(C.c_void_p)
    L.chili_batch_create.argtypes = [C.c_int, PP, C.POINTER(P)]
    L.chili_batch_destroy.argtypes = [P]
    L.chili_batch_iterate.argtypes = [P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint32), C.POINTER(C.c_int64), C.c_int, C.c_int,
                                      C.POINTER(abi.Stats), C.POINTER(C.c_int64), PP, PP]
    L.chili_batch_set_shape.argtypes = [P, C.c_int, C.c_int, C.c_int]
    L.chili_batch_get_shape.argtypes = [P, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int64),
                                        C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    L.chili_batch_tune.argtypes = [P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint32), C.POINTER(C.c_int64), D]
    _lib = L
    return L


def check(rc):
    if rc != 0:
        raise ChiliError(f"{abi.ERRORS.get(rc, rc)}: {load().chili_last_error().decode()}")


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _i(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def device_count():
    return load().chili_device_count()


class Channels:
    """ConstructChannels (src/Channel/ChannelNode.cc:98-291) on the host."""

    def __init__(self, proc):
        L = load()
        h = C.c_void_p()
        check(L.chili_construct_channels(C.byref(proc), C.byref(h)))
        self.h = h
        self.count = L.chili_channels_count(h)
        self.ndims = L.chili_channels_ndims(h)
        self.ntot = proc.nflav

    def kernel_source(self):
        """The specialised kernel source generated for these channels."""
        n = C.c_int64()
        check(load().chili_channels_kernel_source(self.h, None, 0, C.byref(n)))
        buf = C.create_string_buffer(n.value + 1)
        check(load().chili_channels_kernel_source(self.h, buf, n.value + 1, C.byref(n)))
        return buf.value.decode()

    def compile_check(self):
        """Compile the specialised kernels at run time without loading them
        (no GPU needed); returns the compile time in seconds."""
        t = C.c_double()
        check(load().chili_channels_compile_check(self.h, C.byref(t)))
        return t.value

    def desc(self, i):
        d = abi.ChannelDesc()
        check(load().chili_channels_get(self.h, i, C.byref(d)))
        return d

    def descs(self):
        return [self.desc(i).as_tuple() for i in range(self.count)]

    def __del__(self):
        if getattr(self, "h", None):
            try:
                load().chili_channels_free(self.h)
            except TypeError:       # interpreter shutdown: module globals already cleared
                pass
            self.h = None


class MultiChannelGPU:
    """Device context of one multichannel integrand (chili_mc_*)."""

    def __init__(self, channels, nbins, integrand, device=0):
        L = load()
        h = C.c_void_p()
        check(L.chili_mc_create(channels.h, nbins, C.byref(integrand), device, C.byref(h)))
        self.h = h
        nc, nd, nb, nt = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        check(L.chili_mc_info(h, C.byref(nc), C.byref(nd), C.byref(nb), C.byref(nt)))
        self.nchannels, self.ndims, self.nbins, self.ntot = nc.value, nd.value, nb.value, nt.value

    def __del__(self):
        if getattr(self, "h", None):
            try:
                load().chili_mc_destroy(self.h)
            except TypeError:       # interpreter shutdown: module globals already cleared
                pass
            self.h = None

    def set_grids(self, grids):
        g = np.ascontiguousarray(grids, dtype=np.float64)
        check(load().chili_mc_set_grids(self.h, _d(g), g.shape[-1] - 1))
        self.nbins = g.shape[-1] - 1

    def grids(self):
        g = np.zeros((self.nchannels, self.ndims, self.nbins + 1))
        check(load().chili_mc_get_grids(self.h, _d(g)))
        return g

    def set_alpha(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        check(load().chili_mc_set_alpha(self.h, _d(a)))

    def alpha(self):
        a = np.zeros(self.nchannels)
        check(load().chili_mc_get_alpha(self.h, _d(a)))
        return a

    def iterate(self, uni, ncalls, a0=0, train=True, vals_ptr=None, vals_cap=0):
        st = abi.Stats()
        T = np.zeros(self.nchannels)
        hist = np.zeros((self.nchannels, self.ndims, self.nbins))
        na = C.c_int64()
        check(load().chili_mc_iterate(self.h, C.byref(uni), a0, ncalls, 1 if train else 0, C.byref(st), _d(T), _d(hist),
                                      C.byref(na), vals_ptr, vals_cap))
        return st, T, hist, na.value

    def optimize(self, uni, params, summary, best, max_steps=0):
        """chili_mc_optimize (MultiChannel::Optimize with Train and
        RefineChannels): params (abi.MCParams), summary (abi.MCSummary) and
        best ([C] float64, the reference's best_weights) are updated in place;
        returns the StatsData of this call's iterations."""
        if best.dtype != np.float64 or not best.flags.c_contiguous or best.shape != (self.nchannels,):
            raise ChiliError("best must be a contiguous float64 array of nchannels entries")
        cap = int(params.niterations)
        res = (abi.Stats * max(cap, 1))()
        k0 = summary.nresults
        check(load().chili_mc_optimize(self.h, C.byref(uni), C.byref(params), C.byref(summary), _d(best), res, cap, max_steps))
        nc, nd, nb, nt = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        check(load().chili_mc_info(self.h, C.byref(nc), C.byref(nd), C.byref(nb), C.byref(nt)))
        self.nbins = nb.value
        return [res[k] for k in range(k0, min(summary.nresults, cap))]

    def set_host_integrand(self, fn):
        """Host integrand and PreProcess (chili_mc_set_host_integrand): fn(mom)
        takes the momenta [n][ntot][4] of the attempts that passed the device
        cut and returns (f [n], pass [n] booleans); None restores the device
        integrand."""
        if fn is None:
            check(load().chili_mc_set_host_integrand(self.h, HOST_FN(), None))
            self._host_cb = None
            return

        def cb(user, n, ntot, mom, f, ok):
            try:
                m = np.ctypeslib.as_array(mom, shape=(n, ntot, 4))
                fv, pv = fn(m)
                np.ctypeslib.as_array(f, shape=(n,))[:] = fv
                np.ctypeslib.as_array(ok, shape=(n,))[:] = np.asarray(pv, dtype=np.int32)
                return 0
            except Exception:          # reported as an error code, never across the C boundary
                import traceback
                traceback.print_exc()
                return 1

        self._host_cb = HOST_FN(cb)
        check(load().chili_mc_set_host_integrand(self.h, self._host_cb, None))

    def evaluate(self, rans, channel=0):
        """Batch op of the reference's ChiliCPU.evaluate / CallChili on the
        device (chili_mc_evaluate): rans is a 2-D float64 device tensor (torch,
        or any DLPack producer) with unit column stride, [n][d] for a single
        channel's mapping, [n][d+1] for channel=-1 (multichannel); returns a
        torch tensor [n] on the same device."""
        import torch
        t = rans if isinstance(rans, torch.Tensor) else torch.from_dlpack(rans)
        if t.dtype != torch.float64 or not t.is_cuda or t.dim() != 2 or t.stride(1) != 1:
            raise ChiliError("rans must be a 2-D float64 device tensor with contiguous rows")
        if t.shape[1] < self.ndims + (1 if channel < 0 else 0):
            raise ChiliError(f"rans needs {self.ndims + (1 if channel < 0 else 0)} columns")
        out = torch.empty(t.shape[0], dtype=torch.float64, device=t.device)
        # torch.from_dlpack orders a foreign producer on torch's current
        # stream; the library waits for that stream before reading
        stream = torch.cuda.current_stream(t.device).cuda_stream
        check(load().chili_mc_evaluate(self.h, int(channel), C.c_void_p(t.data_ptr()), t.shape[0], t.stride(0),
                                       C.c_void_p(out.data_ptr()), C.c_void_p(stream)))
        return out

    def generate(self, uni, nevents, wmax, ev_ptr, ew_ptr, cap, a0=0):
        """Unweighted events (chili_mc_generate): device buffers ev [4 ntot][cap],
        ew [cap] (raw pointers); returns (stats, attempts)."""
        st = abi.Stats()
        na = C.c_int64()
        check(load().chili_mc_generate(self.h, C.byref(uni), a0, nevents, wmax, ev_ptr, ew_ptr, cap, C.byref(st), C.byref(na)))
        return st, na.value

    def specialize(self, enable=True):
        """Switch to kernels compiled at run time with this context's channel
        programs as constants (chili_mc_specialize): True / 3 both kernels, 1
        the forward kernel, 2 the fused weight kernel, 4 the split weight pass
        (and the learned-proposal path's per-channel geometry kernels), 8 the
        unweighting stage's exact early rejection, bits combined (5: forward +
        split), False / 0 the generic ones.  Returns the compile time."""
        mode = 3 if enable is True else (0 if enable is False else int(enable))
        t = C.c_double()
        check(load().chili_mc_specialize(self.h, mode, C.byref(t)))
        return t.value

    def timing(self, enable=True, reset=False):
        """Weight and forward kernel times since the last reset (chili_mc_timing,
        chili_mc_timing_forward), then enable / disable and optionally reset."""
        ms, fms = C.c_double(), C.c_double()
        nl, npts, fl, fpts = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int64()
        check(load().chili_mc_timing_forward(self.h, C.byref(fl), C.byref(fpts)))
        check(load().chili_mc_timing(self.h, 1 if enable else 0, 1 if reset else 0, C.byref(ms), C.byref(nl), C.byref(npts),
                                     C.byref(fms)))
        return dict(main_ms=ms.value, launches=nl.value, points=npts.value, forward_ms=fms.value,
                    forward_launches=fl.value, forward_attempts=fpts.value)

    def points(self, uni, a0, n):
        d, c, nt = self.ndims, self.nchannels, self.ntot
        out = dict(status=np.zeros(n, np.int32), pick=np.zeros(n, np.int32), w=np.zeros(n), f=np.zeros(n), val=np.zeros(n),
                   ps=np.zeros((c, n)), x=np.zeros((d, n)), bins=np.zeros((d, n), np.int32), mom=np.zeros((nt * 4, n)))
        check(load().chili_mc_points(self.h, C.byref(uni), a0, n, _i(out["status"]), _i(out["pick"]), _d(out["w"]),
                                     _d(out["f"]), _d(out["val"]), _d(out["ps"]), _d(out["x"]), _i(out["bins"]),
                                     _d(out["mom"])))
        # attempt-major views matching the oracle's layout
        out["ps"] = out["ps"].T.copy()
        out["xpick"] = out.pop("x").T.copy()
        out["bins"] = out["bins"].T.copy()
        out["mom"] = out["mom"].T.reshape(n, nt, 4).copy()
        return out

    # -- learned proposal ------------------------------------------------------
    def flow_iterate(self, flow, uni, ncalls, a0=0, vals_ptr=None, vals_cap=0):
        st = abi.Stats()
        na = C.c_int64()
        check(load().chili_mc_flow_iterate(self.h, flow.h if flow else None, C.byref(uni), a0, ncalls, C.byref(st), C.byref(na),
                                           vals_ptr, vals_cap))
        return st, na.value

    def flow_generate(self, flow, uni, nevents, wmax, ev_ptr, ew_ptr, cap, a0=0):
        """Unweighted events on the learned proposal (chili_mc_flow_generate):
        device buffers ev [4 ntot][cap], ew [cap] (raw pointers); returns
        (stats, attempts)."""
        st = abi.Stats()
        na = C.c_int64()
        check(load().chili_mc_flow_generate(self.h, flow.h if flow else None, C.byref(uni), a0, nevents, wmax, ev_ptr, ew_ptr, cap,
                                            C.byref(st), C.byref(na)))
        return st, na.value

    def flow_data(self, flow, uni, ncalls, r_ptr, beta_ptr, w2_ptr, cap, a0=0):
        st = abi.Stats()
        nk, na = C.c_int64(), C.c_int64()
        check(load().chili_mc_flow_data(self.h, flow.h if flow else None, C.byref(uni), a0, ncalls, r_ptr, beta_ptr, w2_ptr, cap,
                                        C.byref(st), C.byref(nk), C.byref(na)))
        return st, nk.value, na.value

    def flow_points(self, flow, uni, a0, n):
        d, c = self.ndims, self.nchannels
        out = dict(status=np.zeros(n, np.int32), pick=np.zeros(n, np.int32), w=np.zeros(n), f=np.zeros(n), val=np.zeros(n),
                   lq=np.zeros((c, n)), r=np.zeros((d, n)))
        check(load().chili_mc_flow_points(self.h, flow.h if flow else None, C.byref(uni), a0, n, _i(out["status"]),
                                          _i(out["pick"]), _d(out["w"]), _d(out["f"]), _d(out["val"]), _d(out["lq"]),
                                          _d(out["r"])))
        out["lq"] = out["lq"].T.copy()
        out["r"] = out["r"].T.copy()
        return out


class FlowGPU:
    """Per-channel coupling-layer spline flows (chili_flow_*); device buffers
    are passed as raw pointers (e.g. torch tensors' data_ptr())."""

    def __init__(self, ndims, nlayers, nflows, seed=1, hidden=64, device=0):
        L = load()
        h = C.c_void_p()
        check(L.chili_flow_create(ndims, nlayers, hidden, nflows, seed, device, C.byref(h)))
        self.h = h
        npar, nd, nl, nf = C.c_int64(), C.c_int(), C.c_int(), C.c_int()
        check(L.chili_flow_info(h, C.byref(npar), C.byref(nd), C.byref(nl), C.byref(nf)))
        self.nparams, self.ndims, self.nlayers, self.nflows = npar.value, nd.value, nl.value, nf.value

    def __del__(self):
        if getattr(self, "h", None):
            try:
                load().chili_flow_destroy(self.h)
            except TypeError:       # interpreter shutdown: module globals already cleared
                pass
            self.h = None

    def masks(self):
        out = []
        for l in range(self.nlayers):
            b = np.zeros(self.ndims, np.int32)
            nb = C.c_int()
            check(load().chili_flow_mask(self.h, l, _i(b), C.byref(nb)))
            out.append(b[:nb.value].tolist())
        return out

    def params(self, flow):
        p = np.zeros(self.nparams, np.float32)
        check(load().chili_flow_get_params(self.h, flow, p.ctypes.data_as(C.POINTER(C.c_float))))
        return p

    def timing(self, enable=True, reset=False):
        """k_couple launch timing (chili_flow_timing): totals so far."""
        ms, fl = C.c_double(), C.c_double()
        nl, nr = C.c_int64(), C.c_int64()
        check(load().chili_flow_timing(self.h, 1 if enable else 0, 1 if reset else 0, C.byref(ms), C.byref(nl), C.byref(nr),
                                       C.byref(fl)))
        return dict(ms=ms.value, launches=nl.value, rows=nr.value, flops=fl.value)

    def set_mixture(self, lam):
        """Proposal (1 - lam) q + lam (uniform on the hypercube) in the
        multichannel flow path (chili_flow_set_mixture)."""
        check(load().chili_flow_set_mixture(self.h, float(lam)))

    def set_moment(self, p):
        """Training loss: the p-th moment of the weights (chili_flow_set_moment)."""
        check(load().chili_flow_set_moment(self.h, int(p)))

    def set_params(self, flow, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        check(load().chili_flow_set_params(self.h, flow, p.ctypes.data_as(C.POINTER(C.c_float))))

    def grads(self, flow):
        g = np.zeros(self.nparams, np.float32)
        check(load().chili_flow_get_grads(self.h, flow, g.ctypes.data_as(C.POINTER(C.c_float))))
        return g

    def log_density(self, flow, n, r_ptr, lq_ptr, z_ptr=None):
        check(load().chili_flow_log_density(self.h, flow, n, r_ptr, lq_ptr, z_ptr))

    def sample(self, flow, n, zr_ptr, lq_ptr):
        check(load().chili_flow_sample(self.h, flow, n, zr_ptr, lq_ptr))

    def loss_grad(self, n, r_ptr, beta_ptr, w2_ptr):
        L = C.c_double()
        check(load().chili_flow_loss_grad(self.h, n, r_ptr, beta_ptr, w2_ptr, C.byref(L)))
        return L.value

    def train(self, n, r_ptr, beta_ptr, w2_ptr, batch, epochs, lr, seed=7):
        hist = np.zeros(max(epochs, 1))
        check(load().chili_flow_train(self.h, n, r_ptr, beta_ptr, w2_ptr, batch, epochs, lr, seed, _d(hist)))
        return hist[:epochs]


class BatchGPU:
    """Many independent multichannel jobs (MultiChannelGPU contexts) in one
    device pass per multiplicity class (chili_batch_*)."""

    def __init__(self, jobs):
        L = load()
        self.jobs = list(jobs)
        arr = (C.c_void_p * len(self.jobs))(*[j.h.value for j in self.jobs])
        h = C.c_void_p()
        check(L.chili_batch_create(len(self.jobs), arr, C.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            try:
                load().chili_batch_destroy(self.h)
            except TypeError:
                pass
            self.h = None

    def _vecs(self, seeds, streams, ncalls):
        n = len(self.jobs)
        s = np.ascontiguousarray(seeds, dtype=np.uint64)
        st = np.ascontiguousarray(streams, dtype=np.uint32)
        nc = np.ascontiguousarray(ncalls, dtype=np.int64)
        assert len(s) == len(st) == len(nc) == n
        return s, st, nc

    def iterate(self, seeds, streams, ncalls, train=True, graph=True):
        s, st, nc = self._vecs(seeds, streams, ncalls)
        n = len(self.jobs)
        stats = (abi.Stats * n)()
        att = np.zeros(n, np.int64)
        Ts = [np.zeros(j.nchannels) for j in self.jobs]
        hs = [np.zeros((j.nchannels, j.ndims, j.nbins)) for j in self.jobs]
        Tp = (C.c_void_p * n)(*[t.ctypes.data for t in Ts])
        hp = (C.c_void_p * n)(*[h.ctypes.data for h in hs])
        check(load().chili_batch_iterate(self.h, s.ctypes.data_as(C.POINTER(C.c_uint64)), st.ctypes.data_as(C.POINTER(C.c_uint32)),
                                         nc.ctypes.data_as(C.POINTER(C.c_int64)), 1 if train else 0, 1 if graph else 0, stats,
                                         att.ctypes.data_as(C.POINTER(C.c_int64)), Tp, hp))
        return [stats[i] for i in range(n)], att, Ts, hs

    def set_shape(self, fwd_wpc, main_wpc, main_rnd):
        check(load().chili_batch_set_shape(self.h, fwd_wpc, main_wpc, main_rnd))

    def shape(self):
        a, b_, c = C.c_int(), C.c_int(), C.c_int()
        cap, rep, top = C.c_int64(), C.c_int64(), C.c_int64()
        check(load().chili_batch_get_shape(self.h, C.byref(a), C.byref(b_), C.byref(c), C.byref(cap), C.byref(rep), C.byref(top)))
        return dict(fwd_wpc=a.value, main_wpc=b_.value, main_rnd=c.value, captures=cap.value, replays=rep.value, topups=top.value)

    def tune(self, seeds, streams, ncalls):
        s, st, nc = self._vecs(seeds, streams, ncalls)
        best = np.zeros(2)
        check(load().chili_batch_tune(self.h, s.ctypes.data_as(C.POINTER(C.c_uint64)), st.ctypes.data_as(C.POINTER(C.c_uint32)),
                                      nc.ctypes.data_as(C.POINTER(C.c_int64)), _d(best)))
        return dict(self.shape(), forward_ms=best[0], weight_ms=best[1])


class VegasGPU:
    """Device context of plain VEGAS (chili_vegas_*)."""

    def __init__(self, ndims, nbins, integrand, device=0):
        L = load()
        h = C.c_void_p()
        check(L.chili_vegas_create(ndims, nbins, C.byref(integrand), device, C.byref(h)))
        self.h, self.ndims, self.nbins = h, ndims, nbins

    def __del__(self):
        if getattr(self, "h", None):
            try:
                load().chili_vegas_destroy(self.h)
            except TypeError:       # interpreter shutdown: module globals already cleared
                pass
            self.h = None

    def set_grid(self, edges):
        e = np.ascontiguousarray(edges, dtype=np.float64)
        check(load().chili_vegas_set_grid(self.h, _d(e), e.shape[1] - 1))
        self.nbins = e.shape[1] - 1

    def grid(self):
        e = np.zeros((self.ndims, self.nbins + 1))
        nb = C.c_int()
        check(load().chili_vegas_get_grid(self.h, _d(e), C.byref(nb)))
        return e

    def iterate(self, uni, ncalls, a0=0, vals_ptr=None, vals_cap=0):
        st = abi.Stats()
        hist = np.zeros((self.ndims, self.nbins))
        check(load().chili_vegas_iterate(self.h, C.byref(uni), a0, ncalls, C.byref(st), _d(hist), vals_ptr, vals_cap))
        return st, hist

    def optimize(self, uni, params, total, nresults, max_steps):
        """chili_vegas_optimize (Vegas::Optimize with Adapt and Refine):
        params (abi.VegasParams) and total (abi.Stats) updated in place;
        returns ([StatsData of this call's iterations], new nresults)."""
        n = C.c_int64(nresults)
        cap = nresults + max_steps
        res = (abi.Stats * max(cap, 1))()
        check(load().chili_vegas_optimize(self.h, C.byref(uni), C.byref(params), C.byref(total), C.byref(n), res, cap, max_steps))
        nb = C.c_int()
        e = np.zeros(self.ndims * 257 * 2)
        check(load().chili_vegas_get_grid(self.h, _d(e), C.byref(nb)))
        self.nbins = nb.value
        return [res[k] for k in range(nresults, n.value)], n.value

    def iterate_static(self, uni, ncalls, a0=0, vals_ptr=None, vals_cap=0):
        """chili_vegas_iterate_static: the same iteration with no device
        allocation (static scratch; Philox uniforms only)."""
        st = abi.Stats()
        hist = np.zeros((self.ndims, self.nbins))
        check(load().chili_vegas_iterate_static(self.h, C.byref(uni), a0, ncalls, C.byref(st), _d(hist), vals_ptr, vals_cap))
        return st, hist


def mc_params(**kw):
    """chili_mc_params with the reference's defaults (MultiChannelParams), overridden by kw."""
    p = abi.MCParams()
    load().chili_mc_params_default(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def mc_summary():
    s = abi.MCSummary()
    load().chili_mc_summary_init(C.byref(s))
    return s


def vegas_params(**kw):
    """chili_vegas_params with the reference's defaults (VegasParams), overridden by kw."""
    p = abi.VegasParams()
    load().chili_vegas_params_default(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def grid_adapt(edges, data, alpha=1.0):
    e = np.array(edges, dtype=np.float64)
    dd = np.ascontiguousarray(data, dtype=np.float64)
    check(load().chili_grid_adapt(e.shape[0], e.shape[1] - 1, _d(e), alpha, _d(dd)))
    return e


def grid_split(edges, nsplit=2):
    e = np.ascontiguousarray(edges, dtype=np.float64)
    out = np.zeros((e.shape[0], nsplit * (e.shape[1] - 1) + 1))
    check(load().chili_grid_split(e.shape[0], e.shape[1] - 1, _d(e), nsplit, _d(out)))
    return out


# ---- checkpoints: results.chili in the reference's layout -------------------
STATE_DEFAULTS = dict(ncalls=10000, niterations=10, rtol=1e-5, nrefine=1, beta=0.25, min_alpha=1e-5, refine_size=1.5,
                      max_bins=200, min_diff=0.0)   # MultiChannelParams defaults, MultiChannel.hh:38-43


def save_checkpoint(path, channels, grids, best_weights, state=None, results=()):
    """chili_checkpoint_write: grids [C][d][B+1], best_weights [C], state
    (MultiChannelParams fields and min_diff, defaults as the reference's),
    results: StatsData of the iterations (abi.Stats)."""
    g = np.ascontiguousarray(grids, dtype=np.float64)
    w = np.ascontiguousarray(best_weights, dtype=np.float64)
    st = abi.CheckpointState(**{**STATE_DEFAULTS, **(state or {})}, nresults=len(results))
    res = (abi.Stats * max(len(results), 1))(*results)
    check(load().chili_checkpoint_write(channels.h, g.shape[-1] - 1, _d(g), _d(w), C.byref(st), res, str(path).encode()))


def load_checkpoint(path, channels):
    """chili_checkpoint_read: (grids [C][d][B+1] and best_weights [C] in the
    channels' order, state dict, list of StatsData)."""
    L = load()
    nc, nd, nb, nr = C.c_int(), C.c_int(), C.c_int(), C.c_int64()
    check(L.chili_checkpoint_info(str(path).encode(), C.byref(nc), C.byref(nd), C.byref(nb), C.byref(nr)))
    g = np.zeros((nc.value, nd.value, nb.value + 1))
    w = np.zeros(nc.value)
    st = abi.CheckpointState()
    res = (abi.Stats * max(nr.value, 1))()
    check(L.chili_checkpoint_read(channels.h, str(path).encode(), _d(g), _d(w), C.byref(st), res, nr.value))
    state = {k: getattr(st, k) for k, _ in abi.CheckpointState._fields_ if k != "nresults"}
    return g, w, state, [res[i] for i in range(nr.value)]


def alpha_adapt(alpha, train, beta=0.25, min_alpha=1e-5):
    a = np.array(alpha, dtype=np.float64)
    t = np.ascontiguousarray(train, dtype=np.float64)
    check(load().chili_alpha_adapt(len(a), _d(a), _d(t), beta, min_alpha))
    return a


def max_difference(train):
    t = np.ascontiguousarray(train, dtype=np.float64)
    return load().chili_max_difference(len(t), _d(t))


def philox_uniform(seed, stream, attempt, k):
    return load().chili_philox_uniform(seed, stream, attempt, k)
