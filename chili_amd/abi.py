"""ctypes mirror of the structs declared in include/chili_gpu.h.

The layouts here must match the C header field for field; the CPU test
suite checks the sizes against the values the library itself reports.
"""
import ctypes as C

MAX_FLAV = 10
MAX_MODEL = 64
MAX_VERTEX = 256
MAX_DR = 256
MAX_DIMS = 3 * (MAX_FLAV - 2) - 2
NPARAMS = 320

CHILI_OK = 0
ERRORS = {
    -1: "CHILI_EINVAL", -2: "CHILI_ERANGE", -3: "CHILI_EHIP",
    -4: "CHILI_ENOMEM", -5: "CHILI_ESTATE", -6: "CHILI_ENODEV",
}

F_UNIT, F_PROPPROD, F_RESONANT, F_EEMUMU = 0, 1, 2, 3
CUTS_NONE, CUTS_PT_ETA_DR, CUTS_KT_JETS = 0, 1, 2


class Process(C.Structure):
    _fields_ = [
        ("sqrts", C.c_double),
        ("nflav", C.c_int32),
        ("flavs", C.c_int32 * MAX_FLAV),
        ("nmodel", C.c_int32),
        ("model_pid", C.c_int32 * MAX_MODEL),
        ("model_mass", C.c_double * MAX_MODEL),
        ("model_width", C.c_double * MAX_MODEL),
        ("nvertex", C.c_int32),
        ("vertex", (C.c_int32 * 3) * MAX_VERTEX),
        ("use_default_cuts", C.c_int32),
        ("ptmin", C.c_double * MAX_FLAV),
        ("etamax", C.c_double * MAX_FLAV),
        ("smin", C.c_double * MAX_FLAV),
        ("ndeltaR", C.c_int32),
        ("deltaR_pid", (C.c_int32 * 2) * MAX_DR),
        ("deltaR", C.c_double * MAX_DR),
        ("smax", C.c_int32),
    ]


class ParticleInfo(C.Structure):
    _fields_ = [("idx", C.c_uint32), ("pid", C.c_int32), ("mass", C.c_double), ("width", C.c_double)]

    def key(self):
        return (self.idx, self.pid, self.mass, self.width)


class ChannelDesc(C.Structure):
    _fields_ = [
        ("ninfo", C.c_int32),
        ("info", ParticleInfo * MAX_FLAV),
        ("ndecay", C.c_int32),
        ("decay", (ParticleInfo * 3) * MAX_FLAV),
    ]

    def as_tuple(self):
        info = tuple(self.info[i].key() for i in range(self.ninfo))
        dec = tuple(tuple(self.decay[i][k].key() for k in range(3)) for i in range(self.ndecay))
        return info, dec


class Integrand(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("cuts", C.c_int32),
        ("ptmin", C.c_double),
        ("etamax", C.c_double),
        ("drmin", C.c_double),
        ("params", C.c_double * NPARAMS),
    ]


class Stats(C.Structure):
    """StatsData moments (Statistics.hh:140 field order)."""
    _fields_ = [("n", C.c_double), ("min", C.c_double), ("max", C.c_double),
                ("sum", C.c_double), ("sum2", C.c_double), ("n_finite", C.c_double)]

    def as_tuple(self):
        return (self.n, self.min, self.max, self.sum, self.sum2, self.n_finite)

    def mean(self):
        return self.sum / self.n

    def error(self):
        m = self.mean()
        return ((self.sum2 / self.n - m * m) / (self.n - 1)) ** 0.5


class MCParams(C.Structure):
    """chili_mc_params: MultiChannelParams (include/Channel/MultiChannel.hh:29-49)."""
    _fields_ = [("ncalls", C.c_int64), ("niterations", C.c_int64), ("rtol", C.c_double), ("nrefine", C.c_int64),
                ("beta", C.c_double), ("min_alpha", C.c_double), ("refine_size", C.c_double), ("max_bins", C.c_int64),
                ("iteration", C.c_int64), ("should_optimize", C.c_int32), ("pad_", C.c_int32), ("vegas_alpha", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "pad_"}


class MCSummary(C.Structure):
    """chili_mc_summary: pooled StatsData, iteration count, min_diff."""
    _fields_ = [("sum", Stats), ("nresults", C.c_int64), ("min_diff", C.c_double)]


class VegasParams(C.Structure):
    """chili_vegas_params: VegasParams (include/Integrator/Vegas.hh:30-38) + refinement counter."""
    _fields_ = [("ncalls", C.c_int64), ("nrefine", C.c_int64), ("rtol", C.c_double), ("atol", C.c_double),
                ("alpha", C.c_double), ("niterations", C.c_int64), ("irefine", C.c_int64)]


class CheckpointState(C.Structure):
    """chili_checkpoint_state: MultiChannelParams, min_diff, result count."""
    _fields_ = [("ncalls", C.c_uint64), ("niterations", C.c_uint64), ("rtol", C.c_double), ("nrefine", C.c_uint64),
                ("beta", C.c_double), ("min_alpha", C.c_double), ("refine_size", C.c_double), ("max_bins", C.c_uint64),
                ("min_diff", C.c_double), ("nresults", C.c_int64)]


class Uniforms(C.Structure):
    _fields_ = [("u", C.POINTER(C.c_double)), ("stride", C.c_int64),
                ("seed", C.c_uint64), ("stream", C.c_uint32), ("pad_", C.c_uint32)]


def philox_uniforms(seed, stream=0):
    return Uniforms(None, 0, seed, stream, 0)


def host_uniforms(arr):
    """SoA uniforms u[k, attempt] (numpy float64, C-contiguous, shape (d+1, n))."""
    import numpy as np
    assert arr.dtype == np.float64 and arr.flags["C_CONTIGUOUS"] and arr.ndim == 2
    return Uniforms(arr.ctypes.data_as(C.POINTER(C.c_double)), arr.shape[1], 0, 0, 0)


def dptr(arr):
    return arr.ctypes.data_as(C.POINTER(C.c_double))


def iptr(arr):
    return arr.ctypes.data_as(C.POINTER(C.c_int32))
