"""The C++ facade (include/chili_gpu_facade.hh), which runs the reference's
MultiChannel::Optimize(Integrand<FourVector>&) on the GPU with the
reference's objects staying authoritative.  CPU: its test program
(tests/facade/facade_driver.cpp, the facade instantiated with test doubles of
the reference's types) compiles and links against the C-ABI library, and the
state-exchange entry points round-trip the reference's byte layouts.  GPU: the
facade's run equals the same optimisation through the Python binding, bit for
bit, with the test Integrand's channels in reversed order (device integrand)
and with the host-callback integrand and PreProcess."""
import ctypes as C
import json
import os
import subprocess

import numpy as np
import pytest

from chili_amd import abi, lib
from chili_amd import processes as P

HERE = os.path.dirname(os.path.abspath(__file__))
FACADE = os.path.join(HERE, "facade")
DRIVER = os.path.join(FACADE, "_build", "facade_driver")


def test_facade_compiles_and_links(tmp_path):
    r = subprocess.run(["make", "-C", FACADE, f"OUT={tmp_path}"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert os.path.exists(tmp_path / "facade_driver")


def test_state_bytes_round_trip():
    """chili_mc_state_write / _read: the MultiChannel::Serialize layout
    (src/Channel/MultiChannel.cc:94-107), and chili_channels_match on an
    Integrand::Serialize stream taken from a results.chili file."""
    L = lib.load()
    st = abi.CheckpointState(ncalls=123, niterations=7, rtol=1e-3, nrefine=2, beta=0.3, min_alpha=1e-4, refine_size=1.25,
                             max_bins=100, min_diff=0.5, nresults=2)
    best = np.array([0.1, 0.2, 0.7])
    res = (abi.Stats * 2)(abi.Stats(10, -1, 2, 3, 4, 9), abi.Stats(20, -2, 3, 4, 5, 19))
    n = C.c_int64()
    lib.check(L.chili_mc_state_write(5, C.byref(st), 3, best.ctypes.data_as(C.POINTER(C.c_double)), res, None, 0, C.byref(n)))
    # size: ndims + 8 params + nbest + 3 + min_diff + nres + 2 x 6 + nbest + 3
    assert n.value == 8 * (1 + 8 + 1 + 3 + 1 + 1 + 12 + 1 + 3)
    buf = C.create_string_buffer(n.value)
    lib.check(L.chili_mc_state_write(5, C.byref(st), 3, best.ctypes.data_as(C.POINTER(C.c_double)), res, buf, n.value, C.byref(n)))
    words = np.frombuffer(buf.raw, dtype=np.float64)
    ints = np.frombuffer(buf.raw, dtype=np.uint64)
    assert ints[0] == 5 and ints[1] == 123 and ints[2] == 7 and words[3] == 1e-3 and ints[9] == 3
    assert list(words[10:13]) == [0.1, 0.2, 0.7] and words[13] == 0.5 and ints[14] == 2
    assert list(words[15:21]) == [10, -1, 2, 3, 4, 9] and ints[27] == 3 and list(words[28:31]) == [0.1, 0.2, 0.7]
    nd, nb = C.c_int(), C.c_int()
    st2 = abi.CheckpointState()
    b2 = np.zeros(8)
    r2 = (abi.Stats * 4)()
    lib.check(L.chili_mc_state_read(buf, n.value, C.byref(nd), C.byref(st2), b2.ctypes.data_as(C.POINTER(C.c_double)), 8,
                                    C.byref(nb), r2, 4))
    assert nd.value == 5 and nb.value == 3 and list(b2[:3]) == [0.1, 0.2, 0.7] and st2.nresults == 2
    assert (st2.ncalls, st2.refine_size, st2.max_bins, st2.min_diff) == (123, 1.25, 100, 0.5)
    assert r2[1].as_tuple() == (20, -2, 3, 4, 5, 19)


def test_channels_match_integrand_stream(tmp_path):
    proc = P.gluon_process(4, smax=3)
    ch = lib.Channels(proc)
    Cn = ch.count
    rng = np.random.default_rng(2)
    grids = np.sort(rng.uniform(0, 1, (Cn, 10, 9)), axis=-1)
    grids[..., 0], grids[..., -1] = 0.0, 1.0
    path = tmp_path / "results.chili"
    lib.save_checkpoint(path, ch, grids, np.full(Cn, 1.0 / Cn))
    data = path.read_bytes()
    L = lib.load()
    perm = np.zeros(Cn, np.int32)
    g = np.zeros_like(grids)
    nb = C.c_int()
    lib.check(L.chili_channels_match(ch.h, data, len(data), perm.ctypes.data_as(C.POINTER(C.c_int32)),
                                     g.ctypes.data_as(C.POINTER(C.c_double)), C.byref(nb)))
    assert nb.value == 8 and list(perm) == list(range(Cn)) and np.array_equal(g, grids)
    # a stream for another process is refused
    other = lib.Channels(P.gluon_process(4, smax=3, cuts=dict(ptmin=30.0, etamax=2.5, deltaR=0.4)))
    assert L.chili_channels_match(other.h, data, len(data), perm.ctypes.data_as(C.POINTER(C.c_int32)), None, C.byref(nb)) != 0


def run_driver(tmp_path, proc, f, mode, ncalls, niter, nrefine, seed, stream):
    ch = lib.Channels(proc)
    gpu = lib.MultiChannelGPU(ch, 16, f)
    lib.save_checkpoint(tmp_path / "results.chili", ch, gpu.grids(), gpu.alpha())
    (tmp_path / "process.bin").write_bytes(C.string_at(C.byref(proc), C.sizeof(proc)))
    (tmp_path / "integrand.bin").write_bytes(C.string_at(C.byref(f), C.sizeof(f)))
    r = subprocess.run([DRIVER, str(tmp_path / "process.bin"), str(tmp_path / "integrand.bin"), str(tmp_path / "results.chili"), mode,
                        str(ncalls), str(niter), str(nrefine), str(seed), str(stream)], capture_output=True, text=True, timeout=300)
    print(r.stderr[-2000:])
    assert r.returncode == 0, r.stderr
    return ch, json.loads(r.stdout)


def direct(ch, f, ncalls, niter, nrefine, seed, stream):
    gpu = lib.MultiChannelGPU(ch, 16, f)
    params = lib.mc_params(ncalls=ncalls, niterations=niter, nrefine=nrefine)
    summ = lib.mc_summary()
    best = np.zeros(gpu.nchannels)
    res = gpu.optimize(abi.philox_uniforms(seed, stream=stream), params, summ, best)
    return gpu, params, summ, best, res


def compare(out, gpu, params, summ, best, res, exact=True):
    """exact: bit for bit; else counts exactly and reals to 1e-8: the host
    callback path runs smaller sub-batches, so the device's histogram sums
    group differently (last-bit differences), and each adaptation carries
    those into the next iteration's grids (measured: 1e-12 after one
    adaptation, 5e-11 after three)."""
    from tests.helpers import assert_rel
    C_ = gpu.nchannels
    grids = gpu.grids()

    def same(a, b, what):
        if exact:
            assert np.array_equal(a, b), what
        else:
            assert_rel(a, b, 1e-8, scale=np.maximum(np.abs(b), 1e-300), what=what)

    assert len(out["results"]) == len(res)
    for k, (r_out, r) in enumerate(zip(out["results"], res)):
        print(f"iteration {k}: facade {r_out} direct {r.as_tuple()}")
        assert r_out[0] == r.n and r_out[5] == r.n_finite
        same(np.array(r_out), np.array(r.as_tuple()), f"results {k}")
    same(np.array(out["best"]), best[::-1], "best weights")
    same(np.array([out["min_diff"]]), np.array([summ.min_diff]), "min_diff")
    assert len(out["grids"]) == C_
    for k in range(C_):            # reference channel k = device channel C - 1 - k
        g = out["grids"][k]
        assert g["bins"] == gpu.nbins and g["dims"] == gpu.ndims
        same(np.array(g["edges"]).reshape(gpu.ndims, gpu.nbins + 1), grids[C_ - 1 - k], f"grid {k}")
    assert out["ncalls"] == params.ncalls and out["iteration"] == params.iteration and out["should_optimize"] == 1


@pytest.mark.gpu
def test_facade_optimize_device_integrand(tmp_path):
    proc, f = P.gluon_process(3, smax=2), P.integrand_propprod()
    seed, stream = 0xC0FFEE30, 3
    ch, out = run_driver(tmp_path, proc, f, "device", 20000, 3, 2, seed, stream)
    gpu, params, summ, best, res = direct(ch, f, 20000, 3, 2, seed, stream)
    assert len(res) == 3 and gpu.nbins == 32 and params.ncalls == 30000
    compare(out, gpu, params, summ, best, res)


@pytest.mark.gpu
def test_facade_optimize_host_callback(tmp_path):
    """The facade's host callback (f = 1, PreProcess: x1, x2 <= 1 and pT > 20
    in C++ on the host) against the device integrand f = 1 with the same cut."""
    proc = P.gluon_process(3, smax=2)
    f_host = P.integrand_unit(None)
    seed, stream = 0xC0FFEE31, 5
    ch, out = run_driver(tmp_path, proc, f_host, "host", 20000, 3, 0, seed, stream)
    f_dev = P.integrand_unit((20.0, 1e300, -1.0))
    gpu, params, summ, best, res = direct(ch, f_dev, 20000, 3, 0, seed, stream)
    compare(out, gpu, params, summ, best, res, exact=False)
