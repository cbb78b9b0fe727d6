"""The algorithmic work per point (tests/golden/opcounts.json, written by
tests/golden/make_opcounts.py): the op-counting oracle reproduces the
committed counts exactly, and its counter counts what the convention says on
a function counted by hand (SqLam, src/Tools/ChannelElements.cc:21-25:
d = s - s1 - s2 is 2 operations, d * d - 4 s1 s2 is 4, then sqrt and the
division: 8 when the argument is positive, 6 otherwise; no transcendental
call)."""
import ctypes as C
import json
import os

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "opcounts.json")


def counts(L):
    a, t = C.c_ulonglong(), C.c_ulonglong()
    L.oracle_count_get(C.byref(a), C.byref(t))
    return a.value, t.value


def test_counter_on_sqlam():
    from oracle import oracle as O
    L = O.lib(True)
    L.oracle_count_reset()
    v = L.oracle_sqlam(10.0, 1.0, 2.0)          # arg = 49 - 8 > 0
    assert abs(v - 41 ** 0.5 / 10.0) < 1e-15
    assert counts(L) == (8, 0)
    L.oracle_count_reset()
    assert L.oracle_sqlam(3.0, 1.0, 1.0) == 0.0  # arg = 1 - 4 < 0
    assert counts(L) == (6, 0)


def test_fixture_reproduces():
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_opcounts", os.path.join(HERE, "golden", "make_opcounts.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    fixture = json.load(open(FIXTURE))["workloads"]
    for name in ("tt_6j", "gg_2g", "gg_3g"):
        got = mod.count(name)
        assert got == fixture[name], name


@pytest.mark.parametrize("name", ["tt_6j", "gg_4g"])
def test_fixture_sane(name):
    w = json.load(open(FIXTURE))["workloads"][name]
    assert w["accepted"] > 0 and w["F_main_per_point"] > w["F_fwd_per_attempt"]
    assert w["F_main_per_point"] == round((w["wgt_arith"] + 20 * w["wgt_trans"]) / w["accepted"], 3)
