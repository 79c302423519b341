"""The C-ABI library without a GPU: it loads, exports every entry point that
include/dms.h declares, and its host-side parts (solar tables, isotope table,
constants, argument checks) agree with the oracle.  Compute entry points are
not called here; they fail loudly on a context without a GPU."""
import ctypes as C
import math
import pathlib
import re

import numpy as np
import pytest

import dmsun
import oracle_lib as O

HEADER = pathlib.Path(__file__).resolve().parent.parent / "include" / "dms.h"


def declared():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^[A-Za-z_][\w\s\*]*?\b(dms_\w+)\s*\(", text, flags=re.M)))


def test_every_declared_symbol_is_exported():
    names = declared()
    assert len(names) >= 20
    lib = C.CDLL(str(dmsun.LIB_PATH))
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(dmsun.EXPORTED) == names


def test_abi_version_matches_header():
    v = int(re.search(r"#define DMS_ABI_VERSION (\d+)", HEADER.read_text()).group(1))
    assert dmsun.lib().dms_abi_version() == v


@pytest.fixture(scope="module")
def host_ctx():
    return dmsun.Context(device=-1)


def test_constants_match_oracle():
    for n in ("cm", "km", "sec", "gram", "Kelvin", "mProton", "mElectron", "mNucleon", "G_Newton", "mSun", "rSun",
              "AU"):
        assert dmsun.lib().dms_constant(n.encode()) == O.const(n), n
    assert math.isnan(dmsun.lib().dms_constant(b"no_such_constant"))


def test_solar_tables_bit_identical_to_oracle(host_ctx, oracle_solar):
    rng = np.random.default_rng(4)
    rSun = O.const("rSun")
    r = np.concatenate([rng.uniform(0, 1.2 * rSun, 3000), [0.0, rSun, 2 * rSun, -1.0, math.nan]])
    for which in [0, 1, 2, 3, 4, 100, 101, 109, 140, 162]:
        got = host_ctx.solar_eval(which, r)
        want = np.array([oracle_solar.eval(which, x) for x in r])
        assert np.array_equal(got.view(np.uint64), want.view(np.uint64)) or \
            np.array_equal(got, want, equal_nan=True), which


def test_isotope_table_matches_oracle(host_ctx, oracle_solar):
    L = dmsun.lib()
    cap = 128
    Z = np.zeros(cap, np.int32)
    A = np.zeros(cap, np.int32)
    m = np.zeros(cap)
    ab = np.zeros(cap)
    n = C.c_size_t()
    P32 = C.POINTER(C.c_int32)
    assert L.dms_isotopes(host_ctx.h, Z.ctypes.data_as(P32), A.ctypes.data_as(P32), dmsun.dptr(m), dmsun.dptr(ab),
                          cap, C.byref(n)) == 0
    assert [(int(Z[i]), int(A[i]), m[i], ab[i]) for i in range(n.value)] == oracle_solar.isotopes()


def test_solar_file_loader_and_bad_input():
    L = dmsun.lib()
    rows = np.zeros((4000, 35))
    n = C.c_size_t()
    assert L.dms_load_solar_file(str(dmsun.DATA_FILE).encode(), dmsun.dptr(rows), 4000, C.byref(n)) == 0
    assert np.array_equal(rows[:n.value], dmsun.load_solar_rows())
    assert L.dms_load_solar_file(b"/nonexistent/model.dat", dmsun.dptr(rows), 4000, C.byref(n)) != 0
    bad = np.zeros((3, 35))
    h = C.c_void_p()
    assert L.dms_create(C.byref(h), -1, dmsun.dptr(bad), 3) == -1
    assert not h.value
    assert b"dms_create" in L.dms_last_error(None)


def test_host_context_refuses_gpu_work(host_ctx):
    """No CPU fallback: a context without a GPU rejects every compute call."""
    with pytest.raises(dmsun.DmsError, match="no GPU"):
        host_ctx.point(dmsun.make_point(1.0, 1e-35, interpolation_points=10))
    with pytest.raises(dmsun.DmsError):
        host_ctx.probe_math(0, np.ones(4))
    with pytest.raises(dmsun.DmsError):
        host_ctx.initial_conditions(dmsun.default_halo(), 1, 0, 0, 4)


def test_null_arguments_are_rejected():
    L = dmsun.lib()
    h = C.c_void_p()
    assert L.dms_job_submit(None, None, C.byref(h)) == -1
    assert L.dms_job_submit_batch(None, None, 0, None) == -1
    assert L.dms_job_wait(None, None) == -1
    assert L.dms_create(None, -1, None, 0) == -1
