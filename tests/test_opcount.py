"""The op-counting build of the oracle (oracle/opcount/), whose per-event
operation counts are the FLOP weights bench.py prices the kernel's work with.

Checks: the counting build computes exactly what the oracle computes (records
bit-identical, so the counts are those of the real arithmetic); its attribution
marks classify every step iteration as the oracle's own step counters do; the
weights it measures match the frozen ones in oracle/opcount/weights.txt, which
bench.py uses."""
import ctypes as C
import importlib.util
import pathlib
import sys

import numpy as np
import pytest

import dmsun
import oracle_lib as O

REPO = pathlib.Path(__file__).resolve().parents[1]
sys.path.insert(0, str(REPO / "oracle" / "opcount"))
import count_ops  # noqa: E402


@pytest.fixture(scope="module")
def counting():
    return count_ops.load()


def frozen_weights():
    out, part = {}, None
    for line in (REPO / "oracle" / "opcount" / "weights.txt").read_text().splitlines():
        line = line.strip()
        if not line or line.startswith("#"):
            continue
        if line.startswith("["):
            part = line.strip("[]")
            out[part] = {}
            continue
        k, v = line.split()
        out[part][k] = float(v)
    return out


@pytest.mark.parametrize("mode", ["parameter", "capture"])
def test_counting_build_computes_what_the_oracle_computes(counting, oracle_solar, mode):
    n = 40
    w, d = count_ops.count(mode, n, counting)
    pdesc, job = count_ops.workload(mode)
    pt = O.Point(oracle_solar, pdesc)
    ev, ok = O.initial_conditions(oracle_solar, job.halo, job.seed, job.point_id, 0, n)
    for i in range(n):
        job.first_trajectory_id = i
        rec = pt.simulate_events(job, (dmsun.Event * 1)(ev[i]))[0]
        assert bytes(rec[0]) == d["records"][i], i
    # every step iteration is classified as the oracle's own counters classify it
    it = d["iterations"]
    assert it[1] == d["n_int"] and it[2] == d["n_ext"] and it[3] == 0
    # a plausible attribution: each part did work
    assert all(d["ops"][k] > 0 for k in (0, 1, 2, 5))


@pytest.mark.parametrize("mode", ["parameter", "capture"])
def test_measured_weights_match_the_frozen_ones(counting, mode):
    w, _ = count_ops.count(mode, 400, counting)
    frozen = frozen_weights()[mode]
    for k in ("rk_attempt", "step_interior", "step_exterior"):
        assert abs(w[k] - frozen[k]) <= 0.03 * frozen[k], (k, w[k], frozen[k])


def test_bench_prices_work_with_the_frozen_weights():
    spec = importlib.util.spec_from_file_location("bench_for_weights", REPO / "bench.py")
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    assert bench.FLOP_WEIGHTS == frozen_weights()


def test_counting_type_counts_each_operation_once(counting):
    """Counts are per executed operation: the same call of
    or_solve_linear_rate counts something, and twice as much when made
    twice (nothing is counted once and cached)."""
    L = counting
    f = L.or_solve_linear_rate
    f.restype = C.c_double
    f.argtypes = [C.c_double] * 4
    buf = (C.c_uint64 * 14)()
    L.or_opcount_read(buf, 1)
    f(1.0, 2.0, 1.0, 0.5)
    L.or_opcount_read(buf, 1)
    once = sum(buf[:7])
    f(1.0, 2.0, 1.0, 0.5)
    f(1.0, 2.0, 1.0, 0.5)
    L.or_opcount_read(buf, 1)
    assert once > 0 and sum(buf[:7]) == 2 * once
