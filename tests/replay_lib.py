"""ctypes loader for tests/harness/libreplay.so (host replay of the kernel's
per-lane phase functions; test infrastructure)."""
import ctypes as C
import pathlib
import subprocess
import sys

import numpy as np

REPO = pathlib.Path(__file__).resolve().parent.parent
sys.path.insert(0, str(REPO / "damascus-earth-evap_amd"))
import dmsun  # noqa: E402

HARNESS = REPO / "tests" / "harness"
SO = HARNESS / "libreplay.so"
_P = C.POINTER
vp = C.c_void_p


def build():
    csrc = REPO / "damascus-earth-evap_amd" / "csrc"
    subprocess.run(["hipcc", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-Wno-unused-result",
                    "-Wno-unused-value", "-o", str(SO), str(HARNESS / "engine_replay.cpp"),
                    str(csrc / "dms_tables.cpp")], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not SO.exists():
            build()
        L = C.CDLL(str(SO))
        sig = {
            "rp_create": (vp, [_P(C.c_double), C.c_size_t]),
            "rp_free": (None, [vp]),
            "rp_solar_eval": (C.c_double, [vp, C.c_int, C.c_double]),
            "rp_point_create": (vp, [vp, _P(dmsun.PointDesc)]),
            "rp_point_free": (None, [vp]),
            "rp_rate": (C.c_double, [vp, C.c_int, C.c_double, C.c_double]),
            "rp_grid": (C.c_int, [vp, _P(C.c_double), C.c_size_t]),
            "rp_initial_conditions": (None, [vp, _P(dmsun.Halo), C.c_uint32, C.c_uint32, C.c_uint64, C.c_size_t,
                                             C.c_double, _P(dmsun.Event), _P(C.c_int32)]),
            "rp_lane_hist": (C.c_int, [vp, _P(dmsun.JobDesc), _P(dmsun.Event), _P(C.c_int32), C.c_size_t,
                                       C.c_size_t, _P(C.c_double), _P(C.c_uint64)]),
            "rp_simulate": (C.c_int, [vp, _P(dmsun.JobDesc), _P(dmsun.Event), C.c_size_t, _P(dmsun.TrajRecord),
                                      _P(C.c_double)]),
            "rp_simulate_pipe": (C.c_int, [vp, _P(dmsun.JobDesc), _P(dmsun.Event), C.c_size_t, _P(dmsun.TrajRecord),
                                           _P(C.c_double), C.c_int, _P(C.c_uint64)]),
            "rp_div_const": (None, [_P(C.c_double), C.c_size_t, C.c_double, C.c_double, _P(C.c_double)]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class Replay:
    def __init__(self, rows=None):
        if rows is None:
            rows = dmsun.load_solar_rows()
        self.rows = np.ascontiguousarray(rows)
        self.h = lib().rp_create(dmsun.dptr(self.rows), self.rows.shape[0])
        assert self.h

    def eval(self, which, r):
        return lib().rp_solar_eval(self.h, which, r)

    def point(self, desc):
        return ReplayPoint(self, desc)

    def initial_conditions(self, halo, seed, point_id, first_id, n, r_final_rsun=2.0):
        ev = (dmsun.Event * n)()
        ok = np.zeros(n, dtype=np.int32)
        lib().rp_initial_conditions(self.h, C.byref(halo), seed, point_id, first_id, n, r_final_rsun, ev,
                                    ok.ctypes.data_as(_P(C.c_int32)))
        return ev, ok


class ReplayPoint:
    def __init__(self, rp, desc):
        self.rp = rp
        self.desc = desc
        self.h = lib().rp_point_create(rp.h, C.byref(desc))
        assert self.h

    def rate(self, which, r, v):
        return lib().rp_rate(self.h, which, r, v)

    def grid(self):
        nr = self.desc.interpolation_points
        nv = self.desc.speed_points or nr
        out = np.empty(nr * nv)
        assert lib().rp_grid(self.h, dmsun.dptr(out), out.size) == 0
        return out.reshape(nr, nv)

    def simulate_events(self, job, events, want_rows=False):
        n = len(events)
        recs = dmsun.record_array(n)
        rows = np.zeros((n, 2, job.n_bins)) if want_rows else None
        assert lib().rp_simulate(self.h, C.byref(job), events, n, recs, dmsun.dptr(rows) if want_rows else None) == 0
        return recs, rows

    def simulate_events_pipe(self, job, events, depth, want_rows=False):
        """simulate_events with every step loop split between the trajectory's
        lane and a helper lane that prepares up to `depth` steps ahead; returns
        (records, rows, steps taken from the helper, steps discarded)."""
        n = len(events)
        recs = dmsun.record_array(n)
        rows = np.zeros((n, 2, job.n_bins)) if want_rows else None
        stats = np.zeros(3, dtype=np.uint64)
        assert lib().rp_simulate_pipe(self.h, C.byref(job), events, n, recs, dmsun.dptr(rows) if want_rows else None,
                                      depth, stats.ctypes.data_as(_P(C.c_uint64))) == 0
        return recs, rows, int(stats[0]), int(stats[1]), int(stats[2])

    def lane_hist(self, job, events, ok, nlanes=3):
        """Class sums through per-lane scratch rows (serial, id order)."""
        nb = job.n_bins
        hist = np.zeros((8, nb))
        counts = np.zeros((2, nb), dtype=np.uint64)
        ok = np.ascontiguousarray(ok, dtype=np.int32)
        st = lib().rp_lane_hist(self.h, C.byref(job), events, ok.ctypes.data_as(_P(C.c_int32)), len(events), nlanes,
                                dmsun.dptr(hist), counts.ctypes.data_as(_P(C.c_uint64)))
        assert st == 0, st
        return hist, counts
