"""dms_simulate_events_noalloc, the entry point for callers that cannot
allocate device memory (no counterpart in the reference), against
dms_simulate_events on the same events: bit-identical records and histogram
rows on two calls in a row, the second smaller and enqueued on a caller's
stream.  The test's own device buffers and stream come from the HIP runtime
the library is linked against (plumbing only)."""
import ctypes as C

import numpy as np
import pytest

import dmsun

pytestmark = pytest.mark.gpu

RK_BUDGET = 20_000_000


class Hip:
    """hipMalloc / hipMemcpy / streams of the HIP runtime libdmsun.so uses."""

    def __init__(self):
        dmsun.lib()  # loads libamdhip64 with the library
        self.L = C.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.L.hipStreamDestroy.argtypes = [C.c_void_p]
        self.L.hipStreamSynchronize.argtypes = [C.c_void_p]
        self.L.hipDeviceSynchronize.argtypes = []

    def ok(self, rc, what):
        assert rc == 0, f"{what} returned {rc}"

    def malloc(self, n):
        p = C.c_void_p()
        self.ok(self.L.hipMalloc(C.byref(p), n), "hipMalloc")
        return p.value

    def h2d(self, dst, src, n):
        self.ok(self.L.hipMemcpy(C.c_void_p(dst), src, n, 1), "hipMemcpy H2D")

    def d2h(self, dst, src, n):
        self.ok(self.L.hipMemcpy(dst, C.c_void_p(src), n, 2), "hipMemcpy D2H")


def test_noalloc_matches_simulate_events(gpu_ctx):
    hip = Hip()
    pt = gpu_ctx.point(dmsun.make_point(0.3, 1e-36, interpolation_points=300))
    rec_size = C.sizeof(dmsun.TrajRecord)
    ev_size = C.sizeof(dmsun.Event)
    nbins = 200
    stream = C.c_void_p()
    hip.ok(hip.L.hipStreamCreate(C.byref(stream)), "hipStreamCreate")
    try:
        for call, (n, first, on_stream) in enumerate([(300, 0, False), (120, 5_000, True)]):
            job = dmsun.make_job(mode=dmsun.MODE_PARAMETER_POINT, max_trajectories=n, max_scatterings=10**14,
                                 n_bins=nbins, rk_attempt_budget=RK_BUDGET, first_trajectory_id=first)
            ev, _ = gpu_ctx.initial_conditions(job.halo, job.seed, job.point_id, first, n)
            print(f"noalloc call {call}: n={n} first id {first} caller stream={on_stream}", flush=True)
            ref_recs, ref_rows = pt.simulate_events(job, ev, want_rows=True)
            d_ev = hip.malloc(n * ev_size)
            d_rec = hip.malloc(n * rec_size)
            d_rows = hip.malloc(n * 2 * nbins * 8)
            try:
                hip.h2d(d_ev, C.cast(ev, C.c_void_p), n * ev_size)
                junk = np.full(n * 2 * nbins, 7.0)  # the call clears the rows
                hip.h2d(d_rows, junk.ctypes.data_as(C.c_void_p), junk.nbytes)
                pt.simulate_events_noalloc(job, d_ev, n, d_rec, d_rows, stream=stream.value if on_stream else None)
                if on_stream:
                    hip.ok(hip.L.hipStreamSynchronize(stream), "hipStreamSynchronize")
                got = np.empty(n * rec_size, dtype=np.uint8)
                hip.d2h(got.ctypes.data_as(C.c_void_p), d_rec, got.nbytes)
                rows = np.empty(n * 2 * nbins)
                hip.d2h(rows.ctypes.data_as(C.c_void_p), d_rows, rows.nbytes)
            finally:
                for p in (d_ev, d_rec, d_rows):
                    hip.L.hipFree(C.c_void_p(p))
            want = np.ctypeslib.as_array(ref_recs).view(np.uint8).reshape(n, rec_size)
            same = (got.reshape(n, rec_size) == want).all(axis=1)
            assert same.all(), f"{(~same).sum()} of {n} records differ"
            assert np.array_equal(rows.reshape(n, 2, nbins).view(np.uint64),
                                  ref_rows.reshape(n, 2, nbins).view(np.uint64))
    finally:
        hip.L.hipStreamDestroy(stream)


def test_noalloc_rejects_oversized_calls(gpu_ctx):
    """More trajectories than the scratch slot counter covers are refused
    before anything is launched (the addresses are never dereferenced)."""
    pt = gpu_ctx.point(dmsun.make_point(0.3, 1e-36, interpolation_points=0))
    job = dmsun.make_job(mode=dmsun.MODE_PARAMETER_POINT, max_trajectories=1)
    with pytest.raises(dmsun.DmsError, match="2\\^31"):
        pt.simulate_events_noalloc(job, 256, 2**31, 256)
