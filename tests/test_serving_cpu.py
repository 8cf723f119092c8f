"""The native serving runtime's host logic on the CPU, through the C ABI in stub mode
(JM_SERVE_STUB: no HIP call; JM_SERVE_MANUAL: the test drives the loop one pass at a
time): admission from one arena in submission order with first fit, release when a
request's last step completes, aging (a request overtaken max_skips times stops
further overtaking until it fits), batching limits, rejection of requests that can
never fit, failure isolation per cohort, and the threaded loop draining everything."""

import ctypes

import pytest

from jamma_amd import _lib

pytestmark = pytest.mark.timeout(120)

T, C, N = 16, 4, 100
SLICE = 32


def _plan():
    p = _lib.SpectrumPlanC()
    p.n, p.n_terms, p.n_clusters, p.n_seg, p.n_slots = N, T, C, 4, 4
    return p


def _desc(max_batch=0):
    return _lib.ServeCohortDesc(n=N, precision=_lib.JM_SERVE_MIXED, n_snps=100_000, ld_g=128,
                                nplanes=2, n_terms=T, n_clusters=C, max_batch=max_batch)


def _cfg(max_batch=SLICE, max_skips=8, manual=True):
    flags = _lib.JM_SERVE_STUB | (_lib.JM_SERVE_MANUAL if manual else 0)
    return _lib.ServeConfig(slice_snps=SLICE, max_batch=max_batch, max_reqs=8, max_inflight=2,
                            max_skips=max_skips, max_cvt=4, flags=flags)


def _req_bytes(cfg, n_cvt):
    return int(_lib.lib.jm_serve_request_ws_bytes(ctypes.byref(cfg), T, n_cvt + 1))


class Runtime:
    def __init__(self, cfg, extra_bytes, n_cohorts=1, max_batch=0):
        self.cfg = cfg
        self.plan, self.desc = _plan(), _desc(max_batch)
        self.ws = int(_lib.lib.jm_serve_cohort_ws_bytes(ctypes.byref(cfg), ctypes.byref(self.desc),
                                                        ctypes.byref(self.plan)))
        assert self.ws > 0
        self.h = ctypes.c_void_p()
        _lib.call("jm_serve_create", ctypes.byref(cfg), None, self.ws * n_cohorts + extra_bytes,
                  ctypes.byref(self.h))
        self.ids = []
        for _ in range(n_cohorts):
            cid = ctypes.c_int()
            _lib.call("jm_serve_add_cohort", self.h, ctypes.byref(self.desc),
                      ctypes.byref(self.plan), ctypes.byref(cid))
            self.ids.append(cid.value)

    def submit(self, n_cvt, n_snps=SLICE, cohort=0, start=0):
        pr = _lib.LmmProblem()
        pr.n, pr.n_cvt, pr.n_terms, pr.n_clusters = N, n_cvt, T, C
        rid = ctypes.c_int64()
        rc = _lib.lib.jm_serve_submit(self.h, cohort, ctypes.byref(pr), None, 0, 4, start,
                                      start + n_snps, None, ctypes.byref(rid))
        return rc, rid.value

    def run(self, n=1):
        for _ in range(n):
            _lib.lib.jm_serve_run_once(self.h)

    def done(self):
        recs = (_lib.ServeDone * 64)()
        got = ctypes.c_int()
        _lib.call("jm_serve_poll", self.h, recs, 64, ctypes.byref(got))
        return [(recs[i].id, recs[i].status, recs[i].message.decode()) for i in range(got.value)]

    def stats(self):
        st = _lib.ServeStats()
        _lib.call("jm_serve_stats", self.h, ctypes.byref(st))
        return st

    def close(self):
        _lib.call("jm_serve_destroy", self.h)


def test_first_fit_in_submission_order_and_release():
    cfg = _cfg()
    a, b, c = _req_bytes(cfg, 2), _req_bytes(cfg, 4), _req_bytes(cfg, 1)
    assert (a, b, c) == (4096 * 4, 4096 * 6, 4096 * 3)  # slice x (nv + 1) x T doubles
    rt = Runtime(cfg, a + c)  # a and c fit together, b needs more than c leaves
    ids = [rt.submit(n)[1] for n in (2, 4, 1)]
    rt.run()  # admits a and c (b does not fit), steps a (one request per step here)
    assert [d[0] for d in rt.done()] == [ids[0]]
    st = rt.stats()
    assert st.queued == 1 and st.active == 1 and st.arena_used == rt.ws + c
    rt.run()  # b still does not fit next to c; c completes
    assert [d[0] for d in rt.done()] == [ids[2]]
    rt.run()
    assert [d for d in rt.done()] == [(ids[1], 0, "")]
    st = rt.stats()
    assert st.queued == 0 and st.active == 0 and st.arena_used == rt.ws
    assert st.steps == 3 and st.snps == 3 * SLICE and st.max_active == 2
    rt.close()


def test_aging_stops_overtaking():
    cfg = _cfg(max_batch=2 * SLICE, max_skips=2)
    small, big = _req_bytes(cfg, 1), _req_bytes(cfg, 4)
    rt = Runtime(cfg, small + big - 4096)  # next to the holder, big never fits
    holder = rt.submit(1, n_snps=100 * SLICE)[1]  # 100 steps long
    rt.run()
    big_id = rt.submit(4)[1]
    for k in range(2):  # two small requests overtake the big one and complete
        sid = rt.submit(1, start=1000 * (k + 1))[1]
        rt.run()
        assert sid in [d[0] for d in rt.done()]
    late = rt.submit(1, start=5000)[1]
    rt.run()
    st = rt.stats()
    assert st.queued == 2 and st.active == 1  # aged: `late` may not pass although it fits
    rt.run(200)
    order = [d[0] for d in rt.done()]
    assert order == [holder, big_id, late]
    rt.close()


def test_step_limits_and_batches():
    cfg = _cfg(max_batch=2 * SLICE)
    rt = Runtime(cfg, 8 * _req_bytes(cfg, 3))
    ids = [rt.submit(3, n_snps=SLICE + 5, start=64 * k)[1] for k in range(4)]
    rt.run(10)
    assert sorted(d[0] for d in rt.done()) == ids
    st = rt.stats()
    # a step holds <= 64 SNPs and <= 32 per request: r0 32 + r1 32 | r0 5 + r1 5 + r2 32
    # + r3 22 | r2 5 + r3 15
    assert st.snps == 4 * (SLICE + 5) and st.steps == 3
    rt.close()


def test_cohort_step_cap():
    """A cohort's own max_batch caps its steps below the server's and shrinks its
    step workspace; a cap below one slice is refused."""
    cfg = _cfg(max_batch=4 * SLICE)
    full = Runtime(cfg, 0)
    rt = Runtime(cfg, 8 * _req_bytes(cfg, 2), max_batch=SLICE)
    assert rt.ws < full.ws
    full.close()
    ids = [rt.submit(2, n_snps=SLICE, start=SLICE * k)[1] for k in range(3)]
    rt.run()  # one slice per step although the server allows four
    assert [d[0] for d in rt.done()] == ids[:1]
    rt.run(2)
    assert [d[0] for d in rt.done()] == ids[1:]
    assert rt.stats().steps == 3
    rt.close()
    h = ctypes.c_void_p()
    _lib.call("jm_serve_create", ctypes.byref(cfg), None, 1 << 30, ctypes.byref(h))
    cid = ctypes.c_int()
    with pytest.raises(RuntimeError, match="at least slice_snps"):
        _lib.call("jm_serve_add_cohort", h, ctypes.byref(_desc(SLICE - 1)), ctypes.byref(_plan()),
                  ctypes.byref(cid))
    _lib.call("jm_serve_destroy", h)


def test_rejections():
    cfg = _cfg()
    rt = Runtime(cfg, _req_bytes(cfg, 2))
    rc, _ = rt.submit(4)  # can never fit beside the cohort's step workspace
    assert rc == -1 and "never fit" in _lib.last_error()
    rc, _ = rt.submit(2, n_snps=10, start=100_000 - 5)  # range past the cohort
    assert rc == -1
    rc, _ = rt.submit(5)  # wider than max_cvt
    assert rc == -1
    rt.close()
    h = ctypes.c_void_p()
    _lib.call("jm_serve_create", ctypes.byref(cfg), None, 1024, ctypes.byref(h))
    cid = ctypes.c_int()
    assert _lib.lib.jm_serve_add_cohort(h, ctypes.byref(_desc()), ctypes.byref(_plan()),
                                        ctypes.byref(cid)) == -1  # workspace does not fit
    _lib.call("jm_serve_destroy", h)


def test_step_failure_fails_its_cohort_alone():
    cfg = _cfg(max_batch=4 * SLICE)
    rt = Runtime(cfg, 8 * _req_bytes(cfg, 2), n_cohorts=2)
    bad = [rt.submit(2, n_snps=3 * SLICE, cohort=0, start=SLICE * k)[1] for k in range(2)]
    good = rt.submit(2, n_snps=3 * SLICE, cohort=1)[1]
    _lib.call("jm_serve_inject_step_error", rt.h, 0)
    rt.run(10)
    recs = {d[0]: d for d in rt.done()}
    assert all(recs[i][1] == -3 and "injected" in recs[i][2] for i in bad)
    assert recs[good][1] == 0
    later = rt.submit(2, cohort=0)[1]
    rt.run(3)
    assert rt.done() == [(later, 0, "")]
    assert rt.stats().arena_used == 2 * rt.ws
    rt.close()


def test_threaded_loop_drains_everything():
    cfg = _cfg(max_batch=4 * SLICE, manual=False)
    rt = Runtime(cfg, 4 * _req_bytes(cfg, 3))
    ids = [rt.submit(1 + k % 4, n_snps=SLICE * (1 + k % 3), start=128 * k)[1] for k in range(40)]
    seen = []
    for _ in range(600):  # at most a minute
        if len(seen) == len(ids):
            break
        _lib.lib.jm_serve_wait(rt.h, 100)
        seen += rt.done()
    assert sorted(d[0] for d in seen) == ids and all(d[1] == 0 for d in seen)
    rt.close()


def test_threaded_destroy_drains_queued_requests():
    cfg = _cfg(max_batch=2 * SLICE, manual=False)
    rt = Runtime(cfg, 2 * _req_bytes(cfg, 4))
    for k in range(10):
        assert rt.submit(4, n_snps=2 * SLICE, start=64 * k)[0] == 0
    rt.close()  # returns once every request has been served


def test_pause_holds_admission_and_steps():
    cfg = _cfg(max_batch=4 * SLICE, manual=False)
    rt = Runtime(cfg, 8 * _req_bytes(cfg, 2))
    _lib.call("jm_serve_pause", rt.h, 1)
    ids = [rt.submit(2, start=SLICE * k)[1] for k in range(3)]
    assert _lib.lib.jm_serve_wait(rt.h, 50) == 0  # nothing runs while paused
    st = rt.stats()
    assert st.queued == 3 and st.steps == 0
    _lib.call("jm_serve_pause", rt.h, 0)
    seen = []
    for _ in range(100):
        if len(seen) == 3:
            break
        _lib.lib.jm_serve_wait(rt.h, 100)
        seen += rt.done()
    assert sorted(d[0] for d in seen) == ids
    assert rt.stats().steps == 1  # admitted together: one step holds all three
    _lib.call("jm_serve_pause", rt.h, 1)
    rt.submit(2)
    rt.close()  # destroy resumes and drains


def test_manual_mode_requires_stub():
    cfg = _lib.ServeConfig(slice_snps=SLICE, max_batch=SLICE, max_reqs=1, max_inflight=1,
                           max_skips=1, max_cvt=1, flags=_lib.JM_SERVE_MANUAL)
    h = ctypes.c_void_p()
    with pytest.raises(RuntimeError, match="requires JM_SERVE_STUB"):
        _lib.call("jm_serve_create", ctypes.byref(cfg), None, 1 << 20, ctypes.byref(h))
