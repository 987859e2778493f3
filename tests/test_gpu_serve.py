"""The serving loop (cu_serve_*) on the GPU.

* a session's frames are the reference renderer's frames: the first is
  fractalRenderMain + compose, the next ones fractalRenderAdvanced with reuse
  and foveation around the centre + compose, each equal to the oracle chain
  (render_main -> render_advanced ... -> compose); a deep frame equals the
  oracle's FP32 perturbation + smooth compose;
* every session's frames are the same alone and in the mix (per-frame image
  checksums), with several frames in flight at once;
* a small HBM budget admits sessions one at a time; all frames still complete,
  with the same results, and the peak reservation stays within the budget;
* shallow frames run in FP64 unless the reference's float/double switch is
  asked for; requests that would fail later are refused at submit; a poll that
  waits does not hold off a submit from another thread.
"""
import importlib
import math
import threading
import time

import numpy as np
import pytest

import _oracle as O

pytestmark = pytest.mark.gpu

sv = importlib.import_module("chaos-ultra_amd.serve")


def rect(centre, width, W, H):
    h = width * H / W
    return (centre[0] - 0.5 * width, centre[1] - 0.5 * h, centre[0] + 0.5 * width, centre[1] + 0.5 * h)


def wait_frame(server, n=1):
    got = []
    for _ in range(20000):
        got += server.poll(wait_ms=5)
        if len(got) >= n:
            return got
    raise TimeoutError("frame did not complete")


def small_sessions(cu, n=4, frames=10, size=(192, 108)):
    ss = sv.zoom_sessions(n, frames, seed=99, resolutions=(size,))
    # one deep session and one that crosses into deep frames
    ss[0]["widths"] = 5e-13 * sv.ZOOM ** np.arange(frames)
    ss[1]["widths"] = 1.3e-12 * sv.ZOOM ** np.arange(frames)
    for s in ss:
        s["max_iter"] = [min(1500, sv.max_iter_for(w)) for w in s["widths"]]
        s["arrivals"] = np.arange(frames) * 1e-3
    return ss


def session_bytes(s):
    pitch = (s["width"] * 16 + 255) // 256 * 256
    b = 2 * pitch * s["height"] + 4 * s["width"] * s["height"] + 256 * 4
    deep = [m for m, w in zip(s["max_iter"], s["widths"]) if w < sv.DEEP_WIDTH]
    if deep:
        b += 13 * s["width"] * s["height"] + (max(deep) + 1) * 32 + 1024 * 8 + 8
    return b


@pytest.mark.parametrize("policy", [0, 1])
def test_frames_match_oracle_chain(gpu, policy):
    """policy 0: every shallow frame in FP64 (the serving configuration's rule);
    policy 1: the reference's switch (FP32 unless the pixel is below the float ulp)."""
    cu = gpu
    W, H, max_ss = 160, 90, 2.0
    server = sv.Server(cu.KERNELS_DIR, 1 << 30, slots=2, max_super_sampling=max_ss,
                       shallow_policy=policy)
    pal = sv.palette256()
    try:
        sid = server.open(W, H, pal, 0)
        centre = (-0.7436438870, 0.1318259042)
        widths = [0.02 * 0.97 ** k for k in range(3)]
        # the float ulp at the view is far below its pixel: the switch picks FP32 there
        assert cu.precision_for_segment(*rect(centre, widths[0], W, H), W, H) == cu.PRECISION_SINGLE
        prev = None
        for k, w in enumerate(widths):
            seg = rect(centre, w, W, H)
            server.submit(sid, centre, w, 600, k)
            done = wait_frame(server)[0]
            got = server.read_rgba(sid, W, H)
            precision = 64
            if policy == 1 and cu.precision_for_segment(*seg, W, H) != cu.PRECISION_DOUBLE:
                precision = 32
            if k == 0:
                assert done.path == 1
                pix = O.render_main(precision, 0, (0.0, 0.0), W, H, seg, 600, max_ss, 1)
            else:
                assert done.path == 0
                pix = O.render_advanced(precision, 0, (0.0, 0.0), W, H, seg, 600, max_ss,
                                        1 | 4 | 8 | 16 | 32, prev_seg, prev, (W // 2, H // 2))
            want = O.compose(pix, pal, max_ss, False)
            assert np.array_equal(got, want), f"frame {k}"
            prev, prev_seg = pix, seg
        # a deep frame: the oracle's FP32 perturbation (bailout 4, rebasing) + smooth compose
        dsid = server.open(96, 54, pal, 3000)
        w = 4e-13
        server.submit(dsid, centre, w, 3000, 99)
        done = wait_frame(server)[0]
        assert done.path == 2 and done.iterations > 0
        got = server.read_rgba(dsid, 96, 54)
        d = cu.make_deep(centre[0], centre[1], w, 96, 54, 3000, 4.0)
        orbit, M = O.reference_orbit(centre[0], centre[1], 3000, 4.0, 32)
        ys, xs = np.mgrid[0:54, 0:96]
        c, mag, st, _ = O.deep_pixels(O.DEEP_PERTURB_F32, xs.ravel(), ys.ravel(), d, orbit, M)
        sm = np.where(st == 1, c + 1.0 - np.log2(np.log(np.where(st == 1, mag, 4.0)) / math.log(4.0)), 0.0)
        assert np.array_equal(got, O.compose_smooth(sm.reshape(54, 96), pal))
    finally:
        server.close()


def run_alone_and_mixed(cu, sessions, budget, slots=4):
    alone = {}
    for s in sessions:
        server = sv.Server(cu.KERNELS_DIR, 1 << 32, slots=slots)
        try:
            recs, _ = sv.run(server, [s], realtime=False)
            alone[s["id"]] = [r["checksum"] for r in recs]
            assert [r["frame"] for r in recs] == list(range(len(s["widths"])))
        finally:
            server.close()
    server = sv.Server(cu.KERNELS_DIR, budget, slots=slots)
    try:
        recs, _ = sv.run(server, sessions, realtime=False)
        rep = server.report()
        mixed = {}
        for r in recs:
            mixed.setdefault(sessions[r["session"]]["id"], []).append(r["checksum"])
        return alone, mixed, rep, recs
    finally:
        server.close()


def test_same_alone_and_in_the_mix(gpu):
    cu = gpu
    sessions = small_sessions(cu)
    alone, mixed, rep, recs = run_alone_and_mixed(cu, sessions, 1 << 32)
    assert alone == mixed
    paths = {r["path"] for r in recs}
    assert paths == {0, 1, 2}
    assert rep.frames == sum(len(s["widths"]) for s in sessions)
    assert rep.iterations > 0
    # frames of different sessions were in flight together
    spans = sorted((r["submit_s"] + r["queue_s"], r["done_s"]) for r in recs)
    overlap = any(spans[i + 1][0] < spans[i][1] for i in range(len(spans) - 1))
    assert overlap


def test_admission_by_hbm_budget(gpu):
    cu = gpu
    sessions = small_sessions(cu)
    need = [session_bytes(s) for s in sessions]
    budget = int(max(need) * 1.5)  # one session at a time
    alone, mixed, rep, recs = run_alone_and_mixed(cu, sessions, budget)
    assert alone == mixed
    assert rep.peak_bytes <= budget
    assert rep.sessions_admitted == len(sessions)
    # sessions waited for admission: their first frame started after an earlier session closed
    firsts = sorted(r["submit_s"] + r["queue_s"] for r in recs if r["frame"] == 0)
    lasts = sorted(r["done_s"] for r in recs if r["frame"] == len(sessions[0]["widths"]) - 1)
    assert firsts[-1] >= lasts[0]


def test_serve_rejects(gpu):
    cu = gpu
    with pytest.raises(cu.IllegalArgumentError, match="slots"):
        sv.Server(cu.KERNELS_DIR, 1 << 30, slots=0)
    server = sv.Server(cu.KERNELS_DIR, 1 << 20, slots=1)
    try:
        with pytest.raises(cu.IllegalArgumentError, match="budget"):
            server.open(3840, 2160, sv.palette256(), 0)
        sid = server.open(64, 36, sv.palette256(), 0)
        with pytest.raises(cu.IllegalArgumentError, match="deep frames"):
            server.submit(sid, (-0.75, 0.1), 1e-13, 1000, 0)
        dsid = server.open(64, 36, sv.palette256(), 1000)
        with pytest.raises(cu.IllegalArgumentError, match="deep_max_iter"):
            server.submit(dsid, (-0.75, 0.1), 1e-13, 2000, 0)
        # a deep frame outside |c| <= 2 is refused at submit, and nothing is queued
        with pytest.raises(cu.IllegalArgumentError, match=r"\|c\| <= 2"):
            server.submit(dsid, (1.9, 0.9), 1e-13, 1000, 1)
        assert server.poll(wait_ms=0) == []
        server.submit(dsid, (-0.75, 0.1), 1e-13, 1000, 2)
        done = wait_frame(server)[0]
        assert done.tag == 2 and done.status == 0
        server.close_session(dsid)
        server.close_session(sid)
        with pytest.raises(cu.IllegalStateError):
            server.submit(sid, (-0.75, 0.1), 1e-3, 1000, 0)
    finally:
        server.close()


def test_deep_frames_with_fp64_deltas(gpu):
    """deep_method CU_DEEP_PERTURB_F64: a served deep frame equals the oracle's
    FP64-delta perturbation (bailout 4, rebasing) + smooth compose, and a second
    frame of the session (narrower, same centre) is still exact."""
    cu = gpu
    server = sv.Server(cu.KERNELS_DIR, 1 << 30, slots=2, deep_method=cu.DEEP_PERTURB_F64)
    pal = sv.palette256()
    centre = (-0.7436438870, 0.1318259042)
    try:
        sid = server.open(80, 48, pal, 2500)
        for k, w in enumerate((6e-13, 6e-13 * sv.ZOOM)):
            server.submit(sid, centre, w, 2500, k)
            done = wait_frame(server)[0]
            assert done.path == 2 and done.iterations > 0
            got = server.read_rgba(sid, 80, 48)
            d = cu.make_deep(centre[0], centre[1], w, 80, 48, 2500, 4.0)
            orbit, M = O.reference_orbit(centre[0], centre[1], 2500, 4.0, 64)
            ys, xs = np.mgrid[0:48, 0:80]
            c, mag, st, _ = O.deep_pixels(O.DEEP_PERTURB_F64, xs.ravel(), ys.ravel(), d, orbit, M)
            sm = np.where(st == 1, c + 1.0 - np.log2(np.log(np.where(st == 1, mag, 4.0)) / math.log(4.0)),
                          0.0)
            assert np.array_equal(got, O.compose_smooth(sm.reshape(48, 80), pal)), f"frame {k}"
    finally:
        server.close()
    with pytest.raises(cu.IllegalArgumentError, match="deep_method"):
        sv.Server(cu.KERNELS_DIR, 1 << 30, slots=1, deep_method=cu.DEEP_DIRECT_F64)


def test_poll_does_not_block_submit(gpu):
    """A thread waiting in poll for a long frame leaves the lock free between its
    checks: another thread's submit returns long before that poll does."""
    cu = gpu
    server = sv.Server(cu.KERNELS_DIR, 1 << 32, slots=1, max_super_sampling=16.0)
    pal = sv.palette256()
    try:
        big = server.open(3840, 2160, pal, 0)
        small = server.open(64, 36, pal, 0)
        times = {}

        def poller():
            t = time.perf_counter()
            got = []
            while not got:
                got = server.poll(wait_ms=5000)
            times["poll"] = (t, time.perf_counter(), [d.tag for d in got])

        # interior pixels run all 10^6 iterations: the frame takes well over 50 ms
        server.submit(big, (-0.743643887037151, 0.131825904205330), 1e-4, 1000000, 1)
        th = threading.Thread(target=poller)
        th.start()
        time.sleep(0.005)
        t = time.perf_counter()
        server.submit(small, (-0.5, 0.0), 3.0, 100, 2)
        times["submit"] = (t, time.perf_counter())
        th.join(60)
        assert not th.is_alive()
        p0, p1, tags = times["poll"]
        s0, s1 = times["submit"]
        assert tags[0] == 1
        assert p1 - p0 > 0.05, "the long frame finished too early to show anything"
        assert s1 - s0 < 0.01 and s1 < p1 - 0.01
        got = []
        while len(got) < 1:
            got += server.poll(wait_ms=100)
        assert got[0].tag == 2 and got[0].status == 0
    finally:
        server.close()


def test_full_hd_sessions_poisson_arrivals(gpu):
    """A 1920x1080 and a 1280x720 session with Poisson arrivals at 30 frames/s,
    replayed in real time: every frame completes, the same alone and in the mix."""
    cu = gpu
    ss = sv.zoom_sessions(2, 12, seed=7, resolutions=((1920, 1080),))
    ss[1]["width"], ss[1]["height"] = 1280, 720
    for s in ss:
        s["widths"] = 1e-3 * sv.ZOOM ** np.arange(12)
        s["max_iter"] = [sv.max_iter_for(w) for w in s["widths"]]
    alone = {}
    for s in ss:
        server = sv.Server(cu.KERNELS_DIR, 1 << 32, slots=4)
        try:
            recs, _ = sv.run(server, [s], realtime=True)
            alone[s["id"]] = [r["checksum"] for r in recs]
        finally:
            server.close()
    server = sv.Server(cu.KERNELS_DIR, 1 << 32, slots=4)
    try:
        recs, span = sv.run(server, ss, realtime=True)
    finally:
        server.close()
    mixed = {}
    for r in recs:
        mixed.setdefault(ss[r["session"]]["id"], []).append(r["checksum"])
    assert alone == mixed
    assert len(recs) == 24 and {r["path"] for r in recs} == {0, 1}
    lat = np.array([r["latency_s"] for r in recs])
    print(f"1080p+720p Poisson: p50 {np.percentile(lat, 50) * 1e3:.2f} ms, "
          f"p99 {np.percentile(lat, 99) * 1e3:.2f} ms")
