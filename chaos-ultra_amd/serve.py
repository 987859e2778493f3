"""Python view of the serving loop (cu_serve_*) and the zoom-session workload.

Server wraps the C ABI.  zoom_sessions() draws a seeded set of sessions: each
picks a resolution from {1280x720, 1920x1080, 3840x2160}, a target point within
1e-3 of the Mandelbrot boundary (bisected along a real-axis row, as for the
uneven tiles), a start width log-uniform in [START_MIN, START_MAX], and frame
arrival times of a Poisson process at `fps`; frame k views width
start * 0.97**k around the target with maxIter = 500 + 200 log10(3 / width).
The start range reaches below 1e-12 for about a quarter of the sessions, so
their last frames run the perturbation form.

run() replays the arrivals in real time against a Server (each session opens
at its first arrival and closes after its last frame) and reports per-frame
arrival -> completion latency.
"""
from __future__ import annotations

import ctypes as C
import math
import time

import numpy as np

from . import abi
from .abi import FrameDone, FrameRequest, ServeConfig, ServeReport

RESOLUTIONS = ((1280, 720), (1920, 1080), (3840, 2160))
START_MIN, START_MAX = 1e-14, 3.0
ZOOM = 0.97
DEEP_WIDTH = 1e-12


def max_iter_for(width: float) -> int:
    return int(500 + 200 * math.log10(3.0 / width))


class Server:
    def __init__(self, kernels_dir: str, hbm_budget: int, slots: int = 4,
                 deep_method: int = abi.DEEP_PERTURB_F32, max_super_sampling: float = 2.0,
                 deep_width: float = DEEP_WIDTH, shallow_policy: int = abi.SERVE_SHALLOW_FP64):
        self.cfg = ServeConfig(int(hbm_budget), deep_width, slots, deep_method, max_super_sampling,
                               shallow_policy)
        self.handle = C.c_void_p()
        abi.call("cu_serve_create", kernels_dir.encode(), C.byref(self.cfg), C.byref(self.handle))
        self._buf = (FrameDone * 64)()

    def open(self, width: int, height: int, palette: np.ndarray, deep_max_iter: int = 0) -> int:
        pal = np.ascontiguousarray(palette, dtype=np.uint32)
        sid = C.c_uint32(0)
        abi.call("cu_serve_open", self.handle, width, height, pal.ctypes.data_as(C.c_void_p), pal.size,
                 deep_max_iter, C.byref(sid))
        return sid.value

    def submit(self, session: int, centre, width: float, max_iter: int, tag: int) -> None:
        r = FrameRequest(centre[0], centre[1], width, tag, max_iter, 0)
        abi.call("cu_serve_submit", self.handle, session, C.byref(r))

    def close_session(self, session: int) -> None:
        abi.call("cu_serve_close", self.handle, session)

    def poll(self, wait_ms: int = 0):
        n = C.c_uint32(0)
        abi.call("cu_serve_poll", self.handle, self._buf, len(self._buf), C.byref(n), wait_ms)
        return [FrameDone.from_buffer_copy(self._buf[k]) for k in range(n.value)]

    def last_failure(self) -> str:
        buf = C.create_string_buffer(512)
        abi.call("cu_serve_last_failure", self.handle, buf, len(buf))
        return buf.value.decode()

    def read_rgba(self, session: int, width: int, height: int) -> np.ndarray:
        out = np.empty(width * height, np.uint32)
        abi.call("cu_serve_read_rgba", self.handle, session, out.ctypes.data_as(C.c_void_p), out.size)
        return out.reshape(height, width)

    def report(self) -> ServeReport:
        r = ServeReport()
        abi.call("cu_serve_stats", self.handle, C.byref(r))
        return r

    def session_times(self, session: int):
        o, a = C.c_double(), C.c_double()
        abi.call("cu_serve_session_times", self.handle, session, C.byref(o), C.byref(a))
        return o.value, a.value

    def close(self) -> None:
        if self.handle and self.handle.value:
            abi.call("cu_serve_destroy", self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def zoom_sessions(n=16, frames=120, seed=20261202, fps=30.0, resolutions=RESOLUTIONS):
    """Seeded sessions: list of dicts with size, target, widths, max_iter, arrivals."""
    from .workloads import _member  # the bisection membership test of the tile generator
    rng = np.random.default_rng(seed)
    sessions = []
    for k in range(n):
        w, h = resolutions[int(rng.integers(len(resolutions)))]
        start = START_MIN * (START_MAX / START_MIN) ** rng.random()
        gaps = rng.exponential(1.0 / fps, frames)
        arrivals = np.cumsum(gaps)
        # target: a bounded/escaping transition on a random row, bisected to ~1e-16
        while True:
            y = rng.uniform(-1.15, 1.15)
            xs = np.linspace(-2.1, 0.6, 1024)
            ins = _member(np.zeros(1, bool), xs[None, :], np.full((1, xs.size), y),
                          np.zeros(1), np.zeros(1))[0]
            edges = np.nonzero(ins[:-1] != ins[1:])[0]
            if edges.size:
                e = int(edges[rng.integers(edges.size)])
                a, b, ia = xs[e], xs[e + 1], ins[e]
                for _ in range(48):
                    m = 0.5 * (a + b)
                    im = _member(np.zeros(1, bool), np.array([[m]]), np.array([[y]]),
                                 np.zeros(1), np.zeros(1))[0, 0]
                    if im == ia:
                        a = m
                    else:
                        b = m
                break
        widths = start * ZOOM ** np.arange(frames)
        sessions.append({"id": k, "width": w, "height": h, "target": (0.5 * (a + b), y),
                         "widths": widths, "max_iter": [max_iter_for(x) for x in widths],
                         "arrivals": arrivals})
    return sessions


def palette256():
    return ((np.arange(256, dtype=np.uint32) * 0x00010305) | 0xFF000000).astype(np.uint32)


def run(server: Server, sessions, realtime=True, palette=None, timeout_s=120.0):
    """Replay the sessions' arrivals against `server`.  Returns per-frame records
    (session index, frame, latency_s, path, checksum) and the wall time span."""
    palette = palette256() if palette is None else palette
    events = sorted((t, si, fi) for si, s in enumerate(sessions) for fi, t in enumerate(s["arrivals"]))
    sid = {}
    remaining = {si: len(s["arrivals"]) for si, s in enumerate(sessions)}
    records = {}
    t0 = time.perf_counter()
    k = 0

    def take(done):
        for d in done:
            si, fi = divmod(int(d.tag), 1 << 20)
            if d.status != 0:
                raise RuntimeError(f"frame {fi} of session {si} failed ({d.status}): "
                                   f"{server.last_failure()}")
            records[(si, fi)] = {"session": si, "frame": fi, "latency_s": d.done_s - d.submit_s,
                                 "queue_s": d.start_s - d.submit_s, "path": d.path,
                                 "checksum": int(d.checksum), "iterations": int(d.iterations),
                                 "done_s": d.done_s, "submit_s": d.submit_s}
            remaining[si] -= 1
            if remaining[si] == 0:
                server.close_session(sid[si])

    while k < len(events) or len(records) < len(events):
        if time.perf_counter() - t0 > timeout_s:
            raise TimeoutError(f"serving run exceeded {timeout_s} s")
        now = time.perf_counter() - t0
        while k < len(events) and (not realtime or events[k][0] <= now):
            _, si, fi = events[k]
            s = sessions[si]
            if si not in sid:
                deep_iter = max((m for m, w in zip(s["max_iter"], s["widths"]) if w < DEEP_WIDTH),
                                default=0)
                sid[si] = server.open(s["width"], s["height"], palette, deep_iter)
            server.submit(sid[si], s["target"], float(s["widths"][fi]), int(s["max_iter"][fi]),
                          (si << 20) | fi)
            k += 1
        take(server.poll(wait_ms=1 if k >= len(events) or not realtime else 0))
    span = max(r["done_s"] for r in records.values()) - min(r["submit_s"] for r in records.values())
    return [records[key] for key in sorted(records)], span


def summarise(records, sessions, report: ServeReport, span_s: float):
    lat = np.array([r["latency_s"] for r in records]) * 1e3
    k4 = np.array([sessions[r["session"]]["width"] == 3840 for r in records])
    paths = np.array([r["path"] for r in records])
    out = {"frames": len(records), "sessions": len(sessions),
           "p50_ms": round(float(np.percentile(lat, 50)), 3),
           "p99_ms": round(float(np.percentile(lat, 99)), 3),
           "max_ms": round(float(lat.max()), 3),
           "giga_iter_s": round(report.iterations / span_s / 1e9, 1),
           "iterations": int(report.iterations), "span_s": round(span_s, 3),
           "frames_reuse": int((paths == 0).sum()), "frames_full": int((paths == 1).sum()),
           "frames_deep": int((paths == 2).sum()),
           "peak_bytes": int(report.peak_bytes)}
    if k4.any():
        out["p50_ms_4k"] = round(float(np.percentile(lat[k4], 50)), 3)
        out["p99_ms_4k"] = round(float(np.percentile(lat[k4], 99)), 3)
    return out
