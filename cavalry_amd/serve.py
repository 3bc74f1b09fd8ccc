"""Render server over the C ABI (cav_serve_*) and a load harness.

``Server`` submits render requests to resident scenes and polls them;
``run_served`` / ``run_plain`` replay one seeded Poisson arrival trace through
the server (continuous batching) and through the plain per-request renderer
(cav_render, one request at a time in arrival order), and report latency
percentiles and throughput for the comparison.
"""
from __future__ import annotations

import ctypes
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from ._lib import RenderParams, ServeConfig, ServeResult, ServeStats, check, lib
from .render import DETERMINISTIC, INTEGRATORS, Film, Scene, render

QUEUED, RUNNING, DONE, CANCELLED, FAILED = range(5)  # CAV_SERVE_*


@dataclass
class Ticket:
    id: int
    film: Film
    integrator: str


class Server:
    def __init__(self, ctx, arena_bytes=0, host_arena_bytes=0, pool_tasks=0, chunk_samples=0,
                 max_steps_ahead=0, use_graphs=False, rl_workers=0):
        self.ctx = ctx
        self.pool_tasks = int(pool_tasks) if pool_tasks > 0 else 4 << 20
        self.chunk_samples = int(chunk_samples) if chunk_samples > 0 else self.pool_tasks // 4
        cfg = ServeConfig(int(arena_bytes), int(host_arena_bytes), int(pool_tasks), int(chunk_samples),
                          int(max_steps_ahead), 1 if use_graphs else 0, int(rl_workers))
        h = ctypes.c_void_p()
        check(ctx.handle, lib().cav_serve_open(ctx.handle, ctypes.byref(cfg), ctypes.byref(h)))
        self.handle = h
        # Films the server may still write (asynchronously, from its own
        # threads): each stays referenced until its request's final status is
        # polled or the server is closed, whatever the caller keeps.
        self._films = {}
        ctx.own(self)

    def rl_round_spp(self, pixels: int, spp: int, spp_per_round: int = 0) -> int:
        """Samples per pixel of each round the server gives an rlpath request
        (as the library chooses them): the film equals render.render's with
        spp_per_round set to this."""
        if spp_per_round > 0:
            return min(spp_per_round, spp)
        return max(1, min(spp, self.chunk_samples // pixels))

    def submit(self, scene: Scene, spp: int, integrator: str = "path", camera=None, max_depth=None,
               q_grid: int = 32, alpha_scale: float = 1.0, deterministic: bool = False,
               spp_per_round: int = 0) -> Ticket:
        """Queues a render (the parameters of render.render); deterministic
        applies to rlpath as there, spp_per_round to rlpath's rounds (0: as
        rl_round_spp chooses)."""
        cam = camera if camera is not None else scene.setup.camera
        depth = max_depth if max_depth is not None else (scene.arrays.max_depth if scene.arrays is not None
                                                         else scene.setup.max_depth)
        prm = RenderParams(INTEGRATORS[integrator], int(spp), int(depth), DETERMINISTIC if deterministic else 0,
                           int(q_grid), float(alpha_scale), int(spp_per_round))
        film = Film(np.zeros((cam.height, cam.width, 3), np.float32), np.zeros((cam.height, cam.width), np.float32))
        rid = ctypes.c_int64()
        check(self.ctx.handle, lib().cav_serve_submit(
            self.handle, scene.handle, ctypes.byref(cam), ctypes.byref(prm),
            film.rgb.ctypes.data_as(ctypes.c_void_p), film.weight.ctypes.data_as(ctypes.c_void_p),
            ctypes.byref(rid)))
        self._films[int(rid.value)] = film
        return Ticket(int(rid.value), film, integrator)

    def poll(self, t: Ticket, wait: bool = True) -> dict:
        r = ServeResult()
        check(self.ctx.handle, lib().cav_serve_poll(self.handle, t.id, 1 if wait else 0, ctypes.byref(r)))
        out = {k: getattr(r, k) for k, _ in r._fields_}
        if out["status"] in (DONE, CANCELLED, FAILED):
            self._films.pop(t.id, None)  # the server no longer writes it
            if out["status"] == FAILED:
                out["message"] = (lib().cav_last_error(self.ctx.handle) or b"").decode(errors="replace")
        return out

    def cancel(self, t: Ticket) -> None:
        check(self.ctx.handle, lib().cav_serve_cancel(self.handle, t.id))

    close_rank = 0  # closed before the scenes it renders (Context.close)

    def close(self) -> dict:
        st = ServeStats()
        if self.handle:
            check(None, lib().cav_serve_close(self.handle, ctypes.byref(st)))
            self.handle = None
            self._films.clear()
        return {k: getattr(st, k) for k, _ in st._fields_}

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------------ harness
@dataclass
class Req:
    scene: int        # index into the resident scene list
    camera: object    # cav_camera
    spp: int
    integrator: str
    max_depth: int
    arrival: float    # seconds from the start of the trace
    deterministic: bool = False  # rlpath: table updates summed in sample order (reproducible films)


@dataclass
class TraceResult:
    latencies_ms: list = field(default_factory=list)
    makespan_s: float = 0.0
    rays: int = 0
    films: list = field(default_factory=list)

    def summary(self) -> dict:
        lat = np.asarray(self.latencies_ms)
        return {"requests": len(lat), "p50_ms": round(float(np.percentile(lat, 50)), 2),
                "p99_ms": round(float(np.percentile(lat, 99)), 2), "mean_ms": round(float(lat.mean()), 2),
                "makespan_s": round(self.makespan_s, 3), "rays": int(self.rays),
                "mrays_s": round(self.rays / self.makespan_s / 1e6, 2)}


def poisson_trace(rng: np.random.Generator, n: int, rate: float, scenes: list, resolutions: list, spps: list,
                  integrators: list, max_depth: int, camera_fn, deterministic: bool = False) -> list:
    """n requests with exponential inter-arrival times of mean 1/rate; each
    picks a scene, resolution, spp and integrator uniformly, and a camera
    from camera_fn(rng, scene_index, width, height). deterministic applies
    to the rlpath requests."""
    t = 0.0
    out = []
    for _ in range(n):
        t += float(rng.exponential(1.0 / rate))
        s = int(rng.integers(len(scenes)))
        w, h = resolutions[int(rng.integers(len(resolutions)))]
        out.append(Req(s, camera_fn(rng, s, w, h), int(spps[int(rng.integers(len(spps)))]),
                       integrators[int(rng.integers(len(integrators)))], max_depth, t, deterministic))
    return out


def run_served(server: Server, scenes: list, trace: list, progress=None) -> TraceResult:
    """Submits each request at its arrival time (from a thread), waits for
    all, and measures arrival -> film-ready latency on the server's clock.
    `rays` stays 0: the requests share the pools' ray tallies, which
    cav_serve_close reports. progress(done, total), if given, is called as
    requests finish (every 16th and the last), and progress(-submitted,
    total) as they are submitted."""
    tickets = [None] * len(trace)
    start = time.perf_counter()

    def submitter():
        for k, r in enumerate(trace):
            delay = start + r.arrival - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            tickets[k] = server.submit(scenes[r.scene], r.spp, r.integrator, camera=r.camera,
                                       max_depth=r.max_depth, deterministic=r.deterministic)
            if progress is not None and ((k + 1) % 16 == 0 or k + 1 == len(trace)):
                progress(-(k + 1), len(trace))

    th = threading.Thread(target=submitter)
    th.start()
    th.join()
    res = TraceResult()
    first_submit, last_done = None, 0.0
    rl_rays = 0
    for k, t in enumerate(tickets):
        out = server.poll(t, wait=True)
        if progress is not None and ((k + 1) % 16 == 0 or k + 1 == len(tickets)):
            progress(k + 1, len(tickets))
        if out["status"] != DONE:
            raise RuntimeError(f"request {t.id} ended with status {out['status']} (error {out['error']}): "
                               f"{out.get('message', '')}")
        res.latencies_ms.append(out["done_ms"] - out["submit_ms"])
        first_submit = out["submit_ms"] if first_submit is None else min(first_submit, out["submit_ms"])
        last_done = max(last_done, out["done_ms"])
        rl_rays += out["closest_rays"] + out["shadow_rays"]
        res.films.append(t.film)
    res.makespan_s = (last_done - first_submit) / 1e3
    res.rays = rl_rays
    return res


def run_plain(ctx, scenes: list, trace: list, rl_round=None, progress=None) -> TraceResult:
    """The plain per-request renderer: one cav_render at a time, in arrival
    order, each starting when it has arrived and the previous one is done.
    rl_round(req) gives an rlpath request's spp_per_round (None: one round,
    as device memory allows); progress as for run_served."""
    res = TraceResult()
    start = time.perf_counter()
    first = None
    for r in trace:
        now = time.perf_counter() - start
        if now < r.arrival:
            time.sleep(r.arrival - now)
        arrived = start + r.arrival
        spr = rl_round(r) if (rl_round is not None and r.integrator == "rlpath") else 0
        film, st = render(ctx, scenes[r.scene], r.spp, integrator=r.integrator, camera=r.camera,
                          max_depth=r.max_depth, spp_per_round=spr, deterministic=r.deterministic)
        done = time.perf_counter()
        first = arrived if first is None else first
        res.latencies_ms.append((done - arrived) * 1e3)
        res.rays += st["closest_rays"] + st["shadow_rays"]
        res.films.append(film)
        if progress is not None and (len(res.films) % 16 == 0 or len(res.films) == len(trace)):
            progress(len(res.films), len(trace))
    res.makespan_s = time.perf_counter() - first
    return res
