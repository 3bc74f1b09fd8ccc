"""Benchmark of the LBVH + traversal hot path on one MI355X.

Workload (one step): the LBVH build + traversal microbenchmark over N = 2^24
random triangles: scene bounds, 30-bit Morton codes, stable radix sort,
Karras radix tree, refit, 3 rounds of 7-leaf treelet restructuring, flatten,
then 2^24 closest-hit rays and 2^24 shadow rays. Inputs are generated once,
uploaded, and resident in HBM before the timed region.

Prints one JSON line. `value` is rays traced per second over the whole step
(build included); the same line carries the LBVH build rate in Mtris/s, the
per-stage and per-kernel device times (HIP events on the library's stream),
the roofline of the dominant LBVH kernel and the CPU-oracle baseline.

  python bench.py [--gpus 1] [--steps 5] [--warmup 2] [--n 16777216] ...
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Hardware queues this process may use (read when HIP initialises, so before
# torch or the library touches the GPU). The server runs one stream per
# resident scene's pool (4 here) beside the library's context stream and
# torch's; with HIP's default of 4 queues, streams share queues and their
# launches serialise. The single-stream paths do not depend on it. Set, not
# defaulted: an environment that already holds HIP's default 4 would
# otherwise keep it.
os.environ["GPU_MAX_HW_QUEUES"] = "16"

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s

# Algorithmic bytes of one launch of each LBVH build kernel (30-bit keys,
# K = 4 B), as functions of N:
#   bounds           reads the 36 B soup;
#   morton           reads the soup, writes key + index, clears the sort's
#                    look-back words (4 passes x 256 x N/4096 words = 1 B/tri);
#   sort_onesweep    one 8-bit digit pass: reads and writes key + index;
#   refit_bottom_up  per leaf: reads the sorted index, the triangle and its
#                    code, writes the 32 B box/area/cost record, the 12 B of
#                    its topology record it owns and its parent's id (4 B);
#                    per internal node: the 32 B record, 12 B of topology, the
#                    parent id and the 4 B height word. The codes of the
#                    boundaries and the boxes passed between arrivals stay
#                    in LDS except at the few splits whose leaves span two
#                    workgroups, which a second launch (refit_crossing)
#                    completes; the formula covers both launches.
ALG_BYTES_LAUNCH_30 = {
    "bounds": lambda n: 36 * n,
    "morton": lambda n: (36 + 4 + 4 + 1) * n,
    "sort_onesweep": lambda n: 2 * (4 + 4) * n,
    "refit_bottom_up": lambda n: (4 + 36 + 4 + 32 + 12 + 4) * n + (n - 1) * (32 + 12 + 4 + 4),
}
ALG_BYTES_PER_TRI_30 = 248  # whole build, canonical model: S1 68 + S2 68 + S3 20 + S4 92
TRACE_RECORD_BYTES = 64  # one traversal record (node or triangle) per step


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(n: int, nrays: int, rounds: int, seed: int) -> dict:
    """The CPU oracle (a C++ restatement of the reference algorithms) on a
    bounded sample of the same pipeline, timed on this host's cores."""
    import oracle
    from cavalry_amd import scenes
    tris = scenes.gmm_triangles(n, seed=seed)
    rays = scenes.random_rays(nrays, seed=seed + 1)
    srays = scenes.random_rays(nrays, seed=seed + 2, stream=2)
    lim = scenes.shadow_limits(nrays, seed=seed + 3)
    oracle.lib()
    t0 = time.perf_counter()
    _, _, nd = oracle.lbvh_build(tris, 30)
    t1 = time.perf_counter()
    log(f"[bench] cpu: LBVH build {t1 - t0:.1f} s")
    for k in range(rounds):  # one round at a time, each a line of progress
        oracle.treelet_optimize(nd, 1)
        log(f"[bench] cpu: treelet round {k + 1}/{rounds}")
    t2 = time.perf_counter()
    oracle.trace_closest(nd, tris, rays)
    t3 = time.perf_counter()
    log(f"[bench] cpu: closest-hit {t3 - t2:.1f} s")
    oracle.trace_occluded(nd, tris, srays, lim)
    t4 = time.perf_counter()
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    total = t4 - t0
    return {
        "value": round(2 * nrays / total / 1e6, 4),
        "unit": "Mrays/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"same pipeline at N=2^{n.bit_length() - 1} triangles with 2^"
                   f"{nrays.bit_length() - 1} closest + 2^{nrays.bit_length() - 1} shadow rays, "
                   f"{rounds} treelet rounds; LBVH/traversal OpenMP on {threads} threads, "
                   f"treelet climb serial; {total:.1f} s"),
        "lbvh_build_mtris_s": round(n / (t1 - t0) / 1e6, 3),
        "stage_s": {"lbvh": round(t1 - t0, 3), "treelet": round(t2 - t1, 3),
                    "closest": round(t3 - t2, 3), "shadow": round(t4 - t3, 3)},
    }


def _time_renders(ctx, scene, cam, integrator: str, renders: int) -> dict:
    import numpy as np
    from cavalry_amd import render as R
    R.render(ctx, scene, 16, integrator=integrator, camera=cam)  # warm-up (allocates)
    stats = []
    t0 = time.perf_counter()
    for _ in range(renders):
        _, st = R.render(ctx, scene, 16, integrator=integrator, camera=cam)
        stats.append(st)
    wall = (time.perf_counter() - t0) / renders
    ms = float(np.mean([st["ms_total"] for st in stats]))
    rays = stats[-1]["closest_rays"] + stats[-1]["shadow_rays"]
    return {
        "mrays_s": round(rays / (ms * 1e-3) / 1e6, 2),
        "ms_per_render": round(ms, 3),
        "wall_ms_per_render": round(wall * 1e3, 3),
        "rays_per_render": int(rays),
        "renders": renders,
    }


def render_bench(renders: int, cpu: bool, tune_launch: bool = True) -> dict:
    """Path tracing on the Cornell-box workload: 512x512, 16 spp, max depth 8,
    no treelet rounds, with the `path` and `rlpath` integrators.
    Mrays/s = (closest + shadow rays) / device time."""
    from cavalry_amd import Context, render as R, scenes
    from cavalry_amd.camera import for_scene
    sc = scenes.cornell_box(seed=1, width=512, height=512, max_depth=8)
    ctx = Context(0)
    scene = R.Scene(ctx, sc, bvh_rounds=0)
    cam = for_scene(sc)
    out = {"workload": "Cornell box (4014 triangles), 512x512, 16 spp, path, max depth 8, Halton"}
    if tune_launch:  # the render loops' traversal launch shape, autotuned on this render
        from cavalry_amd import tune
        res = tune.tune_render(ctx, scene, 16, integrator="path", camera=cam, reps=1, use_cache=False)
        out["trace_launch"] = {"shape": res["shape"], "candidates": len(res["timings"])}
    out.update(_time_renders(ctx, scene, cam, "path", renders))
    rl = {"workload": "same scene and sampling, rlpath (32^3 Q entries x 128 bins, reference learning rate)"}
    rl.update(_time_renders(ctx, scene, cam, "rlpath", renders))
    out["rlpath"] = rl
    scene.close()
    ctx.close()
    if cpu:
        import oracle
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        t0 = time.perf_counter()
        _, _, rays_c = oracle.render_path(sc, 16, cam)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {
            "value": round(rays_c / dt / 1e6, 3), "unit": "Mrays/s", "cores": threads,
            "kind": "port",
            "sample": f"the same 512x512x16 render (incl. BVH build), {dt:.1f} s on {threads} threads",
        }
        small = scenes.cornell_box(seed=1, width=128, height=128, max_depth=8)
        cam_s = for_scene(small)
        t0 = time.perf_counter()
        _, _, rays_r = oracle.render_rlpath(small, 16, cam_s)
        dt = time.perf_counter() - t0
        rl["cpu_baseline"] = {
            "value": round(rays_r / dt / 1e6, 3), "unit": "Mrays/s", "cores": 1,
            "kind": "port",
            "sample": f"same scene at 128x128, 16 spp, rlpath (incl. BVH build), {dt:.1f} s, single thread",
        }
    return out


def serving_bench(n_requests: int, rates: list, seed: int, use_graphs: bool = True,
                  deterministic_rl: bool = True, pool_tasks: int = 32 << 20, steps_ahead: int = 0) -> dict:
    """Serving on four resident interior scenes (about 0.1M, 0.5M, 2M and 8M
    triangles, 1024^2 Kd textures on about a third of the meshes): for each
    arrival rate, one seeded Poisson trace of render requests (resolution
    from 256^2, 512^2, 1280x720, 1920x1080; spp from 1, 4, 16, 64; path /
    rlpath 50/50; max depth 8; the same requests at every rate) replayed
    through the plain per-request renderer (one cav_render at a time, in
    arrival order) and through the server (continuous batching; rlpath
    requests pooled one round at a time; bounce steps replayed from graphs).
    The plain renderer renders each rlpath request with the server's round
    size and both sum rlpath's table updates in sample order (deterministic
    mode), so both make the same film, bit for bit. Latency = arrival ->
    film ready.
    Throughput = traced rays / (first arrival -> last film)."""
    import numpy as np
    from cavalry_amd import Context, render as R, scenes
    from cavalry_amd.camera import look_at
    from cavalry_amd.serve import Server, poisson_trace, run_plain, run_served

    sizes = [100_000, 500_000, 2_000_000, 8_000_000]
    ctx = Context(0)
    log("[bench] serving: generating 4 resident scenes")
    arrays = [scenes.interior(n, seed=seed + k, width=64, height=36) for k, n in enumerate(sizes)]
    res_scenes = [R.Scene(ctx, a, bvh_rounds=3) for a in arrays]

    def camera(rng, s, w, h):
        eye = np.float32([rng.uniform(1, 9), rng.uniform(1, 3), rng.uniform(0.5, 2.0)])
        at = np.float32([rng.uniform(2, 8), rng.uniform(0.3, 1.5), rng.uniform(4, 9)])
        return look_at(eye, at, (0, 1, 0), 60.0, w, h)

    cfg = dict(arena_bytes=64 << 30, host_arena_bytes=4 << 30, pool_tasks=pool_tasks, use_graphs=use_graphs,
               max_steps_ahead=steps_ahead)
    per_rate = []
    for rate in rates:
        rng = np.random.default_rng(seed)
        trace = poisson_trace(rng, n_requests, rate, res_scenes,
                              [(256, 256), (512, 512), (1280, 720), (1920, 1080)], [1, 4, 16, 64],
                              ["path", "rlpath"], 8, camera, deterministic=deterministic_rl)
        srv = Server(ctx, **cfg)
        rl_round = (lambda r, srv=srv: srv.rl_round_spp(r.camera.width * r.camera.height, r.spp))
        log(f"[bench] serving at {rate}/s: plain per-request renderer, {n_requests} requests")
        prog = (lambda k, n, rate=rate: log(f"[bench] serving at {rate}/s: {abs(k)}/{n} requests "
                                             f"{'submitted' if k < 0 else 'done'}"))
        plain = run_plain(ctx, res_scenes, trace, rl_round=rl_round, progress=prog)
        log(f"[bench] serving at {rate}/s: server (continuous batching)")
        served = run_served(srv, res_scenes, trace, progress=prog)
        st = srv.close()
        served.rays += st["pool_closest_rays"] + st["pool_shadow_rays"]
        same = True
        same_rl = True
        worst_rl = 0.0
        for r, a, b in zip(trace, plain.films, served.films):
            ident = np.array_equal(a.rgb, b.rgb) and np.array_equal(a.weight, b.weight)
            if r.integrator == "rlpath":
                # with float-atomic table updates an rlpath film is not
                # bit-reproducible (the deterministic mode's is); compared in
                # the display range so a single firefly does not dominate
                same_rl = same_rl and ident
                ia = np.clip(a.image.astype(np.float64), 0, 1)
                ib = np.clip(b.image.astype(np.float64), 0, 1)
                worst_rl = max(worst_rl, float(np.mean((ib - ia) ** 2 / (ia * ia + 1e-2))))
            else:
                same = same and ident
        p, s_ = plain.summary(), served.summary()
        per_rate.append({
            "rate_per_s": rate, "plain": p, "served": s_,
            "speedup_mrays": round(s_["mrays_s"] / p["mrays_s"], 3),
            "path_films_identical": bool(same),
            "rlpath_films_identical": bool(same_rl),
            "rlpath_max_rel_mse_vs_plain_ldr": worst_rl,
            "server_steps": st["steps"], "graph_pools": st["graph_pools"],
        })
        log(f"[bench] serving at {rate}/s: served/plain {per_rate[-1]['speedup_mrays']}x, "
            f"p50 {s_['p50_ms']} vs {p['p50_ms']} ms, p99 {s_['p99_ms']} vs {p['p99_ms']} ms")
    for sc in res_scenes:
        sc.close()
    ctx.close()
    return {
        "workload": f"{n_requests} requests per rate, Poisson arrivals, 4 resident interior scenes "
                    f"({', '.join(str(len(a.tris)) for a in arrays)} tris, 4 1024^2 textures each), "
                    "res 256^2..1920x1080, spp 1/4/16/64, path/rlpath 50/50, max depth 8"
                    + (", rlpath deterministic" if deterministic_rl else ""),
        "rates": per_rate,
        "speedup_mrays_min": min(r["speedup_mrays"] for r in per_rate),
        "hw_queues": int(os.environ.get("GPU_MAX_HW_QUEUES", "4")),
        "use_graphs": use_graphs,
        "rlpath_deterministic": deterministic_rl,
        "pool_tasks": pool_tasks,
    }


def equal_time_bench(seed: int, n_tris: int, width: int, height: int, ref_spp: int) -> dict:
    """rlpath against path at equal render time on an interior scene (max
    depth 16, 3 treelet rounds). The reference image is a path render of
    ref_spp samples per pixel whose film sums are binary64 (FILM_F64). Each
    time budget is the device time of path at 64, 256 and 1024 spp; rlpath
    (64^3 Q entries x 128 bins as the workload asks, and the reference's 32^3;
    the reference's learning rate) renders at the spp that a linear fit of
    its own time against spp (measured at 64 and 1024 spp) puts within that
    budget. Error = rel-MSE against the reference
    image. Each render pays a fixed device time (about 90 ms on 2M
    triangles at max depth 16: 32 traversal launches, each as long as its
    slowest ray), so budgets below ~64 spp measure that, not the
    integrators."""
    import numpy as np
    from cavalry_amd import Context, render as R, scenes
    from cavalry_amd.camera import for_scene

    sc = scenes.interior(n_tris, seed=seed, width=width, height=height, max_depth=16)
    ctx = Context(0)
    scene = R.Scene(ctx, sc, bvh_rounds=3)
    cam = for_scene(sc)
    log(f"[bench] equal time: reference image, {ref_spp} spp, binary64 film")
    ref_film, ref_st = R.render(ctx, scene, ref_spp, integrator="path", camera=cam, film_f64=True,
                               spp_per_round=min(ref_spp, 1024))
    ref = ref_film.rgb.astype(np.float64) / np.maximum(ref_film.weight.astype(np.float64), 1e-30)[..., None]

    def run(integrator, spp, q_grid=64, alpha_scale=1.0):
        # each integrator with its faster hit order: rlpath's 128-bin BSDF
        # evaluation diverges less over material-sorted hits (-12%), path's
        # shading is too cheap to pay for the sort (+5%)
        film, st = R.render(ctx, scene, spp, integrator=integrator, camera=cam, q_grid=q_grid,
                            alpha_scale=alpha_scale, sort_by_material=integrator == "rlpath")
        return {"spp": int(spp), "ms": round(float(st["ms_total"]), 3),
                "rel_mse": R.rel_mse(film.image, ref),
                "rel_mse_ldr": R.rel_mse(np.clip(film.image, 0, 1), np.clip(ref, 0, 1)),
                "mean": float(film.image.mean()),
                "mrays_s": round(float(st["mrays_per_s"]), 1)}

    run("path", 4)
    run("rlpath", 4)  # warm-up: allocations
    fits = {}
    for q in (64, 32):  # the workload's 64^3 Q entries, and the reference's 32^3
        f = {s: run("rlpath", s, q) for s in (64, 1024)}
        slope = (f[1024]["ms"] - f[64]["ms"]) / 960.0
        fits[q] = (slope, f[64]["ms"] - 64 * slope)
    budgets = []
    for spp in (64, 256, 1024):
        p = run("path", spp)
        b = {"budget_ms": p["ms"], "path": p}
        for q, (slope, icpt) in fits.items():
            rl_spp = max(1, int((p["ms"] - icpt) / slope)) if slope > 0 else spp
            rl = run("rlpath", rl_spp, q)
            b[f"rlpath_q{q}"] = rl
            b[f"rlpath_q{q}_over_path_error"] = round(rl["rel_mse"] / p["rel_mse"], 4)
            log(f"[bench] equal time {p['ms']:.1f} ms: path {spp} spp rel-MSE {p['rel_mse']:.3e}, "
                f"rlpath {q}^3 {rl_spp} spp ({rl['ms']:.1f} ms) rel-MSE {rl['rel_mse']:.3e}")
            if q == 64:
                # the workload's learning rate: 0.3 times the reference's
                # n / (T + n) (the same per-spp cost)
                rl3 = run("rlpath", rl_spp, q, alpha_scale=0.3)
                b["rlpath_q64_alpha0.3"] = rl3
                b["rlpath_q64_alpha0.3_over_path_error"] = round(rl3["rel_mse"] / p["rel_mse"], 4)
                log(f"[bench] equal time {p['ms']:.1f} ms: rlpath 64^3 alpha x0.3 {rl_spp} spp "
                    f"({rl3['ms']:.1f} ms) rel-MSE {rl3['rel_mse']:.3e}")
        budgets.append(b)
    scene.close()
    ctx.close()
    return {
        "workload": (f"interior scene ({len(sc.tris)} tris, 4 1024^2 textures, 64 quad lights), "
                     f"{width}x{height}, max depth 16, 3 treelet rounds; reference: path at {ref_spp} "
                     "spp with binary64 film sums; rlpath with 64^3 and 32^3 Q entries (the reference's "
                     "learning rate n/(T+n)) and with 64^3 entries at 0.3 n/(T+n), material-sorted hits"),
        "reference": {"spp": ref_spp, "ms": round(float(ref_st["ms_total"]), 1),
                      "mrays_s": round(float(ref_st["mrays_per_s"]), 1),
                      # rel-MSE falls as 1 / spp: the reference's own error, from path at 1024 spp
                      "own_rel_mse_est": budgets[-1]["path"]["rel_mse"] * 1024 / ref_spp},
        "reference_mean": float(ref.mean()),
        "rlpath_time_fit_ms": {f"q{q}": {"per_spp": round(a, 4), "fixed": round(c, 3)} for q, (a, c) in fits.items()},
        "budgets": budgets,
    }


def config2_bench(seed: int, sizes: list, width: int, height: int, spp: int, cpu: bool) -> dict:
    """Wavefront path tracing at scale: the procedural interior (displaced
    icospheres and boxes, 64 quad lights, matte / plastic / mirror / glass,
    1024^2 textures) at each triangle count, 3 treelet rounds, max depth 16
    (Russian roulette from bounce 3), width x height at spp samples per
    pixel, with `path` and with `rlpath` (64^3 Q entries x 128 bins,
    material-sorted hits). Mrays/s = (closest + shadow rays) / device time
    of the render (scene upload and BVH build excluded), with the render
    loops' default launch shape (shapes autotuned on short renders of the
    same view measured slower at 64 spp: 2M triangles, path 1517 against
    2200 Mrays/s). The CPU oracle
    renders a 256x256 crop at 4 spp of the first scene with `path`, over the
    GPU's BVH (bit-identical to the oracle's own; its 3 treelet rounds would
    take the oracle ~30 s single-threaded), timed on this host's cores."""
    import numpy as np
    import oracle
    from cavalry_amd import Context, render as R, scenes
    from cavalry_amd.camera import look_at
    ctx = Context(0)
    out = {"workload": (f"interior scenes, {width}x{height}, {spp} spp, max depth 16, 3 treelet rounds; "
                        "path, and rlpath with 64^3 Q entries"), "scenes": []}
    for k, n in enumerate(sizes):
        log(f"[bench] config 2: interior, {n} triangles")
        arr = scenes.interior(n, seed=seed + k, width=width, height=height, max_depth=16)
        sc = R.Scene(ctx, arr, bvh_rounds=3)
        cam = look_at(arr.eye, arr.center, arr.up, arr.fov, width, height)
        rec = {"n_tris": int(len(arr.tris))}
        R.render(ctx, sc, 1, camera=cam, max_depth=16)  # warm-up (allocates)
        for integ, kw in (("path", {}), ("rlpath", {"q_grid": 64, "sort_by_material": True})):
            film, st = R.render(ctx, sc, spp, integrator=integ, camera=cam, max_depth=16, **kw)
            rays = int(st["closest_rays"] + st["shadow_rays"])
            rec[integ] = {"mrays_s": round(rays / (st["ms_total"] * 1e-3) / 1e6, 1),
                          "ms": round(float(st["ms_total"]), 1), "rays": rays,
                          "mean": round(float(film.image.mean()), 5)}
            log(f"[bench] config 2: {n} tris, {integ}: {rec[integ]['mrays_s']} Mrays/s ({rec[integ]['ms']} ms)")
        if cpu and k == 0:
            ex = sc.bvh().export()
            nd = oracle.Nodes(len(arr.tris))
            for key in ("box", "sa", "cost", "left", "right", "parent", "prim"):
                setattr(nd, key, np.ascontiguousarray(ex[key]))
            x0, y0 = width // 2 - 128, height // 2 - 128
            t0 = time.perf_counter()
            _, _, rays_c = oracle.render_path(arr, 4, cam, crop=(x0, y0, x0 + 256, y0 + 256), max_depth=16, bvh=nd)
            dt = time.perf_counter() - t0
            threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
            out["cpu_baseline"] = {
                "value": round(rays_c / dt / 1e6, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
                "sample": (f"path, the 256x256 crop at the image centre at 4 spp, max depth 16, on the "
                           f"{len(arr.tris)}-triangle scene; {dt:.1f} s on {threads} threads")}
            log(f"[bench] config 2: CPU oracle crop {out['cpu_baseline']['value']} Mrays/s")
        out["scenes"].append(rec)
        sc.close()
    ctx.close()
    return out


def cpu_one_thread(n: int, nrays: int, rounds: int, seed: int) -> dict:
    """The CPU oracle's LBVH build and traversal on one thread (a child
    process with OMP_NUM_THREADS=1), on a bounded sample."""
    import subprocess
    code = (f"import json, bench; print(json.dumps(bench.cpu_baseline({n}, {nrays}, {rounds}, {seed})))")
    env = dict(os.environ, OMP_NUM_THREADS="1")
    r = subprocess.run([sys.executable, "-c", code], cwd=ROOT, env=env, capture_output=True, text=True,
                       timeout=600)
    if r.returncode != 0:
        return {"error": r.stderr[-500:]}
    return json.loads(r.stdout.strip().splitlines()[-1])


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1 << 24)
    ap.add_argument("--rays", type=int, default=1 << 24)
    ap.add_argument("--rounds", type=int, default=3)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-n", type=int, default=1 << 20)
    ap.add_argument("--cpu-rays", type=int, default=1 << 20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-only", action="store_true",
                    help="time only the CPU oracle on --cpu-n triangles and --cpu-rays rays, print its JSON and exit")
    ap.add_argument("--renders", type=int, default=3)
    ap.add_argument("--no-render", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=64, help="requests per arrival rate")
    ap.add_argument("--serve-rates", default="5,50", help="Poisson arrival rates (req/s), comma-separated")
    ap.add_argument("--serve-no-graphs", action="store_true")
    ap.add_argument("--serve-pool-tasks", type=int, default=32 << 20, help="task slots of each scene's ray pool")
    ap.add_argument("--serve-steps-ahead", type=int, default=0,
                    help="bounce steps the server's host thread may run ahead of the device (0: the library's 4)")
    ap.add_argument("--serve-rl-atomic", action="store_true",
                    help="rlpath requests with float-atomic table updates (films not bit-reproducible)")
    ap.add_argument("--no-equal-time", action="store_true")
    ap.add_argument("--no-config2", action="store_true")
    ap.add_argument("--c2-tris", default="2000000,8000000", help="triangle counts of the at-scale interiors")
    ap.add_argument("--c2-res", default="1920x1080")
    ap.add_argument("--c2-spp", type=int, default=64)
    ap.add_argument("--et-tris", type=int, default=2_000_000)
    ap.add_argument("--et-res", default="320x180", help="WxH of the equal-time comparison")
    ap.add_argument("--et-ref-spp", type=int, default=16384)
    ap.add_argument("--launch", default="",
                    help="JSON launch shape of the standalone traces instead of autotuning one")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the default traversal launch shapes instead of autotuning them")
    ap.add_argument("--tune-rays", type=int, default=1 << 22,
                    help="rays (a prefix of the workload's) the autotuner times each candidate on")
    ap.add_argument("--probe", action="store_true",
                    help="one LBVH build, one closest-hit and one shadow trace, and one 2^28-byte "
                         "elementwise add, then exit (for rocprofv3 --pmc runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); "
              "this build has no multi-GPU path", file=sys.stderr)
        return 2
    if int(os.environ.get("RANK", "0")) != 0:
        return 0
    if args.cpu_only:
        print(json.dumps(cpu_baseline(args.cpu_n, args.cpu_rays, args.rounds, args.seed)), flush=True)
        return 0

    import numpy as np
    import torch

    from cavalry_amd import BVH, Context, scenes
    from cavalry_amd._lib import check, lib

    n, r = args.n, args.rays
    log(f"[bench] generating {n} triangles and 2x{r} rays (seed {args.seed})")
    tris_h = scenes.gmm_triangles(n, seed=args.seed)
    rays_h = scenes.random_rays(r, seed=args.seed + 1)
    srays_h = scenes.random_rays(r, seed=args.seed + 2, stream=2)
    lim_h = scenes.shadow_limits(r, seed=args.seed + 3)
    dev = torch.device("cuda:0")
    tris = torch.from_numpy(tris_h).to(dev)
    rays = torch.from_numpy(rays_h).to(dev)
    srays = torch.from_numpy(srays_h).to(dev)
    lim = torch.from_numpy(lim_h).to(dev)
    hit_prim = torch.empty(r, dtype=torch.int32, device=dev)
    hit_t = torch.empty(r, dtype=torch.float32, device=dev)
    vis = torch.empty(r, dtype=torch.uint8, device=dev)
    del tris_h, rays_h, srays_h, lim_h
    torch.cuda.synchronize()

    ctx = Context(0)
    stream = torch.cuda.ExternalStream(ctx.stream, device=dev)
    log("[bench] first build (allocates)")
    bvh = BVH.build(ctx, tris, optimization_rounds=args.rounds, morton_bits=30, device=True)
    launch = None
    if args.launch:
        # a fixed launch shape (as the autotuner reports it), e.g. to profile
        # the shape a timed run chose
        from cavalry_amd import tune
        tune.set_launch(ctx, tune.STANDALONE, json.loads(args.launch))
        launch = {"shape": tune.get_launch(ctx, tune.STANDALONE), "fixed": True}
        log(f"[bench] launch shape fixed to {launch['shape']}")
    elif not args.probe and not args.no_tune:
        # the traversal launch shape, autotuned on a prefix of the same rays
        # (before the warmup; not part of the timed steps)
        from cavalry_amd import tune
        tr = min(args.tune_rays, r)
        t0 = time.perf_counter()
        res = tune.tune_trace(ctx, bvh, rays, tr, shadow_rays=srays, limit=lim, reps=1, use_cache=False)
        launch = {"shape": res["shape"], "tuned_on_rays": tr, "candidates": len(res["timings"]),
                  "tune_s": round(time.perf_counter() - t0, 2)}
        log(f"[bench] autotuned launch shape {res['shape']} in {launch['tune_s']} s")
    if args.probe:
        check(ctx.handle, lib().cav_trace_closest_async(
            ctx.handle, bvh.handle, rays.data_ptr(), r, hit_prim.data_ptr(), hit_t.data_ptr()))
        check(ctx.handle, lib().cav_trace_occluded_async(
            ctx.handle, bvh.handle, srays.data_ptr(), lim.data_ptr(), r, vis.data_ptr()))
        ctx.synchronize()
        # an elementwise kernel (a copy may go to a DMA engine, which the
        # counters do not see) calibrates the HBM byte counters: it reads and
        # writes exactly 2^28 B
        src = torch.zeros(1 << 28, dtype=torch.uint8, device=dev)
        dst = torch.empty_like(src)
        torch.add(src, 1, out=dst)
        torch.cuda.synchronize()
        log("[bench] probe done: one build, two traces, one 2^28-byte elementwise kernel")
        return 0

    def step(record: bool):
        bvh.rebuild(tris, optimization_rounds=args.rounds, morton_bits=30)
        e = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
        e[0].record(stream)
        check(ctx.handle, lib().cav_trace_closest_async(
            ctx.handle, bvh.handle, rays.data_ptr(), r, hit_prim.data_ptr(), hit_t.data_ptr()))
        e[1].record(stream)
        check(ctx.handle, lib().cav_trace_occluded_async(
            ctx.handle, bvh.handle, srays.data_ptr(), lim.data_ptr(), r, vis.data_ptr()))
        e[2].record(stream)
        ctx.synchronize()
        if not record:
            return None
        return bvh.stats(), bvh.kernel_times(), e[0].elapsed_time(e[1]), e[1].elapsed_time(e[2])

    for i in range(args.warmup):
        step(False)
        log(f"[bench] warmup {i + 1}/{args.warmup} done")

    # the records each visiting order fetches, on untimed traces with no
    # hand-off to the cooperative passes (the closest-hit pass restarts its
    # rays and visits in another order; the any-hit pass visits the same
    # pairs): closest-hit in the reference's order, any-hit with the limit
    # cull (the default) and along the whole line (the reference's order).
    # Their answers are kept to check the timed traces' at full size.
    from cavalry_amd import tune as _tune
    shape = _tune.get_launch(ctx, _tune.STANDALONE)
    _tune.set_launch(ctx, _tune.STANDALONE, dict(shape, handoff_steps=1 << 30))
    bvh.trace_fetches(reset=True)
    check(ctx.handle, lib().cav_trace_closest_async(
        ctx.handle, bvh.handle, rays.data_ptr(), r, hit_prim.data_ptr(), hit_t.data_ptr()))
    check(ctx.handle, lib().cav_trace_occluded_async(
        ctx.handle, bvh.handle, srays.data_ptr(), lim.data_ptr(), r, vis.data_ptr()))
    ctx.synchronize()
    ref_fetch = bvh.trace_fetches(reset=True)
    ref_prim, ref_t = hit_prim.cpu().numpy(), hit_t.cpu().numpy()
    cull_vis = vis.cpu().numpy()
    _tune.set_launch(ctx, _tune.STANDALONE, dict(shape, handoff_steps=1 << 30, any_hit_whole_line=1))
    check(ctx.handle, lib().cav_trace_occluded_async(
        ctx.handle, bvh.handle, srays.data_ptr(), lim.data_ptr(), r, vis.data_ptr()))
    ctx.synchronize()
    whole_fetch = bvh.trace_fetches(reset=True)
    ref_vis = vis.cpu().numpy()
    _tune.set_launch(ctx, _tune.STANDALONE, shape)

    torch.cuda.synchronize()
    ctx.synchronize()
    bvh.trace_counters(reset=True)  # count stack overflows of the timed traces only
    bvh.trace_fetches(reset=True)
    recs = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        recs.append(step(True))
    ctx.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    log(f"[bench] {args.steps} timed steps in {wall:.3f} s")
    # every timed ray finished inside the reference's 64-entry stack, or the
    # results are partial and the run fails
    trav = bvh.trace_counters()
    fetches = bvh.trace_fetches()
    trav["max_leaf_depth"] = bvh.max_depth()
    trav["stack_entries"] = 64
    # the timed traces' answers (cooperative passes, limit cull) against the
    # untimed reference-order traces, every ray
    trav["closest_mismatch_vs_reference_order"] = int(
        (hit_prim.cpu().numpy() != ref_prim).sum() + (hit_t.cpu().numpy().view(np.uint32) != ref_t.view(np.uint32)).sum())
    tvis = vis.cpu().numpy()
    trav["occluded_mismatch_vs_reference_order"] = int((tvis != ref_vis).sum())
    trav["occluded_mismatch_cull_untimed"] = int((cull_vis != ref_vis).sum())
    del ref_prim, ref_t, ref_vis, cull_vis, tvis
    log(f"[bench] traversal counters {trav}")

    ms_step = wall * 1000.0 / args.steps
    stages = {}
    for k in ("ms_bounds", "ms_morton", "ms_sort", "ms_karras", "ms_refit", "ms_treelet", "ms_flatten"):
        stages[k[3:]] = round(float(np.mean([rec[0][k] for rec in recs])), 4)
    stages["trace_closest"] = round(float(np.mean([rec[2] for rec in recs])), 4)
    stages["trace_shadow"] = round(float(np.mean([rec[3] for rec in recs])), 4)
    # (the radix tree is emitted by the refit's climb: karras is 0)
    lbvh_ms = sum(stages[k] for k in ("bounds", "morton", "sort", "karras", "refit"))

    # per-kernel averages over the timed steps (same launch sequence every step)
    kernels = {}
    for rec in recs:
        for name, ms in rec[1]:
            kernels.setdefault(name, []).append(ms)
    per_launch = {k: float(np.mean(v)) for k, v in kernels.items()}
    launches = {k: len(v) // args.steps for k, v in kernels.items()}
    totals = {k: per_launch[k] * launches[k] for k in per_launch}

    # roofline of the LBVH build's dominant kernel (HBM-bound model); the
    # climb's two passes (within workgroups, then the ~2% of splits that
    # cross them) count as one unit
    unit_ms = dict(per_launch)
    if "refit_bottom_up" in unit_ms:
        unit_ms["refit_bottom_up"] += per_launch.get("refit_crossing", 0.0)
    lbvh_kernels = [k for k in ALG_BYTES_LAUNCH_30 if k in totals]
    dom = max(lbvh_kernels, key=lambda k: unit_ms[k] * launches[k])
    alg = ALG_BYTES_LAUNCH_30[dom](n)
    achieved = alg / (unit_ms[dom] * 1e-3) / 1e9
    lbvh_roofline = {
        "bound": "hbm", "kernel": dom + (" + refit_crossing" if dom == "refit_bottom_up" else ""),
        "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
        "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
        "avg_launch_ms": round(unit_ms[dom], 4), "launches_per_step": launches[dom],
        "alg_bytes_per_launch": alg,
        "build_frac": round(ALG_BYTES_PER_TRI_30 * n / (lbvh_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "build_alg_bytes_per_tri": ALG_BYTES_PER_TRI_30,
    }
    # roofline of the step's dominant kernel, the traversal: algorithmic bytes
    # = the 64-byte records the reference's visiting order fetches (counted by
    # the kernel over the timed steps) + the rays read and the hits written
    fetch = {k: v / args.steps for k, v in fetches.items()}
    trav_bytes = {
        "closest": ref_fetch["closest"] * TRACE_RECORD_BYTES + r * (24 + 8),
        "occluded": ref_fetch["occluded"] * TRACE_RECORD_BYTES + r * (24 + 4 + 1),
    }
    trav_ms = {"closest": stages["trace_closest"], "occluded": stages["trace_shadow"]}
    tdom = max(trav_ms, key=lambda k: trav_ms[k])
    t_ach = trav_bytes[tdom] / (trav_ms[tdom] * 1e-3) / 1e9
    roofline = {
        "bound": "hbm", "kernel": f"k_trace<{'closest' if tdom == 'closest' else 'any'}>",
        "achieved": round(t_ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(t_ach / HBM_PEAK_GBS, 4), "traffic": None,
        "avg_launch_ms": round(trav_ms[tdom], 4), "alg_bytes_per_launch": int(trav_bytes[tdom]),
        "records_per_ray": round(ref_fetch[tdom] / r, 1),
        "records_fetched_per_ray": round(fetch[tdom] / r, 1),
        "note": "algorithmic bytes count every record the visiting order fetches (closest-hit: the "
                "reference's; any-hit: the reference's up to the ray's limit), records_per_ray, counted "
                "on an untimed trace without the cooperative passes, whether L2 or HBM serves it; "
                "records_fetched_per_ray is what the timed traces fetched",
        "lbvh_build_frac": lbvh_roofline["build_frac"],
    }
    traffic_file = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(traffic_file):
        with open(traffic_file) as f:
            tr = json.load(f)
        if tr.get(roofline["kernel"]):
            roofline["traffic"] = tr[roofline["kernel"]]
            roofline["traffic_source"] = tr.get("_source", "profiles/traffic.json")
        if tr.get(lbvh_roofline["kernel"]):
            lbvh_roofline["traffic"] = tr[lbvh_roofline["kernel"]]
            lbvh_roofline["traffic_source"] = tr.get("_source", "profiles/traffic.json")
    trav["records_per_ray_closest"] = round(ref_fetch["closest"] / r, 1)
    trav["records_per_ray_occluded"] = round(ref_fetch["occluded"] / r, 1)
    trav["records_per_ray_occluded_whole_line"] = round(whole_fetch["occluded"] / r, 1)
    trav["records_fetched_per_ray_closest"] = round(fetch["closest"] / r, 1)
    trav["records_fetched_per_ray_occluded"] = round(fetch["occluded"] / r, 1)
    trav["long_closest_rays"] = int(fetches.get("long_closest_rays", 0))
    trav["long_closest_retraced"] = int(fetches.get("long_closest_retraced", 0))
    trav["spilled_any_rays"] = int(fetches.get("spilled_rays", 0))
    trav["spilled_any_retraced"] = int(fetches.get("respilled_rays", 0))

    value = 2 * r / (ms_step * 1e-3) / 1e6
    out = {
        "metric": "Mrays/s path tracing; p50/p99 request latency; LBVH build Mtris/s",
        "value": round(value, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": (f"LBVH build (30-bit Morton, radix sort, Karras, refit) + {args.rounds} "
                         f"treelet rounds + flatten over N=2^{n.bit_length() - 1} Gaussian-mixture "
                         f"triangles, then 2^{r.bit_length() - 1} closest-hit + "
                         f"2^{r.bit_length() - 1} shadow rays"),
            "n_tris": n, "closest_rays": r, "shadow_rays": r, "treelet_rounds": args.rounds,
            "morton_bits": 30, "parallelism": "single GPU",
        },
        "lbvh_build_mtris_s": round(n / (lbvh_ms * 1e-3) / 1e6, 2),
        "lbvh_build_ms": round(lbvh_ms, 4),
        "treelet_ms_per_round": round(stages["treelet"] / max(args.rounds, 1), 4),
        "trace_closest_mrays_s": round(r / (stages["trace_closest"] * 1e-3) / 1e6, 2),
        "trace_shadow_mrays_s": round(r / (stages["trace_shadow"] * 1e-3) / 1e6, 2),
        "stage_ms": stages,
        "kernel_ms": {k: {"per_launch": round(per_launch[k], 4), "launches": launches[k]}
                      for k in per_launch},
        "sah_root": {"lbvh": float(recs[-1][0]["root_cost_lbvh"]),
                     "final": float(recs[-1][0]["root_cost_final"])},
        "hit_fraction": round(float((hit_prim >= 0).float().mean().item()), 4),
        "visible_fraction": round(float(vis.float().mean().item()), 4),
        "roofline": roofline,
        "lbvh_roofline": lbvh_roofline,
        "traversal": trav,
        "trace_launch": launch,
        "cpu_baseline": None,
    }
    bvh.close()
    ctx.close()
    del tris, rays, srays, lim, hit_prim, hit_t, vis
    if not args.no_render:
        log("[bench] path tracing (Cornell box 512x512x16)")
        out["path_tracing"] = render_bench(args.renders, not args.no_cpu, not args.no_tune)
    if not args.no_config2:
        w, h = (int(x) for x in args.c2_res.lower().split("x"))
        sizes = [int(x) for x in args.c2_tris.split(",") if x]
        log(f"[bench] path tracing at scale ({w}x{h}, {args.c2_spp} spp, {sizes} tris)")
        out["path_tracing_at_scale"] = config2_bench(args.seed, sizes, w, h, args.c2_spp, not args.no_cpu)
    if not args.no_equal_time:
        w, h = (int(x) for x in args.et_res.lower().split("x"))
        log(f"[bench] rlpath vs path at equal time ({w}x{h}, {args.et_tris} tris)")
        out["equal_time"] = equal_time_bench(args.seed, args.et_tris, w, h, args.et_ref_spp)
    if not args.no_serve:
        log("[bench] serving")
        rates = [float(x) for x in args.serve_rates.split(",") if x]
        out["serving"] = serving_bench(args.serve_requests, rates, args.seed, not args.serve_no_graphs,
                                       not args.serve_rl_atomic, args.serve_pool_tasks, args.serve_steps_ahead)
    if not args.no_cpu:
        log(f"[bench] CPU oracle baseline at N={args.cpu_n}, rays={args.cpu_rays}")
        out["cpu_baseline"] = cpu_baseline(args.cpu_n, args.cpu_rays, args.rounds, args.seed)
        log("[bench] CPU oracle baseline on one thread at N=2^18, rays=2^18")
        out["cpu_baseline"]["one_thread"] = cpu_one_thread(1 << 18, 1 << 18, args.rounds, args.seed)
    print(json.dumps(out), flush=True)
    if trav["overflow_closest"] or trav["overflow_occluded"]:
        print("bench: traversal stack overflow in the timed steps (results partial)", file=sys.stderr)
        return 3
    return 0


if __name__ == "__main__":
    sys.exit(main())
