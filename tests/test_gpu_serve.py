"""Render server (continuous batching) on the GPU.

Mix invariance: every path / directlighting film the server returns equals,
bit for bit, the same request rendered alone by cav_render, although the
requests are admitted in chunks and share every launch with other requests
at other depths. rlpath requests share the pools too, one whole round (the
server's round size) at a time, each with its own tables: their films match
cav_render with the same spp_per_round within rel-MSE 1e-4 (Q proposals
land through float atomics), and bit for bit in deterministic mode. Also: requests submitted while others run start before those
finish, cancellation, hipGraph replay, and argument errors."""
import time

import numpy as np
import pytest

from cavalry_amd import render as R, scenes
from cavalry_amd.camera import look_at
from cavalry_amd.serve import CANCELLED, DONE, RUNNING, Server

pytestmark = pytest.mark.gpu


def rel_mse(a, b):
    return float(np.mean((a - b) ** 2 / (b * b + 1e-2)))


def cam_for(sc, w, h, shift=0.0):
    return look_at(sc.eye + np.float32([shift, 0, 0]), sc.center, sc.up, sc.fov, w, h)


@pytest.fixture(scope="module")
def two_scenes(ctx):
    a = scenes.cornell_box(seed=11, width=64, height=64, materials="mixed")
    b = scenes.cornell_box(seed=12, width=64, height=64)
    sa, sb = R.Scene(ctx, a), R.Scene(ctx, b)
    yield (a, sa), (b, sb)
    sa.close()
    sb.close()


REQUESTS = [  # scene, width, height, spp, integrator, depth, camera shift
    (0, 64, 48, 4, "path", 8, 0.0),
    (1, 40, 40, 9, "path", 5, 0.05),
    (0, 96, 64, 2, "directlighting", 1, -0.05),
    (1, 64, 64, 3, "rlpath", 6, 0.0),
    (0, 40, 32, 2, "rlpath_det", 5, 0.03),
    (0, 33, 17, 16, "path", 12, 0.1),
    (1, 80, 60, 1, "directlighting", 1, 0.0),
    (0, 48, 48, 6, "path", 3, -0.1),
]


@pytest.mark.parametrize("graphs", [False, True])
def test_served_films_equal_alone(ctx, two_scenes, graphs):
    print(f"[serve] mix invariance, graphs={graphs}", flush=True)
    # a small pool and quantum, so requests are chunked and mixed
    srv = Server(ctx, arena_bytes=1 << 28, host_arena_bytes=1 << 24, pool_tasks=20000, chunk_samples=3000,
                 use_graphs=graphs)
    tickets = []
    for s, w, h, spp, integ, depth, shift in REQUESTS:
        arr, sc = two_scenes[s]
        det = integ == "rlpath_det"
        tickets.append(srv.submit(sc, spp, "rlpath" if det else integ, camera=cam_for(arr, w, h, shift),
                                  max_depth=depth, deterministic=det))
    outs = [srv.poll(t, wait=True) for t in tickets]
    stats = srv.close()
    assert stats["pools"] == 2 and stats["steps"] > 0
    if graphs:  # both pools replayed their bounce steps from graphs
        assert stats["graph_pools"] == 2
    for (s, w, h, spp, integ, depth, shift), t, out in zip(REQUESTS, tickets, outs):
        assert out["status"] == DONE, out
        arr, sc = two_scenes[s]
        det = integ == "rlpath_det"
        spr = srv.rl_round_spp(w * h, spp) if integ.startswith("rlpath") else 0
        alone, _ = R.render(ctx, sc, spp, integrator="rlpath" if det else integ, camera=cam_for(arr, w, h, shift),
                            max_depth=depth, deterministic=det, spp_per_round=spr)
        np.testing.assert_array_equal(t.film.weight, alone.weight)
        if integ == "rlpath":
            err = rel_mse(t.film.image, alone.image)
            print(f"[serve] rlpath served vs alone rel_mse={err:.3e}", flush=True)
            assert err <= 1e-4
        else:
            np.testing.assert_array_equal(t.film.rgb, alone.rgb)


def test_continuous_admission(ctx, two_scenes):
    print("[serve] continuous admission", flush=True)
    (arr, sc), _ = two_scenes
    srv = Server(ctx, arena_bytes=1 << 30, host_arena_bytes=1 << 26, pool_tasks=1 << 16, chunk_samples=1 << 14)
    big = srv.submit(sc, 64, "path", camera=cam_for(arr, 256, 256), max_depth=8)
    while srv.poll(big, wait=False)["status"] != RUNNING:
        time.sleep(0.001)
    late = [srv.submit(sc, 2, "path", camera=cam_for(arr, 32, 32, 0.02 * k), max_depth=4) for k in range(4)]
    outs = [srv.poll(t, wait=True) for t in late]
    ob = srv.poll(big, wait=True)
    srv.close()
    assert ob["status"] == DONE and all(o["status"] == DONE for o in outs)
    # the small requests were admitted while the big one was in flight
    assert all(o["start_ms"] < ob["done_ms"] for o in outs)
    assert min(o["done_ms"] for o in outs) < ob["done_ms"]


def test_cancel_and_errors(ctx, two_scenes):
    from cavalry_amd._lib import CavError
    (arr, sc), _ = two_scenes
    srv = Server(ctx, arena_bytes=1 << 28, host_arena_bytes=1 << 24, pool_tasks=1 << 14)
    blocker = srv.submit(sc, 32, "path", camera=cam_for(arr, 128, 128), max_depth=8)
    t = srv.submit(sc, 4, "path", camera=cam_for(arr, 64, 64), max_depth=8)
    srv.cancel(t)
    out = srv.poll(t, wait=True)
    assert out["status"] in (CANCELLED, DONE)
    if out["status"] == CANCELLED:
        assert not t.film.weight.any()
    assert srv.poll(blocker, wait=True)["status"] == DONE
    with pytest.raises(CavError):
        srv.poll(t, wait=False)  # the id was released with its final status
    with pytest.raises(CavError):
        srv.submit(sc, 0, "path", camera=cam_for(arr, 8, 8))
    with pytest.raises(CavError):  # larger than the arena
        srv.submit(sc, 64, "path", camera=cam_for(arr, 1024, 1024))
    srv.close()


def test_scene_destroy_refused_while_served(ctx):
    """A scene a server still holds (pending request or ray pool) cannot be
    destroyed; after the server closes it can."""
    from cavalry_amd._lib import CavError
    print("[serve] scene lifetime", flush=True)
    arr = scenes.cornell_box(seed=13, width=32, height=32)
    sc = R.Scene(ctx, arr)
    srv = Server(ctx, arena_bytes=1 << 26, host_arena_bytes=1 << 22, pool_tasks=1 << 14)
    t = srv.submit(sc, 2, "path", camera=cam_for(arr, 32, 32), max_depth=4)
    with pytest.raises(CavError) as e:
        sc.close()
    assert e.value.code == 6  # CAV_ERR_BUSY
    assert srv.poll(t, wait=True)["status"] == DONE
    with pytest.raises(CavError):  # the server keeps its ray pool for the scene
        sc.close()
    srv.close()
    sc.close()
    assert not sc.handle


@pytest.fixture(scope="module")
def interior_pair(ctx):
    """Two resident procedural interiors (the serving workload's scene
    family), smaller than the benchmark's and with 256^2 textures."""
    a = scenes.interior(100_000, seed=3, width=64, height=64, max_depth=16, tex_size=256)
    b = scenes.interior(250_000, seed=4, width=64, height=64, max_depth=16, tex_size=256)
    sa, sb = R.Scene(ctx, a, bvh_rounds=3), R.Scene(ctx, b, bvh_rounds=3)
    yield (a, sa), (b, sb)
    sa.close()
    sb.close()


INTERIOR_REQUESTS = [  # scene, width, height, spp, integrator, depth, camera shift
    (0, 96, 64, 4, "path", 16, 0.0),
    (1, 64, 48, 8, "path", 16, 0.3),
    (1, 80, 45, 2, "rlpath_det", 16, 0.0),
    (0, 120, 68, 1, "path", 16, -0.4),
    (0, 48, 48, 4, "rlpath_det", 12, 0.2),
    (1, 128, 72, 2, "path", 16, -0.2),
]


@pytest.mark.parametrize("graphs", [False, True])
def test_interior_served_films_equal_alone(ctx, interior_pair, graphs):
    """Serving on the procedural interiors: chunked admission into a small
    pool mixes requests of both scenes and several depths in every launch;
    every film equals the request rendered alone bit for bit."""
    print(f"[serve] interior mix invariance, graphs={graphs}", flush=True)
    srv = Server(ctx, arena_bytes=1 << 29, host_arena_bytes=1 << 24, pool_tasks=40000, chunk_samples=6000,
                 use_graphs=graphs)
    tickets = []
    for s, w, h, spp, integ, depth, shift in INTERIOR_REQUESTS:
        arr, sc = interior_pair[s]
        det = integ == "rlpath_det"
        tickets.append(srv.submit(sc, spp, "rlpath" if det else integ, camera=cam_for(arr, w, h, shift),
                                  max_depth=depth, deterministic=det))
    outs = [srv.poll(t, wait=True) for t in tickets]
    stats = srv.close()
    print(f"[serve] stats {stats}", flush=True)
    assert stats["pools"] == 2
    if graphs:
        assert stats["graph_pools"] == 2
    for (s, w, h, spp, integ, depth, shift), t, out in zip(INTERIOR_REQUESTS, tickets, outs):
        assert out["status"] == DONE, out
        arr, sc = interior_pair[s]
        det = integ == "rlpath_det"
        spr = srv.rl_round_spp(w * h, spp) if det else 0
        alone, _ = R.render(ctx, sc, spp, integrator="rlpath" if det else integ, camera=cam_for(arr, w, h, shift),
                            max_depth=depth, deterministic=det, spp_per_round=spr)
        np.testing.assert_array_equal(t.film.weight, alone.weight)
        np.testing.assert_array_equal(t.film.rgb, alone.rgb)
