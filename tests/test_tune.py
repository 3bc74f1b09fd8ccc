"""The traversal autotuner's search and cache, with a synthetic timer (no
GPU): coordinate descent finds the fastest shape of a separable cost, every
shape it tries is a valid one, and the cache round-trips."""
import json

from cavalry_amd import tune


def test_descent_finds_separable_minimum(monkeypatch):
    tried, cur = [], {}

    def setter(ctx, which, shape):
        tried.append(dict(shape))
        cur.update(shape)

    monkeypatch.setattr(tune, "set_launch", setter)
    cost = lambda s: ((s["workgroups_per_cu"] - 3) ** 2 + abs(s["ray_chunk"] - 512) / 128
                      + (0 if s["lds_stack"] == 48 else 0.5) + 10)
    best, timings = tune._descend(None, tune.STANDALONE, lambda: cost(cur),
                                  {"workgroups_per_cu": 5, "ray_chunk": 256, "lds_stack": 32}, reps=2)
    assert best == {"workgroups_per_cu": 3, "ray_chunk": 512, "lds_stack": 48, "workgroup_size": 256,
                    "shade_workgroup_size": 256, "handoff_steps": 1024, "any_hit_step": 0,
                    "any_hit_whole_line": 0}
    assert tried[-1] == best  # the winner is left set
    for s in tried:
        assert s["workgroups_per_cu"] in tune.PER_CU and s["ray_chunk"] in tune.CHUNKS and s["lds_stack"] in tune.LDS
        assert s["workgroup_size"] in tune.SIZES and s["shade_workgroup_size"] in tune.SIZES
        assert s["handoff_steps"] in tune.HANDOFFS
        assert s["any_hit_step"] in (0,) + tune.ANY_STEPS
    assert min(timings.values()) == cost(best)


def test_descent_over_workgroup_sizes(monkeypatch):
    """The workgroup size is tried at the same waves per CU, then the
    workgroups per CU at the chosen size, and (render loops) the shading
    workgroup size."""
    tried, cur = [], {}

    def setter(ctx, which, shape):
        tried.append(dict(shape))
        cur.update(shape)

    monkeypatch.setattr(tune, "set_launch", setter)

    def cost(s):
        waves = s["workgroups_per_cu"] * s["workgroup_size"] // 64
        return ((0 if s["workgroup_size"] == 128 else 1) + abs(waves - 16) / 4
                + (0 if s["shade_workgroup_size"] == 64 else 0.25) + 5)

    start = {"workgroups_per_cu": 3, "ray_chunk": 256, "lds_stack": 32, "workgroup_size": 256,
             "shade_workgroup_size": 256}
    best, timings = tune._descend(None, tune.WAVEFRONT, lambda: cost(cur), start, reps=1, shade=True)
    assert best["workgroup_size"] == 128 and best["workgroups_per_cu"] == 8
    assert best["shade_workgroup_size"] == 64
    # the size step keeps the start's 12 waves per CU
    sized = [s for s in tried if s["ray_chunk"] == 256 and s["lds_stack"] == 32 and s["shade_workgroup_size"] == 256]
    assert {(s["workgroup_size"], s["workgroups_per_cu"]) for s in sized[:4]} >= {(64, 12), (128, 6), (256, 3)}
    assert tried[-1] == best
    assert min(timings.values()) == cost(best)
    # without shade, the shading size is never changed
    tried.clear()
    tune._descend(None, tune.STANDALONE, lambda: cost(cur), start, reps=1)
    assert all(s["shade_workgroup_size"] == 256 for s in tried)


def test_cache_round_trip(tmp_path):
    p = str(tmp_path / "t.json")
    tune._store(p, "k1", {"shape": {"workgroups_per_cu": 2, "ray_chunk": 128, "lds_stack": 48}, "ms": 1.5})
    tune._store(p, "k2", {"shape": {"workgroups_per_cu": 4, "ray_chunk": 256, "lds_stack": 32}, "ms": 2.5})
    data = json.load(open(p))
    assert set(data["entries"]) == {"k1", "k2"}
    assert tune._load(p)["entries"]["k1"]["shape"]["lds_stack"] == 48
    assert tune._load(str(tmp_path / "missing.json")) == {}
