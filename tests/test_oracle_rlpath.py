"""Pins for the CPU oracle's `rlpath` integrator (no GPU).

* direction bins (QEntry::dirToCellIndex / sampleDirectionInCell): every
  direction drawn inside a bin maps back to that bin; the bins split the
  sphere into 128 equal areas (8 bands of equal height in z times 16 equal
  azimuth sectors, the z-band then azimuth order of the reference); axis
  directions land in the bins that definition gives;
* Q update (finishUpdateQ with alpha = n / (T + n)): after any sequence of
  batches a bin's Q equals the mean of every proposal it has received, and
  bins never proposed to take the mean Q of the visited bins of their entry
  (averageQ); with a scaled rate one update is the convex combination;
* light-selection training (finishLightingTraining): per-light results are
  the running mean of the luminance samples, and the selection CDF is
  proportional to them;
* end to end: the furnace (radiance Le / (1 - rho) everywhere) and agreement
  with the `path` integrator on the Cornell box within the noise.
"""
import numpy as np
import pytest

import oracle
from cavalry_amd import scenes
from cavalry_amd.camera import look_at

from test_oracle_render import furnace_scene


def test_bin_round_trip():
    offs = np.linspace(0.03, 0.97, 9)
    for c in range(128):
        for rx in offs:
            for ry in offs:
                d = oracle.rl_cell_dir(float(rx), float(ry), c)
                assert abs(float(np.linalg.norm(d.astype(np.float64))) - 1.0) < 1e-6
                assert oracle.rl_dir_to_cell(d) == c, (c, rx, ry, d)


def test_bins_are_equal_area_bands_and_sectors():
    rng = np.random.default_rng(5)
    n = 128 * 800
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    v = v.astype(np.float32)
    cells = np.array([oracle.rl_dir_to_cell(d) for d in v])
    # independent definition: band by z in 8 equal steps, sector by azimuth
    band = np.clip(((v[:, 2].astype(np.float64) + 1) / 2 * 8).astype(int), 0, 7)
    az = np.mod(np.arctan2(v[:, 1], v[:, 0]).astype(np.float64), 2 * np.pi)
    sector = np.clip((az / (2 * np.pi) * 16).astype(int), 0, 15)
    expect = band * 16 + sector
    # boundaries aside (float rounding of the two formulas), they agree
    assert (cells == expect).mean() > 0.999
    counts = np.bincount(cells, minlength=128)
    sigma = np.sqrt(n / 128)
    assert np.abs(counts - n / 128).max() < 5 * sigma, counts


def test_axis_directions():
    # z = 0 -> band 4; azimuth 0, pi/2, pi, 3pi/2 -> sectors 0, 4, 8, 12
    assert oracle.rl_dir_to_cell(np.array([1, 0, 0])) == 64
    assert oracle.rl_dir_to_cell(np.array([0, 1, 0])) == 68
    assert oracle.rl_dir_to_cell(np.array([-1, 0, 0])) == 72
    assert oracle.rl_dir_to_cell(np.array([0, -1, 0])) == 76
    # the poles divide 0 by 0; the NaN azimuth converts to sector 0
    assert oracle.rl_dir_to_cell(np.array([0, 0, 1])) == 112
    assert oracle.rl_dir_to_cell(np.array([0, 0, -1])) == 0


def _batches(rng, E, nb):
    for _ in range(nb):
        newQ = np.zeros((E, 128), np.float32)
        cnt = np.zeros((E, 128), np.int32)
        ent = rng.integers(0, E - 1, size=60)  # the last entry is never visited
        cel = rng.integers(0, 16, size=60)     # bins >= 16 are never visited
        val = rng.uniform(0, 3, size=60).astype(np.float32)
        for e, c, x in zip(ent, cel, val):
            newQ[e, c] += x
            cnt[e, c] += 1
        yield newQ, cnt, ent, cel, val


def test_q_update_is_running_mean_and_average_fills_unvisited():
    rng = np.random.default_rng(11)
    E = 4
    Q = np.ones((E, 128), np.float32)
    total = np.zeros((E, 128), np.int32)
    sums = np.zeros((E, 128))
    counts = np.zeros((E, 128), np.int64)
    for newQ, cnt, ent, cel, val in _batches(rng, E, 7):
        for e, c, x in zip(ent, cel, val):
            sums[e, c] += float(x)
            counts[e, c] += 1
        oracle.rl_finish_q(Q, newQ, cnt, total)
        assert not newQ.any() and not cnt.any()
        np.testing.assert_array_equal(total, counts)
        seen = counts > 0
        np.testing.assert_allclose(Q[seen], sums[seen] / counts[seen], rtol=2e-5)
        for e in range(E):
            if seen[e].any():
                np.testing.assert_allclose(Q[e, ~seen[e]], Q[e, seen[e]].astype(np.float64).mean(), rtol=1e-5)
            else:
                assert (Q[e] == 1.0).all()


def test_q_update_scaled_rate_is_convex_step():
    Q = np.full((1, 128), 2.0, np.float32)
    total = np.zeros((1, 128), np.int32)
    newQ = np.zeros((1, 128), np.float32)
    cnt = np.zeros((1, 128), np.int32)
    newQ[0, 3] = 9.0
    cnt[0, 3] = 3  # mean 3
    oracle.rl_finish_q(Q, newQ, cnt, total, alpha_scale=0.3)
    assert Q[0, 3] == pytest.approx(2.0 * 0.7 + 0.3 * 3.0, rel=1e-6)
    assert (Q[0, np.arange(128) != 3] == Q[0, 3]).all()  # averaged from the one visited bin


def test_light_training_running_mean_and_cdf():
    rng = np.random.default_rng(2)
    E, L = 3, 5
    ldist = np.tile(((np.arange(L) + 1) / L).astype(np.float32), (E, 1))
    lres = np.zeros((E, L), np.float32)
    lcount = np.zeros((E, L), np.int32)
    sums = np.zeros((E, L))
    cnts = np.zeros((E, L), np.int64)
    for _ in range(6):
        lnew = np.zeros((E, L), np.float32)
        lnewc = np.zeros((E, L), np.int32)
        for _ in range(80):
            e = int(rng.integers(0, E - 1))  # the last entry stays untrained
            k = int(rng.integers(0, L))
            x = float(rng.uniform(0.5, 4.0)) * (k + 1)
            lnew[e, k] += np.float32(x)
            lnewc[e, k] += 1
            sums[e, k] += np.float32(x)
            cnts[e, k] += 1
        oracle.rl_finish_light(ldist, lres, lcount, lnew, lnewc)
        assert not lnew.any() and not lnewc.any()
    np.testing.assert_array_equal(lcount, cnts)
    for e in range(E - 1):
        mean = sums[e] / cnts[e]
        np.testing.assert_allclose(lres[e], mean, rtol=2e-5)
        pdf = np.diff(np.concatenate([[0.0], ldist[e].astype(np.float64)]))
        np.testing.assert_allclose(pdf, mean / mean.sum(), rtol=1e-4, atol=1e-6)
        assert ldist[e, -1] == pytest.approx(1.0, abs=1e-6)
    np.testing.assert_array_equal(ldist[E - 1], ((np.arange(L) + 1) / L).astype(np.float32))


def test_rlpath_furnace():
    for rho in (0.3, 0.6):
        sc = furnace_scene(rho=rho)
        cam = look_at(sc.eye, sc.center, sc.up, sc.fov, sc.width, sc.height)
        rgb, w, _ = oracle.render_rlpath(sc, 32, cam)
        img = rgb / w[..., None]
        expect = 1.0 / (1.0 - rho)
        assert abs(float(img.mean()) - expect) < 0.03 * expect, (rho, float(img.mean()), expect)


def test_rlpath_agrees_with_path_on_cornell_box():
    sc = scenes.cornell_box(seed=1, width=32, height=32)
    cam = look_at(sc.eye, sc.center, sc.up, sc.fov, sc.width, sc.height)
    ref_rgb, ref_w, _ = oracle.render_path(sc, 1024, cam)
    ref = ref_rgb / ref_w[..., None]
    rl_rgb, rl_w, _ = oracle.render_rlpath(sc, 64, cam)
    rl = rl_rgb / rl_w[..., None]
    p_rgb, p_w, _ = oracle.render_path(sc, 64, cam)
    pt = p_rgb / p_w[..., None]

    def rel_mse(a):
        return float(np.mean((a - ref) ** 2 / (ref ** 2 + 1e-2)))

    assert abs(float(rl.mean()) - float(ref.mean())) < 0.02 * float(ref.mean())
    # guided sampling is at least in the same noise class as plain path tracing
    assert rel_mse(rl) < 2.0 * rel_mse(pt) + 1e-4, (rel_mse(rl), rel_mse(pt))


def test_cell_sampling_density_is_128_over_4pi():
    """computeQDistributions (RL :580-600) prices a direction drawn by
    sampleDirectionInCell in cell c at 128 p_c / (4 pi): the draw must be
    uniform over the cell's patch of the sphere. The patch is the band
    z0 <= z < z1 times the sector phi0 <= phi < phi1, whose area element is
    dz dphi, so uniform means z and phi uniform on their ranges: density
    1 / (dz dphi) = 1 / ((2/8) (2 pi/16)) = 128 / (4 pi)."""
    rng = np.random.default_rng(11)
    n = 4000
    for c in (0, 37, 64, 90, 127):
        band, sector = divmod(c, 16)
        z0, z1 = -1 + 2 * band / 8, -1 + 2 * (band + 1) / 8
        p0, p1 = 2 * np.pi * sector / 16, 2 * np.pi * (sector + 1) / 16
        assert abs(1.0 / ((z1 - z0) * (p1 - p0)) - 128 / (4 * np.pi)) < 1e-12
        u = rng.uniform(size=(n, 2))
        d = np.array([oracle.rl_cell_dir(float(a), float(b), c) for a, b in u], np.float64)
        z = d[:, 2]
        phi = np.mod(np.arctan2(d[:, 1], d[:, 0]), 2 * np.pi)
        if sector == 0:
            phi = np.where(phi > np.pi, phi - 2 * np.pi, phi)  # the sector next to 0 rad
        for x, a, b in ((z, z0, z1), (phi, p0, p1)):
            assert ((x >= a - 1e-6) & (x <= b + 1e-6)).all()
            # Kolmogorov-Smirnov distance to the uniform CDF (1% critical
            # value 1.63 / sqrt(n))
            s = np.sort((x - a) / (b - a))
            ks = np.max(np.maximum(np.arange(1, n + 1) / n - s, s - np.arange(n) / n))
            assert ks < 1.63 / np.sqrt(n), (c, ks)


def test_furnace_rlpath_excess_over_path():
    """In the furnace rlpath comes out 0.2-0.3% above path at every depth
    (1024 spp: depth 3 +0.24%, 4 +0.26%, 8 +0.29% at rho = 0.6), far
    outside the noise (0.015% per image). That is the reference's estimator:
    computeQDistributions (RL :540-581) evaluates each cell's weight at one
    direction per cell, the sample's random in-cell offset, and the light
    sample's MIS weight takes the Q pdf of the light direction's cell at that
    offset (:580) rather than at the light direction's own offset. In the
    furnace the cell weight grows with cos, as does the light sample's value,
    so the weight the estimator uses is too large on average. path (BSDF pdf
    in the weight) has no such term and stays below the geometric sum only
    by its truncation at max_depth."""
    from test_oracle_render import furnace_scene
    rho = 0.6
    sc = furnace_scene(rho=rho)
    cam = look_at(sc.eye, sc.center, sc.up, sc.fov, sc.width, sc.height)
    means = {}
    for name, fn in (("path", oracle.render_path), ("rlpath", oracle.render_rlpath)):
        rgb, w, _ = fn(sc, 256, cam, max_depth=8)
        means[name] = float((rgb / w[..., None]).mean())
    geometric = sum(rho ** k for k in range(9))
    assert -0.005 < means["path"] / geometric - 1 < 0.0  # truncation only
    excess = means["rlpath"] / means["path"] - 1
    assert 0.0015 < excess < 0.004, means
