"""Where rlpath's small bias comes from: a one-bounce model of the
reference's estimator (Integrators/ReinforcementLearningPathTracing.cu).

computeQDistributions (:540-600) weights each of the 128 direction cells by
|cos| * f * Q at ONE direction per cell, the sample's random in-cell offset,
and samples the next direction at that same offset, so a Q-sampled direction
is priced at its true density, 128 p_c / (4 pi). The light sample's MIS
weight (:576-581) also takes p_c at the sample's offset, for the cell of the
light direction, although the Q strategy would have reached that direction
only from its own offset. Where the cell weight grows with the value being
estimated (cos * f), that weight is too large on average and the estimate
comes out high.

Model: a Lambertian point (normal +z, albedo 0.6) under radiance 1 from the
whole upper hemisphere, light samples uniform over that hemisphere, Q = 1 in
every cell; the exact answer is the albedo. The model is numpy, written
here from the reference's formulas (cells: 8 equal-height z bands x 16
azimuth sectors, the bin definition the oracle's tests pin)."""
import numpy as np

RHO, K, PL = 0.6, 128 / (4 * np.pi), 1 / (2 * np.pi)
CELLS = np.arange(128)


def cell_weights(rx):
    z = -1 + 2 * ((CELLS // 16)[None, :] + rx[:, None]) / 8  # each cell's direction at offset rx
    w = np.maximum(z, 0) * RHO / np.pi  # cos * f * Q, zero below the surface
    return w / w.sum(1, keepdims=True)


def estimate(rng, n, own_offset):
    rx = rng.uniform(size=n)
    p = cell_weights(rx)
    # Q-sampled direction: cell from p, direction at offset rx in it
    c = np.minimum((np.cumsum(p, 1) < rng.uniform(size=n)[:, None]).sum(1), 127)
    zq = -1 + 2 * (c // 16 + rx) / 8
    q = p[np.arange(n), c] * K
    contrib_q = np.where(zq > 0, RHO / np.pi * zq / q * q * q / (q * q + PL * PL), 0.0)
    # light sample, uniform over the upper hemisphere (azimuth irrelevant: the
    # weights depend on the band only)
    zl = rng.uniform(size=n)
    band = np.minimum(((zl + 1) / 2 * 8).astype(int), 7)
    pl_cells = cell_weights((zl + 1) / 2 * 8 - band) if own_offset else p
    ql = pl_cells[np.arange(n), band * 16] * K
    contrib_l = RHO / np.pi * zl / PL * PL * PL / (PL * PL + ql * ql)
    return contrib_q + contrib_l


def test_light_weight_at_the_samples_offset_is_biased_high():
    rng = np.random.default_rng(1)
    res = {}
    for own in (False, True):
        est = np.concatenate([estimate(rng, 200_000, own) for _ in range(10)])
        res[own] = (est.mean() / RHO - 1, est.std() / np.sqrt(len(est)) / RHO)
    ref, ref_se = res[False]
    fixed, fixed_se = res[True]
    assert ref > 4 * ref_se and 0.0005 < ref < 0.0015, res    # the reference's weight: ~ +0.09%
    assert abs(fixed) < 3 * fixed_se, res                     # priced at its own offset: unbiased
