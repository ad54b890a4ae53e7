"""The settled-bed builder (sequential deposition under gravity,
dem_deposit_packing in csrc/scene_rsa.c), which makes the state the 1M
hopper's step is timed on.  Checked on small hoppers against what the
construction fixes: no pair overlaps by more than the preload (plus the
descent's tolerance), no sphere crosses a wall, every sphere rests on
something, almost every sphere off the floor touches three or more
neighbours or walls, and the parallel batched deposition gives exactly the
one-at-a-time deposition, for any batch size and thread count."""
import ctypes

import numpy as np
import pytest

from dem_mbd_cuda_amd import scene

PRELOAD = 1e-4


def _bed(n, batch=32, seed=7):
    width = 0.2 * (n / 1e6) ** 0.5
    rng = np.random.RandomState(2)
    radius = np.clip(1e-3 * np.exp(0.2 * rng.standard_normal(n)), 0.5e-3, 2e-3)
    wo, wn = scene.hopper_walls(0.5 * width, 10.0)
    pos = scene.deposit_packing(radius, (-0.5 * width, -0.5 * width), (0.5 * width, 0.5 * width), wo, wn,
                                seed=seed, preload=PRELOAD, batch=batch)
    return radius, pos, wo, wn


def _pairs(radius, pos):
    """All pairs closer than 2.1 x the largest radius (brute force in slabs)."""
    n = len(radius)
    reach = 2.1 * radius.max()
    order = np.argsort(pos[:, 2])
    zs = pos[order, 2]
    out = []
    for a in range(n):
        i = order[a]
        b = np.searchsorted(zs, zs[a] + reach, side="right")
        js = order[a + 1:b]
        if len(js) == 0:
            continue
        d = np.linalg.norm(pos[js] - pos[i], axis=1)
        keep = d < reach
        out.extend((i, j, dd) for j, dd in zip(js[keep], d[keep]))
    return np.array([(i, j) for i, j, _ in out], int), np.array([d for _, _, d in out])


def test_no_overlap_beyond_preload_and_no_wall_crossing():
    radius, pos, wo, wn = _bed(3000)
    ij, d = _pairs(radius, pos)
    rs = radius[ij[:, 0]] + radius[ij[:, 1]]
    rel = (rs - d) / rs
    # resting contacts overlap by the preload; the descent accepts positions
    # penetrating by at most 1e-4 r, which the final push-out usually clears
    assert rel.max() <= PRELOAD + 0.5e-4 + 1e-9
    assert np.mean(rel > 0) > 0.0  # there are contacts
    wall_gap = np.stack([(pos - o) @ nrm for o, nrm in zip(wo, wn)], 1) - radius[:, None]
    assert wall_gap.min() >= -(PRELOAD + 1e-4) * radius.max()


def test_every_sphere_rests_on_something():
    radius, pos, wo, wn = _bed(3000)
    ij, d = _pairs(radius, pos)
    rs = radius[ij[:, 0]] + radius[ij[:, 1]]
    touching = d <= rs * (1 - PRELOAD) + 4e-4 * rs
    count = np.zeros(len(radius), int)
    np.add.at(count, ij[touching, 0], 1)
    np.add.at(count, ij[touching, 1], 1)
    wall_gap = np.stack([(pos - o) @ nrm for o, nrm in zip(wo, wn)], 1) - radius[:, None] * (1 - PRELOAD)
    walls = (wall_gap <= 4e-4 * radius[:, None]).sum(axis=1)
    on_floor = wall_gap[:, 0] <= 4e-4 * radius
    assert on_floor.sum() > 10
    total = count + walls
    assert total.min() >= 1
    assert np.mean(total[~on_floor] >= 3) > 0.99
    # stored contacts per sphere (both sides of each pair) near the ~5.5 of
    # a deposited bed
    cbar = 2 * np.sum(d < rs) / len(radius)
    assert 4.5 < cbar < 6.5, cbar


@pytest.mark.parametrize("batch", [7, 32, 256])
def test_batched_deposition_is_the_one_at_a_time_deposition(batch):
    ref = _bed(2500, batch=1)[1]
    assert np.array_equal(_bed(2500, batch=batch)[1], ref)


def test_thread_count_does_not_change_the_bed():
    try:
        gomp = ctypes.CDLL("libgomp.so.1")
    except OSError:
        pytest.skip("OpenMP runtime not loadable")
    ref = _bed(2500)[1]
    gomp.omp_set_num_threads(1)
    try:
        one = _bed(2500)[1]
    finally:
        gomp.omp_set_num_threads(8)
    assert np.array_equal(one, ref)
