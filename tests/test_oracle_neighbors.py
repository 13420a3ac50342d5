"""Oracle neighbour grid and lists vs brute force (runner.hpp:1827-1911).

Neighbour sets are checked against a KD-tree candidate search followed by the
reference's exact FP32 predicate |x_i - x_j|^2 < h^2 (strict, p_j != p_i);
the periodic case uses the minimum image.  Cell counts must be the histogram
of floor(x/h) - offset (z-fastest numbering) and offsets its exclusive scan.
The list order is canonical: 27 cells in the reference order (dx outer,
dz inner), ascending particle id inside a cell.
"""
import numpy as np
from scipy.spatial import cKDTree

import oracle_ffi as of
from alluvion_amd import scenes


def f32_pred_pairs(x, h, wrap=None):
    """Neighbour sets with the FP32 predicate; wrap = box length or None."""
    x = x.astype(np.float32)
    h32 = np.float32(h)
    h2 = np.float32(h32 * h32)
    if wrap is None:
        tree = cKDTree(x.astype(np.float64))
    else:
        tree = cKDTree((x.astype(np.float64) + wrap / 2) % wrap, boxsize=wrap)
    pairs = tree.query_pairs(h * 1.0001, output_type="ndarray")
    i, j = pairs[:, 0], pairs[:, 1]
    d = x[i] - x[j]
    if wrap is not None:
        L = np.float32(wrap)
        hi, lo = np.float32(L * 0.5), np.float32(L * -0.5)
        d = np.where(d >= hi, d - L, np.where(d < lo, d + L, d)).astype(np.float32)
    r2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
    keep = r2 < h2
    i, j = i[keep], j[keep]
    n = x.shape[0]
    sets = [[] for _ in range(n)]
    for a, b in zip(i.tolist(), j.tolist()):
        sets[a].append(b)
        sets[b].append(a)
    return [sorted(s) for s in sets]


def cell_key(x, scene):
    h = np.float32(scene.kernel_radius)
    ip = np.floor(x.astype(np.float32) / h).astype(np.int64) - np.array(scene.grid_offset)
    res = np.array(scene.grid_res)
    inside = np.all((ip >= 0) & (ip < res), axis=1)
    key = (ip[:, 0] * res[1] + ip[:, 1]) * res[2] + ip[:, 2]
    return np.where(inside, key, -1), ip


def check_scene(scene):
    orc = of.Oracle(scene, precision=4)
    orc.update_particle_neighbors()
    ids, cnt = orc.neighbors()
    want = f32_pred_pairs(scene.x, scene.kernel_radius,
                          wrap=scene.wrap_length[0] if scene.wrap_length[0] > 0 else None)
    key, ip = cell_key(scene.x, scene)
    res = np.array(scene.grid_res)
    for i in range(scene.n):
        got = ids[i, : cnt[i]]
        assert sorted(got.tolist()) == want[i], i
        # canonical order: non-decreasing cell rank, ascending ids in a cell
        rank = []
        for j in got:
            dcell = ip[j] - ip[i]
            if scene.wrap_length[0] > 0:
                dcell = (dcell + 1) % res - 1
            rank.append((dcell[0] + 1) * 9 + (dcell[1] + 1) * 3 + (dcell[2] + 1))
        order = list(zip(rank, got.tolist()))
        assert order == sorted(order), i
    C = int(np.prod(scene.grid_res))
    counts, start, anomaly = orc.grid()
    hist = np.bincount(key[key >= 0], minlength=C)
    np.testing.assert_array_equal(counts, hist)
    np.testing.assert_array_equal(start[:-1], np.concatenate([[0], np.cumsum(hist)[:-1]]))
    assert start[-1] == hist.sum()
    return orc, cnt, anomaly


def test_dam_break_lists_bit_exact():
    sc = scenes.dam_break_small()
    _, cnt, anomaly = check_scene(sc)
    assert anomaly.tolist() == [0, 0, 0]
    assert 20 < cnt.mean() < 40


def test_periodic_random_lists_bit_exact():
    # same density as the 8M-particle periodic stress scene (8.39 per h^3 cell)
    n = int(round(8_388_608 * 0.2 ** 3))
    sc = scenes.periodic_random(n)
    sc.x = (sc.x * np.float32(0.2)).astype(np.float32)
    sc.wrap_length = (0.2, 0.2, 0.2)
    sc.grid_offset = (-10, -10, -10)
    sc.grid_res = (20, 20, 20)
    _, cnt, anomaly = check_scene(sc)
    assert anomaly.tolist() == [0, 0, 0]
    assert 30 < cnt.mean() < 40  # expected 35.1


def test_anomaly_flags_and_caps():
    sc = scenes.dam_break_small()
    x = sc.x.copy()
    x[0] = [5.0, 5.0, 5.0]  # outside the grid
    sc.x = x
    sc.max_num_particles_per_cell = 4  # lattice cells hold 8
    sc.max_num_neighbors = 16  # interior particles have ~33
    orc = of.Oracle(sc, precision=4)
    orc.update_particle_neighbors()
    counts, start, anomaly = orc.grid()
    assert anomaly.tolist() == [1, 1, 1]
    ids, cnt = orc.neighbors()
    assert cnt.max() == 16
    # the out-of-grid particle is in no cell and in nobody's list
    assert counts.sum() == sc.n - 1
    assert not np.any(ids[1:][np.arange(16)[None, :] < cnt[1:, None]] == 0)
