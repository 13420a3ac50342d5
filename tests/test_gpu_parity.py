"""GPU (HIP, through the C ABI) vs the CPU oracle on the same seeded inputs.

Integer data (cell counts, offsets, anomaly flags, neighbour lists in
canonical order) must be bit-exact.  Floating-point state after one step from
an identical state must agree to 1e-5 relative to the field's scale
(max-norm), the per-step tolerance of the FP32 solver; iteration counts must
be identical.
"""
import numpy as np
import pytest

import oracle_ffi as of
from alluvion_amd import scenes

pytestmark = pytest.mark.gpu

REL = 1e-5


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    scale = max(np.abs(b).max(), 1e-30)
    return np.abs(a - b).max() / scale


def gpu_scene(scene, kind="df"):
    dp, pile, solver, cn, cni = scenes.build_solver(scene, kind)
    return dp, pile, solver


def oracle_like(scene, dp, pile, kind="df", precision=4):
    orc = of.Oracle(scene, precision=precision, kind=kind)
    for b, var in enumerate(pile.volume_grids):
        orc.set_volume_map(b, dp.coat(var).get())
    return orc


def sync_state(dp, solver, orc, kind="df"):
    n = solver.num_particles
    orc.set("x", dp.coat(solver.particle_x).get()[:n])
    orc.set("v", dp.coat(solver.particle_v).get()[:n])
    if kind == "df":
        orc.set("kappa", dp.coat(solver.particle_kappa).get()[:n])
        orc.set("kappa_v", dp.coat(solver.particle_kappa_v).get()[:n])
    else:
        orc.set("pressure", dp.coat(solver.particle_pressure).get()[:n])


def test_volume_map_matches_oracle():
    sc = scenes.dam_break_small()
    sc.boundaries[0].resolution = (8, 8, 8)
    dp, pile, solver = gpu_scene(sc)
    gpu = dp.coat(pile.volume_grids[0]).get()
    orc = of.Oracle(sc, precision=4)
    orc.build_volume_map(0)
    ref = orc.volume_map(0)
    scale = 4 * np.pi / 3 * sc.kernel_radius ** 3
    assert gpu.shape == ref.shape
    err = np.abs(gpu.astype(np.float64) - ref).max() / scale
    print("volume map max error / ball volume: %.3g" % err)
    assert err <= 1e-5, err


def test_field_op_and_field_gen_kats():
    import alluvion_amd._alluvion as al  # noqa: F401
    import ctypes
    from alluvion_amd import LIB_PATH
    L = ctypes.CDLL(LIB_PATH)
    fp = ctypes.POINTER(ctypes.c_float)
    up = ctypes.POINTER(ctypes.c_uint)
    # field_op: node positions on the reference's 11^3 grid
    dmin = np.array([-23.1, 12.23, -31.44], np.float32)
    dmax = np.array([45.1, 103.2, 12.4], np.float32)
    res = np.array([11, 11, 11], np.uint32)
    cs = ((dmax - dmin) / res.astype(np.float32)).astype(np.float32)
    nn = of.num_nodes(res)
    out = np.zeros((nn, 3), np.float32)
    n_out = ctypes.c_uint(0)
    L.alv_node_positions.argtypes = [fp, up, fp, fp, up]
    L.alv_node_positions.restype = ctypes.c_int
    rc = L.alv_node_positions(dmin.ctypes.data_as(fp), res.ctypes.data_as(up),
                              cs.ctypes.data_as(fp), out.ctypes.data_as(fp), ctypes.byref(n_out))
    assert rc == 0 and n_out.value == nn
    want = np.array([of.node_position(dmin.astype(np.float64), res, cs.astype(np.float64), l)
                     for l in range(nn)])
    np.testing.assert_allclose(out, want, rtol=1e-5, atol=1e-4)
    # field_gen: ellipsoid volume field, r = 1.25, sign 1
    r = 1.25
    margin = 4 * r
    dmin = np.array([-30 - margin, -20 - margin, -40 - margin], np.float32)
    dmax = np.array([30 + margin, 20 + margin, 40 + margin], np.float32)
    cs = ((dmax - dmin) / res.astype(np.float32)).astype(np.float32)
    pos = np.array([of.node_position(dmin.astype(np.float64), res, cs.astype(np.float64), l)
                    for l in range(nn)])
    dist = (pos[:, 0] ** 2 / 625 + pos[:, 1] ** 2 / 225 + pos[:, 2] ** 2 / 1225 - 1.0).astype(np.float32)

    class AlvConst(ctypes.Structure):
        _fields_ = [(k, ctypes.c_float) for k in (
            "kernel_radius", "kernel_radius_sqr", "cubic_kernel_k", "cubic_kernel_l",
            "cubic_kernel_zero", "particle_radius", "particle_vol", "particle_mass", "density0",
            "viscosity", "boundary_viscosity")] + [("gravity", ctypes.c_float * 3)] + [
            (k, ctypes.c_float) for k in ("boundary_epsilon", "dfsph_factor_epsilon",
                                          "contact_tolerance")] + [
            ("wrap_length", ctypes.c_float * 3)]

    cn = AlvConst()
    assert ctypes.sizeof(AlvConst) == 20 * 4  # alv_const: 20 floats
    cp = ctypes.POINTER(AlvConst)
    L.alv_const_set_kernel_radius.argtypes = [cp, ctypes.c_float]
    L.alv_const_set_kernel_radius.restype = None
    L.alv_const_set_particle_attr.argtypes = [cp, ctypes.c_float, ctypes.c_float, ctypes.c_float]
    L.alv_const_set_particle_attr.restype = None
    L.alv_volume_from_distance_map.argtypes = [cp, fp, fp, up, fp, ctypes.c_float, fp, fp,
                                               ctypes.c_uint]
    L.alv_volume_from_distance_map.restype = ctypes.c_int
    L.alv_const_set_kernel_radius(ctypes.byref(cn), r)
    L.alv_const_set_particle_attr(ctypes.byref(cn), 0.25, 0.2, 1.0)
    vol = np.zeros(nn, np.float32)
    rc = L.alv_volume_from_distance_map(
        ctypes.byref(cn), dmin.ctypes.data_as(fp), dmax.ctypes.data_as(fp), res.ctypes.data_as(up),
        cs.ctypes.data_as(fp), 1.0, dist.ctypes.data_as(fp), vol.ctypes.data_as(fp), nn)
    assert rc == 0
    ref = of.volume_from_distance_map(4, r, dmin, dmax, res, cs, 1.0, dist.astype(np.float64))
    scale = 4 * np.pi / 3 * r ** 3
    # doctest::Approx default: relative 1.19e-5 of the magnitude (field_gen.cu:93-100)
    assert np.all(np.abs(vol - ref) <= 1.2e-5 * np.maximum(np.abs(ref), scale * 1e-2))


def neighbor_parity(scene):
    dp, pile, solver = gpu_scene(scene)
    solver.update_particle_neighbors()
    ids, cnt = solver.get_neighbors()
    gcount, gstart, ganom = solver.get_grid()
    orc = oracle_like(scene, dp, pile)
    # alv_solver_update_particle_neighbors computes the boundary records first
    # (Solver::compute_all_boundaries, then update_particle_neighbors)
    orc.compute_all_boundaries()
    orc.update_particle_neighbors()
    oids, ocnt = orc.neighbors()
    ocount, ostart, oanom = orc.grid()
    n = scene.n
    np.testing.assert_array_equal(gcount, ocount)
    np.testing.assert_array_equal(gstart, ostart)
    np.testing.assert_array_equal(ganom, oanom)
    np.testing.assert_array_equal(cnt[:n], ocnt[:n])
    K = ids.shape[1]
    mask = np.arange(K)[None, :] < ocnt[:n, None]
    np.testing.assert_array_equal(np.where(mask, ids[:n], 0), np.where(mask, oids[:n], 0))
    return dp, solver, orc


def test_neighbor_search_dam_break_bit_exact():
    sc = scenes.dam_break_small()
    dp, solver, orc = neighbor_parity(sc)
    dens = dp.coat(solver.particle_density).get()[: sc.n]
    assert rel_err(dens, orc.get("density")[: sc.n]) < REL


def test_neighbor_search_periodic_bit_exact():
    n = int(round(8_388_608 * 0.2 ** 3))
    sc = scenes.periodic_random(n)
    sc.x = (sc.x * np.float32(0.2)).astype(np.float32)
    sc.wrap_length = (0.2, 0.2, 0.2)
    sc.grid_offset = (-10, -10, -10)
    sc.grid_res = (20, 20, 20)
    neighbor_parity(sc)


def test_neighbor_search_far_from_origin_bit_exact():
    """A random cloud 300 m from the origin (30,000 cells of h): the rounding
    of x/h and of the cell faces is about 3e-5 m there, far above a fixed
    margin of 1e-4 h^2, so the search's skip of cells farther than h must
    scale its margin with the coordinates.  Lists, counts and offsets equal
    the oracle's (which scans all 27 cells), and every list equals brute force
    with the FP32 predicate ((dx dx + dy dy) + dz dz < h^2)."""
    from scipy.spatial import cKDTree

    rng = np.random.default_rng(11)
    origin = np.array([301.27, -187.9, 256.3])
    edge = 0.15
    n = 24_000
    x = (origin + rng.random((n, 3)) * edge).astype(np.float32)
    h = 0.01
    off = tuple(int(np.floor(o / h)) - 1 for o in origin)
    res = int(np.ceil(edge / h)) + 3
    sc = scenes.Scene(name="far_cloud", x=x, v=np.zeros_like(x), gravity=(0.0, 0.0, 0.0),
                      grid_offset=off, grid_res=(res, res, res), boundaries=[])
    dp, solver, orc = neighbor_parity(sc)
    ids, cnt = solver.get_neighbors()
    assert int(np.asarray(solver.get_grid()[2])[0]) == 0  # nothing outside the grid
    h32 = np.float32(h)
    tree = cKDTree(x.astype(np.float64))
    pairs = tree.query_ball_point(x.astype(np.float64), r=h * 1.001)
    for i in range(n):
        cand = np.array([j for j in pairs[i] if j != i], np.int64)
        d = x[i][None, :] - x[cand]
        d2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        want = np.sort(cand[d2 < h32 * h32])
        got = np.sort(ids[i, : cnt[i]].astype(np.int64))
        assert np.array_equal(got, want), i


@pytest.mark.parametrize("warm", [0, 40])
def test_dfsph_step_parity_dam_break(warm):
    sc = scenes.dam_break_small()
    dp, pile, solver = gpu_scene(sc)
    for _ in range(warm):
        solver.step()
    orc = oracle_like(sc, dp, pile)
    sync_state(dp, solver, orc)
    solver.step()
    orc.step()
    n = sc.n
    rep = orc.report()
    assert solver.num_divergence_solve == rep["num_divergence_solve"]
    assert solver.num_density_solve == rep["num_density_solve"]
    for name, gvar, oname in (("x", solver.particle_x, "x"), ("v", solver.particle_v, "v"),
                              ("density", solver.particle_density, "density"),
                              ("kappa", solver.particle_kappa, "kappa"),
                              ("kappa_v", solver.particle_kappa_v, "kappa_v")):
        g = dp.coat(gvar).get()[:n]
        o = orc.get(oname)[:n]
        if name == "x":
            # positions: relative to the displacement scale of one step
            err = np.abs(g - o).max() / max(np.abs(dp.coat(solver.particle_v).get()[:n]).max() * sc.dt, 1e-12)
            assert err < 1e-3, name
            assert np.abs(g - o).max() < 1e-6
        else:
            assert rel_err(g, o) < REL, (name, rel_err(g, o))
    assert abs(solver.mean_density_error - rep["mean_density_error"]) <= 1e-6


def test_dfsph_without_divergence_solve_carries_the_factor():
    """Without the divergence solve the reference does not compute the DFSPH
    factor: its density solve rescales the factor each particle ended the
    last step with (solver_df.hpp:81-91, runner.hpp:3525-3541).  The HIP
    path carries it across the re-sort, so after such a step the factor
    equals the oracle's particle by particle.  (The step's velocities are not
    compared: with the factor rescaled twice the reference's solve diverges.)
    After 40 steps the column falls fast enough that particles cross cell
    faces every step, so the sorted order changes under the factor."""
    sc = scenes.dam_break_small()
    dp, pile, solver = gpu_scene(sc)
    for _ in range(40):
        solver.step()
    orc = oracle_like(sc, dp, pile)
    sync_state(dp, solver, orc)
    solver.step()
    orc.step()
    solver.enable_divergence_solve = False
    orc.update_settings(enable_divergence_solve=0)
    solver.step()
    orc.step()
    n = sc.n
    g = dp.coat(solver.particle_dfsph_factor).get()[:n]
    o = orc.get("dfsph_factor")[:n]
    assert np.all(np.isfinite(g)) and np.abs(o).max() > 0
    assert rel_err(g, o) < REL, rel_err(g, o)


@pytest.mark.parametrize("kind", ["ii", "i"])
def test_iisph_and_isph_step_parity_dam_break(kind):
    sc = scenes.dam_break_small()
    dp, pile, solver = gpu_scene(sc, kind)
    for _ in range(20):
        solver.step()
    orc = oracle_like(sc, dp, pile, kind=kind)
    sync_state(dp, solver, orc, kind=kind)
    solver.step()
    orc.step()
    n = sc.n
    rep = orc.report()
    assert solver.num_density_solve == rep["num_density_solve"]
    for name, gvar in (("v", solver.particle_v), ("density", solver.particle_density),
                       ("pressure", solver.particle_pressure)):
        g = dp.coat(gvar).get()[:n]
        o = orc.get(name)[:n]
        assert rel_err(g, o) < REL, (name, rel_err(g, o))


def periodic_lattice(kind):
    """Fully periodic 24^3 jittered lattice (spacing 2r) with small random
    velocities and gravity: every neighbour loop takes its wrapped form."""
    m = 24
    r = 0.0025
    L = m * 2 * r
    x = scenes.lattice_block(m, m, m, r, -0.5 * L - r, 11, 0.05 * r)
    v = (0.05 * (scenes.uniform(12, 3 * x.shape[0]).reshape(-1, 3) - 0.5)).astype(np.float32)
    sc = scenes.Scene(name="periodic_lattice", x=x, v=v, wrap_length=(L, L, L),
                      grid_offset=(-7, -7, -7), grid_res=(14, 14, 14), boundaries=[])
    return sc


@pytest.mark.parametrize("kind", ["df", "ii", "i"])
def test_step_parity_periodic(kind):
    sc = periodic_lattice(kind)
    dp, pile, solver = gpu_scene(sc, kind)
    for _ in range(5):
        solver.step()
    orc = oracle_like(sc, dp, pile, kind=kind)
    sync_state(dp, solver, orc, kind=kind)
    solver.step()
    orc.step()
    n = sc.n
    rep = orc.report()
    assert solver.num_density_solve == rep["num_density_solve"]
    if kind == "df":
        assert solver.num_divergence_solve == rep["num_divergence_solve"]
    for name, gvar in (("v", solver.particle_v), ("density", solver.particle_density)):
        g = dp.coat(gvar).get()[:n]
        o = orc.get(name)[:n]
        assert rel_err(g, o) < REL, (name, rel_err(g, o))
    gx = dp.coat(solver.particle_x).get()[:n]
    ox = orc.get("x")[:n]
    assert np.abs(gx - ox).max() < 1e-6


def scaled_periodic(frac=0.2):
    """The 8M-particle periodic stress scene cut to a periodic cube of edge
    `frac` at the same number density (same neighbour statistics)."""
    n = int(round(8_388_608 * frac ** 3))
    sc = scenes.periodic_random(n)
    sc.x = (sc.x * np.float32(frac)).astype(np.float32)
    sc.wrap_length = (frac, frac, frac)
    c = int(round(frac / sc.kernel_radius))
    sc.grid_offset = (-c // 2,) * 3
    sc.grid_res = (c,) * 3
    return sc


def test_stress_path_parity_periodic():
    """Neighbour search + density + density-change sums + wrapped advection,
    three steps, against the oracle doing the same calls."""
    sc = scaled_periodic()
    dp, pile, solver = gpu_scene(sc)
    orc = oracle_like(sc, dp, pile)
    orc.set("x", sc.x)
    orc.set("v", sc.v)
    n = sc.n
    for step in range(3):
        solver.update_particle_neighbors()
        solver.compute_density_change()
        orc.compute_all_boundaries()
        orc.update_particle_neighbors()
        orc.compute_density_change()
        ids, cnt = solver.get_neighbors()
        oids, ocnt = orc.neighbors()
        np.testing.assert_array_equal(cnt[:n], ocnt[:n])
        mask = np.arange(ids.shape[1])[None, :] < ocnt[:n, None]
        np.testing.assert_array_equal(np.where(mask, ids[:n], 0), np.where(mask, oids[:n], 0))
        dens = dp.coat(solver.particle_density).get()[:n]
        assert rel_err(dens, orc.get("density")[:n]) < REL, step
        dadv = dp.coat(solver.particle_density_adv).get()[:n]
        assert rel_err(dadv, orc.get("density_adv")[:n]) < REL, step
        solver.integrate_velocity()
        orc.integrate_velocity()
        gx = dp.coat(solver.particle_x).get()[:n]
        np.testing.assert_array_equal(gx, orc.get("x")[:n])


def _f32_pair_r2(xi, xj, L):
    """|x_i - x_j|^2 with the minimum image, in the device's FP32 operation
    order: d = x_i - x_j; d -= L if d >= L/2; d += L if d < -L/2;
    r2 = (dx*dx + dy*dy) + dz*dz."""
    f = np.float32
    d = (xi[None, :] - xj).astype(f)
    lo, hi = f(L) * f(-0.5), f(L) * f(0.5)
    d = np.where(d >= hi, d - f(L), np.where(d < lo, d + f(L), d)).astype(f)
    return ((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]).astype(f)


def test_stress_8m_sampled_neighbors_brute_force():
    """Full-size periodic stress scene (8,388,608 particles): for 2,000 sampled
    particles the neighbour set equals brute force with the same FP32
    predicate (|x_ij|^2 < h^2, minimum image), and the density equals an FP64
    sum over that set."""
    from scipy.spatial import cKDTree

    sc = scenes.periodic_random(8_388_608)
    dp, pile, solver = gpu_scene(sc)
    solver.update_particle_neighbors()
    n = sc.n
    ids, cnt = solver.get_neighbors()
    K = ids.shape[1]
    dens = dp.coat(solver.particle_density).get()[:n]
    x = sc.x
    tree = cKDTree((x.astype(np.float64) + 0.5) % 1.0, boxsize=1.0)
    rng = np.random.default_rng(7)
    sample = rng.choice(n, 2000, replace=False)
    h = np.float32(sc.kernel_radius)
    h2 = h * h
    kc = 8.0 / (np.pi * float(h) ** 3)
    vol = sc.particle_mass / sc.density0
    checked = 0
    for i in sample:
        cand = np.array(tree.query_ball_point((x[i].astype(np.float64) + 0.5) % 1.0,
                                              float(h) * 1.001), dtype=np.int64)
        cand = cand[cand != i]
        r2 = _f32_pair_r2(x[i], x[cand], 1.0)
        true = np.sort(cand[r2 < h2])
        if len(true) > K:
            assert cnt[i] == K
            continue
        assert cnt[i] == len(true), i
        np.testing.assert_array_equal(np.sort(ids[i, :cnt[i]]), true)
        q = np.sqrt(r2[r2 < h2].astype(np.float64)) / float(h)
        w = np.where(q <= 0.5, 6 * q ** 3 - 6 * q ** 2 + 1, 2 * (1 - q) ** 3) * kc
        want = sc.density0 * vol * (kc + w.sum())
        assert abs(dens[i] - want) <= 1e-5 * want, i
        checked += 1
    assert checked > 1900
