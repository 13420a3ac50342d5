"""Per-step parity helpers: copy the GPU solver's state into an oracle, step
both, compare.

Parity is judged one step at a time from a synchronised state: before each
compared step the oracle takes the GPU's positions, velocities, stored
stiffnesses (DFSPH) or pressures (IISPH, ISPH), the bodies' kinematics and
the current time step, so differences never accumulate across steps.
"""
import numpy as np

import oracle_ffi as of

REL = 1e-5


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    scale = max(np.abs(b).max(), 1e-30)
    return np.abs(a - b).max() / scale


def body_state(pile, b):
    q = pile.get_q(b)
    return {"x": np.array(list(pile.get_x(b)), np.float64),
            "v": np.array(list(pile.get_v(b)), np.float64),
            "q": np.array([q.x, q.y, q.z, q.w], np.float64),
            "omega": np.array(list(pile.get_omega(b)), np.float64)}


def oracle_for(scene, dp, piles_grids, kind="df", precision=4, max_particles=None,
               distance_grids=None):
    """An oracle of `scene` using the GPU-built volume maps (device
    variables, one per boundary) and, for sampled boundaries, the GPU's
    distance maps."""
    orc = of.Oracle(scene, precision=precision, kind=kind, max_particles=max_particles)
    for b, var in enumerate(piles_grids):
        orc.set_volume_map(b, dp.coat(var).get())
    for b, spec in enumerate(scene.boundaries):
        if getattr(spec, "sampled", False):
            orc.set_distance_map(b, dp.coat(distance_grids[b]).get())
    return orc


def sync(dp, solver, pile, orc, kind="df", first=0, dt=None):
    """Copies the GPU state of rows [first, first + orc.n) (a batch scene's
    range, or the whole single-scene solver) into the oracle, with the time
    step the next step starts from (`dt`, default the solver's)."""
    n = orc.n
    orc.update_settings(dt=float(solver.dt if dt is None else dt))
    full = np.zeros((orc.nt, 3), np.float32)
    full[:n] = dp.coat(solver.particle_x).get()[first:first + n]
    orc.set("x", full)
    full = np.zeros((orc.nt, 3), np.float32)
    full[:n] = dp.coat(solver.particle_v).get()[first:first + n]
    orc.set("v", full)
    if kind == "df":
        k = np.zeros(orc.max_n, np.float32)
        k[:n] = dp.coat(solver.particle_kappa).get()[first:first + n]
        orc.set("kappa", k)
        k = np.zeros(orc.max_n, np.float32)
        k[:n] = dp.coat(solver.particle_kappa_v).get()[first:first + n]
        orc.set("kappa_v", k)
    else:
        p = np.zeros(orc.nt, np.float32)
        pg = dp.coat(solver.particle_pressure).get()
        p[:n] = pg[first:first + n]
        if orc.max_p:  # pellets' rows (their extrapolated pressures)
            p[orc.max_n:] = pg[orc.max_n:orc.nt]
        orc.set("pressure", p)
    if pile is not None:
        for b in range(pile.get_size()):
            s = body_state(pile, b)
            orc.set_rigid(b, s["x"], s["v"], s["q"], s["omega"])


def compare_fields(dp, solver, orc, kind="df", first=0, rel=REL, tag=""):
    """Max-norm relative errors of the step's fields against the oracle;
    asserts each is within `rel` and returns them."""
    n = orc.n
    errs = {}
    names = ["v", "density"] + (["kappa", "kappa_v"] if kind == "df" else ["pressure"])
    for name in names:
        g = dp.coat(getattr(solver, "particle_" + name)).get()[first:first + n]
        errs[name] = rel_err(g, orc.get(name)[:n])
    gx = dp.coat(solver.particle_x).get()[first:first + n]
    errs["x_abs"] = float(np.abs(gx - orc.get("x")[:n]).max())
    print(tag, {k: "%.3g" % v for k, v in errs.items()}, flush=True)
    for name in names:
        assert errs[name] < rel, (tag, name, errs[name])
    assert errs["x_abs"] < 1e-6, (tag, "x", errs["x_abs"])
    return errs
