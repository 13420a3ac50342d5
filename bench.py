"""Throughput benchmark of the SPH solver core on one MI355X.

Default workload (the headline metric): DFSPH dam break with 1,048,576 fluid
particles (128 x 128 x 64 jittered lattice, r = 2.5 mm) in a 1 m cube
container plus four dynamic rigid bodies (two spheres r = 0.05 m, two cubes of
edge 0.08 m, 128^3 volume maps) floating partly submerged in the fluid,
dt = 2e-4 s, divergence and density tolerance 0.1 %.  One step is
SolverDf::step (neighbour search, boundaries, both Jacobi solves, viscosity,
integration, rigid coupling).  The timed window starts after --start-step
simulated steps (400 of the scene's 500 by default: pre-roll, autotuning and
warm-up steps together), when the dam break has developed.

``--workload periodic_8m`` is the roofline stress run: 8,388,608 particles
uniform in a periodic unit cube with velocities uniform in +-0.5 m/s; one step
is the neighbour search with the density sum, the density-change
(divergence) sum, and x += dt v with wrapping.

Inputs are resident on the GPU when the timed region starts; volume-map
construction is setup.  Prints one JSON line.  ``value`` is particle-steps/s
of the timed steps.  ``roofline`` is the dominant kernel's algorithmic bytes
per launch divided by its average duration (HIP events around every launch
during the timed steps) against the 8 TB/s HBM peak; ``traffic`` is that
kernel's HBM bytes per launch from the committed rocprofv3 PMC summary.
``cpu_baseline`` times the CPU oracle (C++/OpenMP restatement of the
reference algorithm) on a bounded window of the same workload, started from
the GPU state at the start of the timed region.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic.json")
CALIBRATION_MIB = 256
METRIC = "particle-steps/s @1M DFSPH 0.1% ρ-err; learned vs plain warm start"
METRIC_STRESS = "particle-steps/s @8M periodic: neighbour search + density + divergence sums"


def kernel_bytes(name, K, nb, df_fused=False):
    """Algorithmic bytes per particle of one launch (minimal widths: FP32
    scalar 4 B, vector 12 B, u32 index 4 B; neighbour-state gathers counted
    once per particle as cache-resident; boundary records 16 B each only for
    (particle, boundary) pairs with non-zero volume, nb per particle).  The
    neighbour list is counted at 4 B per entry.  df_fused: the DFSPH step with
    a divergence solve, whose neighbour search also runs the divergence warm
    start (its own velocity, the stored stiffness read and written, the
    stiffness copy into x.w, and the boundary points) and whose divergence
    passes also compute the viscosity (rho_j and the acceleration written),
    the bodies' reactions and the solve's finish scaling (factor and stored
    stiffness read and written, the boundary mask) moving to k_integrate_np."""
    nbr = 4 + 4 * K  # count + neighbour list
    table = {
        "k_neighbors": 12 + 4 * K + 4 + 8 + 4 + 16 * nb + ((12 + 8 + 4 + 16 * nb) if df_fused else 0),
        "k_ws_div0": nbr + 12 + 12 + 8 + 4 + 32 * nb,
        "k_ws_div1": nbr + 4 + 12 + 24 + 4 + 16 * nb,
        "k_vel_density_change": nbr + 12 + 12 + 8 + 4 + 4 + 4 + 32 * nb + (16 if df_fused else 0),
        "k_div_iter": nbr + 4 + 12 + 8 + 24 + 4 + 16 * nb,
        "k_div_err": nbr + 12 + 12 + 4 + 4 + 4 + 4 + 32 * nb + (16 if df_fused else 0),
        "k_viscosity": nbr + 12 + 12 + 12 + 16 + 4 + 16 * nb,
        "k_integrate_np": 36 + (20 if df_fused else 0),
        "k_ws_p0": nbr + 12 + 12 + 4 + 8 + 4 + 32 * nb,
        "k_ws_p1": nbr + 4 + 12 + 24 + 4 + 16 * nb,
        "k_rho_adv": nbr + 12 + 12 + 4 + 8 + 4 + 4 + 4 + 32 * nb,
        "k_p_iter": nbr + 4 + 12 + 8 + 24 + 4 + 16 * nb,
        "k_p_err": nbr + 12 + 12 + 4 + 4 + 4 + 4 + 4 + 32 * nb,
        "k_df_finish": 4 + 12 + 12 + 8 + 12 + 12 + 8,
        "k_gather_boundary": 4 + 12 + 12 + 8 + 12 + 12 + 8 + 4,
        "k_cell_count": 12 + 4 + 4,
        "k_scatter": 4 + 4 + 4,
        "k_density_change": nbr + 12 + 12 + 4 + 32 * nb,
        "k_integrate_velocity": 12 + 12 + 12,
    }
    return table.get(name)


T0 = time.perf_counter()


def log(msg):
    print("[bench %8.2fs] %s" % (time.perf_counter() - T0, msg), file=sys.stderr, flush=True)


def measured_traffic(workload, kernel):
    """HBM bytes per launch of a kernel from the committed rocprofv3 PMC
    summary (FETCH_SIZE + WRITE_SIZE, each scaled by the ratio of known bytes to
    counter reading for a plain device-to-device copy), or None."""
    try:
        with open(TRAFFIC_FILE) as f:
            table = json.load(f)
    except (OSError, ValueError):
        return None
    ent = table.get(workload, {}).get("kernels", {}).get(kernel)
    if not ent:
        return None
    return int(ent["fetch_bytes"] + ent["write_bytes"])


def stream_copy_gbs(dp, mib=1024, reps=10):
    """Measured copy ceiling: device-to-device copies of `mib` MiB
    (Variable.set_from, the runtime's copy kernel), 2 x bytes per copy over
    the mean time of `reps` copies after one untimed copy."""
    import alluvion_amd._alluvion as al

    n = mib * (1 << 20) // 4
    src = dp.create((n,), 1)
    dst = dp.create((n,), 1)
    src.set_same(1)  # any bytes: only the traffic matters
    dst.set_from(src)
    al.device_synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        dst.set_from(src)
    al.device_synchronize()
    el = (time.perf_counter() - t0) / reps
    dp.remove(src)
    dp.remove(dst)
    gbs = 2.0 * n * 4 / el / 1e9
    log("stream copy: %d MiB in %.3f ms = %.1f GB/s" % (mib, el * 1e3, gbs))
    return round(gbs, 1)


def host_cpu():
    """lscpu's model name, sockets, cores per socket and threads per core,
    and the CPUs this process may run on."""
    import subprocess

    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=30).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            info[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass

    def num(key):
        try:
            return int(info.get(key, ""))
        except ValueError:
            return None

    sockets, cores, threads = num("Socket(s)"), num("Core(s) per socket"), num("Thread(s) per core")
    return {"model": info.get("Model name"), "sockets": sockets, "cores_per_socket": cores,
            "threads_per_core": threads,
            "physical_cores": sockets * cores if sockets and cores else None,
            "affinity_cpus": len(os.sched_getaffinity(0))}


def copy_calibration(dp):
    """Plain device-to-device copies of a known size (Variable.set_from), so
    that a PMC run of this program can scale its HBM counters."""
    import alluvion_amd._alluvion as al

    n = CALIBRATION_MIB * (1 << 20) // 4
    src = dp.create((n,), 1)
    dst = dp.create((n,), 1)
    dp.coat(src).set(np.ones(n, np.float32))
    for _ in range(3):
        dst.set_from(src)
    al.device_synchronize()
    dp.remove(src)
    dp.remove(dst)
    log("copy calibration: 3 x %d MiB device-to-device" % CALIBRATION_MIB)


class DamBreak:
    name = "dam_break_1m_rigid"
    df_fused = True  # DFSPH: warm start in the search, viscosity in the divergence passes
    metric = METRIC
    default_steps = 100
    default_warmup = 20
    default_cpu_steps = 48
    default_start_step = 400

    def __init__(self, args):
        from alluvion_amd import scenes

        s = args.scale
        self.sc = scenes.dam_break_rigid(nx=128 // s, ny=128 // s, nz=64 // s,
                                         body_resolution=args.body_resolution)
        log("scene: %d particles, %d boundaries" % (self.sc.n, len(self.sc.boundaries)))
        t0 = time.perf_counter()
        self.dp, self.pile, self.solver, _, _ = scenes.build_solver(self.sc, log=log)
        self.setup_s = time.perf_counter() - t0
        self.n_div = self.n_dens = 0

    def step(self):
        self.solver.step()  # ends with a stream synchronisation (report read-back)
        self.n_div += self.solver.num_divergence_solve
        self.n_dens += self.solver.num_density_solve

    def body_forces(self):
        """Force sums of the last step on the dynamic bodies, and each as a
        fraction of the body's weight."""
        out = []
        for b, spec in enumerate(self.sc.boundaries):
            if spec.mass == 0:
                continue
            f = np.array(list(self.pile.get_force(b)), np.float64)
            out.append({"body": b, "kind": spec.kind, "force_N": [round(float(c), 5) for c in f],
                        "lift_over_weight": round(float(f[1] / (spec.mass * 9.81)), 4)})
        return out

    def snapshot(self):
        dp, solver, pile, n = self.dp, self.solver, self.pile, self.sc.n
        return {
            "x": dp.coat(solver.particle_x).get()[:n].copy(),
            "v": dp.coat(solver.particle_v).get()[:n].copy(),
            "kappa": dp.coat(solver.particle_kappa).get()[:n].copy(),
            "kappa_v": dp.coat(solver.particle_kappa_v).get()[:n].copy(),
            "rigid": [(tuple(pile.get_x(b)), tuple(pile.get_v(b)), tuple(pile.get_q(b)),
                       tuple(pile.get_omega(b))) for b in range(pile.get_size())],
        }

    def oracle(self, state):
        import oracle_ffi as of

        orc = of.Oracle(self.sc, precision=4)
        for b, var in enumerate(self.pile.volume_grids):
            orc.set_volume_map(b, self.dp.coat(var).get())
        for k in ("x", "v", "kappa", "kappa_v"):
            orc.set(k, state[k])
        for b, kin in enumerate(state["rigid"]):
            orc.set_rigid(b, *kin)
        return orc, orc.step

    def sample_text(self, steps, cores):
        return ("%d DFSPH steps of the same 1,048,576-particle scene on the host CPU "
                "(C++/OpenMP FP32 oracle, %d threads), started from the GPU state at the "
                "start of the timed region, after 1 warm-up step" % (steps, cores))

    def config(self, steps, K, nb):
        return {
            "workload": "DFSPH dam break, 1,048,576 particles (128x128x64 lattice, r=2.5mm) in a "
                        "1 m cube + 4 dynamic rigid bodies floating partly submerged (128^3 "
                        "volume maps), dt=2e-4, tolerance 0.1%, plain warm start",
            "particles": self.sc.n,
            "mean_neighbors": round(K, 2),
            "boundary_pairs_per_particle": round(nb, 4),
            "mean_divergence_iterations": round(self.n_div / steps, 3),
            "mean_density_iterations": round(self.n_dens / steps, 3),
            "setup_s": round(self.setup_s, 2),
        }


class PeriodicStress:
    name = "periodic_8m"
    metric = METRIC_STRESS
    default_steps = 100
    default_warmup = 10
    default_cpu_steps = 3

    def __init__(self, args):
        from alluvion_amd import scenes

        n = 8_388_608 // (args.scale ** 3)
        self.sc = scenes.periodic_random(n)
        log("scene: %d particles, periodic unit cube" % n)
        t0 = time.perf_counter()
        self.dp, self.pile, self.solver, _, _ = scenes.build_solver(self.sc, log=log)
        self.setup_s = time.perf_counter() - t0

    def step(self):
        s = self.solver
        s.update_particle_neighbors()  # grid, neighbour list, density
        s.compute_density_change()  # divergence sum
        s.integrate_velocity()  # x += dt v, wrapped

    def snapshot(self):
        dp, solver, n = self.dp, self.solver, self.sc.n
        return {"x": dp.coat(solver.particle_x).get()[:n].copy(),
                "v": dp.coat(solver.particle_v).get()[:n].copy()}

    def oracle(self, state):
        import oracle_ffi as of

        orc = of.Oracle(self.sc, precision=4)
        orc.set("x", state["x"])
        orc.set("v", state["v"])

        def step():
            orc.compute_all_boundaries()
            orc.update_particle_neighbors()
            orc.compute_density_change()
            orc.integrate_velocity()

        return orc, step

    def sample_text(self, steps, cores):
        return ("%d steps (neighbour search + density + density-change sums + advection) of the "
                "same 8,388,608-particle periodic scene on the host CPU (C++/OpenMP FP32 oracle, "
                "%d threads), after 1 warm-up step" % (steps, cores))

    def config(self, steps, K, nb):
        return {
            "workload": "periodic unit cube, 8,388,608 uniform random particles, v in +-0.5 m/s, "
                        "h=0.01 (100^3 cells): neighbour search + density + density-change sums "
                        "+ wrapped advection",
            "particles": self.sc.n,
            "mean_neighbors": round(K, 2),
            "setup_s": round(self.setup_s, 2),
        }


class BatchedScenes:
    name = "batched_256"
    df_fused = True
    metric = "particle-steps/s @256 batched dam breaks (4k-120k particles each), DFSPH 0.1% ρ-err"
    default_steps = 500
    default_warmup = 10
    default_cpu_steps = 30
    cpu_scenes = 32

    def __init__(self, args):
        from alluvion_amd import scenes

        self.scs = scenes.batched_dam_breaks(256 // (args.scale ** 3) or 1,
                                             map_res=args.batch_map_res)
        n = sum(sc.n for sc in self.scs)
        log("scene: %d scenes, %d particles" % (len(self.scs), n))
        t0 = time.perf_counter()
        self.dp, self.piles, self.solver, _, _ = scenes.build_batch(self.scs, log=log)
        self.setup_s = time.perf_counter() - t0
        self.sc = type("Batch", (), {"n": n})()
        self.iters = 0

    def step(self):
        self.solver.step()
        self.iters += self.solver.num_divergence_solve + self.solver.num_density_solve

    def ranges(self):
        return [self.solver.scene_range(k) for k in range(len(self.scs))]

    def snapshot(self):
        X = self.dp.coat(self.solver.particle_x).get()
        V = self.dp.coat(self.solver.particle_v).get()
        out = []
        for k in range(min(self.cpu_scenes, len(self.scs))):
            first, count = self.solver.scene_range(k)
            out.append((X[first:first + count].copy(), V[first:first + count].copy()))
        return out

    def oracle(self, state):
        """The CPU oracle has no batch: it steps the first scenes one after
        another (their own volume maps, same state)."""
        import oracle_ffi as of

        orcs = []
        for k, (x, v) in enumerate(state):
            orc = of.Oracle(self.scs[k], precision=4)
            grids = self.piles[k].volume_grids
            for b, var in enumerate(grids):
                orc.set_volume_map(b, self.dp.coat(var).get())
            orc.set("x", x)
            orc.set("v", v)
            orcs.append(orc)
        self.cpu_particles = sum(len(x) for x, _ in state)

        def step():
            for orc in orcs:
                orc.step()

        return orcs, step

    def sample_text(self, steps, cores):
        return ("%d DFSPH steps of the first %d of the 256 scenes (%d particles) on the host "
                "CPU, one after another (C++/OpenMP FP32 oracle, %d threads), after 1 warm-up "
                "step" % (steps, self.cpu_scenes, self.cpu_particles, cores))

    def config(self, steps, K, nb):
        return {
            "workload": "256 independent DFSPH dam breaks in one batch solver: particle counts "
                        "log-uniform in 4k-120k, cube edges 0.2-0.6 m, own %d^3 volume map each, "
                        "dt=2e-4, tolerance 0.1%%" % self.scs[0].boundaries[0].resolution[0],
            "scenes": len(self.scs),
            "particles": self.sc.n,
            "mean_neighbors": round(K, 2),
            "iterations_per_step_max_over_scenes": round(self.iters / steps, 3),
            "setup_s": round(self.setup_s, 2),
        }


WORKLOADS = {"dam_break_1m_rigid": DamBreak, "periodic_8m": PeriodicStress,
             "batched_256": BatchedScenes}


def learned_comparison(scale, body_resolution, replay_steps=300):
    """Plain vs learned warm start, measured by replay.

    The network is trained on the GPU on pairs logged from plain runs with
    the CFL-limited step (cfl 0.4, dt <= 2e-3): 300 steps of the
    16,384-particle dam break, and every 5th of 350 steps of a 1M-particle
    floating-body dam break drawn with other seeds (lattice jitter 12, bodies
    13), so that the pairs cover the whole flow and none comes from the run
    the comparison replays.  Then, on each scene (the 16,384-particle run
    continued; the benchmark's own 1M scene after 50 plain steps), every one
    of `replay_steps` further steps is taken twice from the same state, with
    the network and with the reference's warm start
    (alluvion_amd.warm_start.replay_compare), first with the CFL-limited step,
    then with the fixed dt = 2e-4 of the configs.  Tolerance 0.1 % and the
    reference's minimum iteration counts (1 divergence, 2 density)
    throughout, so no step can take fewer than 3 iterations: the largest
    possible ratio is plain / 3."""
    import alluvion_amd._alluvion as al
    from alluvion_amd import Depot, scenes, warm_start

    def big(seed, body_seed):
        return scenes.adaptive_step(scenes.dam_break_rigid(
            seed=seed, body_seed=body_seed, nx=128 // scale, ny=128 // scale, nz=64 // scale,
            body_resolution=body_resolution))

    t0 = time.perf_counter()
    logs = []
    sc_small = scenes.adaptive_step(scenes.dam_break_small())
    dp_s, pile_s, solver_s, _, _ = scenes.build_solver(sc_small)
    logs.append((dp_s,) + warm_start.collect_pairs(dp_s, solver_s, 300))
    log("warm start: %d pairs logged on the 16,384-particle dam break" % logs[-1][-1])
    sc_train = big(12, 13)
    dp_t, pile_t, solver_t, _, _ = scenes.build_solver(sc_train)
    logs.append((dp_t,) + warm_start.collect_pairs(dp_t, solver_t, 350, stride=5))
    log("warm start: %d pairs logged on the 1M-particle training run" % logs[-1][-1])
    # one data set: the logged pairs back to back
    count = sum(r[-1] for r in logs)
    D = Depot(np.float32)
    X = D.create((count * 8,), 1)
    Y = D.create((count * 2,), 1)
    off = 0
    for dp, x, y, c in logs:
        X.set_from(x, c * 8, off * 8)
        Y.set_from(y, c * 2, off * 2)
        off += c
        dp.remove(x)
        dp.remove(y)
    del solver_t, pile_t
    al.device_synchronize()
    t1 = time.perf_counter()
    net, losses = warm_start.train(X, Y, count)
    t2 = time.perf_counter()
    D.remove(X)
    D.remove(Y)
    log("warm start: %d pairs logged in %.2f s, trained in %.2f s (loss %.3f -> %.3f)" % (
        count, t1 - t0, t2 - t1, float(np.mean(losses[:10])), float(np.mean(losses[-100:]))))
    sc_big = big(2, 3)
    dp_b, pile_b, solver_b, _, _ = scenes.build_solver(sc_big)
    for _ in range(50):
        solver_b.step()
    runs = [("dam break 16,384 particles", dp_s, pile_s, solver_s),
            ("dam break 1,048,576 particles + 4 floating bodies", dp_b, pile_b, solver_b)]
    results = []
    floor = 3  # min_divergence_solve + min_density_solve
    for name, dp, pile, solver in runs:
        for rule in ("CFL-limited, cfl 0.4, dt <= 2e-3", "fixed dt = 2e-4"):
            if rule.startswith("fixed"):
                solver.min_dt = solver.max_dt = solver.dt = 2e-4
            lit, pit, lms, pms = warm_start.replay_compare(dp, solver, pile, net, replay_steps)
            res = {"scene": name, "time_step": rule, "steps": replay_steps,
                   "plain_iterations_per_step": round(float(pit.mean()), 3),
                   "learned_iterations_per_step": round(float(lit.mean()), 3),
                   "iteration_ratio": round(float(pit.mean() / lit.mean()), 3),
                   "ratio_ceiling_at_min_counts": round(float(pit.mean() / floor), 3),
                   "steps_learned_fewer_or_equal": round(float((lit <= pit).mean()), 3),
                   "plain_ms_per_step": round(float(np.median(pms)), 4),
                   "learned_ms_per_step": round(float(np.median(lms)), 4),
                   "speedup": round(float(np.median(pms) / np.median(lms)), 3)}
            results.append(res)
            log("warm start replay %s, %s: %s" % (name, rule, res))
    return {"tolerance": "0.1% (divergence, density)",
            "method": "replay: each step taken from one state with the network and with the "
                      "reference warm start; the trajectory continues from the plain one",
            "target": "correction to the reference warm start (stored stiffness)",
            "train": {"pairs": int(count),
                      "source": "300 plain CFL steps of the 16,384-particle dam break and every "
                                "5th of 350 plain CFL steps of a 1M-particle floating-body dam "
                                "break drawn with other seeds (12, 13)",
                      "adam_steps": len(losses), "batch": 65536, "seconds": round(t2 - t1, 3)},
            "runs": results}


def cpu_baseline(wl, state, steps):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ffi as of

    native = of.use_native_build()  # -march=native for this host, built here
    log("cpu oracle build: %s" % (native or "portable (x86-64-v3)"))
    orc, step = wl.oracle(state)
    log("cpu oracle ready")
    step()  # warm-up (first touch of the neighbour arrays)
    t0 = time.perf_counter()
    for i in range(steps):
        step()
        log("cpu step %d" % i)
    dt = time.perf_counter() - t0
    cores = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
    particles = getattr(wl, "cpu_particles", wl.sc.n)  # the sample's particle count
    return {
        "value": particles * steps / dt,
        "unit": "particle-steps/s",
        "cores": cores,
        "kind": "port",
        "sample": wl.sample_text(steps, cores),
        "host": host_cpu(),
        "build": "g++ -O3 %s -fopenmp -ffp-contract=off, FP32" % (
            "-march=native" if native else "-march=x86-64-v3"),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="dam_break_1m_rigid")
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--cpu-steps", type=int, default=None)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--scale", type=int, default=1,
                    help="divide the lattice dimensions (testing only)")
    ap.add_argument("--body-resolution", type=int, default=128)
    ap.add_argument("--batch-map-res", type=int, default=64)
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep direct launches instead of tuning the step execution")
    ap.add_argument("--no-learned", action="store_true",
                    help="skip the plain-vs-learned warm start comparison")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run 3 device copies of 256 MiB before the steps (PMC scaling)")
    ap.add_argument("--start-step", type=int, default=None,
                    help="simulated steps before the timed window (pre-roll, autotuning and "
                         "warm-up together); default 400 for the dam break, 0 otherwise")
    ap.add_argument("--no-stream", action="store_true",
                    help="skip the measured device-copy (STREAM copy) ceiling")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on exactly one GPU (one-GPU solver core); "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import alluvion_amd._alluvion as al

    wl = WORKLOADS[args.workload](args)
    steps = args.steps if args.steps is not None else wl.default_steps
    warmup = args.warmup if args.warmup is not None else wl.default_warmup
    cpu_steps = args.cpu_steps if args.cpu_steps is not None else wl.default_cpu_steps
    dp, solver, n = wl.dp, wl.solver, wl.sc.n
    if args.copy_calibration:
        copy_calibration(dp)
    start_step = args.start_step if args.start_step is not None else getattr(
        wl, "default_start_step", 0)
    do_tune = not args.no_autotune and hasattr(wl.solver, "set_graph") and \
        args.workload != "periodic_8m"
    from alluvion_amd import autotune
    tune_steps = autotune.steps_taken() if do_tune else 0
    preroll = max(0, start_step - tune_steps - warmup)
    start_step = preroll + tune_steps + warmup  # the first timed step
    for i in range(preroll):
        wl.step()
        if i % 50 == 0:
            log("pre-roll step %d of %d" % (i, preroll))
    tuned = None
    if do_tune:
        best, timing = autotune.tune_step(wl.solver, log=log)
        tuned = {"lds_tiles": best[0], "graph": best[1], "slack": best[2],
                 "ms_per_step": {"%s/%s/%d" % ("tiles" if ti else "gathers",
                                               "graph" if g else "direct", s): round(t, 4)
                                 for (ti, g, s), t in timing.items()}}
    for i in range(warmup):
        wl.step()
        log("warmup step %d" % i)
    state = None if args.no_cpu_baseline else wl.snapshot()
    if hasattr(wl, "n_div"):
        wl.n_div = wl.n_dens = 0
    if hasattr(wl, "iters"):
        wl.iters = 0
    al.device_synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        wl.step()
        if i % 25 == 0:
            log("timed step %d" % i)
    al.device_synchronize()
    elapsed = time.perf_counter() - t0
    cfg = wl.config(steps, 0.0, 0.0)  # iteration statistics of the timed steps
    # Per-kernel times come from a separate window after the timed one: the
    # HIP events around every launch need direct launches (no graphs).
    stats = {}
    if not args.no_profile:
        al.profiling_enable(True)
        al.profiling_reset()
        for _ in range(min(steps, 50)):
            wl.step()
        al.device_synchronize()
        stats = al.profiling_stats()
        al.profiling_enable(False)
    value = n * steps / elapsed
    ranges = wl.ranges() if hasattr(wl, "ranges") else [(0, n)]
    nnb = dp.coat(solver.particle_num_neighbors).get()
    nnb = np.concatenate([nnb[a:a + c] for a, c in ranges])
    K = float(nnb.mean())
    mask = dp.coat(solver.boundary_mask).get().astype(np.uint64)
    mask = np.concatenate([mask[a:a + c] for a, c in ranges])
    nb = float(sum(int(((mask >> np.uint64(b)) & np.uint64(1)).sum()) for b in range(32)) / n)
    cfg.update({k: v for k, v in wl.config(steps, K, nb).items()
                if k in ("mean_neighbors", "boundary_pairs_per_particle")})

    roofline = None
    kernels = {}
    if stats:
        total_ms = sum(v[0] for v in stats.values())
        for name, (ms, launches) in stats.items():
            ent = {"total_ms": round(ms, 3), "launches": int(launches),
                   "avg_us": round(1000.0 * ms / max(launches, 1), 2)}
            bpp = kernel_bytes(name, K, nb, getattr(wl, "df_fused", False))
            if bpp is not None and launches and ms > 0:
                ent["algorithmic_gbs"] = round(bpp * n / (ms / 1000.0 / launches) / 1e9, 1)
            kernels[name] = ent
        dom = max(stats.items(), key=lambda kv: kv[1][0])
        name, (ms, launches) = dom
        bpp = kernel_bytes(name, K, nb, getattr(wl, "df_fused", False))
        if bpp is not None and launches:
            avg_s = ms / 1000.0 / launches
            achieved = bpp * n / avg_s / 1e9
            roofline = {
                "bound": "hbm",
                "kernel": name,
                "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS,
                "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": measured_traffic(wl.name, name),
                "bytes_per_launch": int(bpp * n),
                "avg_launch_us": round(avg_s * 1e6, 2),
                "share_of_gpu_time": round(ms / total_ms, 3),
            }
    log("timed region: %.3f s" % elapsed)
    cfg["timed_steps"] = "%d-%d" % (start_step, start_step + steps)
    if hasattr(wl, "body_forces"):
        cfg["body_forces_last_step"] = wl.body_forces()
    if roofline is not None and not args.no_stream:
        roofline["stream_copy_gbs"] = stream_copy_gbs(dp)
        roofline["frac_of_stream_copy"] = round(roofline["achieved"] /
                                                roofline["stream_copy_gbs"], 4)
    learned = None
    if args.workload == "dam_break_1m_rigid" and not args.no_learned:
        learned = learned_comparison(args.scale, args.body_resolution)
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(wl, state, cpu_steps)
        log("cpu baseline done")

    out = {
        "metric": wl.metric,
        "value": round(value, 1),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": round(1000.0 * elapsed / steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": cfg,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "learned_warm_start": learned,
        "step_execution": tuned,
        "kernels": kernels,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
