"""Throughput benchmark of the SPH solver core on one MI355X.

Workload (default): DFSPH dam break with 1,048,576 fluid particles
(128 x 128 x 64 jittered lattice, r = 2.5 mm) in a 1 m cube container plus
four dynamic rigid bodies (two spheres r = 0.05 m, two cubes of edge 0.08 m,
128^3 volume maps), dt = 2e-4 s, divergence and density tolerance 0.1 %.
One step is SolverDf::step (neighbour search, boundaries, both Jacobi solves,
viscosity, integration, rigid coupling).  Inputs are resident on the GPU when
the timed region starts; volume-map construction is setup.

Prints one JSON line.  ``value`` is particle-steps/s of the timed steps.
``roofline`` is the dominant kernel's algorithmic bytes per launch divided by
its average duration (HIP events around every launch during the timed steps)
against the 8 TB/s HBM peak.  ``cpu_baseline`` times the CPU oracle
(C++/OpenMP restatement of the reference algorithm) on a bounded window of the
same workload, started from the GPU state at the start of the timed region.
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
METRIC = "particle-steps/s @1M DFSPH 0.1% ρ-err; learned vs plain warm start"


def kernel_bytes(name, K, nb):
    """Algorithmic bytes per particle of one launch (minimal widths: FP32
    scalar 4 B, vector 12 B, u32 index 4 B; neighbour-state gathers counted
    once per particle as cache-resident; boundary records 16 B each only for
    (particle, boundary) pairs with non-zero volume, nb per particle)."""
    nbr = 4 + 4 * K  # count + neighbour list
    table = {
        "k_neighbors": 12 + 4 * K + 4 + 8 + 4 + 16 * nb,
        "k_ws_div0": nbr + 12 + 12 + 8 + 4 + 32 * nb,
        "k_ws_div1": nbr + 4 + 12 + 24 + 4 + 16 * nb,
        "k_vel_density_change": nbr + 12 + 12 + 8 + 4 + 4 + 4 + 32 * nb,
        "k_div_iter": nbr + 4 + 12 + 8 + 24 + 4 + 16 * nb,
        "k_div_err": nbr + 12 + 12 + 4 + 4 + 4 + 4 + 32 * nb,
        "k_viscosity": nbr + 12 + 12 + 12 + 16 + 4 + 16 * nb,
        "k_integrate_np": 36,
        "k_ws_p0": nbr + 12 + 12 + 4 + 8 + 4 + 32 * nb,
        "k_ws_p1": nbr + 4 + 12 + 24 + 4 + 16 * nb,
        "k_rho_adv": nbr + 12 + 12 + 4 + 8 + 4 + 4 + 4 + 32 * nb,
        "k_p_iter": nbr + 4 + 12 + 8 + 24 + 4 + 16 * nb,
        "k_p_err": nbr + 12 + 12 + 4 + 4 + 4 + 4 + 4 + 32 * nb,
        "k_df_finish": 4 + 12 + 12 + 8 + 12 + 12 + 8,
        "k_gather_boundary": 4 + 12 + 12 + 8 + 12 + 12 + 8 + 4,
        "k_cell_count": 12 + 4 + 4,
        "k_scatter": 4 + 4 + 4,
    }
    return table.get(name)


def log(msg):
    print("[bench %8.2fs] %s" % (time.perf_counter() - T0, msg), file=sys.stderr, flush=True)


T0 = time.perf_counter()


def build(args):
    from alluvion_amd import scenes

    s = args.scale
    sc = scenes.dam_break_rigid(nx=128 // s, ny=128 // s, nz=64 // s,
                                body_resolution=args.body_resolution)
    log("scene: %d particles, %d boundaries" % (sc.n, len(sc.boundaries)))
    t0 = time.perf_counter()
    dp, pile, solver, cn, cni = scenes.build_solver(sc, log=log)
    setup_s = time.perf_counter() - t0
    log("setup done in %.2f s" % setup_s)
    return sc, dp, pile, solver, setup_s


def cpu_baseline(sc, dp, pile, solver, state, steps):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ffi as of

    orc = of.Oracle(sc, precision=4)
    for b, var in enumerate(pile.volume_grids):
        orc.set_volume_map(b, dp.coat(var).get())
    orc.set("x", state["x"])
    orc.set("v", state["v"])
    orc.set("kappa", state["kappa"])
    orc.set("kappa_v", state["kappa_v"])
    for b, kin in enumerate(state["rigid"]):
        orc.set_rigid(b, *kin)
    log("cpu oracle ready")
    orc.step()  # warm-up (first touch of the neighbour arrays)
    t0 = time.perf_counter()
    for i in range(steps):
        orc.step()
        log("cpu step %d" % i)
    dt = time.perf_counter() - t0
    cores = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
    return {
        "value": sc.n * steps / dt,
        "unit": "particle-steps/s",
        "cores": cores,
        "kind": "port",
        "sample": "%d DFSPH steps of the same 1,048,576-particle scene on the host CPU "
                  "(C++/OpenMP FP32 oracle, %d threads), started from the GPU state at the "
                  "start of the timed region, after 1 warm-up step" % (steps, cores),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-steps", type=int, default=48)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--scale", type=int, default=1,
                    help="divide the lattice dimensions (testing only)")
    ap.add_argument("--body-resolution", type=int, default=128)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on exactly one GPU (one-GPU solver core); "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import alluvion_amd._alluvion as al

    sc, dp, pile, solver, setup_s = build(args)
    n = sc.n
    for i in range(args.warmup):
        solver.step()
        log("warmup step %d: div %d dens %d" % (i, solver.num_divergence_solve,
                                                 solver.num_density_solve))
    state = None
    if not args.no_cpu_baseline:
        state = {
            "x": dp.coat(solver.particle_x).get()[:n].copy(),
            "v": dp.coat(solver.particle_v).get()[:n].copy(),
            "kappa": dp.coat(solver.particle_kappa).get()[:n].copy(),
            "kappa_v": dp.coat(solver.particle_kappa_v).get()[:n].copy(),
            "rigid": [(tuple(pile.get_x(b)), tuple(pile.get_v(b)), tuple(pile.get_q(b)),
                       tuple(pile.get_omega(b))) for b in range(pile.get_size())],
        }
    if not args.no_profile:
        al.profiling_enable(True)
        al.profiling_reset()
    n_div = n_dens = 0
    al.device_synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        solver.step()  # ends with a stream synchronisation (report read-back)
        n_div += solver.num_divergence_solve
        n_dens += solver.num_density_solve
        if i % 25 == 0:
            log("timed step %d" % i)
    al.device_synchronize()
    elapsed = time.perf_counter() - t0
    stats = {}
    if not args.no_profile:
        stats = al.profiling_stats()
        al.profiling_enable(False)
    value = n * args.steps / elapsed
    nnb = dp.coat(solver.particle_num_neighbors).get()[:n]
    K = float(nnb.mean())
    mask = dp.coat(solver.boundary_mask).get()[:n].astype(np.uint64)
    nb = float(sum(int(((mask >> np.uint64(b)) & np.uint64(1)).sum()) for b in range(32)) / n)

    roofline = None
    kernels = {}
    if stats:
        total_ms = sum(v[0] for v in stats.values())
        for name, (ms, launches) in stats.items():
            kernels[name] = {"total_ms": round(ms, 3), "launches": int(launches),
                             "avg_us": round(1000.0 * ms / max(launches, 1), 2)}
        dom = max(stats.items(), key=lambda kv: kv[1][0])
        name, (ms, launches) = dom
        bpp = kernel_bytes(name, K, nb)
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
                "traffic": None,
                "bytes_per_launch": int(bpp * n),
                "avg_launch_us": round(avg_s * 1e6, 2),
                "share_of_gpu_time": round(ms / total_ms, 3),
            }
    log("timed region: %.3f s" % elapsed)
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(sc, dp, pile, solver, state, args.cpu_steps)
        log("cpu baseline done")

    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000.0 * elapsed / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "DFSPH dam break, 1,048,576 particles (128x128x64 lattice, r=2.5mm) in a "
                        "1 m cube + 4 dynamic rigid bodies (128^3 volume maps), dt=2e-4, "
                        "tolerance 0.1%, plain warm start",
            "particles": n,
            "mean_neighbors": round(K, 2),
            "boundary_pairs_per_particle": round(nb, 4),
            "mean_divergence_iterations": round(n_div / args.steps, 3),
            "mean_density_iterations": round(n_dens / args.steps, 3),
            "setup_s": round(setup_s, 2),
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernels": kernels,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
