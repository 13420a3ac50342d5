"""Autotuner for the step execution of a solver.

Runs a few steps of the live simulation under each candidate execution
setting and keeps the fastest.  The settings are the neighbour passes'
source (each 256-particle block's neighbourhood tile staged in LDS, or
gathers from global memory), and the step execution: direct kernel launches
or the graph-replayed step with 0, 1 or 2 slack Jacobi iterations per solve
(extra iterations are no-ops once a scene has converged; fewer slack means
more host round trips when a solve needs more iterations than the step
before).  Every candidate computes the same results, so tuning only moves
time; the tuning steps are ordinary simulation steps.
"""
import time

# (LDS tiles, graph, slack)
CANDIDATES = ((False, False, 0), (False, True, 0), (False, True, 1), (False, True, 2),
              (True, False, 0), (True, True, 0), (True, True, 1))


def steps_taken(steps=5, candidates=CANDIDATES):
    """Simulation steps that tune_step runs with these arguments."""
    return len(candidates) * (steps + 1)


def tune_step(solver, steps=5, candidates=CANDIDATES, log=None):
    """Returns (best (tiles, graph, slack), {(tiles, graph, slack): ms per step})."""
    timing = {}
    for tiles, graph, slack in candidates:
        solver.set_tiles(tiles)
        solver.set_graph(graph, slack)
        solver.step()  # capture and first replay stay outside the timing
        t0 = time.perf_counter()
        for _ in range(steps):
            solver.step()  # each step ends with a synchronisation
        key = (tiles, graph, slack)
        timing[key] = 1e3 * (time.perf_counter() - t0) / steps
        if log:
            log("autotune: tiles=%s graph=%s slack=%d: %.3f ms/step" % (key + (timing[key],)))
    best = min(timing, key=timing.get)
    solver.set_tiles(best[0])
    solver.set_graph(best[1], best[2])
    return best, timing
