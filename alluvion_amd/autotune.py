"""Autotuner for the step execution of a solver.

Runs a few steps of the live simulation under each candidate execution
setting and keeps the fastest.  The candidates are direct kernel launches
and the graph-replayed step with 0, 1 or 2 slack Jacobi iterations per solve
(extra iterations are no-ops once a scene has converged; fewer slack means
more host round trips when a solve needs more iterations than the step
before).  Every candidate computes the same results, so tuning only moves
time; the tuning steps are ordinary simulation steps.
"""
import time

CANDIDATES = ((False, 0), (True, 0), (True, 1), (True, 2))


def tune_step(solver, steps=5, candidates=CANDIDATES, log=None):
    """Returns (best (graph, slack), {(graph, slack): ms per step})."""
    timing = {}
    for graph, slack in candidates:
        solver.set_graph(graph, slack)
        solver.step()  # capture and first replay stay outside the timing
        t0 = time.perf_counter()
        for _ in range(steps):
            solver.step()  # each step ends with a synchronisation
        timing[(graph, slack)] = 1e3 * (time.perf_counter() - t0) / steps
        if log:
            log("autotune: graph=%s slack=%d: %.3f ms/step" % (graph, slack, timing[(graph, slack)]))
    best = min(timing, key=timing.get)
    solver.set_graph(*best)
    return best, timing
