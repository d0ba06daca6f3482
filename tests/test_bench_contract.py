"""bench.py's contract on the CPU: one GPU only (another count exits with
status 2 and says why), and the algorithmic bytes of the solve-units kernel
against the formula written out term by term for one case."""
import importlib.util
import os
import subprocess
import sys
from types import SimpleNamespace

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(REPO, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120, cwd=REPO)
    assert r.returncode == 2
    assert "one GPU" in (r.stdout + r.stderr)


def test_solver_bytes_terms():
    b = _bench()
    # 2 steps executed with 10 awake bodies each (one solve-units launch per step), 20 contact points, 5 manifolds,
    # 4 substeps, 8 velocity iterations, 1 executed position iteration
    st = SimpleNamespace(steps_executed=2, awake_bodies_stepped=20, solver_points=20,
                         solver_manifolds=5, position_iterations_run=1)
    N, K, M, S, I = 10, 20, 5, 4, 8
    per_substep = N * 208 + K * 120 + M * 184 + (I + 1) * (K * 112 + M * 176)
    one_step = S * per_substep + (K * 72 + M * 176) + (K * 40 + M * 192) + K * 20
    # two launches of the kernel in that frame: two steps' bytes
    assert b.solver_bytes(st, S, I) == 2 * one_step
    # no manifolds: only the bodies' substeps; nothing executed: nothing
    st0 = SimpleNamespace(steps_executed=1, awake_bodies_stepped=7, solver_points=0,
                          solver_manifolds=0, position_iterations_run=0)
    assert b.solver_bytes(st0, S, I) == S * 7 * 208
    st1 = SimpleNamespace(steps_executed=0, awake_bodies_stepped=0, solver_points=0,
                          solver_manifolds=0, position_iterations_run=0)
    assert b.solver_bytes(st1, S, I) == 0
