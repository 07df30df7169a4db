"""bench.py's command line on the CPU: one GPU only.

``--gpus`` above 1, or a launch with WORLD_SIZE above 1, ends with status 2
and a message before anything touches a GPU or loads the library.
"""
import importlib.util
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench_cli", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_is_refused(bench, capsys, monkeypatch):
    monkeypatch.delenv("WORLD_SIZE", raising=False)
    assert bench.main(["--gpus", "2"]) == 2
    assert "one GPU only" in capsys.readouterr().err


def test_world_size_above_one_is_refused(bench, capsys, monkeypatch):
    monkeypatch.setenv("WORLD_SIZE", "4")
    assert bench.main([]) == 2
    assert "WORLD_SIZE 4" in capsys.readouterr().err


def test_cpu_baseline_scales_one_solves_counts(bench, monkeypatch):
    """The CPU estimate multiplies the sample's unit costs by the counts of ONE
    GPU solve (not divided by the timed steps), single-threaded and all-core."""
    from goma_amd import problems

    calls = []

    def fake_sample(prob, out, key, gmres_its=10):
        # 1 ms per element, 2 ns per nonzero to factor, 0.5 ns per nonzero per
        # iteration on one thread; 4x slower per copy when run concurrently
        calls.append(key)
        slow = 1.0 if len(calls) == 1 else 4.0
        ne = prob.mesh.num_elems
        out[key] = dict(fill=1e-3 * ne * slow, fact=2e-9 * 1000 * slow, it=0.5e-9 * 1000 * slow,
                        its=10, elems=ne, nnz=1000)

    monkeypatch.setattr(bench, "_cpu_sample", fake_sample)

    class Ctx:
        nval = 1_000_001

    prob = problems.lid_cavity(n=64, Re=100.0)
    counts = dict(assemblies=12, residual_fills=22, factorisations=11, gmres_its=129)
    out = bench.cpu_baseline("cavity2d", Ctx(), prob, counts, threads=8)
    ne, nnz = prob.mesh.num_elems, 1_000_000
    single = 34 * 1e-3 * ne + 11 * 2e-9 * nnz + 129 * 0.5e-9 * nnz
    assert abs(out["single_thread"]["value"] - round(single, 1)) <= 0.05
    # 8 copies at 4x the time each: twice the single-thread throughput
    assert abs(out["value"] - round(single / 2, 1)) <= 0.05
    assert out["cores"] == 8 and out["counts"] == counts
    assert "12 Jacobian and 22 residual assemblies" in out["sample"]
