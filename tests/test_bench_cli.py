"""bench.py on the CPU: the one-GPU limit, the CPU baseline leg and the
default launch shapes. No GPU compute is called here."""
import importlib.util
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_is_refused(bench, capsys):
    assert bench.main(["--gpus", "2"]) == 2
    assert "one GPU" in capsys.readouterr().err


def test_cpu_baseline_respects_the_thread_cap():
    # In a child process: the timed CPU restatement is built with the
    # reference's -Ofast, whose start-up code sets flush-to-zero for the whole
    # process, which the strict oracle tests must not inherit.
    code = ("import json, sys; sys.path.insert(0, sys.argv[1]); import bench; "
            "a = bench.cpu_baseline(64, 300); b = bench.cpu_baseline(64, 300, threads=1); "
            "c = bench.cpu_baseline(64, 10); print(json.dumps([a, b, c]))")
    out = subprocess.run([sys.executable, "-c", code, ROOT], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    (v_all, t_all, s_all), (v_one, t_one, s_one), (_, t_again, _) = json.loads(out.stdout.strip().splitlines()[-1])
    assert t_one == 1 and t_all >= 1
    assert v_all > 0 and v_one > 0 and s_all > 0 and s_one > 0
    # the cap is undone afterwards: a later uncapped call uses every core again
    assert t_again == t_all


def test_default_launch_shapes_are_library_configurations(bench):
    capi = bench.load_capi()
    table = capi.launch_configs()
    assert len(table) == 20
    for option, form in (("--launch-fp32", "fp32"), ("--launch-fp64", "fp64")):
        default = next(a.default for a in _parser_actions(bench) if option in a.option_strings)
        shape = tuple(int(x) for x in (default + "x1").split("x")[:3])
        assert shape in table, (form, shape)
    # the kernel name reported for the FP32-local default is the refilled lane kernel
    assert bench.kernel_name({"lanes_per_cru": 1, "block": 64, "crus_per_lane": 4}, True) == "k_step_lane<0, 64, 4>"
    assert bench.kernel_name({"lanes_per_cru": 4, "block": 64, "crus_per_lane": 1}, False) == "k_step<0, false, double, 64>"


def _parser_actions(bench):
    """The argparse actions of bench.main's parser, captured without running it."""
    import argparse

    captured = {}
    real = argparse.ArgumentParser.parse_args

    def grab(self, argv=None, namespace=None):
        captured["actions"] = list(self._actions)
        raise SystemExit(0)

    argparse.ArgumentParser.parse_args = grab
    try:
        with pytest.raises(SystemExit):
            bench.main([])
    finally:
        argparse.ArgumentParser.parse_args = real
    return captured["actions"]


def test_timed_window_starts_at_the_stated_time(bench):
    import argparse

    ns = argparse.Namespace(warmup=5, window_start_ms=1.0)
    assert bench.advance_steps(ns) == 1000          # a short --warmup still reaches the window
    ns = argparse.Namespace(warmup=4000, window_start_ms=1.0)
    assert bench.advance_steps(ns) == 4000          # a longer one moves it later


def test_cpu_baseline_runs_the_batch_first_cell_over_the_window(bench):
    """The CPU sample takes the first cell's parameters, seed and pacing period
    and the same untimed advance as the GPU leg (child process: -Ofast build)."""
    capi = bench.load_capi()
    cells, period = bench.make_cells(capi, "large1024", 0, 0)
    assert len(cells) == 1024 and period == 500.0
    c0 = cells[0]
    assert c0.params.PCaL != capi.Params.default().PCaL   # scaled conductances
    code = ("import json, sys; sys.path.insert(0, sys.argv[1]); import bench; capi = bench.load_capi(); "
            "cells, period = bench.make_cells(capi, 'large1024', 0, 0); c0 = cells[0]; "
            "print(json.dumps(bench.cpu_baseline(32, 200, 100, seed=c0.seed, params=c0.params, period=period)))")
    out = subprocess.run([sys.executable, "-c", code, ROOT], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    v, threads, secs = json.loads(out.stdout.strip().splitlines()[-1])
    assert v > 0 and threads >= 1 and secs > 0


def test_whole_beat_modes_and_steps_per_launch_parse_before_the_gpu(bench):
    """The whole-beat modes and the steps-per-launch choice are accepted
    options; more than one GPU is still refused before any GPU work."""
    for extra in (["--ensemble-beats", "5", "--workload", "seeds64"], ["--alternans-beats", "3"],
                  ["--steps-per-launch", "0"], ["--steps-per-launch", "2"], ["--steps-per-launch", "1"],
                  ["--latency"], ["--latency", "--latency-steps", "100"]):
        assert bench.main(["--gpus", "2"] + extra) == 2
    with pytest.raises(SystemExit):
        bench.main(["--steps-per-launch", "3"])
