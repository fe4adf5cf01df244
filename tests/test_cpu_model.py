"""The library's kernels and host runtime on the CPU model, under
AddressSanitizer and UndefinedBehaviorSanitizer (no GPU needed).

tests/cpu_model/build_model.py compiles the unchanged csrc/ sources with g++
against a CPU model of the HIP device language and runtime: workgroup threads
run as fibers with real barriers, LDS is poisoned before every workgroup and
bounds-checked, buffer-resource accesses are checked against their extent,
and hipMalloc / launches / copies can be made to fail on demand.  Each suite
runs in a subprocess with the ASan runtime preloaded and checks every output
against the CPU oracle.  The full sweep of autotuner schedules takes several
minutes on 8 cores; SFNTT_MODEL_FULL=1 runs it, otherwise one NTT and one LDE
shape of it run.
"""
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
MODEL = os.path.join(HERE, "cpu_model")
sys.path.insert(0, MODEL)


@pytest.fixture(scope="module")
def model_env():
    import build_model

    build_model.build()
    env = dict(os.environ)
    env["LD_PRELOAD"] = build_model.asan_runtime()
    env["ASAN_OPTIONS"] = "detect_leaks=0:halt_on_error=1:abort_on_error=1"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    env.pop("SFNTT_MODEL_FAIL", None)
    return env


SUITES = ["ntt", "lde", "plans", "server", "faults", "cache", "noalloc",
          "schedules_full" if os.environ.get("SFNTT_MODEL_FULL") == "1" else "schedules"]


def _run_suite(env, suite):
    r = subprocess.run([sys.executable, os.path.join(MODEL, "run_model.py"), suite], env=env,
                       capture_output=True, text=True, timeout=1800)
    tail = "\n".join((r.stdout + r.stderr).splitlines()[-40:])
    assert r.returncode == 0, f"suite {suite} failed (rc {r.returncode}):\n{tail}"
    assert f"suite {suite}: ok" in r.stdout, tail


@pytest.mark.parametrize("suite", SUITES)
def test_cpu_model_suite(model_env, suite):
    _run_suite(model_env, suite)


@pytest.mark.parametrize("suite", ["lde", "plans"])
def test_cpu_model_reverse_thread_order(model_env, suite):
    """Workgroup threads run last to first between barriers: LDS data passed
    between threads without a barrier, or overwritten before another thread
    has read it, would give results that differ from the oracle.  With
    SFNTT_MODEL_FULL=1 every suite also runs in a shuffled order."""
    _run_suite(dict(model_env, SFNTT_MODEL_ORDER="reverse"), suite)


@pytest.mark.skipif(os.environ.get("SFNTT_MODEL_FULL") != "1", reason="full model sweep only")
@pytest.mark.parametrize("suite", ["ntt", "lde", "plans", "noalloc", "schedules"])
def test_cpu_model_shuffled_thread_order(model_env, suite):
    _run_suite(dict(model_env, SFNTT_MODEL_ORDER="shuffle:7"), suite)


@pytest.mark.parametrize("suite", ["plans", "server"])
def test_cpu_model_few_compute_units(model_env, suite):
    """Four compute units: the many-job launches, whose grid is sized to the
    device, run several tiles per workgroup."""
    _run_suite(dict(model_env, SFNTT_MODEL_CUS="4"), suite)
