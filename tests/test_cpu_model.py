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
FULL = os.environ.get("SFNTT_MODEL_FULL") == "1"

# environment overrides of each test function's runs
VARIANTS = {
    "test_cpu_model_suite": {},
    "test_cpu_model_reverse_thread_order": {"SFNTT_MODEL_ORDER": "reverse"},
    "test_cpu_model_shuffled_thread_order": {"SFNTT_MODEL_ORDER": "shuffle:7"},
    "test_cpu_model_few_compute_units": {"SFNTT_MODEL_CUS": "4"},
}


@pytest.fixture(scope="module")
def model_runs(model_env, request, tmp_path_factory):
    """Starts every model run the session selected from this module at once:
    the runs are independent processes, so the module takes about as long as
    its longest run instead of their sum.  Each test then waits for its own."""
    logs = tmp_path_factory.mktemp("cpu_model")
    runs = {}
    for item in request.session.items:
        name = getattr(item, "originalname", item.name)
        if item.module.__name__ != __name__ or name not in VARIANTS or not hasattr(item, "callspec"):
            continue
        if item.get_closest_marker("skipif") and any(m.args and m.args[0] for m in item.iter_markers("skipif")):
            continue
        suite = item.callspec.params["suite"]
        key = (name, suite)
        if key in runs:
            continue
        log = logs / f"{name}_{suite}.log"
        fh = open(log, "w")
        proc = subprocess.Popen([sys.executable, os.path.join(MODEL, "run_model.py"), suite],
                                env=dict(model_env, **VARIANTS[name]), stdout=fh, stderr=subprocess.STDOUT,
                                text=True)
        runs[key] = (proc, log, fh)
    yield runs
    for proc, _, fh in runs.values():
        if proc.poll() is None:
            proc.kill()
        proc.wait()
        fh.close()


def _check(runs, name, suite):
    proc, log, fh = runs[(name, suite)]
    rc = proc.wait(timeout=1800)
    fh.flush()
    out = open(log).read()
    tail = "\n".join(out.splitlines()[-40:])
    assert rc == 0, f"suite {suite} failed (rc {rc}):\n{tail}"
    assert f"suite {suite}: ok" in out, tail


@pytest.mark.parametrize("suite", SUITES)
def test_cpu_model_suite(model_runs, suite):
    _check(model_runs, "test_cpu_model_suite", suite)


@pytest.mark.parametrize("suite", ["lde", "plans"])
def test_cpu_model_reverse_thread_order(model_runs, suite):
    """Workgroup threads run last to first between barriers: LDS data passed
    between threads without a barrier, or overwritten before another thread
    has read it, would give results that differ from the oracle.  With
    SFNTT_MODEL_FULL=1 every suite also runs in a shuffled order."""
    _check(model_runs, "test_cpu_model_reverse_thread_order", suite)


@pytest.mark.skipif(not FULL, reason="full model sweep only")
@pytest.mark.parametrize("suite", ["ntt", "lde", "plans", "noalloc", "schedules"])
def test_cpu_model_shuffled_thread_order(model_runs, suite):
    _check(model_runs, "test_cpu_model_shuffled_thread_order", suite)


@pytest.mark.parametrize("suite", ["plans", "server"])
def test_cpu_model_few_compute_units(model_runs, suite):
    """Four compute units: the many-job launches, whose grid is sized to the
    device, run several tiles per workgroup."""
    _check(model_runs, "test_cpu_model_few_compute_units", suite)
