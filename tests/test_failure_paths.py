"""Failure paths and step tuning without a GPU.

- The step search of edgeMC::optimal_step (em_bisection.cu:49-153) over synthetic acceptance
  curves through the C ABI (le_step_search makes no GPU call of its own): it converges to
  target +- tol on monotone curves from either side, and returns LE_ERR_BISECTION where the
  reference loops forever or exits (a target the curve never crosses; a curve that jumps over
  the tolerance band), passes a measurement error through, and rejects impossible acceptances.
- Capacity and argument errors of le_engine_create come back as codes before any device work.
- The drop-in C++ shells end the process (exit 1 with the reference's message) or throw
  std::runtime_error exactly where the reference does: capacity beyond sys_params.h
  (em_spectrum.cu:235-252), dsf_compute(dL = 0) (em_dsf.cu:241-244), occupation dL beyond
  ansatzMaxAngularMomentum (gf_occupation_numbers.cu:316-319), a missing or malformed TSV
  (dynamics.cpp:164-166, 182-184) and a first sector other than dL = 0 (dynamics.cpp:211-214);
  a negative TSV index, undefined in the reference, is a parsing error here.
- The host code (oracle, basis, step search, TSV reader) under ASan + UBSan."""
import math
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CPP = os.path.join(HERE, "cpp")


def _curve(scale, floor=0.0):
    # acceptance falls from 1 to 0 as the step grows
    return lambda s: max(floor, 1.0 / (1.0 + (s / scale) ** 2))


@pytest.mark.parametrize("target", [0.2, 0.5, 0.8])
@pytest.mark.parametrize("step0", [1e-3, 0.8, 50.0])
def test_step_search_converges(target, step0):
    from laughlinedge_amd.engine import step_search
    f = _curve(1.7)
    step, n = step_search(f, step0, target, 0.01)
    assert abs(f(step) - target) < 0.01
    if abs(f(step0) - target) >= 0.01:
        assert n > 1


def test_step_search_keeps_a_good_start():
    from laughlinedge_amd.engine import step_search
    f = _curve(1.0)
    step, n = step_search(f, 1.0, 0.5, 0.01)
    assert step == 1.0 and n == 1


def test_step_search_error_codes():
    from laughlinedge_amd.engine import step_search, LEError, LE_ERR_BISECTION, LE_ERR_ARG
    # never reaches the target: the doubling search gives up after 41 measurements
    with pytest.raises(LEError) as ei:
        step_search(_curve(1.0, floor=0.7), 0.3, 0.5, 0.01)
    assert ei.value.code == LE_ERR_BISECTION and "bracket" in str(ei.value)
    assert ei.value.measurements == 42
    # a jump across the tolerance band: the bracket shrinks but no guess is within tol
    with pytest.raises(LEError) as ei:
        step_search(lambda s: 0.7 if s < math.pi / 3 else 0.3, 0.2, 0.5, 0.01)
    assert ei.value.code == LE_ERR_BISECTION and "did not converge" in str(ei.value)
    # a failed measurement comes back unchanged
    with pytest.raises(LEError) as ei:
        step_search(lambda s: (3, 0.0), 0.5, 0.5, 0.01)
    assert ei.value.code == 3
    with pytest.raises(LEError) as ei:
        step_search(lambda s: 1.5, 0.5, 0.5, 0.01)
    assert ei.value.code == LE_ERR_ARG
    with pytest.raises(LEError) as ei:
        step_search(_curve(1.0), -1.0, 0.5, 0.01)
    assert ei.value.code == LE_ERR_ARG


@pytest.mark.parametrize("N,sectors,msg", [(30, [25], "degree 25 exceeds 24"),
                                           (30, [21], "sector larger than 640"),
                                           (10, [41], "dL must be in [0, 40]")])
def test_engine_capacity_codes(N, sectors, msg):
    from laughlinedge_amd import Engine, LEError
    from laughlinedge_amd.engine import LE_ERR_CAPACITY, LE_ERR_ARG
    with pytest.raises(LEError) as ei:
        Engine(N, 3, sectors, walkers=64)
    assert msg in str(ei.value)
    assert ei.value.code == (LE_ERR_ARG if "must be" in msg else LE_ERR_CAPACITY)


def _driver():
    exe = os.path.join(CPP, "failure_driver")
    if not os.path.exists(exe):
        pytest.skip("tests/cpp/failure_driver is not built (python __graft_entry__.py)")
    return exe


@pytest.mark.parametrize("mode,rc,text", [
    ("capacity", 1, "the maximal single-coordinate degree is 13"),
    ("dsf0", 1, "Select dL>0"),
    ("occupation_dl", 1, "ansatzMaxAngularMomentum is smaller than dL"),
    ("tsv_missing", 3, "Cannot open file"),
    ("tsv_malformed", 3, "Parsing error"),
    ("tsv_negative", 3, "negative index"),
    ("sectors0", 1, "The Laughlin state (assumed to be the initial state) is not included"),
])
def test_shell_failure_paths(tmp_path, mode, rc, text):
    exe = _driver()
    cwd = tmp_path / "build"
    cwd.mkdir()
    env = dict(os.environ, HIP_VISIBLE_DEVICES="")
    p = subprocess.run([exe, mode], cwd=cwd, capture_output=True, text=True, timeout=60, env=env)
    assert p.returncode == rc, (p.returncode, p.stdout, p.stderr)
    assert text in p.stdout, p.stdout


def test_host_code_under_sanitizers(tmp_path):
    exe = str(tmp_path / "sanitized_host")
    b = subprocess.run(["bash", os.path.join(CPP, "build_sanitized.sh"), exe], capture_output=True, text=True,
                       timeout=300)
    if b.returncode != 0 and "cannot find -lasan" in (b.stderr + b.stdout):
        pytest.skip("no libasan in this toolchain")
    assert b.returncode == 0, b.stderr
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")
    p = subprocess.run([exe], cwd=tmp_path, capture_output=True, text=True, timeout=300, env=env)
    assert p.returncode == 0 and "SANITIZED OK" in p.stdout, (p.stdout, p.stderr[-3000:])
