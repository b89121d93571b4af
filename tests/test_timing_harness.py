"""The timing harness vis_timing_test (vis_timing_test/main.cpp:24-93) without a GPU: its
usage text for another argument count (exit status 0, as the reference's) and the refusal of
an unknown form or a negative count before anything is rendered.  Rendering runs in
test_gpu_timing_harness.py."""
import subprocess

import pytest

from libmeshlessvis_amd import build as B

PROG = B.PROGRAMS["vis_timing_test"]


def run(args):
    if not PROG.exists():
        pytest.fail(f"{PROG} is not built (python -m libmeshlessvis_amd.build)")
    return subprocess.run([str(PROG), *map(str, args)], capture_output=True, text=True, timeout=60)


@pytest.mark.parametrize("args", [[], ["1"], ["1", "2", "3", "4", "5", "6"], list("123456789") + ["x"]])
def test_usage_for_other_argument_counts(args):
    r = run(args)
    assert r.returncode == 0
    assert "vis_timing_test number_of_runs number_of_terms cutoff_frequency block_length" in r.stdout
    assert "a good value for block length is 256" in r.stdout


def test_unknown_form_is_refused():
    r = run([1, 16, 8, 64, 64, 0, 1, "fp8"])
    assert r.returncode == 1
    assert "unknown form fp8" in r.stderr


def test_negative_count_is_refused():
    r = run([1, -5, 8, 64, 64, 0, 1])
    assert r.returncode == 1
    assert "negative count" in r.stderr
