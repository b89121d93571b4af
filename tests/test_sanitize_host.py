"""AddressSanitizer / UndefinedBehaviorSanitizer runs of the host-only input code (no GPU).

The .sph reader (meshless.cpp:86-145 of the reference, csrc/meshless.cpp here) and the batch
program's --from-raw conversion read untrusted files.  Both are built here with g++
-fsanitize=address,undefined -fno-sanitize-recover=all and fed well-formed and malformed inputs:
truncated files, negative and huge counts (headers claiming more terms than the file holds),
unknown bases, non-numeric tokens, short raw images.  Every run must end normally, report a
malformed file through the error code (VIS_ERROR_INVALID_VALUE / VIS_ERROR_OUT_OF_MEMORY) and
leave no datasets behind, with no sanitizer report.
"""
import os
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from libmeshlessvis_amd import build as B

ROOT = Path(__file__).resolve().parent.parent
CSRC = ROOT / "libmeshlessvis_amd" / "csrc"
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]
HIP_HOST = ["-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", f"-I{ROOT / 'include'}"]


def gxx():
    g = shutil.which("g++")
    if not g or not Path("/opt/rocm/include/hip/hip_vector_types.h").exists():
        pytest.skip("g++ or the HIP headers are missing")
    return g


@pytest.fixture(scope="module")
def sph_driver(tmp_path_factory):
    out = tmp_path_factory.mktemp("san") / "sph_driver"
    cmd = [gxx(), "-std=c++17", *SAN, *HIP_HOST, str(Path(__file__).parent / "sanitize" / "sph_driver.cpp"),
           str(CSRC / "meshless.cpp"), str(CSRC / "errors.cpp"), "-o", str(out)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    return out


def run_driver(exe, path, out=None):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    args = [str(exe), str(path)] + ([str(out)] if out else [])
    res = subprocess.run(args, capture_output=True, text=True, timeout=120, env=env)
    assert "Sanitizer" not in res.stderr and "runtime error" not in res.stderr, res.stderr
    assert res.returncode == 0, (res.returncode, res.stderr)
    code, n, terms = map(int, res.stdout.split())
    return code, n, terms


GOOD = """2
2
3 sph unused 0
2 wendland_d3_c2 unused 0
0 0 0
1 1 1
0.5 0.25 -1
2 2 2
-3 0 1e-3
1
2
3
4
-5
1
0.1
0.2
0.3
0.05
0.06
1
1 gaussian unused 0
0.5 0.5 0.5
7
0
"""

MALFORMED = {
    "empty": "",
    "word": "abc",
    "negative_datasets": "-1\n",
    "huge_datasets": "2000000000\n",
    "missing_groups": "1\n",
    "negative_groups": "1\n-3\n",
    "huge_groups": "1\n2000000000\n",
    "huge_terms": "1\n1\n2000000000 sph unused 0\n",
    "terms_beyond_file": "1\n1\n100 sph unused 0\n0 0 0\n1\n0\n",
    "negative_terms": "1\n1\n-5 sph unused 0\n",
    "unknown_basis": "1\n1\n1 cubic unused 0\n0 0 0\n1\n0\n",
    "truncated_positions": "1\n1\n2 sph unused 0\n0 0 0\n1 1\n",
    "missing_weights": "1\n1\n2 sph unused 0\n0 0 0\n1 1 1\n",
    "non_numeric_weight": "1\n1\n1 sph unused 0\n0 0 0\nx\n0\n",
    "missing_radii_flag": "1\n1\n1 sph unused 0\n0 0 0\n1\n",
    "truncated_radii": "1\n1\n2 sph unused 0\n0 0 0\n1 1 1\n1\n2\n1\n0.1\n",
    "second_dataset_truncated": GOOD.replace("1\n1 gaussian unused 0\n0.5 0.5 0.5\n7\n0\n", "1\n1 gaussian unused 0\n0.5"),
}


def test_sanitized_reader_accepts_good_file(sph_driver, tmp_path):
    src = tmp_path / "good.sph"
    src.write_text(GOOD)
    code, n, terms = run_driver(sph_driver, src, tmp_path / "out.sph")
    assert (code, n, terms) == (0, 2, 6)
    # The written file (shifted positions) reads back with the same counts.
    assert run_driver(sph_driver, tmp_path / "out.sph") == (0, 2, 6)


@pytest.mark.parametrize("name", sorted(MALFORMED))
def test_sanitized_reader_rejects_malformed(sph_driver, tmp_path, name):
    src = tmp_path / f"{name}.sph"
    src.write_text(MALFORMED[name])
    code, n, terms = run_driver(sph_driver, src)
    if name == "empty":
        assert code == 1  # no dataset count
    else:
        assert code in (1, 4), (name, code)
    assert n == 0 and terms == 0


def test_sanitized_reader_missing_file(sph_driver, tmp_path):
    code, n, _ = run_driver(sph_driver, tmp_path / "absent.sph")
    assert code == 1 and n == 0


@pytest.fixture(scope="module")
def sanitized_cli(tmp_path_factory):
    if not B.LIB.exists():
        pytest.fail(f"{B.LIB} is not built (python -m libmeshlessvis_amd.build)")
    out = tmp_path_factory.mktemp("san_cli") / "vis_noninteractive"
    cmd = [gxx(), "-std=c++17", *SAN, *HIP_HOST, str(CSRC / "vis_noninteractive.cpp"), f"-L{B.PKG}",
           "-lmeshless_vis", f"-Wl,-rpath,{B.PKG}", "-o", str(out)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    return out


def run_cli(exe, args, cwd):
    # The HIP runtime the library links is not instrumented: its allocations are not checked for leaks.
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0", UBSAN_OPTIONS="print_stacktrace=1")
    res = subprocess.run([str(exe), *map(str, args)], capture_output=True, text=True, timeout=120, cwd=cwd, env=env)
    assert "Sanitizer" not in res.stderr and "runtime error" not in res.stderr, res.stderr
    return res


def test_sanitized_cli_from_raw(sanitized_cli, tmp_path):
    img = np.linspace(-1.0, 3.0, 6 * 5, dtype=np.float32)
    raw = tmp_path / "img.f32"
    img.tofile(raw)
    res = run_cli(sanitized_cli, ["--from-raw", raw, 6, 5, "--depth", "16"], tmp_path)
    assert res.returncode == 0, res.stderr
    assert (tmp_path / "image_000.pgm").exists()
    # A file shorter than the image it claims, and an image far larger than the file.
    for nx, ny in ((6, 6), (100000, 100000)):
        res = run_cli(sanitized_cli, ["--from-raw", raw, nx, ny], tmp_path)
        assert res.returncode == 1 and "cannot read" in res.stderr
    for bad in (["--from-raw", raw, 0, 5], ["--from-raw", raw, -2, 5], ["--from-raw", raw], ["--depth", "12"],
                ["--nonsense"]):
        assert run_cli(sanitized_cli, bad, tmp_path).returncode == 2


def test_sanitized_cli_malformed_sph(sanitized_cli, tmp_path):
    for name in ("huge_terms", "unknown_basis", "truncated_radii"):
        src = tmp_path / f"{name}.sph"
        src.write_text(MALFORMED[name])
        res = run_cli(sanitized_cli, [src, tmp_path / "x_"], tmp_path)
        assert res.returncode == 1 and "load_meshless_datasets_from_file failed" in res.stderr, res.stderr
