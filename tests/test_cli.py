"""The batch renderer vis_noninteractive (vis_noninteractive/main.cpp:37-93): option
parsing, .sph loading errors and the image normalisation / flip / PGM encoding of
--from-raw, all without a GPU.  The rendering path is in test_gpu_cli.py."""
import subprocess

import numpy as np
import pytest

from libmeshlessvis_amd import build as B

CLI = B.CLI


def run(args, cwd):
    if not CLI.exists():
        pytest.fail(f"{CLI} is not built (python -m libmeshlessvis_amd.build)")
    return subprocess.run([str(CLI), *map(str, args)], cwd=cwd, capture_output=True, text=True, timeout=60)


def read_pgm(path):
    data = path.read_bytes()
    head = data.split(b"\n", 3)
    assert head[0] == b"P5"
    w, h = map(int, head[1].split())
    maxval = int(head[2])
    pix = np.frombuffer(head[3], dtype=">u2" if maxval > 255 else np.uint8)
    assert pix.size == w * h
    return pix.reshape(h, w).astype(np.int64), maxval


def expected(img_memory, width, height, maxval):
    """The reference's normalisation (main.cpp:70-84) and ImageMagick's flip (:28-34),
    restated: memory read as `height` rows of `width`, max over the raw values,
    negatives to 0, divided by the max, rows bottom to top."""
    a = img_memory.reshape(height, width).astype(np.float32)
    m = a.max()
    v = np.where(a < 0, np.float32(0), a)
    v = v / m if m > 0 else np.zeros_like(v)
    return np.rint(np.minimum(v, 1).astype(np.float64) * maxval).astype(np.int64)[::-1]


@pytest.mark.parametrize("depth", [8, 16])
def test_from_raw_matches_the_reference_normalisation(tmp_path, depth):
    rng = np.random.default_rng(depth)
    nx, ny = 7, 5  # width 7, height 5: a non-square image exercises the row order
    img = rng.normal(0.3, 1.0, nx * ny).astype(np.float32)
    img[3] = 4.0  # the maximum, off the first row
    (tmp_path / "img.f32").write_bytes(img.tobytes())
    r = run(["--from-raw", "img.f32", nx, ny, "--depth", depth, "x", "frame_"], tmp_path)
    assert r.returncode == 0, r.stderr
    assert r.stdout.split() == ["frame_000.pgm"]
    got, maxval = read_pgm(tmp_path / "frame_000.pgm")
    assert maxval == (255 if depth == 8 else 65535)
    want = expected(img, nx, ny, maxval)
    # lround (half away from zero) vs rint (half to even) differ only on exact halves.
    assert np.abs(got - want).max() <= 1
    assert (got == want).mean() > 0.9
    assert got[-1 - 0, 3] == maxval  # the maximum lands in the last output row (flip)


def test_from_raw_nonpositive_image_is_black(tmp_path):
    img = -np.abs(np.arange(12, dtype=np.float32))
    (tmp_path / "neg.f32").write_bytes(img.tobytes())
    r = run(["--from-raw", "neg.f32", 4, 3, "x", "n_"], tmp_path)
    assert r.returncode == 0, r.stderr
    got, _ = read_pgm(tmp_path / "n_000.pgm")
    assert not got.any()


def test_errors(tmp_path):
    assert run(["--bogus"], tmp_path).returncode == 2
    assert run(["--depth", "12"], tmp_path).returncode == 2
    assert run(["--from-raw", "missing.f32", 4, 4], tmp_path).returncode == 1
    (tmp_path / "short.f32").write_bytes(np.zeros(3, np.float32).tobytes())
    assert run(["--from-raw", "short.f32", 4, 4], tmp_path).returncode == 1
    r = run(["missing.sph"], tmp_path)
    assert r.returncode == 1 and "cannot open" in r.stderr
    (tmp_path / "bad.sph").write_text("1\n1\n3 nosuchbasis unused 0\n")
    r = run(["bad.sph"], tmp_path)
    assert r.returncode == 1 and "unknown basis" in r.stderr
    (tmp_path / "empty.sph").write_text("0\n")
    r = run(["empty.sph"], tmp_path)
    assert r.returncode == 0 and r.stdout == ""
