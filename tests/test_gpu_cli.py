"""The batch renderer vis_noninteractive on the GPU: a multi-dataset .sph file in,
one PGM per dataset out (vis_noninteractive/main.cpp:37-93).  Each image equals the
library's own render of the shifted dataset put through the reference's
normalisation (within one LSB: rounding of exact halves), and the FP64 pipeline of
the oracle within 1 LSB (8 bits) or 3 LSB (16 bits)."""
import subprocess

import numpy as np
import pytest

import oracle
from libmeshlessvis_amd import host

from test_cli import CLI, expected, read_pgm

pytestmark = pytest.mark.gpu

CUTOFF, SAMPLES, STEP = (32, 32), (64, 64), (1.0, 1.0)
U, V = (1.0, 0.0, 0.0), (0.0, 1.0, 0.0)  # vis_config_get_default's axes


def write_sph(path, datasets):
    """The text format of meshless.cpp:86-187: counts, one 'N basis unused 0' line per
    group, positions, weights, a radii flag per dataset, radii."""
    out = [str(len(datasets))]
    for groups in datasets:
        out.append(str(len(groups)))
        out += [f"{len(g[1])} {g[3]} unused 0" for g in groups]
        for g in groups:
            out += [" ".join(f"{c:.9g}" for c in p) for p in g[0]]
        for g in groups:
            out += [f"{w:.9g}" for w in g[1]]
        has = groups[0][2] is not None
        out.append("1" if has else "0")
        if has:
            for g in groups:
                out += [f"{h:.9g}" for h in g[2]]
    path.write_text("\n".join(out) + "\n")


def make_group(rng, n, basis):
    pos = rng.uniform(4.6, 5.4, (n, 3)).astype(np.float32)
    w = rng.uniform(0.1, 1.0, n).astype(np.float32)
    h = rng.uniform(0.02, 0.08, n).astype(np.float32)
    return pos, w, h, basis


def library_image(groups, form, phi_mode):
    shifted = [((g[0] + np.float32(-5.0)).astype(np.float32), g[1], g[2], g[3].upper()) for g in groups]
    ds = host.Dataset(shifted)
    with host.Config(cutoff=CUTOFF, samples=SAMPLES, step=STEP, u=U, v=V, form=form, phi_mode=phi_mode) as cfg:
        cfg.register(ds)
        try:
            cfg.render(ds)
            img = cfg.image()
        finally:
            cfg.unregister(ds)
    return img, shifted


@pytest.mark.parametrize(
    "case",
    [
        # defaults of the reference program: plain form, reference Phi, asynchronous uploads
        dict(args=[], form="plain", phi="reference", depth=8,
             datasets=[["wendland_d3_c2"], ["sph", "wendland_d3_c2"]]),
        dict(args=["--form", "mfma_split_f16", "--phi", "analytic", "--sync", "--depth", "16"],
             form="mfma_split_f16", phi="analytic", depth=16, datasets=[["gaussian"], ["wendland_d3_c2", "sph"]]),
    ],
)
def test_cli_renders_every_dataset(tmp_path, case):
    rng = np.random.default_rng(11)
    datasets = [[make_group(rng, n, b) for n, b in zip((3000, 1700), bases)] for bases in case["datasets"]]
    write_sph(tmp_path / "clouds.sph", datasets)
    print(f"[case] cli {case['form']} {case['phi']} depth {case['depth']}", flush=True)
    r = subprocess.run(
        [str(CLI), "clouds.sph", "img_", "--samples", str(SAMPLES[0]), "--cutoff", str(CUTOFF[0]), "--step",
         str(STEP[0]), *case["args"]], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout.split() == ["img_000.pgm", "img_001.pgm"]
    for k, groups in enumerate(datasets):
        got, maxval = read_pgm(tmp_path / f"img_{k:03d}.pgm")
        assert maxval == (2 ** case["depth"] - 1)
        img, shifted = library_image(groups, case["form"], case["phi"])
        want = expected(img.reshape(-1), SAMPLES[0], SAMPLES[1], maxval)
        assert np.abs(got - want).max() <= 1
        img64 = oracle.image64(shifted, CUTOFF, SAMPLES, STEP, U, V, case["phi"])
        want64 = expected(img64.reshape(-1), SAMPLES[0], SAMPLES[1], maxval)
        d = np.abs(got - want64).max()
        print(f"dataset {k}: max |pgm - oracle pgm| = {d} LSB")
        assert d <= (1 if case["depth"] == 8 else 3)
