"""The reference's timing harness on the GPU (vis_timing_test/main.cpp:24-93).

* The program itself: every form, factor 0 (radius 1e-5) and 1 (no radii), the reference's
  progress lines and the added elapsed-time line, exit status 0.
* Its dataset through the library: every particle at the origin with weight 1, SPH basis,
  radius 1e-5 or none, step 1, the axis-aligned view.  All phases are 0, so the spectrum is
  N * Phi(|k| r0) (reference Phi formulas), checked against the FP64 oracle (relative L2
  <= 1e-5, split BF16 3e-5, as the other parity tests) and against that closed form.
* Reference parity on the same dataset: the FP32 restatement of fourier_transform_cpu.cpp
  (e_ref against FP64), the GPU (e_build) and the GPU against the restatement (e_par), with
  e_par <= e_ref + e_build + 1e-6 as in the full-size parity tests.
"""
import re
import subprocess

import numpy as np
import pytest

import oracle
from libmeshlessvis_amd import build as B
from libmeshlessvis_amd import host
from libmeshlessvis_amd import workloads as W

pytestmark = pytest.mark.gpu

PROG = B.PROGRAMS["vis_timing_test"]


@pytest.mark.parametrize("form", ["plain", "mfma_split_f16", "mfma_split_bf16", "mfma_f16"])
@pytest.mark.parametrize("factor", [0, 1])
def test_program_runs(form, factor):
    print(f"[case] vis_timing_test {form} factor={factor}", flush=True)
    args = [3, 20000, 32, 64, 64, factor, 2, form]
    r = subprocess.run([str(PROG), *map(str, args)], capture_output=True, text=True, timeout=120)
    print(r.stdout[-400:], r.stderr[-400:])
    assert r.returncode == 0, r.stderr
    for line in ("number of terms: 20000", "Valid Configuration", "loading data", "finished running tests"):
        assert line in r.stdout
    m = re.search(r"elapsed: ([0-9.e+-]+) s for 3 renders", r.stdout)
    assert m and float(m.group(1)) > 0.0


def test_invalid_configuration_is_reported():
    # 2 * cutoff^2 = 2 * 9 * 9 is not a multiple of block length 64: the reference warns and
    # renders anyway; the library refuses the render, and the program ends with status 1.
    r = subprocess.run([str(PROG), "1", "100", "9", "64", "64", "0", "1"], capture_output=True, text=True, timeout=120)
    assert "Warning: Invalid Configuration" in r.stdout
    assert r.returncode == 1 and "vis_fourier_volume_rendering" in r.stderr


@pytest.mark.parametrize("form,tol", [("plain", 1e-5), ("mfma_split_f16", 1e-5), ("mfma_split_bf16", 3e-5)])
@pytest.mark.parametrize("factor", [0, 1])
def test_dataset_parity(form, tol, factor):
    print(f"[case] harness dataset {form} factor={factor}", flush=True)
    wl = W.timing_test(4096, cutoff=32, samples=64, factor=factor)
    vw = wl.views[0]
    ds = host.Dataset([wl.group()])
    with host.Config(cutoff=wl.cutoff, samples=wl.samples, step=wl.step, u=vw.u, v=vw.v, form=form,
                     phi_mode=wl.phi_mode, block_length=64) as cfg:
        cfg.register(ds)
        cfg.render(ds)
        F = cfg.spectrum()
        img = cfg.image()
        cfg.unregister(ds)
    F64 = oracle.spectrum64([wl.group()], wl.cutoff, wl.step, vw.u, vw.v, wl.phi_mode)
    e = oracle.rel_l2(F, F64)
    # The closed form: N * Phi(|k| r0) at every stored frequency (x0 > cx is x = x0 - 2cx).
    cx, cy = wl.cutoff
    x0 = np.arange(2 * cx)
    x = np.where(x0 > cx, x0 - 2 * cx, x0).astype(np.float64)
    y = np.arange(cy, dtype=np.float64)
    r = np.hypot(x[None, :], y[:, None])
    r0 = 1.0 if factor else 1e-5
    phi = oracle.phi64("SPH", wl.phi_mode, r * r0).reshape(r.shape)
    closed = (wl.n * phi).astype(np.complex128)
    print(f"rel-L2 {e:.2e} vs FP64, {oracle.rel_l2(F64, closed):.2e} oracle vs closed form")
    assert oracle.rel_l2(F64.reshape(closed.shape), closed) <= 1e-12
    assert e <= tol
    img64 = oracle.c2r64(oracle.arrange64(F64, *wl.cutoff, *wl.samples), *wl.samples)
    assert oracle.rel_l2(img, img64) <= tol


@pytest.mark.parametrize("form", ["plain", "mfma_split_f16"])
@pytest.mark.parametrize("factor", [0, 1])
def test_dataset_reference_parity(form, factor):
    print(f"[case] harness reference parity {form} factor={factor}", flush=True)
    wl = W.timing_test(4096, cutoff=32, samples=64, factor=factor)
    vw = wl.views[0]
    ds = host.Dataset([wl.group()])
    with host.Config(cutoff=wl.cutoff, samples=wl.samples, step=wl.step, u=vw.u, v=vw.v, form=form,
                     phi_mode=wl.phi_mode, block_length=64) as cfg:
        cfg.register(ds)
        cfg.render(ds)
        F = cfg.spectrum()
        cfg.unregister(ds)
    F64 = oracle.spectrum64([wl.group()], wl.cutoff, wl.step, vw.u, vw.v, wl.phi_mode)
    Fport = oracle.port_spectrum32([wl.group()], wl.cutoff, wl.step, vw.u, vw.v, block_length=64,
                                   phi_mode=wl.phi_mode)
    e_ref, e_build, e_par = oracle.rel_l2(Fport, F64), oracle.rel_l2(F, F64), oracle.rel_l2(F, Fport)
    print(f"e_ref {e_ref:.2e} e_build {e_build:.2e} e_par {e_par:.2e}")
    assert e_par <= e_ref + e_build + 1e-6
