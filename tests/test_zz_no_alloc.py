"""vis_fourier_volume_rendering_noalloc against vis_fourier_volume_rendering (plain form).

The allocation-free render keeps its packets, split-K partials and FP64 accumulator in the
library's static 64 MiB device array instead of buffers it allocates.  Its spectrum and image
must equal the allocating render's: bit for bit when the scratch holds the render's usual
number of split-K slots, else within FP64 reassociation of the partial sums (rel-L2 <= 1e-6).
Two calls in a row, the second on a smaller dataset, check that nothing of the first call's
scratch contents reaches the second.  Renders beyond the scratch, other forms and groups not
registered are refused with their error codes, and the image is left as it was.
"""
import numpy as np
import pytest

import oracle
from libmeshlessvis_amd import abi, host
from libmeshlessvis_amd import workloads as W

pytestmark = pytest.mark.gpu


def noalloc(ds, cfg):
    abi.lib().vis_fourier_volume_rendering_noalloc(ds.ptr, cfg.p)
    abi.check("vis_fourier_volume_rendering_noalloc")


def cloud(rng, n, basis, radii="var"):
    pos = rng.uniform(-0.5, 0.5, (n, 3)).astype(np.float32)
    w = rng.uniform(-1.0, 1.0, n).astype(np.float32)
    h = {"var": rng.uniform(0.005, 0.05, n).astype(np.float32), "const": np.full(n, 0.02, np.float32),
         "none": None}[radii]
    return (pos, w, h, basis)


def both(cfg, ds):
    """(spectrum, image) of the allocating render, then of the allocation-free one."""
    cfg.render(ds)
    ref = (cfg.spectrum(), cfg.image())
    noalloc(ds, cfg)
    return ref, (cfg.spectrum(), cfg.image())


def compare(ref, got, what):
    same = np.array_equal(ref[0], got[0]) and np.array_equal(ref[1], got[1])
    e_f, e_i = oracle.rel_l2(got[0], ref[0]), oracle.rel_l2(got[1], ref[1])
    print(f"{what}: identical={same} spectrum rel-L2 {e_f:.2e} image rel-L2 {e_i:.2e}", flush=True)
    assert np.isfinite(got[1]).all()
    assert e_f <= 1e-6 and e_i <= 1e-6, what
    return same


def test_noalloc_matches_render_two_calls_second_smaller():
    print("[case] noalloc: 20000 particles then 3000 (two groups)", flush=True)
    rng = np.random.default_rng(41)
    vw = W.random_view(rng)
    big = host.Dataset([cloud(rng, 20000, "WENDLAND_D3_C2")])
    small = host.Dataset([cloud(rng, 3000, "GAUSSIAN", "const"), cloud(rng, 700, "SPH", "none")])
    with host.Config(cutoff=(96, 80), samples=(256, 192), step=(0.5, 0.5), u=vw.u, v=vw.v, form="plain",
                     phi_mode="analytic") as cfg:
        cfg.register(big)
        cfg.register(small)
        ref_big, got_big = both(cfg, big)
        assert compare(ref_big, got_big, "20000 particles")
        # Second call, smaller input, straight after the first: nothing left over from it.
        noalloc(small, cfg)
        got_small = (cfg.spectrum(), cfg.image())
        cfg.render(small)
        ref_small = (cfg.spectrum(), cfg.image())
        assert compare(ref_small, got_small, "3000 + 700 particles after the 20000-particle call")
        cfg.unregister(big)
        cfg.unregister(small)


def test_noalloc_large_grid_fewer_slots():
    """262,144 particles at 256^2: the scratch holds fewer split-K slots than the render's 128."""
    print("[case] noalloc: 262144 particles, cutoff 128", flush=True)
    rng = np.random.default_rng(42)
    vw = W.random_view(rng)
    ds = host.Dataset([cloud(rng, 262_144, "GAUSSIAN")])
    with host.Config(cutoff=(128, 128), samples=(256, 256), step=(0.5, 0.5), u=vw.u, v=vw.v, form="plain",
                     phi_mode="analytic") as cfg:
        cfg.register(ds)
        for rep in range(2):
            ref, got = both(cfg, ds)
            compare(ref, got, f"262144 particles, call {rep}")
        cfg.unregister(ds)


def test_noalloc_refusals():
    print("[case] noalloc refusals", flush=True)
    rng = np.random.default_rng(43)
    L = abi.lib()
    small = host.Dataset([cloud(rng, 1000, "WENDLAND_D3_C2")])
    with host.Config(cutoff=(32, 32), samples=(64, 64), step=(0.5, 0.5), form="plain") as cfg:
        # Not registered.
        L.vis_fourier_volume_rendering_noalloc(small.ptr, cfg.p)
        assert L.vis_get_last_error_code() == 6  # VIS_ERROR_NOT_REGISTERED
        L.vis_clear_last_error()
        cfg.register(small)
        noalloc(small, cfg)
        before = cfg.image()
        # Beyond the scratch: 4,194,304 particles need 64 MiB of packets alone.
        n = 4_194_304
        huge = host.Dataset([(np.zeros((n, 3), np.float32), np.ones(n, np.float32), None, "GAUSSIAN")])
        cfg.register(huge)
        L.vis_fourier_volume_rendering_noalloc(huge.ptr, cfg.p)
        assert L.vis_get_last_error_code() == 1  # VIS_ERROR_INVALID_VALUE
        assert b"64 MiB" in L.vis_get_last_error()
        L.vis_clear_last_error()
        assert np.array_equal(cfg.image(), before)
        cfg.unregister(huge)
        # Another form.
        cfg.form = "mfma_split_f16"
        L.vis_fourier_volume_rendering_noalloc(small.ptr, cfg.p)
        assert L.vis_get_last_error_code() == 2  # VIS_ERROR_INVALID_CONFIG
        L.vis_clear_last_error()
        cfg.form = "plain"
        noalloc(small, cfg)
        assert np.array_equal(cfg.image(), before)
        cfg.unregister(small)
