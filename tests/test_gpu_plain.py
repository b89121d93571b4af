"""GPU parity of the plain direct sum + C2R against the CPU oracle (through the C ABI).

Tolerances (relative L2):
  spectrum vs FP64 oracle of the same Phi formula       <= 1e-5  (the accuracy target)
  spectrum vs the FP32 restatement of the reference CPU <= e_ref + e_build + 1e-6
  image vs the FP64 pipeline (oracle spectrum, arrange, C2R) <= 1e-5
"""
import numpy as np
import pytest

import oracle
from libmeshlessvis_amd import host
from libmeshlessvis_amd import workloads as W

pytestmark = pytest.mark.gpu


def gpu_render(groups, cutoff, samples, step, u, v, phi_mode="reference", form="plain", block_length=256,
               partial_sums=1):
    ds = host.Dataset(groups)
    with host.Config(cutoff=cutoff, samples=samples, step=step, u=u, v=v, block_length=block_length,
                     partial_sums=partial_sums, form=form, phi_mode=phi_mode) as cfg:
        cfg.register(ds)
        try:
            cfg.render(ds)
            F = cfg.spectrum()
            img = cfg.image()
        finally:
            cfg.unregister(ds)
    return F, img


def test_uniform_gaussian_4096_matches_oracle_and_reference_port():
    wl = W.uniform_gaussian_4096()
    vw = wl.views[0]
    print(f"[case] {wl.name} N={wl.n} cutoff={wl.cutoff} samples={wl.samples}", flush=True)
    F, img = gpu_render([wl.group()], wl.cutoff, wl.samples, wl.step, vw.u, vw.v, wl.phi_mode)
    F64 = oracle.spectrum64([wl.group()], wl.cutoff, wl.step, vw.u, vw.v, wl.phi_mode)
    Fref = oracle.port_spectrum32([wl.group()], wl.cutoff, wl.step, vw.u, vw.v)
    e_build = oracle.rel_l2(F, F64)
    e_ref = oracle.rel_l2(Fref, F64)
    e_par = oracle.rel_l2(F, Fref)
    print(f"e_build={e_build:.3e} e_ref={e_ref:.3e} e_par={e_par:.3e}")
    assert e_build <= 1e-5
    assert e_par <= e_ref + e_build + 1e-6
    img64 = oracle.c2r64(oracle.arrange64(F64, *wl.cutoff, *wl.samples), *wl.samples)
    e_img = oracle.rel_l2(img, img64)
    print(f"e_img={e_img:.3e}")
    assert e_img <= 1e-5
    # The reference's C2R applied to the GPU's own spectrum (numpy's pocketfft as the library routine).
    A = oracle.arrange64(F.astype(np.complex128), *wl.cutoff, *wl.samples)
    np_img = np.fft.irfft2(A, s=wl.samples, norm="forward")
    assert oracle.rel_l2(img, np_img) <= 2e-6


def _small_groups(rng, n, basis, radii_kind):
    pos = rng.uniform(-0.5, 0.5, (n, 3)).astype(np.float32)
    w = rng.uniform(-1.0, 1.0, n).astype(np.float32)
    if radii_kind == "none":
        radii = None
    elif radii_kind == "const":
        radii = np.full(n, 0.07, np.float32)
    else:
        radii = rng.uniform(0.01, 0.12, n).astype(np.float32)
    return [(pos, w, radii, basis)]


@pytest.mark.parametrize("basis", ["SPH", "GAUSSIAN", "WENDLAND_D3_C2"])
@pytest.mark.parametrize("radii_kind", ["none", "const", "var"])
@pytest.mark.parametrize("phi_mode", ["reference", "analytic"])
def test_bases_radii_modes(basis, radii_kind, phi_mode):
    rng = np.random.default_rng(11)
    n = 1000  # not a multiple of the 256-particle tile
    groups = _small_groups(rng, n, basis, radii_kind)
    if basis == "GAUSSIAN" and phi_mode == "reference" and radii_kind == "none":
        groups = [(g[0], g[1], np.full(n, 0.02, np.float32), g[3]) for g in groups]  # keep exp(+pi r^2) finite
    u = np.array([0.6, 0.8, 0.0])
    v = np.array([-0.48, 0.36, 0.8])
    cutoff, samples, step = (16, 8), (40, 32), (0.7, 0.9)
    print(f"[case] {basis} {radii_kind} {phi_mode}", flush=True)
    F, img = gpu_render(groups, cutoff, samples, step, u, v, phi_mode)
    F64 = oracle.spectrum64(groups, cutoff, step, u, v, phi_mode)
    e_build = oracle.rel_l2(F, F64)
    print(f"e_build={e_build:.3e}")
    assert e_build <= 1e-5
    if phi_mode == "reference":
        Fref = oracle.port_spectrum32(groups, cutoff, step, u, v)
        e_ref = oracle.rel_l2(Fref, F64)
        assert oracle.rel_l2(F, Fref) <= e_ref + e_build + 1e-6
    img64 = oracle.c2r64(oracle.arrange64(F64, *cutoff, *samples), *samples)
    assert oracle.rel_l2(img, img64) <= 1e-5


def test_multigroup_partial_sums_and_empty_group():
    rng = np.random.default_rng(5)
    g1 = _small_groups(rng, 700, "WENDLAND_D3_C2", "var")[0]
    g2 = _small_groups(rng, 0, "SPH", "none")[0]
    g3 = _small_groups(rng, 300, "GAUSSIAN", "const")[0]
    groups = [g1, g2, g3]
    u, v = (1.0, 0.0, 0.0), (0.0, 0.0, 1.0)
    cutoff, samples, step = (24, 16), (64, 48), (1.0, 1.0)
    F, img = gpu_render(groups, cutoff, samples, step, u, v, "analytic", block_length=128, partial_sums=3)
    F64 = oracle.spectrum64(groups, cutoff, step, u, v, "analytic")
    assert oracle.rel_l2(F, F64) <= 1e-5
    parts = sum(oracle.spectrum64([g], cutoff, step, u, v, "analytic") for g in groups)
    assert oracle.rel_l2(F64, parts) <= 1e-12
    img64 = oracle.c2r64(oracle.arrange64(F64, *cutoff, *samples), *samples)
    assert oracle.rel_l2(img, img64) <= 1e-5


@pytest.mark.parametrize("form", ["plain", "mfma_split_f16"])
def test_dataset_without_groups_renders_zero(form):
    """A dataset with no groups renders a zero spectrum and image, also right after a non-empty
    render on the same config (the spectrum buffer is cleared on the config's stream)."""
    rng = np.random.default_rng(8)
    full = host.Dataset(_small_groups(rng, 200, "GAUSSIAN", "var"))
    empty = host.Dataset([])
    with host.Config(cutoff=(16, 16), samples=(64, 64), form=form) as cfg:
        cfg.register(full)
        cfg.render(full)
        assert np.abs(cfg.spectrum()).max() > 0
        cfg.unregister(full)
        cfg.register(empty)
        cfg.render(empty)
        F, img = cfg.spectrum(), cfg.image()
        cfg.unregister(empty)
    assert not np.any(F) and not np.any(img)


def fft_cutoff(samples, limit=(512, 512), block_length=64):
    """The largest cutoff within `limit` and half the image with 2 cx cy a multiple of the block
    length (the reference's vis_config_check rules, meshless_vis.cu:191-210)."""
    best = (0, 0)
    for cx in range(min(limit[0], samples[0] // 2), -1, -1):
        for cy in range(min(limit[1], samples[1] // 2), -1, -1):
            if (2 * cx * cy) % block_length == 0 and cx * cy > best[0] * best[1]:
                best = (cx, cy)
    return best


# Powers of two, smooth mixed-radix lengths (2, 3, 5, 7, 11, 13), lengths with a large prime
# factor (Bluestein: 17, 97, 131, 1018 = 2 * 509, 8191), and the longest lines (16384).
FFT_SIZES = [(96, 80), (64, 256), (1024, 1024), (600, 600), (768, 512), (8192, 64), (64, 8192), (17, 9),
             (25, 33), (97, 131), (1000, 1018), (8191, 66), (16384, 64), (64, 16384), (1001, 715)]


@pytest.mark.parametrize("samples", FFT_SIZES)
def test_fft_sizes(samples):
    rng = np.random.default_rng(3)
    groups = _small_groups(rng, 512, "GAUSSIAN", "var")
    cutoff = fft_cutoff(samples)
    print(f"[case] fft {samples} cutoff {cutoff}", flush=True)
    F, img = gpu_render(groups, cutoff, samples, (0.8, 0.8), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0), "analytic",
                        block_length=64)
    assert np.abs(F).max() > 0
    A = oracle.arrange64(F.astype(np.complex128), *cutoff, *samples)
    np_img = np.fft.irfft2(A, s=samples, norm="forward")
    e = oracle.rel_l2(img, np_img)
    print(f"fft {samples}: {e:.3e}")
    assert e <= 3e-6


@pytest.mark.parametrize("samples", [(16385, 64), (64, 20000), (8209, 64), (16383, 64)])
def test_fft_sizes_beyond_the_build_are_refused_by_check(samples):
    """Lines above 16384 values, and above 8192 with a prime factor above 13 (8209 is prime,
    16383 = 3 * 43 * 127), are outside this build: the config is created (as the reference's
    vis_config_create, which validates nothing), vis_config_check says no, and a render reports
    VIS_ERROR_INVALID_CONFIG."""
    from libmeshlessvis_amd.abi import VisError

    rng = np.random.default_rng(4)
    ds = host.Dataset(_small_groups(rng, 64, "GAUSSIAN", "var"))
    with host.Config(cutoff=(16, 16), samples=samples) as cfg:
        assert not cfg.check()
        cfg.register(ds)
        with pytest.raises(VisError) as e:
            cfg.render(ds)
        assert e.value.name == "VIS_ERROR_INVALID_CONFIG"
        cfg.unregister(ds)


def test_errors_are_reported_not_fatal():
    from libmeshlessvis_amd.abi import VisError

    rng = np.random.default_rng(0)
    ds = host.Dataset(_small_groups(rng, 10, "SPH", "none"))
    cfg = host.Config(cutoff=(64, 64), samples=(64, 64))  # 2*cutoff > samples: invalid
    assert not cfg.check()
    cfg.register(ds)
    with pytest.raises(VisError) as e:
        cfg.render(ds)
    assert e.value.name == "VIS_ERROR_INVALID_CONFIG"
    cfg.unregister(ds)
    cfg.close()
    ok = host.Config(cutoff=(16, 16), samples=(64, 64))
    assert ok.check()
    with pytest.raises(VisError) as e:
        ok.render(ds)  # not registered
    assert e.value.name == "VIS_ERROR_NOT_REGISTERED"
    ok.close()
