"""GPU parity of the MFMA complex-GEMM forms against the FP64 oracle (through the C ABI).

Tolerances (relative L2 of the spectrum against the FP64 oracle of the same Phi):
  hi/lo split FP16 (3 products)      <= 1e-5   (the accuracy target)
  hi/lo split BF16 (3 products)      <= 3e-5   (16 significant bits per operand)
  single FP16 product                <= 5e-3   (11 bits; reported for comparison)
"""
import numpy as np
import pytest

import oracle
from libmeshlessvis_amd import host
from libmeshlessvis_amd import workloads as W

pytestmark = pytest.mark.gpu

TOL = {"mfma_split_f16": 1e-5, "mfma_split_bf16": 3e-5, "mfma_f16": 5e-3}


def gpu_spectrum(groups, cutoff, samples, step, u, v, phi_mode, form):
    ds = host.Dataset(groups)
    with host.Config(cutoff=cutoff, samples=samples, step=step, u=u, v=v, form=form, phi_mode=phi_mode) as cfg:
        cfg.register(ds)
        try:
            cfg.render(ds)
            return cfg.spectrum(), cfg.image()
        finally:
            cfg.unregister(ds)


def cloud(rng, n, basis, radii):
    pos = rng.uniform(-0.5, 0.5, (n, 3)).astype(np.float32)
    w = rng.uniform(-1.0, 1.0, n).astype(np.float32)
    return [(pos, w, radii, basis)]


U = np.array([0.6, 0.8, 0.0])
V = np.array([-0.48, 0.36, 0.8])


@pytest.mark.parametrize("form", ["mfma_split_f16", "mfma_split_bf16", "mfma_f16"])
@pytest.mark.parametrize("kind", ["gauss_var_analytic", "gauss_var_reference", "wendland_const", "sph_none",
                                  "gauss_const_reference"])
def test_forms_small(form, kind):
    rng = np.random.default_rng(7)
    n = 3000
    if kind == "gauss_var_analytic":
        basis, radii, mode = "GAUSSIAN", rng.uniform(0.005, 0.05, n).astype(np.float32), "analytic"
    elif kind == "gauss_var_reference":
        basis, radii, mode = "GAUSSIAN", rng.uniform(0.002, 0.01, n).astype(np.float32), "reference"
    elif kind == "wendland_const":
        basis, radii, mode = "WENDLAND_D3_C2", np.full(n, 0.03, np.float32), "analytic"
    elif kind == "sph_none":
        basis, radii, mode = "SPH", None, "reference"
    else:
        basis, radii, mode = "GAUSSIAN", np.full(n, 0.01, np.float32), "reference"
    groups = cloud(rng, n, basis, radii)
    cutoff, samples, step = (64, 48), (256, 128), (0.5, 0.5)
    print(f"[case] {form} {kind}", flush=True)
    F, img = gpu_spectrum(groups, cutoff, samples, step, U, V, mode, form)
    F64 = oracle.spectrum64(groups, cutoff, step, U, V, mode)
    e = oracle.rel_l2(F, F64)
    print(f"{form} {kind}: rel-L2 {e:.3e}")
    assert np.isfinite(F).all()
    assert e <= TOL[form]
    img64 = oracle.c2r64(oracle.arrange64(F64, *cutoff, *samples), *samples)
    assert oracle.rel_l2(img, img64) <= TOL[form] * 1.5


def test_split_f16_matches_plain_on_clustered_subset():
    wl = W.lognormal_field(n=65536, grid=64)
    vw = wl.views[0]
    Fm, _ = gpu_spectrum([wl.group()], wl.cutoff, wl.samples, wl.step, vw.u, vw.v, wl.phi_mode, "mfma_split_f16")
    Fp, _ = gpu_spectrum([wl.group()], wl.cutoff, wl.samples, wl.step, vw.u, vw.v, wl.phi_mode, "plain")
    e = oracle.rel_l2(Fm, Fp)
    print(f"lognormal 65536 split-f16 vs plain: {e:.3e}")
    assert e <= 1e-5
    rng = np.random.default_rng(1)
    idx = rng.choice(Fm.size, 256, replace=False)
    F64 = oracle.spectrum64([wl.group()], wl.cutoff, wl.step, vw.u, vw.v, wl.phi_mode, idx=idx)
    full = np.concatenate([Fm.reshape(-1)[idx]])
    assert oracle.rel_l2(full, F64) <= 1e-5


@pytest.mark.parametrize("form", ["mfma_split_f16", "mfma_split_bf16"])
@pytest.mark.parametrize("basis", ["WENDLAND_D3_C2", "SPH"])
@pytest.mark.parametrize("nodes", [2, 3, 4])
def test_binned_variable_radii(form, basis, nodes):
    rng = np.random.default_rng(9)
    n = 5000
    radii = np.exp(rng.uniform(np.log(0.004), np.log(0.15), n)).astype(np.float32)
    groups = cloud(rng, n, basis, radii)
    cutoff, samples, step = (64, 48), (256, 128), (0.5, 0.5)
    ds = host.Dataset(groups)
    with host.Config(cutoff=cutoff, samples=samples, step=step, u=U, v=V, form=form, phi_mode="analytic",
                     nodes=nodes, h_bins=96 if nodes == 3 else 0) as cfg:
        cfg.register(ds)
        cfg.render(ds)
        F = cfg.spectrum()
        cfg.unregister(ds)
    F64 = oracle.spectrum64(groups, cutoff, step, U, V, "analytic")
    e = oracle.rel_l2(F, F64)
    print(f"binned {form} {basis} T={nodes}: rel-L2 {e:.3e}")
    assert e <= TOL[form]


def test_binned_multigroup_and_reference_mode():
    rng = np.random.default_rng(10)
    g1 = cloud(rng, 2000, "WENDLAND_D3_C2", np.exp(rng.uniform(np.log(0.01), np.log(0.1), 2000)).astype(np.float32))[0]
    g2 = cloud(rng, 1500, "GAUSSIAN", np.exp(rng.uniform(np.log(0.01), np.log(0.1), 1500)).astype(np.float32))[0]
    g3 = cloud(rng, 700, "SPH", np.full(700, 0.05, np.float32))[0]
    groups = [g1, g2, g3]
    cutoff, samples, step = (32, 24), (96, 64), (0.7, 0.6)
    for mode, tol in (("analytic", 1e-5), ("reference", 1e-5)):
        F, _ = gpu_spectrum(groups, cutoff, samples, step, U, V, mode, "mfma_split_f16")
        F64 = oracle.spectrum64(groups, cutoff, step, U, V, mode)
        e = oracle.rel_l2(F, F64)
        print(f"multigroup {mode}: rel-L2 {e:.3e}")
        assert e <= tol


@pytest.mark.parametrize("basis", ["GAUSSIAN", "WENDLAND_D3_C2"])
def test_weights_spanning_decades(basis):
    """w ~ h^3 over two decades of h (SPH-like m/rho weights): per-bin weight scaling keeps FP16 exact enough."""
    rng = np.random.default_rng(12)
    n = 6000
    h = np.exp(rng.uniform(np.log(0.002), np.log(0.2), n)).astype(np.float32)
    w = (h.astype(np.float64) ** 3 * rng.uniform(0.5, 2.0, n)).astype(np.float32)
    pos = rng.uniform(-0.5, 0.5, (n, 3)).astype(np.float32)
    groups = [(pos, w, h, basis)]
    cutoff, samples, step = (64, 64), (128, 128), (0.5, 0.5)
    F, _ = gpu_spectrum(groups, cutoff, samples, step, U, V, "analytic", "mfma_split_f16")
    F64 = oracle.spectrum64(groups, cutoff, step, U, V, "analytic")
    e = oracle.rel_l2(F, F64)
    print(f"decade weights {basis}: rel-L2 {e:.3e}")
    assert e <= 1e-5


@pytest.mark.parametrize("basis", ["WENDLAND_D3_C2", "GAUSSIAN"])
def test_autotune_meets_tolerance(basis):
    """vis_autotune picks (nodes, h_bin_ratio) whose spectrum is within the tolerance of the plain
    kernel's; the chosen setting then renders within that tolerance of the FP64 oracle too."""
    rng = np.random.default_rng(13)
    n = 6000
    radii = np.exp(rng.uniform(np.log(0.004), np.log(0.12), n)).astype(np.float32)
    groups = cloud(rng, n, basis, radii)
    cutoff, samples, step = (64, 48), (128, 128), (0.5, 0.5)
    ds = host.Dataset(groups)
    F64 = oracle.spectrum64(groups, cutoff, step, U, V, "analytic")
    with host.Config(cutoff=cutoff, samples=samples, step=step, u=U, v=V, form="mfma_split_f16",
                     phi_mode="analytic", h_bins=40) as cfg:
        cfg.register(ds)
        ms = cfg.autotune(ds, 1e-5)
        assert ms > 0.0
        assert cfg.form == "mfma_split_f16" and cfg.cs.number_of_h_bins == 0
        assert cfg.cs.h_bin_ratio > 1.0
        cfg.render(ds)
        e = oracle.rel_l2(cfg.spectrum(), F64)
        print(f"autotune {basis}: T={cfg.cs.number_of_interpolation_nodes} ratio={cfg.cs.h_bin_ratio:.3f} "
              f"tile={cfg.cs.tile} {ms:.3f} ms rel-L2 {e:.3e}")
        assert e <= 1.2e-5
        assert 0 <= cfg.cs.tile <= 3
        # An unreachable tolerance keeps the most accurate candidate, still a valid setting.
        assert cfg.autotune(ds, 1e-12) > 0.0
        cfg.render(ds)
        assert oracle.rel_l2(cfg.spectrum(), F64) <= 1.2e-5
        cfg.form = "plain"
        with pytest.raises(host.abi.VisError) as ei:
            cfg.autotune(ds, 1e-5)
        assert ei.value.name == "VIS_ERROR_INVALID_CONFIG"
        cfg.form = "mfma_split_f16"
        cfg.cs.h_bin_ratio = 0.5
        with pytest.raises(host.abi.VisError) as ei:
            cfg.render(ds)
        assert ei.value.name == "VIS_ERROR_INVALID_CONFIG"
        cfg.unregister(ds)


@pytest.mark.parametrize("form,basis,radii_kind,nodes", [
    ("plain", "WENDLAND_D3_C2", "var", 0),
    ("mfma_split_f16", "GAUSSIAN", "var", 0),         # sum kernel, Gaussian bins
    ("mfma_split_f16", "WENDLAND_D3_C2", "const", 0),  # sum kernel, uniform radius
    ("mfma_split_f16", "WENDLAND_D3_C2", "var", 2),    # two-node binned kernel
    ("mfma_split_f16", "SPH", "var", 3),               # binned kernel, 3 nodes
])
def test_renders_are_deterministic(form, basis, radii_kind, nodes):
    """Every form sums in a fixed order (FP64 chunk reductions in a fixed order, the x = 0 column
    by shuffles, no floating-point atomics): a render repeated on the same data is bit-identical,
    on the same config and on a fresh one."""
    print(f"[case] {form} {basis} {radii_kind} nodes={nodes}", flush=True)
    rng = np.random.default_rng(21)
    n = 6000
    radii = rng.uniform(0.01, 0.06, n).astype(np.float32) if radii_kind == "var" else np.full(n, 0.03, np.float32)
    ds = host.Dataset(cloud(rng, n, basis, radii))
    out = []
    for _ in range(2):
        with host.Config(cutoff=(80, 48), samples=(256, 128), step=(0.5, 0.5), u=U, v=V, form=form,
                         phi_mode="analytic", nodes=nodes) as cfg:
            cfg.register(ds)
            for _ in range(2):
                cfg.render(ds)
                out.append((cfg.spectrum(), cfg.image()))
            cfg.unregister(ds)
    for F, img in out[1:]:
        assert np.array_equal(F, out[0][0]) and np.array_equal(img, out[0][1])


def test_autotune_tile_on_a_small_view():
    """On a 4096-particle view with a 64 x 64 image (cutoff 32: a 32 x 33 half plane) the default
    128 x 128 tile of the sum kernel is mostly empty; vis_autotune times every VisTile variant and
    keeps the fastest within the tolerance.  The kept tile is a valid variant and its render is
    within 1e-5 of the FP64 oracle."""
    wl = W.uniform_gaussian_4096()
    vw = wl.views[0]
    ds = host.Dataset([wl.group()])
    F64 = oracle.spectrum64([wl.group()], wl.cutoff, wl.step, vw.u, vw.v, wl.phi_mode)
    with host.Config(cutoff=wl.cutoff, samples=wl.samples, step=wl.step, u=vw.u, v=vw.v, form="mfma_split_f16",
                     phi_mode=wl.phi_mode) as cfg:
        cfg.register(ds)
        default_ms = cfg.autotune(ds, 1e-5)  # warms the kernels; then time once more
        cfg.cs.tile = 0
        ms = cfg.autotune(ds, 1e-5)
        cfg.render(ds)
        e = oracle.rel_l2(cfg.spectrum(), F64)
        print(f"tile {cfg.cs.tile}: {ms:.4f} ms (first pass {default_ms:.4f} ms), rel-L2 {e:.2e}")
        assert 0 <= cfg.cs.tile <= 3
        assert e <= 1e-5
        cfg.unregister(ds)
