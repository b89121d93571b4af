"""Entry points and caller-side patterns of the C ABI beyond register / render / copy.

  * vis_config_manual_d_image (meshless_vis.cu:244-247): rendering into a caller's device
    buffer gives the automatic image bit for bit;
  * vis_config_change_number_of_partial_sums (meshless_vis.cu:236-242): a render after the
    change gives the same image (the partial-sum count sets allocation and padding only);
  * fourier_transform (fourier_transform.h:33, the internal seam): the spectrum equals the
    render's;
  * groups whose device arrays the caller filled without vis_register_meshless_dataset (the
    reference's kernels read only d_constraints / d_radii and the counts): every form, against
    the plain form and the FP64 oracle, also after the arrays are rewritten in place;
  * radii of 0 (Phi(0) for every frequency) in the binned MFMA forms, against the FP64 oracle;
  * two batched calls queued back to back on a non-blocking stream with one synchronisation at
    the end: the second call's tables must not reach the first call's launches.
"""
import ctypes as C

import numpy as np
import pytest

import oracle
from libmeshlessvis_amd import abi, host
from libmeshlessvis_amd import workloads as W

pytestmark = pytest.mark.gpu


def cloud(rng, n, h_lo=0.005, h_hi=0.08, zero_frac=0.0, signed=False):
    pos = rng.uniform(-0.5, 0.5, (n, 3)).astype(np.float32)
    h = np.exp(rng.uniform(np.log(h_lo), np.log(h_hi), n)).astype(np.float32)
    if zero_frac:
        h[rng.random(n) < zero_frac] = 0.0
    w = rng.uniform(-1 if signed else 0.2, 1, n).astype(np.float32)
    return pos, w, h


def cfg_for(vw, form, phi="analytic", cutoff=(48, 40), samples=(128, 96), **kw):
    return host.Config(cutoff=cutoff, samples=samples, step=(0.5, 0.5), u=vw.u, v=vw.v, form=form, phi_mode=phi, **kw)


def test_manual_d_image_matches_automatic():
    print("[case] manual_d_image", flush=True)
    rng = np.random.default_rng(1)
    vw = W.random_view(rng)
    pos, w, h = cloud(rng, 3000)
    ds = host.Dataset([(pos, w, h, "WENDLAND_D3_C2")])
    for form in ("plain", "mfma_split_f16"):
        with cfg_for(vw, form) as auto, cfg_for(vw, form, automatic_d_image=False) as man:
            nx, ny = man.samples
            buf = host.DeviceBuffer(4 * nx * ny)
            try:
                abi.lib().vis_config_manual_d_image(man.p, buf.ptr)
                assert man.cs._d_image == buf.ptr.value
                auto.register(ds)
                auto.render(ds)
                man.render(ds)
                man.synchronize()
                a = auto.image()
                b = buf.download(np.float32, nx * ny).reshape(nx, ny)
                assert np.array_equal(a, b), form
                assert np.array_equal(man.image(), a), form
                auto.unregister(ds)
            finally:
                man.close()
                buf.free()


def test_change_number_of_partial_sums_keeps_image():
    print("[case] change_number_of_partial_sums", flush=True)
    rng = np.random.default_rng(2)
    vw = W.random_view(rng)
    pos, w, h = cloud(rng, 2500)
    ds = host.Dataset([(pos, w, h, "SPH")])
    for form in ("plain", "mfma_split_f16"):
        with cfg_for(vw, form, phi="reference") as cfg:
            cfg.register(ds)
            cfg.render(ds)
            before, F0 = cfg.image(), cfg.spectrum()
            abi.lib().vis_config_change_number_of_partial_sums(cfg.p, 3)
            abi.check("vis_config_change_number_of_partial_sums")
            assert cfg.cs._number_of_partial_sums == 3 and cfg.check()
            cfg.render(ds)
            assert np.array_equal(cfg.image(), before), form
            assert np.array_equal(cfg.spectrum(), F0), form
            # Registered again under P = 3 (padding to a multiple of 3 * block_length).
            cfg.unregister(ds)
            cfg.register(ds)
            assert ds.c.groups[0].d_number_of_terms == oracle.padded_count(2500, 256, 3)
            cfg.render(ds)
            assert oracle.rel_l2(cfg.image(), before) <= 1e-6, form
            cfg.unregister(ds)


def test_fourier_transform_seam_equals_render_spectrum():
    print("[case] fourier_transform seam", flush=True)
    rng = np.random.default_rng(3)
    vw = W.random_view(rng)
    pos, w, h = cloud(rng, 2000)
    ds = host.Dataset([(pos, w, h, "GAUSSIAN"), (pos[:500] * 0.5, w[:500], None, "WENDLAND_D3_C2")])
    for form in ("plain", "mfma_split_f16"):
        with cfg_for(vw, form) as cfg:
            cfg.register(ds)
            cfg.render(ds)
            F_render = cfg.spectrum()
            cfg.fourier_transform(ds)
            assert np.array_equal(cfg.spectrum(), F_render), form
            assert cfg.cs._scale == np.float32(0.25)
            F64 = oracle.spectrum64([(pos, w, h, "GAUSSIAN"), (pos[:500] * 0.5, w[:500], None, "WENDLAND_D3_C2")],
                                    cfg.cutoff, (0.5, 0.5), vw.u, vw.v, "analytic")
            assert oracle.rel_l2(F_render, F64) <= 1e-5, form
            cfg.unregister(ds)


class FilledGroup:
    """A one-group dataset whose device arrays are allocated and filled here (hipMalloc +
    hipMemcpy), with no vis_register_meshless_dataset."""

    def __init__(self, pos, w, h, basis):
        n = pos.shape[0]
        self.cons = np.zeros((n, 4), np.float32)
        self.cons[:, :3], self.cons[:, 3] = pos, w
        self.dc = host.DeviceBuffer(self.cons.nbytes, self.cons)
        self.dr = host.DeviceBuffer(4 * n, np.ascontiguousarray(h, np.float32)) if h is not None else None
        self.groups = (abi.Group * 1)()
        g = self.groups[0]
        g.number_of_terms = n
        g.d_number_of_terms = n
        g.basis_function_id = abi.BASIS[basis]
        g.d_constraints = C.cast(self.dc.ptr, C.POINTER(abi.Constraint))
        g.d_radii = C.cast(self.dr.ptr, C.POINTER(C.c_float)) if self.dr else None
        self.c = abi.MeshlessDataset()
        self.c.groups = C.cast(self.groups, C.POINTER(abi.Group))
        self.c.number_of_groups = 1

    @property
    def ptr(self):
        return C.pointer(self.c)

    def rewrite(self, w, h):
        self.cons[:, 3] = w
        self.dc.upload(self.cons)
        if self.dr is not None:
            self.dr.upload(np.ascontiguousarray(h, np.float32))

    def free(self):
        self.dc.free()
        if self.dr:
            self.dr.free()


@pytest.mark.parametrize("basis,radii", [("GAUSSIAN", "varying"), ("WENDLAND_D3_C2", "varying"),
                                         ("SPH", "varying"), ("WENDLAND_D3_C2", "uniform"), ("GAUSSIAN", "none")])
def test_caller_filled_group_every_form(basis, radii):
    print(f"[case] caller-filled group {basis} {radii}", flush=True)
    rng = np.random.default_rng(4)
    vw = W.random_view(rng)
    pos, w, h = cloud(rng, 3000)
    h = {"varying": h, "uniform": np.full_like(h, 0.03), "none": None}[radii]
    fg = FilledGroup(pos, w, h, basis)
    try:
        for phase in range(2):
            F64 = oracle.spectrum64([(pos, w, h, basis)], (48, 40), (0.5, 0.5), vw.u, vw.v, "analytic")
            spectra = {}
            for form, tol in (("plain", 1e-5), ("mfma_split_f16", 1e-5), ("mfma_split_bf16", 3e-5)):
                with cfg_for(vw, form) as cfg:
                    abi.lib().vis_fourier_volume_rendering(fg.ptr, cfg.p)
                    abi.check("render of a caller-filled group")
                    spectra[form] = cfg.spectrum()
                e = oracle.rel_l2(spectra[form], F64)
                print(f"{basis} {radii} phase {phase} {form}: rel-L2 vs FP64 {e:.3e}")
                assert e <= tol, (form, phase, e)
            assert oracle.rel_l2(spectra["mfma_split_f16"], spectra["plain"]) <= 1e-5
            # Rewrite the arrays in place (new weights, radii scaled by 1.5): the next renders must
            # follow the new data, not facts or bins kept from the first.
            w = rng.uniform(0.2, 3.0, w.size).astype(np.float32)
            h = None if h is None else (h * 1.5).astype(np.float32)
            fg.rewrite(w, h)
    finally:
        fg.free()


@pytest.mark.parametrize("basis", ["WENDLAND_D3_C2", "SPH", "GAUSSIAN"])
@pytest.mark.parametrize("phi", ["analytic", "reference"])
def test_zero_radii_get_their_own_bin(basis, phi):
    print(f"[case] zero radii {basis} {phi}", flush=True)
    rng = np.random.default_rng(5)
    vw = W.random_view(rng)
    pos, w, h = cloud(rng, 4000, zero_frac=0.1, h_lo=0.004, h_hi=0.02 if phi == "reference" else 0.08)
    ds = host.Dataset([(pos, w, h, basis)])
    F64 = oracle.spectrum64([(pos, w, h, basis)], (48, 40), (0.5, 0.5), vw.u, vw.v, phi)
    for form, nodes in (("plain", 0), ("mfma_split_f16", 0), ("mfma_split_f16", 3)):
        if basis == "GAUSSIAN" and nodes:
            continue
        with cfg_for(vw, form, phi=phi, nodes=nodes) as cfg:
            cfg.register(ds)
            cfg.render(ds)
            e = oracle.rel_l2(cfg.spectrum(), F64)
            cfg.unregister(ds)
        print(f"zero radii {basis} {phi} {form} nodes {nodes}: rel-L2 vs FP64 {e:.3e}")
        assert e <= 1e-5, (form, nodes, e)


def test_batches_back_to_back_on_non_blocking_stream():
    print("[case] batches on a non-blocking stream", flush=True)
    rng = np.random.default_rng(6)
    pos, w, h = cloud(rng, 20000)
    ds = host.Dataset([(pos, w, h, "WENDLAND_D3_C2")])
    vws = [W.random_view(rng) for _ in range(3)]
    cfgs = [cfg_for(vw, "mfma_split_f16", cutoff=(64, 64), samples=(128, 128)) for vw in vws]
    stream = host.create_stream(non_blocking=True)
    try:
        for c in cfgs:
            abi.lib().vis_set_stream(c.p, stream)
        cfgs[0].register(ds)
        eager = []
        for c in cfgs:
            c.render(ds)
            eager.append(c.image())
        for rep in range(3):
            # [c0, c1] then [c0, c2]: one batch state (c0's), the second call's tables uploaded
            # while the first call may still run; nothing waits in between.
            host.render_batched([ds, ds], [cfgs[0], cfgs[1]])
            host.render_batched([ds, ds], [cfgs[0], cfgs[2]])
            host.hip().hipStreamSynchronize(stream)
            for k in (1, 2):
                e = oracle.rel_l2(cfgs[k].image(), eager[k])
                assert e <= 1e-6, (rep, k, e)
        cfgs[0].unregister(ds)
    finally:
        for c in cfgs:
            c.close()
        host.hip().hipStreamDestroy(stream)
