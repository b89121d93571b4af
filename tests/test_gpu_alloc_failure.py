"""Out-of-memory paths, driven by the allocation-failure hook vis_debug_fail_device_allocations.

The reference aborts on any CUDA error (cutil, meshless_vis.cu:73-81); here a failed device
allocation is reported as VIS_ERROR_OUT_OF_MEMORY and leaves the library usable.  Each scenario -
create a config, register a dataset, render (plain form, split-FP16 MFMA form over a Gaussian and a
two-node Wendland group, and a batched call captured on its second run), change the image size and
render again, copy the image - is first run without the hook; then, for k = 0, 1, 2, ..., with the
k-th device allocation and every later one failing, until a run completes.  Every failure must be
reported as VIS_ERROR_OUT_OF_MEMORY (or, for a render after a failed allocation left the config
without a buffer or a group unregistered, VIS_ERROR_INVALID_CONFIG / VIS_ERROR_NOT_REGISTERED);
a failed registration must leave the groups' device pointers cleared; the cleanup calls must
succeed; and with the hook off again the same scenario must reproduce the first images bit for
bit."""
import ctypes as C

import numpy as np
import pytest

from libmeshlessvis_amd import abi, host

pytestmark = pytest.mark.gpu

ALLOWED_AFTER_FAILURE = {"VIS_ERROR_OUT_OF_MEMORY", "VIS_ERROR_INVALID_CONFIG", "VIS_ERROR_NOT_REGISTERED",
                         "VIS_ERROR_INVALID_VALUE"}


def make_dataset():
    rng = np.random.default_rng(17)
    groups = []
    for n, basis in ((3000, "GAUSSIAN"), (2500, "WENDLAND_D3_C2")):
        pos = rng.uniform(-0.5, 0.5, (n, 3)).astype(np.float32)
        w = rng.uniform(0.1, 1.0, n).astype(np.float32)
        h = np.exp(rng.uniform(np.log(0.01), np.log(0.06), n)).astype(np.float32)
        groups.append((pos, w, h, basis))
    return host.Dataset(groups)


def code_name(L):
    return abi.ERRORS.get(L.vis_get_last_error_code(), "?")


def scenario(form, batched):
    """Runs the scenario; returns (images or None, list of (step, error name) of failed steps)."""
    L = abi.lib()
    ds = make_dataset()
    errors = []
    cfgs = []

    def step(name, fn, *args):
        L.vis_clear_last_error()
        fn(*args)
        e = code_name(L)
        if e != "VIS_SUCCESS":
            errors.append((name, e))
        return e == "VIS_SUCCESS"

    def create():
        p = L.vis_config_create(True, abi.float2(0.5, 0.5), abi.int2(32, 24), abi.float3(0.6, 0.8, 0.0),
                                abi.float3(-0.48, 0.36, 0.8), abi.int2(64, 64), 64, 1)
        if p:
            p.contents.form = abi.FORM[form]
            p.contents.phi_mode = abi.PHI_MODE["analytic"]
            cfgs.append(p)
        return p

    images = None
    try:
        L.vis_clear_last_error()
        c0 = create()
        if not c0:
            errors.append(("create", code_name(L)))
            return None, errors
        c1 = create() if batched else None
        if batched and not c1:
            errors.append(("create", code_name(L)))
            return None, errors
        if not step("register", L.vis_register_meshless_dataset, c0, ds.ptr):
            for g in range(2):
                grp = ds.ptr.contents.groups[g]
                assert not grp.d_constraints and not grp.d_radii, "failed registration left device pointers"
        renders_ok = True
        for _ in range(2):  # batched: the first call runs eagerly, the second captures
            if batched:
                pairs = (abi.DatasetP * 2)(ds.ptr, ds.ptr)
                confs = (abi.VisConfigP * 2)(c0, c1)
                renders_ok &= step("render_batched", L.vis_fourier_volume_rendering_batched, pairs, confs, 2)
            else:
                renders_ok &= step("render", L.vis_fourier_volume_rendering, ds.ptr, c0)
        step("change_samples", L.vis_config_change_number_of_samples, c0, abi.int2(96, 64))
        renders_ok &= step("render_after_change", L.vis_fourier_volume_rendering, ds.ptr, c0)
        if renders_ok:
            out = np.empty(96 * 64, np.float32)
            if step("copy", L.vis_copy_to_host, c0, out.ctypes.data_as(C.POINTER(C.c_float))):
                images = out
    finally:
        L.vis_debug_fail_device_allocations(-1)
        if cfgs:
            L.vis_clear_last_error()
            L.vis_unregister_meshless_dataset(cfgs[0], ds.ptr)
            assert code_name(L) == "VIS_SUCCESS"
        for p in cfgs:
            L.vis_config_destroy(p)
    return images, errors


@pytest.mark.parametrize("form,batched", [("plain", False), ("mfma_split_f16", False), ("mfma_split_f16", True)])
def test_every_allocation_failure_is_reported_and_recovered(form, batched):
    L = abi.lib()
    print(f"[case] allocation failures {form} batched={batched}", flush=True)
    ref, errs = scenario(form, batched)
    assert not errs and ref is not None and np.isfinite(ref).all() and np.abs(ref).max() > 0
    for k in range(200):
        L.vis_debug_fail_device_allocations(k)
        try:
            images, errs = scenario(form, batched)
        finally:
            L.vis_debug_fail_device_allocations(-1)
        print(f"k={k}: {errs}", flush=True)
        for i, (name, e) in enumerate(errs):
            assert e in ALLOWED_AFTER_FAILURE, (k, name, e)
            if i == 0:
                assert e == "VIS_ERROR_OUT_OF_MEMORY", (k, name, e)
        again, errs2 = scenario(form, batched)
        assert not errs2 and np.array_equal(again, ref), k
        if not errs:
            assert np.array_equal(images, ref)
            break
    else:
        pytest.fail("the scenario never completed with the hook on")
