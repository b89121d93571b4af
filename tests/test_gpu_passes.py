"""MFMA renders that run in passes (vis_set_partial_budget): a budget far below the partial
spectra of a render makes it run its chunks in passes, each reduced into the FP64 accumulator
before the next reuses the slots.  The FP32 partial of every chunk is the same as in one pass
(same items, same chunk boundaries), so a render in passes differs from the one-pass render only
by the FP64 order of the chunk sums: <= 1e-9 relative here, and both within 1e-5 of the FP64
oracle.  A batch in passes runs through the per-render finish; its graph replays must equal its
first call bit for bit.  The default budget is restored after every test."""
import numpy as np
import pytest

import oracle
from libmeshlessvis_amd import abi
from libmeshlessvis_amd import host
from libmeshlessvis_amd import workloads as W

pytestmark = pytest.mark.gpu

STEP = (0.5, 0.5)
CASES = {
    # name: (particles, basis, nodes, cutoff, samples)
    "gauss_sum": (60_000, "GAUSSIAN", 0, (128, 96), (256, 192)),       # sum kernel
    "wendland_t2": (30_000, "WENDLAND_D3_C2", 2, (96, 64), (192, 128)),  # two-node kernel
    "sph_t3": (20_000, "SPH", 3, (64, 64), (128, 128)),                  # T >= 3 kernel
}


def cloud(name, seed=3):
    n, basis, nodes, cutoff, samples = CASES[name]
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-0.5, 0.5, (n, 3)).astype(np.float32)
    h = np.exp(rng.uniform(np.log(0.005), np.log(0.05), n)).astype(np.float32)
    w = rng.uniform(-1.0, 1.0, n).astype(np.float32)
    return [(pos, w, h, basis)], nodes, cutoff, samples


def render_spectrum(groups, nodes, cutoff, samples, vw, budget):
    abi.lib().vis_set_partial_budget(budget)
    abi.check("vis_set_partial_budget")
    ds = host.Dataset(groups)
    with host.Config(cutoff=cutoff, samples=samples, step=STEP, u=vw.u, v=vw.v, form="mfma_split_f16",
                     phi_mode="analytic", nodes=nodes) as cfg:
        cfg.register(ds)
        try:
            cfg.render(ds)
            return cfg.spectrum(), cfg.image()
        finally:
            cfg.unregister(ds)


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("budget", [1, 256 << 10, 1 << 20])  # partials of 0.7-24 MB in one pass; 1: a pass per chunk
def test_render_in_passes(name, budget):
    groups, nodes, cutoff, samples = cloud(name)
    vw = W.random_view(np.random.default_rng(5))
    print(f"[case] {name} budget {budget}", flush=True)
    try:
        F1, img1 = render_spectrum(groups, nodes, cutoff, samples, vw, 0)
        Fp, imgp = render_spectrum(groups, nodes, cutoff, samples, vw, budget)
    finally:
        abi.lib().vis_set_partial_budget(0)
    F64 = oracle.spectrum64(groups, cutoff, STEP, vw.u, vw.v, "analytic")
    e_pass = oracle.rel_l2(Fp, F1)
    e1, ep = oracle.rel_l2(F1, F64), oracle.rel_l2(Fp, F64)
    print(f"{name} budget {budget}: passes vs one pass {e_pass:.2e}; vs FP64 {e1:.2e} / {ep:.2e}")
    assert np.isfinite(Fp).all() and np.isfinite(imgp).all()
    assert e_pass <= 1e-9
    assert oracle.rel_l2(imgp, img1) <= 1e-9
    assert e1 <= 1e-5 and ep <= 1e-5


def test_batch_in_passes_replays():
    """Three views of two clouds (sum and two-node kernels) as one batch under a 1 MiB budget:
    every call (eager, captured, replayed) gives the same images bit for bit, each within 1e-6 of
    the one-at-a-time render under the default budget."""
    rng = np.random.default_rng(8)
    vws = [W.random_view(rng) for _ in range(3)]
    sets, cfgs, pairs = [], [], []
    for name in ("gauss_sum", "wendland_t2"):
        groups, nodes, cutoff, samples = cloud(name, seed=9)
        sets.append((host.Dataset(groups), nodes))
    cutoff, samples = (64, 48), (128, 96)
    for ds, nodes in sets:
        for vw in vws:
            cfgs.append(host.Config(cutoff=cutoff, samples=samples, step=STEP, u=vw.u, v=vw.v,
                                    form="mfma_split_f16", phi_mode="analytic", nodes=nodes))
            pairs.append(ds)
    try:
        for ds, _ in sets:
            cfgs[0].register(ds)
        single = []
        for ds, cfg in zip(pairs, cfgs):
            cfg.render(ds)
            single.append(cfg.image())
        abi.lib().vis_set_partial_budget(1 << 20)
        first = None
        for rep in range(4):  # eager, capture, replay, replay
            print(f"[case] batch in passes, call {rep}", flush=True)
            host.render_batched(pairs, cfgs)
            imgs = [cfg.image() for cfg in cfgs]
            if first is None:
                first = imgs
                for a, b in zip(imgs, single):
                    assert oracle.rel_l2(a, b) <= 1e-6
            else:
                for a, b in zip(imgs, first):
                    assert np.array_equal(a, b), rep
    finally:
        abi.lib().vis_set_partial_budget(0)
        for ds, _ in sets:
            cfgs[0].unregister(ds)
        for cfg in cfgs:
            cfg.close()


def test_default_budget_splits_a_large_render():
    """8,388,608 uniformly placed particles (Gaussian, h in [0.001, 0.004]) at 2048^2: 512 chunks
    of 2,097,152 frequencies, 8.6 GB of partial spectra, so the default 8 GiB budget runs two
    passes.  Bit-identical to one pass (a 64 GiB budget), and within 1e-5 of the FP64 oracle on a
    stratified sample (a spectrum that keeps its energy up to the highest frequencies)."""
    n, c = 8_388_608, 1024
    rng = np.random.default_rng(1)
    pos = rng.uniform(-0.5, 0.5, (n, 3)).astype(np.float32)
    w = rng.uniform(-1, 1, n).astype(np.float32)
    h = np.exp(rng.uniform(np.log(0.001), np.log(0.004), n)).astype(np.float32)
    groups = [(pos, w, h, "GAUSSIAN")]
    vw = W.random_view(rng)
    cutoff, samples = (c, c), (2 * c, 2 * c)
    ds = host.Dataset(groups)
    spectra = {}
    with host.Config(cutoff=cutoff, samples=samples, step=STEP, u=vw.u, v=vw.v, form="mfma_split_f16",
                     phi_mode="analytic") as cfg:
        cfg.register(ds)
        try:
            for budget in (0, 64 << 30):
                print(f"[case] 8.4M particles at 2048^2, budget {budget}", flush=True)
                abi.lib().vis_set_partial_budget(budget)
                cfg.render(ds)
                spectra[budget] = cfg.spectrum()
        finally:
            abi.lib().vis_set_partial_budget(0)
            cfg.unregister(ds)
    assert np.array_equal(spectra[0], spectra[64 << 30])
    idx, weight = oracle.stratified_sample(cutoff, np.random.default_rng(2), low=4, n_random=64)
    F64 = oracle.spectrum64(groups, cutoff, STEP, vw.u, vw.v, "analytic", idx=idx)
    e = oracle.sampled_rel_l2(spectra[0], F64, idx, weight)
    print(f"8.4M particles at 2048^2 in two passes: sampled rel-L2 vs FP64 {e:.3e} ({idx.size} frequencies)")
    assert e <= 1e-5
