"""The drifting-cluster animation at its full size, through vis_fourier_volume_rendering_batched.

16 timesteps of clustered particles (N log-uniform in [1e4, 5e5], 2.1M in all, kNN(32) radii,
Wendland, analytic Phi), 32 random views each, 256^2: 512 renders in one call, the per-frame
loop of the reference's batch program (vis_noninteractive/main.cpp:61-67) over
vis_fourier_volume_rendering (meshless_vis.cu:334-350).

  * The batch's first call runs its launches without a graph; the second captures them into a
    HIP graph and the third replays it.  Every image and spectrum of the captured and replayed
    calls equals the first call's bit for bit.
  * Every render equals its one-at-a-time render to FP32 rounding (relative L2 <= 1e-6 on the
    image; a batch cuts its work items for the whole batch, so its FP32 partial sums end at
    other particles than a single render's).
  * Eight renders spread over the timesteps' sizes (smallest to largest N) are within 1e-5 of the
    FP64 oracle on a stratified frequency sample (every |x|, y <= 8 plus 384 others, weighted).
"""
import numpy as np
import pytest

import oracle
from libmeshlessvis_amd import host
from libmeshlessvis_amd import workloads as W

pytestmark = pytest.mark.gpu


def test_drifting_clusters_full_batch():
    print("[case] drifting clusters 16 x 32, 256^2, split FP16", flush=True)
    wls = W.drifting_clusters()
    assert len(wls) == 16 and all(len(wl.views) == 32 for wl in wls)
    sets = [host.Dataset([wl.group()]) for wl in wls]
    pairs, cfgs, meta = [], [], []
    for t, (ds, wl) in enumerate(zip(sets, wls)):
        for k, vw in enumerate(wl.views):
            cfgs.append(host.Config(cutoff=wl.cutoff, samples=wl.samples, step=wl.step, u=vw.u, v=vw.v,
                                    form="mfma_split_f16", phi_mode=wl.phi_mode))
            pairs.append(ds)
            meta.append((t, k))
    single = host.Config(cutoff=wls[0].cutoff, samples=wls[0].samples, step=wls[0].step, form="mfma_split_f16",
                         phi_mode=wls[0].phi_mode)
    try:
        for ds in sets:
            cfgs[0].register(ds)
        first = None
        for call in range(3):  # no graph, capture + launch, replay
            host.render_batched(pairs, cfgs)
            imgs = [c.image() for c in cfgs]
            spectra = [c.spectrum() for c in cfgs]
            assert all(np.isfinite(a).all() for a in imgs)
            if first is None:
                first = (imgs, spectra)
                continue
            same = sum(np.array_equal(a, b) for a, b in zip(imgs + spectra, first[0] + first[1]))
            print(f"call {call}: {same} of {2 * len(cfgs)} images/spectra equal to the first call", flush=True)
            assert same == 2 * len(cfgs), call
        imgs, spectra = first
        # One-at-a-time renders of the same 512 views.
        worst = 0.0
        for i, (t, k) in enumerate(meta):
            vw = wls[t].views[k]
            single.set_view(vw.u, vw.v)
            single.render(sets[t])
            worst = max(worst, oracle.rel_l2(single.image(), imgs[i]))
        print(f"batched vs one-at-a-time renders: max image rel-L2 {worst:.3e} over {len(meta)}", flush=True)
        assert worst <= 1e-6
        # FP64 oracle on renders spread over the timesteps' sizes.
        order = np.argsort([wl.n for wl in wls])
        rng = np.random.default_rng(16)
        errs = []
        for t in order[[0, 2, 4, 7, 9, 11, 13, 15]]:
            k = int(rng.integers(32))
            wl, vw = wls[t], wls[t].views[k]
            idx, weight = oracle.stratified_sample(wl.cutoff, rng, n_random=384)
            F64 = oracle.spectrum64([wl.group()], wl.cutoff, wl.step, vw.u, vw.v, wl.phi_mode, idx=idx)
            e = oracle.sampled_rel_l2(spectra[int(t) * 32 + k], F64, idx, weight)
            errs.append(e)
            print(f"timestep {t} (N={wl.n}) view {k}: sampled rel-L2 vs FP64 {e:.3e}", flush=True)
        assert max(errs) <= 1e-5
        for ds in sets:
            cfgs[0].unregister(ds)
    finally:
        for c in cfgs + [single]:
            c.close()
