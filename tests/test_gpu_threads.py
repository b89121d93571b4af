"""Host threads rendering at the same time, each config on its own non-blocking stream.

The library keeps per-config state (plans, workspaces, FFT plans, batch graphs) in the config,
process-wide state (the dataset registry, the per-group bins, device limits, the partial budget)
behind locks or atomics, and its last error per thread.  ctypes releases the GIL around every
library call, so the threads below really run in the library together.  Each render must equal
bit for bit the same render made alone beforehand (same plan, same kernels, same ordered
reductions), also when two threads render one registered dataset whose bins neither has built
yet."""
import threading

import numpy as np
import pytest

from libmeshlessvis_amd import abi, host
from libmeshlessvis_amd import workloads as W

pytestmark = pytest.mark.gpu

JOBS = [  # (basis, form, particles)
    ("WENDLAND_D3_C2", "mfma_split_f16", 20000),
    ("GAUSSIAN", "mfma_split_f16", 30000),
    ("SPH", "mfma_split_bf16", 15000),
    ("WENDLAND_D3_C2", "plain", 4000),
]


def make(rng, basis, n):
    pos = rng.uniform(-0.5, 0.5, (n, 3)).astype(np.float32)
    w = rng.uniform(-1, 1, n).astype(np.float32)
    h = np.exp(rng.uniform(np.log(0.005), np.log(0.05), n)).astype(np.float32)
    return host.Dataset([(pos, w, h, basis)])


def config(vw, form):
    return host.Config(cutoff=(64, 48), samples=(128, 96), step=(0.5, 0.5), u=vw.u, v=vw.v, form=form,
                       phi_mode="analytic")


def run_threads(work):
    errors = []

    def body(fn):
        try:
            fn()
        except BaseException as e:  # reported by the main thread
            errors.append(e)

    threads = [threading.Thread(target=body, args=(fn,)) for fn in work]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in threads), "a rendering thread did not finish"
    if errors:
        raise errors[0]


def test_threads_render_their_own_datasets():
    print("[case] four threads, own datasets and streams", flush=True)
    rng = np.random.default_rng(31)
    sets = [make(rng, b, n) for b, _, n in JOBS]
    vws = [W.random_view(rng) for _ in JOBS]
    cfgs = [config(vw, f) for vw, (_, f, _) in zip(vws, JOBS)]
    streams = [host.create_stream(non_blocking=True) for _ in JOBS]
    try:
        for ds, cfg, s in zip(sets, cfgs, streams):
            abi.lib().vis_set_stream(cfg.p, s)
            cfg.register(ds)
        alone = []
        for ds, cfg in zip(sets, cfgs):
            cfg.render(ds)
            alone.append(cfg.image())
        got = [[] for _ in JOBS]

        def worker(k):
            def fn():
                for _ in range(5):
                    cfgs[k].render(sets[k])
                    got[k].append(cfgs[k].image())
            return fn

        run_threads([worker(k) for k in range(len(JOBS))])
        for k in range(len(JOBS)):
            assert len(got[k]) == 5
            for img in got[k]:
                assert np.array_equal(img, alone[k]), JOBS[k]
    finally:
        for ds, cfg in zip(sets, cfgs):
            cfg.unregister(ds)
            cfg.close()
        for s in streams:
            host.hip().hipStreamDestroy(s)


def test_threads_share_one_dataset():
    """Two threads render views of one registered dataset, whose binned form's bins are built at
    the first render (under the bin cache's lock), then batches of their own views."""
    print("[case] two threads, one dataset", flush=True)
    rng = np.random.default_rng(32)
    ds = make(rng, "WENDLAND_D3_C2", 25000)
    vws = [W.random_view(rng) for _ in range(4)]
    streams = [host.create_stream(non_blocking=True) for _ in range(2)]
    cfgs = [config(vw, "mfma_split_f16") for vw in vws]
    try:
        for k, cfg in enumerate(cfgs):
            abi.lib().vis_set_stream(cfg.p, streams[k % 2])
        # Reference images from a separately registered copy (its own bins), rendered alone.
        ref_cfgs = [config(vw, "mfma_split_f16") for vw in vws]
        ref = []
        src = make(np.random.default_rng(32), "WENDLAND_D3_C2", 25000)  # the same particles again
        ref_cfgs[0].register(src)
        for c in ref_cfgs:
            c.render(src)
            ref.append(c.image())
        ref_cfgs[0].unregister(src)
        for c in ref_cfgs:
            c.close()
        cfgs[0].register(ds)
        out = {}

        def worker(k):
            def fn():
                for rep in range(3):
                    for j in (k, k + 2):
                        cfgs[j].render(ds)
                        out[(j, rep)] = cfgs[j].image()
                host.render_batched([ds, ds], [cfgs[k], cfgs[k + 2]])
                out[(k, "batch")] = cfgs[k].image()
                out[(k + 2, "batch")] = cfgs[k + 2].image()
            return fn

        run_threads([worker(0), worker(1)])
        for j in range(4):
            for rep in range(3):
                assert np.array_equal(out[(j, rep)], ref[j]), (j, rep)
            assert np.abs(out[(j, "batch")] - ref[j]).max() <= 1e-6 * np.abs(ref[j]).max(), j
        cfgs[0].unregister(ds)
    finally:
        for c in cfgs:
            c.close()
        for s in streams:
            host.hip().hipStreamDestroy(s)
