"""vis_opengl_fourier_volume_rendering (meshless_vis.cu:357-364) on a box without a display.

The render needs a current GL context on the GPU; this box has none (no X server, no EGL), so the
interop call must fail cleanly: an error recorded (VIS_ERROR_HIP from hipGraphicsGLRegisterBuffer),
nothing rendered, the config's own _d_image unchanged, and the next ordinary render equal to one
made before the call.  Run in a child process so that the HIP runtime's GL lookup cannot take the
test runner down with it.  The mapped-buffer path itself (register, map, render into the buffer,
restore _d_image, unmap, unregister) needs a GL context and is not exercised here."""
import subprocess
import sys
import textwrap
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]

CHILD = textwrap.dedent("""
    import sys
    sys.path.insert(0, {root!r})
    import numpy as np
    from libmeshlessvis_amd import host
    from libmeshlessvis_amd.abi import lib
    rng = np.random.default_rng(0)
    pos = rng.random((2000, 3)).astype(np.float32)
    w = rng.random(2000).astype(np.float32)
    ds = host.Dataset.simple(pos, w, np.full(2000, 0.05, np.float32), "GAUSSIAN")
    with host.Config(cutoff=(32, 32), samples=(64, 64), step=(1.0, 1.0)) as cfg:
        cfg.register(ds)
        cfg.render(ds)
        before = cfg.image()
        img_ptr = cfg.cs._d_image
        L = lib()
        L.vis_opengl_fourier_volume_rendering(ds.ptr, cfg.p, 1)
        code = L.vis_get_last_error_code()
        print("code", code, L.vis_get_last_error().decode(), flush=True)
        assert code != 0, "interop without a GL context reported success"
        assert cfg.cs._d_image == img_ptr
        cfg.render(ds)
        after = cfg.image()
        assert np.array_equal(before, after)
        cfg.unregister(ds)
    print("ok", flush=True)
""")


def test_without_gl_context_reports_an_error():
    r = subprocess.run([sys.executable, "-c", CHILD.format(root=str(ROOT))], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout, r.stderr[-2000:])
    assert r.returncode == 0
    assert "ok" in r.stdout
