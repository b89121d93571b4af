"""numpy convenience layer over the C ABI (libmeshless_vis.so).

``Dataset`` builds MeshlessDataset/Group structures over numpy arrays kept
alive by the Python object (the arrays are never handed to
delete_meshless_dataset).  ``Config`` owns one VisConfig and exposes the
reference's calls: register / render / copy, plus the additive ones.
"""
from __future__ import annotations

import ctypes as C
from typing import Iterable, Optional, Sequence

import numpy as np

from . import abi
from .abi import check, lib


def _basis_id(b) -> int:
    if isinstance(b, str):
        return abi.BASIS[b.upper()]
    return int(b)


class Dataset:
    """A MeshlessDataset whose host arrays live in numpy.

    groups: iterable of (positions (N,3), weights (N,), radii (N,) or None, basis).
    """

    def __init__(self, groups: Iterable[Sequence]):
        groups = list(groups)
        self._keep = []
        self._groups = (abi.Group * max(1, len(groups)))()
        for gi, (pos, w, radii, basis) in enumerate(groups):
            pos = np.asarray(pos, dtype=np.float32).reshape(-1, 3)
            n = pos.shape[0]
            cons = np.zeros((max(n, 1), 4), dtype=np.float32)
            cons[:n, :3] = pos
            cons[:n, 3] = np.asarray(w, dtype=np.float32).reshape(-1)
            self._keep.append(cons)
            g = self._groups[gi]
            g.number_of_terms = n
            g.d_number_of_terms = 0
            g.basis_function_id = _basis_id(basis)
            g.h_constraints = cons.ctypes.data_as(C.POINTER(abi.Constraint))
            if radii is not None:
                r = np.ascontiguousarray(np.asarray(radii, dtype=np.float32).reshape(-1))
                if r.shape[0] != n:
                    raise ValueError("radii length differs from positions")
                r = np.concatenate([r, np.zeros(1, np.float32)]) if n == 0 else r
                self._keep.append(r)
                g.h_radii = r.ctypes.data_as(C.POINTER(C.c_float))
            else:
                g.h_radii = None
        self.c = abi.MeshlessDataset()
        self.c.groups = C.cast(self._groups, C.POINTER(abi.Group))
        self.c.number_of_groups = len(groups)

    @classmethod
    def simple(cls, positions, weights, radii=None, basis="GAUSSIAN") -> "Dataset":
        return cls([(positions, weights, radii, basis)])

    @property
    def ptr(self):
        return C.pointer(self.c)

    @property
    def number_of_terms(self) -> int:
        return int(lib().get_number_of_terms(self.ptr))


class Config:
    """One VisConfig (vis_config_create ... vis_config_destroy)."""

    def __init__(
        self,
        cutoff=(64, 64),
        samples=(128, 128),
        step=(1.0, 1.0),
        u=(1.0, 0.0, 0.0),
        v=(0.0, 1.0, 0.0),
        block_length=256,
        partial_sums=1,
        automatic_d_image=True,
        form="plain",
        phi_mode="reference",
        h_bins=0,
        nodes=0,
        h_bin_ratio=0.0,
        cull=False,
        tile=0,
    ):
        L = lib()
        L.vis_clear_last_error()
        p = L.vis_config_create(
            bool(automatic_d_image),
            abi.float2(*step),
            abi.int2(*cutoff),
            abi.float3(*u),
            abi.float3(*v),
            abi.int2(*samples),
            int(block_length),
            int(partial_sums),
        )
        check("vis_config_create")
        if not p:
            raise abi.VisError(1, "vis_config_create returned NULL")
        self.p = p
        self.form = form
        self.phi_mode = phi_mode
        self.cs.number_of_h_bins = int(h_bins)
        self.cs.number_of_interpolation_nodes = int(nodes)
        self.cs.h_bin_ratio = float(h_bin_ratio)
        self.cs.tile = abi.TILE[tile] if isinstance(tile, str) else int(tile)
        self.cs.cull_fully_aliased_terms = bool(cull)

    # -- fields ------------------------------------------------------------
    @property
    def cs(self) -> abi.VisConfig:
        return self.p.contents

    @property
    def form(self) -> str:
        return {v: k for k, v in abi.FORM.items()}[self.cs.form]

    @form.setter
    def form(self, f):
        self.cs.form = abi.FORM[f] if isinstance(f, str) else int(f)

    @property
    def phi_mode(self) -> str:
        return {v: k for k, v in abi.PHI_MODE.items()}[self.cs.phi_mode]

    @phi_mode.setter
    def phi_mode(self, m):
        self.cs.phi_mode = abi.PHI_MODE[m] if isinstance(m, str) else int(m)

    @property
    def cutoff(self):
        return (self.cs._cutoff_frequency.x, self.cs._cutoff_frequency.y)

    @property
    def samples(self):
        return (self.cs._number_of_samples.x, self.cs._number_of_samples.y)

    def set_view(self, u, v, step=None):
        self.cs.u_axis = abi.float3(*u)
        self.cs.v_axis = abi.float3(*v)
        if step is not None:
            self.cs.step_size = abi.float2(*step)

    # -- calls ---------------------------------------------------------------
    def check(self) -> bool:
        return bool(lib().vis_config_check(self.p))

    def register(self, ds: Dataset) -> None:
        lib().vis_register_meshless_dataset(self.p, ds.ptr)
        check("vis_register_meshless_dataset")

    def register_async(self, ds: Dataset) -> None:
        """vis_register_meshless_dataset_async: ``ds``'s host arrays must stay alive and
        unchanged until this config's stream has passed the copies."""
        lib().vis_register_meshless_dataset_async(self.p, ds.ptr)
        check("vis_register_meshless_dataset_async")

    def unregister(self, ds: Dataset) -> None:
        lib().vis_unregister_meshless_dataset(self.p, ds.ptr)
        check("vis_unregister_meshless_dataset")

    def render(self, ds: Dataset) -> None:
        lib().vis_fourier_volume_rendering(ds.ptr, self.p)
        check("vis_fourier_volume_rendering")

    def fourier_transform(self, ds: Dataset) -> None:
        lib().fourier_transform(ds.ptr, self.p)
        check("fourier_transform")

    def autotune(self, ds: Dataset, tolerance: float = 1e-5) -> float:
        """vis_autotune: choose nodes per bin, bin ratio and tile for the MFMA form on ``ds``
        (registered); returns the chosen setting's spectrum time in ms."""
        ms = lib().vis_autotune(ds.ptr, self.p, float(tolerance))
        check("vis_autotune")
        return ms

    def synchronize(self) -> None:
        lib().vis_synchronize(self.p)
        check("vis_synchronize")

    def image(self) -> np.ndarray:
        nx, ny = self.samples
        out = np.empty((nx, ny), dtype=np.float32)
        lib().vis_copy_to_host(self.p, out.ctypes.data_as(C.POINTER(C.c_float)))
        check("vis_copy_to_host")
        return out

    def spectrum(self) -> np.ndarray:
        """Slot-0 spectrum as complex64 [cy, 2cx] (x0 fastest, as _d_freq_image)."""
        cx, cy = self.cutoff
        out = np.empty((cy, 2 * cx), dtype=np.complex64)
        if out.size:
            lib().vis_copy_spectrum_to_host(self.p, out.ctypes.data_as(C.POINTER(abi.float2)))
            check("vis_copy_spectrum_to_host")
        return out

    def set_timing(self, on: bool) -> None:
        lib().vis_set_kernel_timing(self.p, 1 if on else 0)

    def kernel_timing(self):
        ms = C.c_double(0.0)
        n = lib().vis_get_kernel_timing(self.p, C.byref(ms))
        check("vis_get_kernel_timing")
        return n, ms.value

    def reset_timing(self) -> None:
        lib().vis_reset_kernel_timing(self.p)

    def close(self) -> None:
        if getattr(self, "p", None):
            lib().vis_config_destroy(self.p)
            self.p = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def render_batched(datasets: Sequence[Dataset], configs: Sequence[Config]) -> None:
    n = len(datasets)
    ds_arr = (abi.DatasetP * n)(*[d.ptr for d in datasets])
    cf_arr = (abi.VisConfigP * n)(*[c.p for c in configs])
    lib().vis_fourier_volume_rendering_batched(ds_arr, cf_arr, n)
    check("vis_fourier_volume_rendering_batched")


def render(positions, weights, radii=None, basis="GAUSSIAN", **config) -> np.ndarray:
    """One-shot: register, render, copy, unregister; returns the [Nx, Ny] image."""
    ds = Dataset.simple(positions, weights, radii, basis)
    with Config(**config) as cfg:
        cfg.register(ds)
        try:
            cfg.render(ds)
            return cfg.image()
        finally:
            cfg.unregister(ds)


# ------------------------------------------------------------------ HIP runtime ---
# Device memory and streams for callers that hand the library their own device
# arrays (filled groups, a manual _d_image) or their own stream.
_HIP = None


def hip() -> C.CDLL:
    """The HIP runtime the library is linked against (libamdhip64), with the few calls used here."""
    global _HIP
    if _HIP is None:
        lib()  # loads libamdhip64 as the library's dependency
        H = C.CDLL("libamdhip64.so", mode=C.RTLD_GLOBAL)
        H.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        H.hipFree.argtypes = [C.c_void_p]
        H.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        H.hipStreamCreateWithFlags.argtypes = [C.POINTER(C.c_void_p), C.c_uint]
        H.hipStreamDestroy.argtypes = [C.c_void_p]
        H.hipStreamSynchronize.argtypes = [C.c_void_p]
        for f in ("hipMalloc", "hipFree", "hipMemcpy", "hipStreamCreateWithFlags", "hipStreamDestroy",
                  "hipStreamSynchronize", "hipDeviceSynchronize"):
            getattr(H, f).restype = C.c_int
        _HIP = H
    return _HIP


def _hip_check(err: int, what: str) -> None:
    if err != 0:
        raise RuntimeError(f"{what} failed with HIP error {err}")


class DeviceBuffer:
    """hipMalloc'd bytes, optionally filled from a numpy array (hipMemcpy host to device)."""

    H2D, D2H = 1, 2

    def __init__(self, nbytes: int, init: Optional[np.ndarray] = None):
        self.ptr = C.c_void_p()
        self.nbytes = int(nbytes)
        _hip_check(hip().hipMalloc(C.byref(self.ptr), max(1, self.nbytes)), "hipMalloc")
        if init is not None:
            self.upload(init)

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        _hip_check(hip().hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, self.H2D), "hipMemcpy")

    def download(self, dtype, count: int) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        assert out.nbytes <= self.nbytes
        _hip_check(hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, self.D2H), "hipMemcpy")
        return out

    def free(self) -> None:
        if self.ptr.value:
            hip().hipFree(self.ptr)
            self.ptr = C.c_void_p()


def create_stream(non_blocking: bool = True) -> C.c_void_p:
    """A HIP stream of default priority (hipStreamNonBlocking when non_blocking)."""
    s = C.c_void_p()
    _hip_check(hip().hipStreamCreateWithFlags(C.byref(s), 1 if non_blocking else 0), "hipStreamCreateWithFlags")
    return s
