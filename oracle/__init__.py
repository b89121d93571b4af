"""CPU oracle of the Fourier-projection hot path -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package, and only to check results or to time the CPU baseline.  The
product (libmeshlessvis_amd / libmeshless_vis.so) never imports, links or
calls it.

numpy wrappers over liboracle.so (oracle.c, built by oracle/Makefile):

  phi64(basis, mode, r)          Phi in FP64 (reference formulas or exact)
  phi32_port(basis, r)           Phi in FP32 as fourier_transform_cpu.cpp:37-66
  spectrum64(groups, view, ...)  FP64 stored spectrum (fourier_transform_cpu.cpp:73-171)
  port_spectrum32(groups, view)  FP32/OpenMP restatement of the same loop nest
  arrange64(F, ...)              arrange_samples (meshless_vis_cpu.cpp:209-234)
  c2r64(A, Nx, Ny)               unnormalised 2-D C2R (meshless_vis_cpu.cpp:62-64,261)

Pinned by tests/test_oracle_phi.py, tests/test_oracle_sum.py and
tests/test_oracle_fft.py (quadrature, closed forms, invariants, numpy).
The reference's own CPU path cannot be built in this image (its headers need
fftw3.h, GL/glew.h and CUDA's vector_types.h), so there is no oracle/_ref.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
SO = HERE / "liboracle.so"
BASIS = {"SPH": 0, "GAUSSIAN": 1, "WENDLAND_D3_C2": 2}
PHI_MODE = {"reference": 0, "analytic": 1}

_LIB = None


def build(force: bool = False) -> Path:
    """Compile liboracle.so with gcc (always when force)."""
    if force or not SO.exists() or SO.stat().st_mtime < (HERE / "oracle.c").stat().st_mtime:
        subprocess.run(["make", "-s", *(["-B"] if force else []), "-C", str(HERE)], check=True)
    return SO


def _lib() -> C.CDLL:
    global _LIB
    if _LIB is None:
        build()
        L = C.CDLL(str(SO))
        d, f, i = C.c_double, C.c_float, C.c_int
        P = C.POINTER
        L.mvo_phi64.restype = d
        L.mvo_phi64.argtypes = [i, i, d]
        L.mvo_phi64_array.restype = None
        L.mvo_phi64_array.argtypes = [i, i, P(d), P(d), i]
        L.mvo_phi32_port.restype = f
        L.mvo_phi32_port.argtypes = [i, f]
        L.mvo_phi32_port_array.restype = None
        L.mvo_phi32_port_array.argtypes = [i, P(f), P(f), i]
        L.mvo_spectrum64.restype = None
        L.mvo_spectrum64.argtypes = [i, P(i), P(i), P(P(f)), P(P(f)), f, f, P(f), P(f), i, i, i, P(i), i, P(d), i]
        L.mvo_port_spectrum32.restype = None
        L.mvo_port_spectrum32.argtypes = [i, P(i), P(i), P(i), P(P(f)), P(P(f)), f, f, P(f), P(f), i, i, P(f), i, i, i,
                                          i, P(i), i]
        L.mvo_arrange64.restype = None
        L.mvo_arrange64.argtypes = [P(d), i, i, i, i, P(d)]
        L.mvo_c2r64.restype = None
        L.mvo_c2r64.argtypes = [P(d), i, i, P(d), i]
        _LIB = L
    return _LIB


def _ptr(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def _basis(b) -> int:
    return BASIS[b.upper()] if isinstance(b, str) else int(b)


def _mode(m) -> int:
    return PHI_MODE[m] if isinstance(m, str) else int(m)


def phi64(basis, mode, r) -> np.ndarray:
    r = np.ascontiguousarray(np.asarray(r, dtype=np.float64).reshape(-1))
    out = np.empty_like(r)
    _lib().mvo_phi64_array(_basis(basis), _mode(mode), _ptr(r, C.c_double), _ptr(out, C.c_double), r.size)
    return out


def phi32_port(basis, r) -> np.ndarray:
    r = np.ascontiguousarray(np.asarray(r, dtype=np.float32).reshape(-1))
    out = np.empty_like(r)
    _lib().mvo_phi32_port_array(_basis(basis), _ptr(r, C.c_float), _ptr(out, C.c_float), r.size)
    return out


class _Groups:
    """Marshals [(positions, weights, radii|None, basis), ...] for the C oracle."""

    def __init__(self, groups):
        self.keep = []
        n = len(groups)
        self.counts = np.zeros(max(n, 1), np.int32)
        self.bases = np.zeros(max(n, 1), np.int32)
        self.cons = (C.POINTER(C.c_float) * max(n, 1))()
        self.radii = (C.POINTER(C.c_float) * max(n, 1))()
        for g, (pos, w, radii, basis) in enumerate(groups):
            pos = np.asarray(pos, np.float32).reshape(-1, 3)
            cons = np.zeros((max(pos.shape[0], 1), 4), np.float32)
            cons[: pos.shape[0], :3] = pos
            cons[: pos.shape[0], 3] = np.asarray(w, np.float32).reshape(-1)
            self.keep.append(cons)
            self.counts[g] = pos.shape[0]
            self.bases[g] = _basis(basis)
            self.cons[g] = _ptr(cons, C.c_float)
            if radii is not None:
                r = np.ascontiguousarray(np.asarray(radii, np.float32).reshape(-1))
                r = r if r.size else np.zeros(1, np.float32)
                self.keep.append(r)
                self.radii[g] = _ptr(r, C.c_float)
            else:
                self.radii[g] = None
        self.n = n


def spectrum64(groups, cutoff, step, u, v, phi_mode="reference", idx=None, nthreads=0) -> np.ndarray:
    """FP64 stored spectrum.  Shape (cy, 2cx) complex128, or (len(idx),) at flat indices idx."""
    G = _Groups(groups)
    cx, cy = int(cutoff[0]), int(cutoff[1])
    u = np.asarray(u, np.float32)
    v = np.asarray(v, np.float32)
    if idx is None:
        n_out = 2 * cx * cy
        idx_arr = None
    else:
        idx_arr = np.ascontiguousarray(np.asarray(idx, np.int32).reshape(-1))
        n_out = idx_arr.size
    out = np.zeros(2 * max(n_out, 1), np.float64)
    _lib().mvo_spectrum64(
        G.n, _ptr(G.counts, C.c_int), _ptr(G.bases, C.c_int), G.cons, G.radii,
        float(step[0]), float(step[1]), _ptr(u, C.c_float), _ptr(v, C.c_float), cx, cy, _mode(phi_mode),
        _ptr(idx_arr, C.c_int) if idx_arr is not None else None, n_out, _ptr(out, C.c_double), int(nthreads),
    )
    z = out[: 2 * n_out].view(np.complex128)
    return z.reshape(cy, 2 * cx) if idx is None else z


def padded_count(n: int, block_length: int, partial_sums: int) -> int:
    """d_number_of_terms of load_into_device (meshless_vis.cu:274-275)."""
    m = block_length * partial_sums
    return m * (n // m + min(1, n % m))


def port_spectrum32(groups, cutoff, step, u, v, block_length=256, partial_sums=1, row_stride=1, nthreads=0,
                    phi_mode="reference", idx=None) -> np.ndarray:
    """FP32 restatement of the reference's CPU direct sum; (cy, 2cx) complex64, or (len(idx),)
    at the flat indices idx.  phi_mode "analytic" swaps the reference's Phi formulas for the exact
    transforms (rounded to FP32), keeping the reference's loop, FP32 phase and FP32 sums."""
    G = _Groups(groups)
    cx, cy = int(cutoff[0]), int(cutoff[1])
    padded = np.array([padded_count(int(c), block_length, partial_sums) for c in G.counts], np.int32)
    u = np.asarray(u, np.float32)
    v = np.asarray(v, np.float32)
    idx_arr = None if idx is None else np.ascontiguousarray(np.asarray(idx, np.int32).reshape(-1))
    n_out = 2 * cx * cy if idx_arr is None else idx_arr.size
    out = np.zeros(2 * max(n_out, 1), np.float32)
    _lib().mvo_port_spectrum32(
        G.n, _ptr(G.counts, C.c_int), _ptr(padded, C.c_int), _ptr(G.bases, C.c_int), G.cons, G.radii,
        float(step[0]), float(step[1]), _ptr(u, C.c_float), _ptr(v, C.c_float), cx, cy,
        _ptr(out, C.c_float), int(block_length), int(row_stride), int(nthreads), _mode(phi_mode),
        _ptr(idx_arr, C.c_int) if idx_arr is not None else None, n_out,
    )
    z = out[: 2 * n_out].view(np.complex64)
    return z.reshape(cy, 2 * cx) if idx_arr is None else z


def arrange64(F, cx, cy, Nx, Ny) -> np.ndarray:
    F = np.ascontiguousarray(np.asarray(F, np.complex128).reshape(-1))
    out = np.zeros((Nx, Ny // 2 + 1), np.complex128)
    _lib().mvo_arrange64(_ptr(F.view(np.float64), C.c_double), int(cx), int(cy), int(Nx), int(Ny),
                         _ptr(out.view(np.float64), C.c_double))
    return out


def c2r64(A, Nx, Ny, nthreads=0) -> np.ndarray:
    A = np.ascontiguousarray(np.asarray(A, np.complex128))
    img = np.zeros((Nx, Ny), np.float64)
    _lib().mvo_c2r64(_ptr(A.view(np.float64), C.c_double), int(Nx), int(Ny), _ptr(img, C.c_double), int(nthreads))
    return img


def image64(groups, cutoff, samples, step, u, v, phi_mode="reference", nthreads=0) -> np.ndarray:
    """The whole render in FP64: spectrum -> arrange -> C2R; (Nx, Ny) float64."""
    F = spectrum64(groups, cutoff, step, u, v, phi_mode, nthreads=nthreads)
    A = arrange64(F, cutoff[0], cutoff[1], samples[0], samples[1])
    return c2r64(A, samples[0], samples[1], nthreads)


def rel_l2(a, b) -> float:
    """||a - b|| / ||b||."""
    a = np.asarray(a, np.complex128 if np.iscomplexobj(a) or np.iscomplexobj(b) else np.float64)
    b = np.asarray(b, a.dtype)
    den = np.linalg.norm(b.ravel())
    return float(np.linalg.norm((a - b).ravel()) / (den if den > 0 else 1.0))


def stratified_sample(cutoff, rng, low=8, n_random=512):
    """Flat spectrum indices [y][x0] for checks at sizes where the whole-plane FP64 sum is out of
    reach: every index with |x| <= low and y <= low (where the energy sits) plus n_random drawn
    from the rest, and weights such that sum(weight * |z[idx]|^2) estimates the whole plane's."""
    cx, cy = int(cutoff[0]), int(cutoff[1])
    x0 = np.arange(2 * cx)
    x = np.where(x0 > cx, x0 - 2 * cx, x0)
    low_idx = np.array([y * 2 * cx + i for y in range(min(cy, low + 1)) for i in range(2 * cx) if abs(x[i]) <= low],
                       dtype=np.int64)
    rest = np.setdiff1d(np.arange(2 * cx * cy), low_idx)
    pick = rng.choice(rest, size=min(n_random, rest.size), replace=False)
    weight = np.concatenate([np.ones(low_idx.size), np.full(pick.size, rest.size / max(1, pick.size))])
    return np.concatenate([low_idx, pick]), weight


def sampled_rel_l2(F, F64_at, idx, weight) -> float:
    """Relative L2 of a spectrum F (any shape, flat [y][x0] order) against FP64 values at idx."""
    err = np.abs(np.asarray(F).reshape(-1)[idx].astype(np.complex128) - F64_at) ** 2
    return float(np.sqrt((weight * err).sum() / (weight * np.abs(F64_at) ** 2).sum()))
