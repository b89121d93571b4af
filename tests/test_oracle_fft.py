"""Pins of the oracle's arrange and C2R (the reference's arrange_samples + FFTW/cuFFT C2R).

  * arrange: every spectrum entry lands where the closed-form index map puts it
    (x0 == cx dropped, x0 > cx -> row x0 - 2cx + Nx, column y), the rest is zero;
  * C2R: equal to numpy.fft.irfft2(norm="forward") (pocketfft, an independent
    library routine) on non-Hermitian inputs, even and odd sizes; the sum of the
    image is Nx*Ny*Re(A[0,0]); a single axis-0 frequency gives a cosine.
"""
import numpy as np
import pytest

import oracle


@pytest.mark.parametrize("cx,cy,Nx,Ny", [(3, 2, 8, 6), (4, 4, 8, 8), (2, 3, 7, 9)])
def test_arrange_index_map(cx, cy, Nx, Ny):
    F = (np.arange(2 * cx * cy) + 1).astype(np.complex128) * (1 + 0.5j)
    A = oracle.arrange64(F, cx, cy, Nx, Ny)
    expect = np.zeros((Nx, Ny // 2 + 1), np.complex128)
    for y in range(cy):
        for x0 in range(2 * cx):
            if x0 == cx:
                continue
            row = x0 if x0 < cx else x0 - 2 * cx + Nx
            expect[row, y] = F[y * 2 * cx + x0]
    assert np.array_equal(A, expect)


@pytest.mark.parametrize("Nx,Ny", [(8, 6), (5, 7), (16, 16), (6, 9)])
def test_c2r_matches_numpy_irfft2(Nx, Ny):
    rng = np.random.default_rng(Nx * 100 + Ny)
    A = rng.normal(size=(Nx, Ny // 2 + 1)) + 1j * rng.normal(size=(Nx, Ny // 2 + 1))
    img = oracle.c2r64(A, Nx, Ny)
    ref = np.fft.irfft2(A, s=(Nx, Ny), norm="forward")
    assert np.max(np.abs(img - ref)) <= 1e-11 * np.abs(ref).max()


def test_c2r_sum_and_single_frequency():
    Nx, Ny = 12, 10
    A = np.zeros((Nx, Ny // 2 + 1), np.complex128)
    A[0, 0] = 2.5 + 7.0j  # imaginary part of the DC term is ignored
    img = oracle.c2r64(A, Nx, Ny)
    assert img.sum() == pytest.approx(Nx * Ny * 2.5, rel=1e-13)
    A[:] = 0
    A[1, 0] = 1.0
    img = oracle.c2r64(A, Nx, Ny)
    r = np.arange(Nx)[:, None]
    assert np.max(np.abs(img - np.cos(2 * np.pi * r / Nx) * np.ones((1, Ny)))) <= 1e-13
