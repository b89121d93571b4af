"""Host-side dataset helpers of meshless.h through the C ABI (no GPU needed):
ownership, term counting, translation and the .sph text format
(meshless.cpp:27-187)."""
import ctypes as C

import numpy as np
import pytest

from libmeshlessvis_amd import abi


def make_simple(L, n, with_radii, basis=2, seed=0):
    rng = np.random.default_rng(seed)
    cons = L.meshless_new_constraints(n)
    radii = L.meshless_new_radii(n) if with_radii else None
    arr = np.ctypeslib.as_array(C.cast(cons, C.POINTER(C.c_float)), shape=(max(n, 1), 4))
    arr[:n] = rng.uniform(-1, 1, (n, 4)).astype(np.float32)
    if with_radii:
        np.ctypeslib.as_array(radii, shape=(max(n, 1),))[:n] = rng.uniform(0.01, 0.1, n).astype(np.float32)
    return L.get_simple_meshless_dataset(n, cons, radii, basis), arr[:n].copy()


def test_simple_dataset_count_shift_delete():
    L = abi.lib()
    ds, arr = make_simple(L, 37, True)
    assert ds.number_of_groups == 1
    assert L.get_number_of_terms(C.byref(ds)) == 37
    L.shift_meshless_dataset(C.byref(ds), 1.0, -2.0, 0.5)
    got = np.ctypeslib.as_array(C.cast(ds.groups[0].h_constraints, C.POINTER(C.c_float)), shape=(37, 4))
    np.testing.assert_array_equal(got[:, :3], (arr[:, :3] + np.array([1.0, -2.0, 0.5], np.float32)).astype(np.float32))
    np.testing.assert_array_equal(got[:, 3], arr[:, 3])
    L.delete_meshless_dataset(ds)


def test_sph_round_trip(tmp_path):
    L = abi.lib()
    a, arr_a = make_simple(L, 20, True, basis=0, seed=1)
    b, arr_b = make_simple(L, 7, False, basis=1, seed=2)
    pair = (abi.MeshlessDataset * 2)(a, b)
    path = str(tmp_path / "x.sph").encode()
    L.vis_clear_last_error()
    L.save_meshless_datasets_to_file(path, pair, 2)
    assert L.vis_get_last_error_code() == 0
    out = abi.DatasetP()
    n = C.c_int(0)
    L.load_meshless_datasets_from_file(path, C.byref(out), C.byref(n))
    assert L.vis_get_last_error_code() == 0
    assert n.value == 2
    for i, (orig, arr, nr, basis) in enumerate([(a, arr_a, True, 0), (b, arr_b, False, 1)]):
        g = out[i].groups[0]
        assert out[i].number_of_groups == 1
        assert g.number_of_terms == arr.shape[0]
        assert g.basis_function_id == basis
        got = np.ctypeslib.as_array(C.cast(g.h_constraints, C.POINTER(C.c_float)), shape=(arr.shape[0], 4))
        np.testing.assert_array_equal(got, arr)  # max_digits10 text is exact
        assert bool(g.h_radii) == nr
    L.delete_meshless_datasets(out, n.value)
    L.delete_meshless_dataset(a)
    L.delete_meshless_dataset(b)


def test_malformed_files_are_errors(tmp_path):
    L = abi.lib()
    out = abi.DatasetP()
    n = C.c_int(-1)
    L.load_meshless_datasets_from_file(str(tmp_path / "missing.sph").encode(), C.byref(out), C.byref(n))
    assert L.vis_get_last_error_code() == 1 and n.value == 0 and not out
    bad = tmp_path / "bad.sph"
    bad.write_text("1\n1\n3 cubic unused 0\n")
    L.load_meshless_datasets_from_file(str(bad).encode(), C.byref(out), C.byref(n))
    assert L.vis_get_last_error_code() == 1 and b"unknown basis" in L.vis_get_last_error()
    trunc = tmp_path / "trunc.sph"
    trunc.write_text("1\n1\n3 sph unused 0\n0 0 0\n1 1\n")
    L.load_meshless_datasets_from_file(str(trunc).encode(), C.byref(out), C.byref(n))
    assert L.vis_get_last_error_code() == 1 and n.value == 0
    empty = tmp_path / "empty.sph"
    empty.write_text("0\n")
    L.vis_clear_last_error()
    L.load_meshless_datasets_from_file(str(empty).encode(), C.byref(out), C.byref(n))
    assert L.vis_get_last_error_code() == 0 and n.value == 0 and not out
    L.vis_clear_last_error()
