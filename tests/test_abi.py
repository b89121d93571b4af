"""The C ABI: the library loads, exports every function the headers declare, the
ctypes mirror has the headers' struct layouts, and failures come back as error
codes (no GPU is needed for any of this)."""
import re
import shutil
import subprocess
from pathlib import Path

import ctypes as C
import pytest

from libmeshlessvis_amd import abi

ROOT = Path(__file__).resolve().parent.parent
HEADERS = [ROOT / "include" / "meshless.h", ROOT / "include" / "meshless_vis.h"]


def declared_functions():
    names = []
    for h in HEADERS:
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        for m in re.finditer(r"^\s*[A-Za-z_][\w\s\*]*?\b([a-z_][a-z0-9_]*)\s*\([^;{]*\)\s*;", text, flags=re.M):
            names.append(m.group(1))
    return sorted(set(names))


def test_every_declared_function_is_exported_and_bound():
    names = declared_functions()
    assert len(names) >= 30, names
    L = abi.lib()
    for n in names:
        assert hasattr(L, n), n
        assert n in abi.PROTOTYPES, n
    assert set(abi.PROTOTYPES) == set(names)


LAYOUT_PROG = r"""
#include <cstddef>
#include <cstdio>
#include "meshless_vis.h"
#define F(s, f) std::printf("%s.%s %zu\n", #s, #f, offsetof(s, f));
int main() {
  std::printf("sizeof.Constraint %zu\nsizeof.Group %zu\nsizeof.MeshlessDataset %zu\nsizeof.VisConfig %zu\n",
              sizeof(Constraint), sizeof(Group), sizeof(MeshlessDataset), sizeof(VisConfig));
  F(Constraint, position) F(Constraint, weight)
  F(Group, number_of_terms) F(Group, d_number_of_terms) F(Group, basis_function_id) F(Group, h_constraints)
  F(Group, d_constraints) F(Group, h_radii) F(Group, d_radii)
  F(MeshlessDataset, groups) F(MeshlessDataset, number_of_groups)
  F(VisConfig, step_size) F(VisConfig, _cutoff_frequency) F(VisConfig, u_axis) F(VisConfig, v_axis)
  F(VisConfig, cull_fully_aliased_terms) F(VisConfig, _automatic_d_image) F(VisConfig, _plan)
  F(VisConfig, _number_of_samples) F(VisConfig, block_length) F(VisConfig, _scale) F(VisConfig, _d_image)
  F(VisConfig, _d_freq_image) F(VisConfig, _number_of_partial_sums) F(VisConfig, _d_freq_image_arranged)
  F(VisConfig, form) F(VisConfig, phi_mode) F(VisConfig, number_of_h_bins) F(VisConfig, number_of_interpolation_nodes) F(VisConfig, h_bin_ratio) F(VisConfig, tile)
  return 0;
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_ctypes_layout_matches_headers(tmp_path):
    src = tmp_path / "layout.cpp"
    src.write_text(LAYOUT_PROG)
    exe = tmp_path / "layout"
    subprocess.run(["g++", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", f"-I{ROOT / 'include'}",
                    str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    structs = {"Constraint": abi.Constraint, "Group": abi.Group, "MeshlessDataset": abi.MeshlessDataset,
               "VisConfig": abi.VisConfig}
    for line in filter(None, out):
        key, val = line.split()
        s, f = key.split(".")
        if s == "sizeof":
            assert C.sizeof(structs[f]) == int(val), line
        else:
            assert getattr(structs[s], f).offset == int(val), line


def test_failures_are_error_codes_not_aborts():
    L = abi.lib()
    L.vis_clear_last_error()
    # negative cutoff is rejected before any device call
    p = L.vis_config_create(True, abi.float2(1, 1), abi.int2(-1, 4), abi.float3(1, 0, 0), abi.float3(0, 1, 0),
                            abi.int2(16, 16), 256, 1)
    assert not p
    assert L.vis_get_last_error_code() == 1
    assert b"negative" in L.vis_get_last_error()
    L.vis_clear_last_error()
    assert L.vis_get_last_error_code() == 0
    # null arguments
    L.vis_fourier_volume_rendering(None, None)
    assert L.vis_get_last_error_code() == 1
    L.vis_opengl_fourier_volume_rendering(None, None, 0)
    assert L.vis_get_last_error_code() == 1
    L.vis_debug_batch_graph_counts(None, None, None)
    assert L.vis_get_last_error_code() == 1
    L.vis_clear_last_error()


def test_config_create_without_device_reports_hip_error():
    L = abi.lib()
    L.vis_clear_last_error()
    p = L.vis_config_create(True, abi.float2(1, 1), abi.int2(4, 4), abi.float3(1, 0, 0), abi.float3(0, 1, 0),
                            abi.int2(16, 16), 256, 1)
    if p:  # a GPU is present: nothing to check here
        L.vis_config_destroy(p)
        pytest.skip("GPU present")
    assert L.vis_get_last_error_code() == 3
    L.vis_clear_last_error()
