// Host-only driver of the .sph reader and writer (meshless.cpp), built with
// AddressSanitizer and UndefinedBehaviorSanitizer by tests/test_sanitize_host.py.
//
//   sph_driver IN.sph [OUT.sph]
//
// Loads IN; prints "<error code> <dataset count> <total terms>"; on success
// shifts every dataset, writes OUT (if given) and frees everything.
#include "meshless_vis.h"

#include <cstdio>

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  MeshlessDataset* ds = nullptr;
  int n = -1;
  load_meshless_datasets_from_file(argv[1], &ds, &n);
  const int code = vis_get_last_error_code();
  long long terms = 0;
  if (code == VIS_SUCCESS) {
    for (int i = 0; i < n; ++i) terms += get_number_of_terms(ds + i);
    shift_meshless_datasets(ds, n, 1.0f, -2.0f, 0.5f);
    if (argc > 2) {
      save_meshless_datasets_to_file(argv[2], ds, n);
      if (vis_get_last_error_code() != VIS_SUCCESS) return 3;
    }
    delete_meshless_datasets(ds, n);
  } else if (ds != nullptr || n != 0) {
    return 4;  // a failed load must leave no datasets behind
  }
  std::printf("%d %d %lld\n", code, n, terms);
  return 0;
}
