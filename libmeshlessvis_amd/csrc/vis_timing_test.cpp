// Timing harness: one dataset rendered number_of_runs times.
//
// The reference's timing program, vis_timing_test/main.cpp:24-93, on this
// library's C ABI, with the same seven positional arguments and meaning:
//
//   vis_timing_test number_of_runs number_of_terms cutoff_frequency block_length
//                   number_of_samples factor number_of_partial_sums [form]
//
// A square configuration (step 1, axes (1,0,0) and (0,1,0), cutoff c x c,
// samples n x n); every particle at the origin with weight 1 and the SPH basis,
// with radius 1e-5 (factor 0) or no radii (factor 1: Phi is hoisted out of the
// sum).  Registered once, rendered number_of_runs times, unregistered.  With
// another argument count the usage text is printed and the program returns 0, as
// the reference's does.
//
// Additions: an eighth argument names the form (plain, mfma_f16,
// mfma_split_bf16, mfma_split_f16; default plain, the reference's kernel); the
// elapsed time of the runs (after one untimed render, synchronised on the
// config's stream), the time per render and the particle x frequency terms per
// second are printed after the reference's last line; an error reported by the
// library ends the program with status 1.
#include "meshless_vis.h"

#include <chrono>
#include <cstdio>
#include <cstring>
#include <iostream>

namespace {

bool failed(const char* what) {
  if (vis_get_last_error_code() == VIS_SUCCESS) return false;
  std::cerr << "vis_timing_test: " << what << ": " << vis_get_last_error() << std::endl;
  return true;
}

int form_of(const char* s) {
  if (!std::strcmp(s, "plain")) return VIS_FORM_PLAIN;
  if (!std::strcmp(s, "mfma_f16")) return VIS_FORM_MFMA_F16;
  if (!std::strcmp(s, "mfma_split_bf16")) return VIS_FORM_MFMA_SPLIT_BF16;
  if (!std::strcmp(s, "mfma_split_f16")) return VIS_FORM_MFMA_SPLIT_F16;
  return -1;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 8 && argc != 9) {
    std::cout << std::endl
              << "vis_timing_test number_of_runs number_of_terms cutoff_frequency block_length number_of_samples "
                 "factor number_of_partial_sums [form]"
              << std::endl;
    std::cout << "a good value for block length is 256" << std::endl;
    std::cout << "factor should usually be 0, 1 is a special case that should be faster" << std::endl;
    std::cout << "form: plain (default), mfma_f16, mfma_split_bf16 or mfma_split_f16" << std::endl << std::endl;
    return 0;
  }
  int number_of_runs = 0, number_of_terms = 0, block_length = 0, factor = 0, number_of_partial_sums = 0;
  int2 cutoff_frequency = make_int2(0, 0), number_of_samples = make_int2(0, 0);
  std::sscanf(argv[1], "%d", &number_of_runs);
  std::sscanf(argv[2], "%d", &number_of_terms);
  std::sscanf(argv[3], "%d", &cutoff_frequency.x);
  std::sscanf(argv[4], "%d", &block_length);
  std::sscanf(argv[5], "%d", &number_of_samples.x);
  std::sscanf(argv[6], "%d", &factor);
  std::sscanf(argv[7], "%d", &number_of_partial_sums);
  const int form = argc == 9 ? form_of(argv[8]) : VIS_FORM_PLAIN;
  if (form < 0) {
    std::cerr << "vis_timing_test: unknown form " << argv[8] << std::endl;
    return 1;
  }
  if (number_of_terms < 0 || number_of_runs < 0) {
    std::cerr << "vis_timing_test: negative count" << std::endl;
    return 1;
  }

  std::cout << "number of runs: " << number_of_runs << std::endl;
  std::cout << "number of terms: " << number_of_terms << std::endl;
  std::cout << "cutoff frequency: " << cutoff_frequency.x << std::endl;
  std::cout << "block length: " << block_length << std::endl;
  std::cout << "number of samples: " << number_of_samples.x << std::endl;
  std::cout << "factoring out fourier transform: " << factor << std::endl;
  std::cout << "number of partial sums: " << number_of_partial_sums << std::endl;

  number_of_samples.y = number_of_samples.x;
  cutoff_frequency.y = cutoff_frequency.x;

  std::cout << "initializing vis config" << std::endl;
  VisConfig* vis_config = vis_config_create(true, make_float2(1.0f, 1.0f), cutoff_frequency, make_float3(1.0f, 0.0f, 0.0f),
                                            make_float3(0.0f, 1.0f, 0.0f), number_of_samples, block_length,
                                            number_of_partial_sums);
  if (!vis_config || failed("vis_config_create")) return 1;
  vis_config->form = form;
  if (!vis_config_check(vis_config))
    std::cout << "Warning: Invalid Configuration" << std::endl;
  else
    std::cout << "Valid Configuration" << std::endl;

  std::cout << "initializing data" << std::endl;
  Constraint* h_constraints = meshless_new_constraints(number_of_terms);
  float* h_radii = factor ? nullptr : meshless_new_radii(number_of_terms);
  if (!h_constraints || (!factor && !h_radii)) {
    std::cerr << "vis_timing_test: out of host memory" << std::endl;
    return 1;
  }
  for (int i = 0; i != number_of_terms; ++i) {
    h_constraints[i].weight = 1.0f;
    h_constraints[i].position = make_float3(0.0f, 0.0f, 0.0f);
    if (h_radii) h_radii[i] = 1e-5f;
  }
  MeshlessDataset meshless_dataset = get_simple_meshless_dataset(number_of_terms, h_constraints, h_radii, SPH);
  std::cout << "loading data" << std::endl;
  vis_register_meshless_dataset(vis_config, &meshless_dataset);
  bool ok = !failed("vis_register_meshless_dataset");

  std::cout << "running code to measure performance" << std::endl;
  double seconds = 0.0;
  // Each call's error is read before the next call (every entry point starts by clearing it).
  if (ok && number_of_runs > 0) {
    vis_fourier_volume_rendering(&meshless_dataset, vis_config);  // untimed: plan, buffers
    ok = !failed("vis_fourier_volume_rendering");
    if (ok) {
      vis_synchronize(vis_config);
      ok = !failed("vis_synchronize");
    }
    const auto t0 = std::chrono::steady_clock::now();
    for (int k = 0; ok && k != number_of_runs; ++k) {
      vis_fourier_volume_rendering(&meshless_dataset, vis_config);
      ok = !failed("vis_fourier_volume_rendering");
    }
    if (ok) {
      vis_synchronize(vis_config);
      ok = !failed("vis_synchronize");
    }
    seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  std::cout << "finished running tests" << std::endl << std::endl;
  if (ok && number_of_runs > 0) {
    const double terms = static_cast<double>(number_of_terms) * 2.0 * cutoff_frequency.x * cutoff_frequency.y;
    std::printf("elapsed: %.6f s for %d renders (%.4f ms per render, %.4e terms/s)\n", seconds, number_of_runs,
                1e3 * seconds / number_of_runs, terms * number_of_runs / seconds);
  }

  vis_unregister_meshless_dataset(vis_config, &meshless_dataset);
  vis_config_destroy(vis_config);
  delete_meshless_dataset(meshless_dataset);
  return ok ? 0 : 1;
}
