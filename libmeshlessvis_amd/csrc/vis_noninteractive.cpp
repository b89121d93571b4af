// Batch renderer: every dataset of a .sph file -> one grey image each.
//
// The command-line program of the reference, vis_noninteractive/main.cpp:37-93:
// load the file, shift the positions by (-5, -5, -5), render every dataset with
// the default configuration (register, render, unregister, copy to host),
// clamp negative intensities to 0, divide by the image's maximum and write the
// image flipped top to bottom, named <prefix><3-digit index>.  The reference
// writes JPEG through ImageMagick; this program writes binary PGM (8 or 16
// bits), which needs no library.  Additions: the form / Phi mode / samples /
// cutoff / step of the configuration can be set, registration is asynchronous
// by default (the upload of dataset k + 1 overlaps the render of dataset k),
// and --from-raw converts a float32 image file without rendering (no GPU).
//
// Deliberate difference: an image whose maximum is <= 0 (the reference divides
// by it) is written as all zeros.
#include "meshless_vis.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

struct Options {
  std::string file = "../data/cartwheel.sph";  // vis_noninteractive/main.cpp:41
  std::string prefix = "image_";               // vis_noninteractive/main.cpp:42
  bool shift = true;
  float sx = -5.0f, sy = -5.0f, sz = -5.0f;  // vis_noninteractive/main.cpp:46
  int depth = 8;
  bool async = true;
  int form = VIS_FORM_PLAIN;
  int phi = VIS_PHI_REFERENCE;
  int samples = 0, cutoff = 0;
  float step = 0.0f;
  std::string raw;
  int raw_nx = 0, raw_ny = 0;
};

void usage() {
  std::fprintf(stderr,
               "usage: vis_noninteractive [file.sph [prefix]] [options]\n"
               "  --form plain|mfma_f16|mfma_split_bf16|mfma_split_f16   (default plain)\n"
               "  --phi reference|analytic                               (default reference)\n"
               "  --samples N --cutoff C --step S      square image / cutoff, isotropic step\n"
               "  --shift X Y Z | --no-shift           (default: shift by -5 -5 -5)\n"
               "  --depth 8|16                         PGM bits per pixel (default 8)\n"
               "  --sync                               synchronous registration\n"
               "  --from-raw FILE NX NY                convert a float32 [NX][NY] image, no rendering\n");
}

bool parse(int argc, char** argv, Options* o) {
  int positional = 0;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    auto need = [&](int k) {
      if (i + k >= argc) {
        std::fprintf(stderr, "vis_noninteractive: %s needs %d argument(s)\n", a.c_str(), k);
        return false;
      }
      return true;
    };
    if (a == "--form") {
      if (!need(1)) return false;
      const std::string v = argv[++i];
      if (v == "plain") o->form = VIS_FORM_PLAIN;
      else if (v == "mfma_f16") o->form = VIS_FORM_MFMA_F16;
      else if (v == "mfma_split_bf16") o->form = VIS_FORM_MFMA_SPLIT_BF16;
      else if (v == "mfma_split_f16") o->form = VIS_FORM_MFMA_SPLIT_F16;
      else return false;
    } else if (a == "--phi") {
      if (!need(1)) return false;
      const std::string v = argv[++i];
      if (v == "reference") o->phi = VIS_PHI_REFERENCE;
      else if (v == "analytic") o->phi = VIS_PHI_ANALYTIC;
      else return false;
    } else if (a == "--samples") {
      if (!need(1)) return false;
      o->samples = std::atoi(argv[++i]);
    } else if (a == "--cutoff") {
      if (!need(1)) return false;
      o->cutoff = std::atoi(argv[++i]);
    } else if (a == "--step") {
      if (!need(1)) return false;
      o->step = static_cast<float>(std::atof(argv[++i]));
    } else if (a == "--shift") {
      if (!need(3)) return false;
      o->shift = true;
      o->sx = static_cast<float>(std::atof(argv[++i]));
      o->sy = static_cast<float>(std::atof(argv[++i]));
      o->sz = static_cast<float>(std::atof(argv[++i]));
    } else if (a == "--no-shift") {
      o->shift = false;
    } else if (a == "--depth") {
      if (!need(1)) return false;
      o->depth = std::atoi(argv[++i]);
      if (o->depth != 8 && o->depth != 16) return false;
    } else if (a == "--sync") {
      o->async = false;
    } else if (a == "--from-raw") {
      if (!need(3)) return false;
      o->raw = argv[++i];
      o->raw_nx = std::atoi(argv[++i]);
      o->raw_ny = std::atoi(argv[++i]);
      if (o->raw_nx <= 0 || o->raw_ny <= 0) return false;
    } else if (a == "-h" || a == "--help") {
      return false;
    } else if (!a.empty() && a[0] == '-') {
      std::fprintf(stderr, "vis_noninteractive: unknown option %s\n", a.c_str());
      return false;
    } else if (positional == 0) {
      o->file = a;
      ++positional;
    } else if (positional == 1) {
      o->prefix = a;
      ++positional;
    } else {
      return false;
    }
  }
  return true;
}

// vis_noninteractive/main.cpp:70-84 and :28-34: the image memory read as `height`
// rows of `width` pixels (width = samples.x), negatives clamped to 0, divided by the
// maximum, rows written bottom to top (ImageMagick's flip), quantised to the depth.
bool write_pgm(const std::string& name, const float* img, int width, int height, int depth) {
  const size_t n = static_cast<size_t>(width) * height;
  float mmax = img[0];
  for (size_t i = 0; i < n; ++i)
    if (img[i] > mmax) mmax = img[i];
  const int maxval = depth == 16 ? 65535 : 255;
  FILE* f = std::fopen(name.c_str(), "wb");
  if (!f) {
    std::fprintf(stderr, "vis_noninteractive: cannot write %s\n", name.c_str());
    return false;
  }
  std::fprintf(f, "P5\n%d %d\n%d\n", width, height, maxval);
  std::vector<unsigned char> row(static_cast<size_t>(width) * (depth == 16 ? 2 : 1));
  for (int r = 0; r < height; ++r) {
    const float* src = img + static_cast<size_t>(height - 1 - r) * width;
    for (int c = 0; c < width; ++c) {
      float v = src[c] < 0.0f ? 0.0f : src[c];
      v = mmax > 0.0f ? v / mmax : 0.0f;
      const long q = std::lround(std::fmin(1.0f, v) * static_cast<float>(maxval));
      if (depth == 16) {
        row[2 * c] = static_cast<unsigned char>(q >> 8);
        row[2 * c + 1] = static_cast<unsigned char>(q & 0xff);
      } else {
        row[c] = static_cast<unsigned char>(q);
      }
    }
    if (std::fwrite(row.data(), 1, row.size(), f) != row.size()) {
      std::fclose(f);
      return false;
    }
  }
  return std::fclose(f) == 0;
}

std::string frame_name(const std::string& prefix, int k) {
  char idx[32];
  std::snprintf(idx, sizeof idx, "%03d", k);  // vis_noninteractive/main.cpp:53-58
  return prefix + idx + ".pgm";
}

bool fail(const char* what) {
  std::fprintf(stderr, "vis_noninteractive: %s failed: %s\n", what, vis_get_last_error());
  return false;
}

bool convert_raw(const Options& o) {
  const size_t n = static_cast<size_t>(o.raw_nx) * o.raw_ny;
  FILE* f = std::fopen(o.raw.c_str(), "rb");
  // The file must hold the whole image before anything is allocated for it.
  long bytes = -1;
  if (f && std::fseek(f, 0, SEEK_END) == 0) bytes = std::ftell(f);
  if (f && (bytes < 0 || static_cast<unsigned long>(bytes) / sizeof(float) < n || std::fseek(f, 0, SEEK_SET) != 0)) {
    std::fclose(f);
    f = nullptr;
  }
  std::vector<float> img(f ? n : 0);
  if (!f || std::fread(img.data(), sizeof(float), n, f) != n) {
    if (f) std::fclose(f);
    std::fprintf(stderr, "vis_noninteractive: cannot read %zu floats from %s\n", n, o.raw.c_str());
    return false;
  }
  std::fclose(f);
  const std::string name = frame_name(o.prefix, 0);
  std::printf("%s\n", name.c_str());
  return write_pgm(name, img.data(), o.raw_nx, o.raw_ny, o.depth);
}

bool render_all(const Options& o) {
  MeshlessDataset* ds = nullptr;
  int n = 0;
  vis_clear_last_error();
  load_meshless_datasets_from_file(o.file.c_str(), &ds, &n);
  if (vis_get_last_error_code() != 0) return fail("load_meshless_datasets_from_file");
  if (n == 0) return true;  // an empty file: nothing to render
  if (o.shift) shift_meshless_datasets(ds, n, o.sx, o.sy, o.sz);
  VisConfig* c = vis_config_get_default();
  if (!c) return fail("vis_config_get_default");
  c->form = o.form;
  c->phi_mode = o.phi;
  if (o.step > 0.0f) c->step_size = make_float2(o.step, o.step);
  bool ok = true;
  if (o.cutoff > 0) {
    vis_config_change_cutoff_frequency(c, make_int2(o.cutoff, o.cutoff));
    ok = vis_get_last_error_code() == 0 || fail("vis_config_change_cutoff_frequency");
  }
  if (ok && o.samples > 0) {
    vis_config_change_number_of_samples(c, make_int2(o.samples, o.samples));
    ok = vis_get_last_error_code() == 0 || fail("vis_config_change_number_of_samples");
  }
  if (ok && !vis_config_check(c)) ok = fail("vis_config_check");
  const int w = c->_number_of_samples.x, h = c->_number_of_samples.y;
  std::vector<float> img(static_cast<size_t>(w) * h);
  auto reg = [&](int k) {
    if (o.async)
      vis_register_meshless_dataset_async(c, &ds[k]);
    else
      vis_register_meshless_dataset(c, &ds[k]);
    return vis_get_last_error_code() == 0 || fail("vis_register_meshless_dataset");
  };
  if (ok && n > 0) ok = reg(0);
  for (int k = 0; ok && k < n; ++k) {
    const std::string name = frame_name(o.prefix, k);
    std::printf("%s\n", name.c_str());
    vis_fourier_volume_rendering(&ds[k], c);
    if (vis_get_last_error_code() != 0) {
      ok = fail("vis_fourier_volume_rendering");
      break;
    }
    // The next dataset's upload is enqueued behind this render and overlaps it.
    if (k + 1 < n && !reg(k + 1)) {
      ok = false;
      break;
    }
    // Checked before the unregister, which clears the last error.
    vis_copy_to_host(c, img.data());
    if (vis_get_last_error_code() != 0) {
      ok = fail("vis_copy_to_host");
      vis_unregister_meshless_dataset(c, &ds[k]);
      break;
    }
    vis_unregister_meshless_dataset(c, &ds[k]);
    ok = write_pgm(name, img.data(), w, h, o.depth);
  }
  vis_config_destroy(c);
  delete_meshless_datasets(ds, n);
  return ok;
}

}  // namespace

int main(int argc, char** argv) {
  Options o;
  if (!parse(argc, argv, &o)) {
    usage();
    return 2;
  }
  const bool ok = o.raw.empty() ? render_all(o) : convert_raw(o);
  return ok ? 0 : 1;
}
