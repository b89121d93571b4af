// Host-side dataset helpers of include/meshless.h.
//
// Reference counterparts: meshless_vis/meshless.cpp:27-187.  Ownership follows
// the reference: a dataset adopts new[] arrays and delete_meshless_dataset
// releases them with delete[].
//
// The .sph text format (meshless.cpp:86-187): number of datasets; per dataset
// the number of groups, one "count basis unused parameters" line per group,
// then all positions, then all weights, then a radii flag and, if set, all
// radii.  Deviation: the reference's writer takes the radii flag from dataset
// 0 for every dataset (meshless.cpp:178); here each dataset writes its own.
#include "../../include/meshless_vis.h"
#include "errors.h"

#include <fstream>
#include <limits>
#include <new>
#include <string>

using namespace mv;

extern "C" {

MeshlessDataset get_simple_meshless_dataset(int number_of_terms, Constraint* h_constraints, float* h_radii,
                                            BasisFunctionId basis_function_id) {
  MeshlessDataset ds;
  ds.number_of_groups = 1;
  ds.groups = new Group[1]();
  ds.groups[0].basis_function_id = basis_function_id;
  ds.groups[0].number_of_terms = number_of_terms;
  ds.groups[0].d_number_of_terms = 0;
  ds.groups[0].h_constraints = h_constraints;
  ds.groups[0].h_radii = h_radii;
  ds.groups[0].d_constraints = nullptr;
  ds.groups[0].d_radii = nullptr;
  return ds;
}

void delete_meshless_dataset(MeshlessDataset ds) {
  for (int i = 0; i < ds.number_of_groups; ++i) {
    delete[] ds.groups[i].h_constraints;
    delete[] ds.groups[i].h_radii;
  }
  delete[] ds.groups;
}

void delete_meshless_datasets(MeshlessDataset* datasets, int number_of_datasets) {
  if (!datasets) return;
  for (int i = 0; i < number_of_datasets; ++i) delete_meshless_dataset(datasets[i]);
  delete[] datasets;
}

int get_number_of_terms(MeshlessDataset* ds) {
  if (!ds) return 0;
  int n = 0;
  for (int i = 0; i < ds->number_of_groups; ++i) n += ds->groups[i].number_of_terms;
  return n;
}

void shift_meshless_dataset(MeshlessDataset* ds, float x, float y, float z) {
  if (!ds) return;
  for (int j = 0; j < ds->number_of_groups; ++j) {
    Group& g = ds->groups[j];
    for (int k = 0; k < g.number_of_terms; ++k) {
      g.h_constraints[k].position.x += x;
      g.h_constraints[k].position.y += y;
      g.h_constraints[k].position.z += z;
    }
  }
}

void shift_meshless_datasets(MeshlessDataset* ds, int number_of_datasets, float x, float y, float z) {
  for (int i = 0; i < number_of_datasets; ++i) shift_meshless_dataset(ds + i, x, y, z);
}

static bool basis_from_name(const std::string& s, BasisFunctionId* id) {
  if (s == "sph") *id = SPH;
  else if (s == "gaussian") *id = GAUSSIAN;
  else if (s == "wendland_d3_c2") *id = WENDLAND_D3_C2;
  else return false;
  return true;
}

static const char* basis_name(BasisFunctionId id) {
  switch (id) {
    case SPH: return "sph";
    case GAUSSIAN: return "gaussian";
    default: return "wendland_d3_c2";
  }
}

void load_meshless_datasets_from_file(const char* filename, MeshlessDataset** out, int* number_of_datasets) {
  clear_error();
  if (!out || !number_of_datasets) {
    set_error(VIS_ERROR_INVALID_VALUE, "load_meshless_datasets_from_file: null output");
    return;
  }
  *out = nullptr;
  *number_of_datasets = 0;
  std::ifstream in(filename ? filename : "");
  if (!in) {
    set_error(VIS_ERROR_INVALID_VALUE, std::string("cannot open ") + (filename ? filename : "(null)"));
    return;
  }
  // Bytes left in the file bound every count read from it: a dataset takes at least one
  // character, a group header or a term at least four tokens (seven characters).  A header
  // claiming more than the file can hold is malformed and allocates nothing.
  in.seekg(0, std::ios::end);
  const long long file_bytes = static_cast<long long>(in.tellg());
  in.seekg(0, std::ios::beg);
  auto left = [&]() {
    const long long at = static_cast<long long>(in.tellg());
    return at < 0 ? 0LL : file_bytes - at;
  };
  int nd = 0;
  if (!(in >> nd) || nd < 0 || nd > left()) {
    set_error(VIS_ERROR_INVALID_VALUE, "malformed dataset count");
    return;
  }
  if (nd == 0) return;
  // Counts come from the file: every array is allocated without throwing, and a count the
  // host cannot hold is reported (VIS_ERROR_OUT_OF_MEMORY) instead of ending the process.
  MeshlessDataset* ds = new (std::nothrow) MeshlessDataset[nd]();
  if (!ds) {
    set_error(VIS_ERROR_OUT_OF_MEMORY, "load_meshless_datasets_from_file: cannot allocate " + std::to_string(nd) +
                                           " datasets");
    return;
  }
  int built = 0;
  auto fail_code = [&](int code, const std::string& why) {
    for (int i = 0; i < built; ++i) delete_meshless_dataset(ds[i]);
    delete[] ds;
    set_error(code, why);
  };
  auto fail = [&](const std::string& why) { fail_code(VIS_ERROR_INVALID_VALUE, "malformed .sph file: " + why); };
  auto no_memory = [&](const std::string& what) {
    fail_code(VIS_ERROR_OUT_OF_MEMORY, "load_meshless_datasets_from_file: cannot allocate " + what);
  };
  for (int i = 0; i < nd; ++i) {
    int ng = 0;
    if (!(in >> ng) || ng < 0 || 4LL * ng > left()) return fail("group count");
    ds[i].groups = new (std::nothrow) Group[ng]();
    if (!ds[i].groups) return no_memory(std::to_string(ng) + " groups");
    ds[i].number_of_groups = ng;
    built = i + 1;
    for (int j = 0; j < ng; ++j) {
      std::string name, unused;
      int params = 0;
      Group& g = ds[i].groups[j];
      if (!(in >> g.number_of_terms >> name >> unused >> params) || g.number_of_terms < 0)
        return fail("group header");
      if (!basis_from_name(name, &g.basis_function_id)) return fail("unknown basis '" + name + "'");
      if (4LL * g.number_of_terms > left()) return fail("group header: more terms than the file holds");
      g.h_constraints = new (std::nothrow) Constraint[g.number_of_terms]();
      if (!g.h_constraints) return no_memory(std::to_string(g.number_of_terms) + " constraints");
    }
    for (int j = 0; j < ng; ++j) {
      Group& g = ds[i].groups[j];
      for (int k = 0; k < g.number_of_terms; ++k)
        if (!(in >> g.h_constraints[k].position.x >> g.h_constraints[k].position.y >> g.h_constraints[k].position.z))
          return fail("positions");
    }
    for (int j = 0; j < ng; ++j) {
      Group& g = ds[i].groups[j];
      for (int k = 0; k < g.number_of_terms; ++k)
        if (!(in >> g.h_constraints[k].weight)) return fail("weights");
    }
    int has_radii = 0;
    if (!(in >> has_radii)) return fail("radii flag");
    for (int j = 0; j < ng; ++j) {
      Group& g = ds[i].groups[j];
      if (has_radii) {
        g.h_radii = new (std::nothrow) float[g.number_of_terms]();
        if (!g.h_radii) return no_memory(std::to_string(g.number_of_terms) + " radii");
        for (int k = 0; k < g.number_of_terms; ++k)
          if (!(in >> g.h_radii[k])) return fail("radii");
      } else {
        g.h_radii = nullptr;
      }
    }
  }
  *out = ds;
  *number_of_datasets = nd;
}

void save_meshless_datasets_to_file(const char* filename, MeshlessDataset* ds, int number_of_datasets) {
  clear_error();
  std::ofstream out(filename ? filename : "");
  if (!out) {
    set_error(VIS_ERROR_INVALID_VALUE, std::string("cannot write ") + (filename ? filename : "(null)"));
    return;
  }
  out.precision(std::numeric_limits<float>::max_digits10);
  out << number_of_datasets << "\n";
  for (int i = 0; i < number_of_datasets; ++i) {
    const MeshlessDataset& d = ds[i];
    out << d.number_of_groups << "\n";
    for (int j = 0; j < d.number_of_groups; ++j)
      out << d.groups[j].number_of_terms << " " << basis_name(d.groups[j].basis_function_id) << " unused 0\n";
    for (int j = 0; j < d.number_of_groups; ++j)
      for (int k = 0; k < d.groups[j].number_of_terms; ++k) {
        const float3& p = d.groups[j].h_constraints[k].position;
        out << p.x << " " << p.y << " " << p.z << "\n";
      }
    for (int j = 0; j < d.number_of_groups; ++j)
      for (int k = 0; k < d.groups[j].number_of_terms; ++k) out << d.groups[j].h_constraints[k].weight << "\n";
    const bool radii = d.number_of_groups > 0 && d.groups[0].h_radii != nullptr;
    out << (radii ? 1 : 0) << "\n";
    if (radii)
      for (int j = 0; j < d.number_of_groups; ++j)
        for (int k = 0; k < d.groups[j].number_of_terms; ++k)
          out << (d.groups[j].h_radii ? d.groups[j].h_radii[k] : 0.0f) << "\n";
  }
  if (!out) set_error(VIS_ERROR_INVALID_VALUE, "write failed");
}

Constraint* meshless_new_constraints(int number_of_terms) {
  if (number_of_terms < 0) return nullptr;
  return new (std::nothrow) Constraint[number_of_terms > 0 ? number_of_terms : 1]();
}

float* meshless_new_radii(int number_of_terms) {
  if (number_of_terms < 0) return nullptr;
  return new (std::nothrow) float[number_of_terms > 0 ? number_of_terms : 1]();
}

}  // extern "C"
