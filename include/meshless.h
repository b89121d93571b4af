/*
 * meshless.h -- dataset types and host-side dataset helpers of the MI355X
 * Fourier-projection library (libmeshless_vis.so).
 *
 * Drop-in for the reference's include/meshless.h: same type names, field
 * names, field order and function signatures.  The only differences are the
 * vector types, which come from HIP (<hip/hip_vector_types.h>, the HIP
 * counterpart of CUDA's <vector_types.h>; float3 is 12 bytes, 4-byte aligned,
 * trivially copyable, exactly as in CUDA), and a handful of additive helpers
 * that are marked as such.
 *
 * Like the reference header this is a C++ header (the reference's Group
 * names its enum without the `enum` keyword).  Build callers with
 * -D__HIP_PLATFORM_AMD__ -I/opt/rocm/include (hipcc does this itself).
 */
#ifndef MESHLESS_H_
#define MESHLESS_H_

#include <hip/hip_vector_types.h>

#ifdef __cplusplus
extern "C"
{
#endif

/* Basis of every particle in a Group.  Replaces include/meshless.h:30-35. */
enum BasisFunctionId
{
	SPH,             /* M4 cubic spline, support 2*r0                  */
	GAUSSIAN,        /* exp(-pi |x|^2 / r0^2) (see phi_mode below)      */
	WENDLAND_D3_C2   /* Wendland C2 on the unit ball scaled by r0       */
};

/* One particle: position and weight, 16 bytes.  Replaces include/meshless.h:37-41. */
typedef struct alignas(16)
{
	float3 position;
	float weight;
} Constraint;

/*
 * A set of particles sharing one basis.  Replaces include/meshless.h:43-50.
 * h_* arrays are host memory owned by the dataset (allocated with new[]);
 * d_* arrays and d_number_of_terms are written by vis_register_meshless_dataset
 * (device memory owned by the library until vis_unregister_meshless_dataset).
 * d_radii == 0 means "no radii": every particle uses Phi(|k|).
 */
typedef struct
{
	int number_of_terms, d_number_of_terms;
	BasisFunctionId basis_function_id;

	Constraint* h_constraints, *d_constraints;
	float*  h_radii,   *d_radii;
} Group;

/* Replaces include/meshless.h:52-57. */
typedef struct
{
	Group* groups;
	int number_of_groups;

} MeshlessDataset;

/* One-group dataset that adopts h_constraints / h_radii (both new[]).  meshless.h:59, meshless.cpp:27-37 */
MeshlessDataset get_simple_meshless_dataset(int number_of_terms, Constraint* h_constraints, float* h_radii, BasisFunctionId basis_function_id);

/* delete[] the host arrays and the group array.  meshless.h:61-62, meshless.cpp:39-56 */
void delete_meshless_dataset(MeshlessDataset meshless_dataset);
void delete_meshless_datasets(MeshlessDataset* meshless_datasets, int number_of_datasets);

/* Sum of number_of_terms over groups.  meshless.h:64, meshless.cpp:58-66 */
int get_number_of_terms(MeshlessDataset* meshless_dataset);

/* Translate host positions in place.  meshless.h:66-67, meshless.cpp:68-84 */
void shift_meshless_dataset(MeshlessDataset* meshless_dataset, float x, float y, float z);
void shift_meshless_datasets(MeshlessDataset* meshless_dataset, int number_of_datasets, float x, float y, float z);

/* Text (.sph) dataset I/O.  meshless.h:69-70, meshless.cpp:86-187 */
void load_meshless_datasets_from_file(const char* filename, MeshlessDataset** meshless_datasets, int* number_of_datasets);
void save_meshless_datasets_to_file(const char* filename, MeshlessDataset* meshless_datasets, int number_of_datasets);

/*
 * Additive (no counterpart in the reference): allocate the host arrays with
 * the library's own new[] so that callers outside C++ (ctypes, cgo) can hand
 * them to get_simple_meshless_dataset, whose ownership rule is new[]/delete[].
 * Return 0 on failure; the arrays are zero-filled.
 */
Constraint* meshless_new_constraints(int number_of_terms);
float* meshless_new_radii(int number_of_terms);

#ifdef __cplusplus
}
#endif

#endif /*MESHLESS_H_*/
