/*
 * meshless_vis.h -- Fourier volume rendering of meshless datasets on MI355X
 * (gfx950), C ABI of libmeshless_vis.so.
 *
 * Drop-in for the reference's include/meshless_vis.h: the VisConfig fields
 * the reference declares keep their names, types and order, and every
 * function keeps its name and signature, except where noted below.  A render
 * is: direct nonuniform sum of the particles' Fourier transform over the
 * Hermitian half of the view's slice plane (hand-written HIP kernels), then an
 * unnormalised 2-D C2R inverse FFT (hand-written HIP Stockham kernels) into
 * _d_image (float[Nx*Ny], row-major, first index along u_axis).
 *
 * Differences from the reference, all deliberate:
 *  - _plan is an opaque library handle (the reference holds a cufftHandle or
 *    fftwf_plan there, meshless_vis.h:129-133).
 *  - errors never abort the process: every entry point records a code and a
 *    message, read with vis_get_last_error_code()/vis_get_last_error().  The
 *    reference aborts through cutil macros (meshless_vis.cu:141-146).
 *  - vis_fourier_volume_rendering validates the config with the
 *    vis_config_check rules and renders nothing (error VIS_ERROR_INVALID_CONFIG)
 *    when they fail; the reference silently skips an unsupported
 *    block_length (fourier_transform.cu:167-171).
 *  - vis_config_manual_d_image takes float*, as the reference's definitions
 *    do (meshless_vis.cu:244); its header said float2*.
 *  - vis_opengl_fourier_volume_rendering renders through HIP's GL interop
 *    into the buffer object and then restores the config's own _d_image (the
 *    reference left _d_image pointing into the unmapped buffer).
 *  - fields after _d_freq_image_arranged are additive (form, phi_mode, ...).
 */
#ifndef MESHLESS_VISh
#define MESHLESS_VISh

#include "meshless.h"
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C"
{
#endif

/* Additive: which implementation of the direct sum runs. */
enum VisForm
{
	VIS_FORM_PLAIN = 0,            /* FP32 direct sum, fixed-point (FP64-reduced) phase     */
	VIS_FORM_MFMA_F16 = 1,         /* one FP16 MFMA complex GEMM, FP32 accumulate           */
	VIS_FORM_MFMA_SPLIT_BF16 = 2,  /* hi/lo BF16 operands, 3 products, FP32 accumulate      */
	VIS_FORM_MFMA_SPLIT_F16 = 3    /* hi/lo FP16 operands, 3 products, FP32 accumulate      */
};

/* Additive: which Phi (Fourier transform of the basis) is evaluated. */
enum VisPhiMode
{
	/* The reference's formulas (fourier_transform_cpu.cpp:37-66): quadratic fits
	   below r = 0.06 (SPH) / 0.24 (Wendland), and exp(+pi r^2) for GAUSSIAN. */
	VIS_PHI_REFERENCE = 0,
	/* Exact transforms: series about 0 then closed forms, exp(-pi r^2) for GAUSSIAN
	   (the transform of exp(-pi |x|^2)). */
	VIS_PHI_ANALYTIC = 1
};

/* Additive: error codes of vis_get_last_error_code(). */
enum VisError
{
	VIS_SUCCESS = 0,
	VIS_ERROR_INVALID_VALUE = 1,
	VIS_ERROR_INVALID_CONFIG = 2,
	VIS_ERROR_HIP = 3,
	VIS_ERROR_OUT_OF_MEMORY = 4,
	VIS_ERROR_UNSUPPORTED = 5,
	VIS_ERROR_NOT_REGISTERED = 6
};

/* Additive: output tile of the MFMA direct-sum kernels (|x| rows x y columns per workgroup),
   relative to each kernel's default: 128 x 128 for the Gaussian / uniform-radius kernel, 64 x 64
   for the two-node SPH/Wendland kernel.  The split forms build every variant; the single-FP16
   form and the kernel for 3 or more interpolation nodes (128 x 64) keep their default. */
enum VisTile
{
	VIS_TILE_DEFAULT = 0,
	VIS_TILE_HALF_X = 1,   /* half the |x| rows:   64 x 128, 32 x 64 */
	VIS_TILE_HALF_Y = 2,   /* half the y columns:  128 x 64, 64 x 32 */
	VIS_TILE_QUARTER = 3   /* both halved:         64 x 64,  32 x 32 */
};

typedef struct VisPlan VisPlan; /* opaque, owned by the library */

/* Replaces include/meshless_vis.h:39-79. */
typedef struct
{
	// variables prefixed with an underscore are private and are only changed by vis_config_* functions

	float2 step_size;
	int2 _cutoff_frequency;
	float3 u_axis;
	float3 v_axis;

	bool cull_fully_aliased_terms;   /* The reference never reads it (its culling, meshless_vis.cu:39-127,
	                                    is dead code).  Here, when true, every render drops the terms
	                                    whose projected footprint - the square of half-side k * r0
	                                    (k = 2 SPH, 1 Wendland, 3 Gaussian; r0 = 1 without radii)
	                                    around (p.u_axis, p.v_axis) - lies wholly outside the image's
	                                    period window |u| <= 1/(2 step.x), |v| <= 1/(2 step.y).       */

	bool _automatic_d_image;

	VisPlan* _plan;

	int2 _number_of_samples;
	int block_length;
	float _scale;
	float* _d_image;

	float2* _d_freq_image;           /* device, P * 2*cx*cy complex, [p][y][x0] */

	int _number_of_partial_sums;

	float2* _d_freq_image_arranged;  /* device, Nx * (Ny/2+1) complex: FFT work buffer */

	/* ---- additive fields (no counterpart in the reference) ---- */
	int form;                           /* enum VisForm, default VIS_FORM_PLAIN        */
	int phi_mode;                       /* enum VisPhiMode, default VIS_PHI_REFERENCE  */
	int number_of_h_bins;               /* MFMA form, SPH/Wendland radii: 0 = auto     */
	int number_of_interpolation_nodes;  /* MFMA form, SPH/Wendland radii: nodes per bin, 0 = auto */
	float h_bin_ratio;                  /* MFMA form: ratio h_hi/h_lo of the radius bins, 0 = auto
	                                       (used when number_of_h_bins is 0).  Auto for SPH/Wendland:
	                                       2 nodes, ratio 1.01 (1.0025 below 1024 particles); the
	                                       interpolation error grows with (|f| h (ratio - 1))^2, so
	                                       spectra with much of their energy at |f| h above ~5 want
	                                       narrower bins: vis_autotune measures and picks them.     */
	int tile;                           /* MFMA form: enum VisTile, default VIS_TILE_DEFAULT      */
} VisConfig;

/* meshless_vis.h:81 / meshless_vis.cu:129-176.  Returns 0 (and sets the last error) on failure. */
VisConfig* vis_config_create(bool automatic_d_image, float2 step_size, int2 cutoff_frequency, float3 u_axis, float3 v_axis, int2 number_of_samples, int block_length, int number_of_partial_sums);
/* meshless_vis.h:82 / meshless_vis.cu:178-189: step 0.1, cutoff 128, 512x512, block 256, 1 partial sum. */
VisConfig* vis_config_get_default();
/* meshless_vis.h:83 / meshless_vis.cu:191-210 (wavefront size is 64 on gfx950). */
bool vis_config_check(VisConfig* vis_config);
/* meshless_vis.h:84 / meshless_vis.cu:244-247 */
void vis_config_manual_d_image(VisConfig* vis_config, float* d_image);
/* meshless_vis.h:85-87 / meshless_vis.cu:212-242 */
void vis_config_change_number_of_samples(VisConfig* vis_config, int2 number_of_samples);
void vis_config_change_cutoff_frequency(VisConfig* vis_config, int2 cutoff_frequency);
void vis_config_change_number_of_partial_sums(VisConfig* vis_config, int number_of_partial_sums);
/* meshless_vis.h:88 / meshless_vis.cu:249-256 */
void vis_config_destroy(VisConfig* vis_config);
/* meshless_vis.h:89 / meshless_vis.cu:258-262: _scale = step_size.x * step_size.y */
void vis_config_compute_scale(VisConfig* vis_config);
/* meshless_vis.h:90-91 / meshless_vis.cu:264-299: H2D copy, zero padding to a multiple of P*block_length */
void vis_register_meshless_dataset(VisConfig* vis_config, MeshlessDataset* meshless_dataset);
void vis_unregister_meshless_dataset(VisConfig* vis_config, MeshlessDataset* meshless_dataset);
/* meshless_vis.h:92 / meshless_vis.cu:357-364: registers and maps the GL buffer object
   (hipGraphicsGLRegisterBuffer, on the config's stream), renders the image into it (at least
   Nx*Ny floats, else VIS_ERROR_INVALID_VALUE), unmaps and unregisters it.  Needs a current GL
   context on the GPU; interop failures are reported as VIS_ERROR_HIP. */
void vis_opengl_fourier_volume_rendering(MeshlessDataset* meshless_dataset, VisConfig* vis_config, unsigned int buffer_object);
/* meshless_vis.h:93 / meshless_vis.cu:334-350: asynchronous on the config's stream. */
void vis_fourier_volume_rendering(MeshlessDataset* meshless_dataset, VisConfig* vis_config);
/* meshless_vis.h:94 / meshless_vis.cu:352-355: synchronising D2H copy of Nx*Ny floats. */
void vis_copy_to_host(VisConfig* vis_config, float* h_image);

/* The reference's internal seam (fourier_transform.h:33, fourier_transform.cu:161-181):
   spectrum only, into _d_freq_image slot 0, no FFT. */
void fourier_transform(MeshlessDataset* meshless_dataset, VisConfig* vis_config);

/* ---- additive entry points (no counterpart in the reference) ---- */

/* Stream every later call on this config runs on (default: the null stream). */
void vis_set_stream(VisConfig* vis_config, hipStream_t stream);
hipStream_t vis_get_stream(VisConfig* vis_config);
/* Wait for the config's stream. */
void vis_synchronize(VisConfig* vis_config);
/* Synchronising D2H copy of the 2*cx*cy complex spectrum (slot 0, layout [y][x0]). */
void vis_copy_spectrum_to_host(VisConfig* vis_config, float2* h_spectrum);

/* vis_register_meshless_dataset with the host-to-device copies on a library-owned
   copy stream: returns without waiting for them, and later work on the config's
   stream waits for them (stream-ordered), so uploading the next timestep overlaps
   the render of the current one.  The host arrays must stay unchanged until the
   config's stream has passed this point (vis_synchronize, vis_copy_to_host of a
   later render, or vis_unregister_meshless_dataset). */
void vis_register_meshless_dataset_async(VisConfig* vis_config, MeshlessDataset* meshless_dataset);

/* Render many (dataset, config) pairs; same results as calling
   vis_fourier_volume_rendering on each pair.  When every config has its own
   entry and they share a stream, an MFMA form, the Phi mode, cutoff, samples,
   steps and bin settings, the whole batch runs as one persistent MFMA launch per
   kernel variant (with batched prep, reductions and C2R): the first time at once,
   the second time captured into a HIP graph, replayed while the batch is
   unchanged; otherwise the renders run one after the other. */
void vis_fourier_volume_rendering_batched(MeshlessDataset** meshless_datasets, VisConfig** vis_configs, int number_of_renders);

/* Device memory the MFMA forms may hold at once in FP32 partial spectra (one per chunk of
   up to 16,384 particles: 8 B per half-plane frequency each), process-wide; <= 0 restores
   the default, 8 GiB.  A render or batch needing more runs in passes over its chunks, each
   reduced into the FP64 accumulator before the next reuses the memory, so the particle
   count is bounded by the particles' own storage rather than by the partials.  Takes effect
   at the next plan (a changed dataset, grid or setting, or this call). */
void vis_set_partial_budget(long long bytes);

/* No counterpart in the reference.  For callers that cannot allocate device memory
   (code under HIP graph capture, tight serving loops, harnesses that forbid it): the
   render of vis_fourier_volume_rendering for a config with form VIS_FORM_PLAIN,
   with every scratch buffer (packets, split-K partials, FP64 accumulator) in a
   static 64 MiB device array that each GPU the library runs on holds until the
   process ends.  Inputs (registered datasets) and outputs (the config's buffers)
   are the caller's as usual; the call allocates nothing.  A render needing more
   scratch (16 B per particle of the largest group, plus 16 B per frequency and
   16 B per frequency per split-K slot, at least one) is refused with
   VIS_ERROR_INVALID_VALUE; other forms with VIS_ERROR_INVALID_CONFIG, groups not
   registered through vis_register_meshless_dataset with VIS_ERROR_NOT_REGISTERED.
   All calls on one GPU share its array: calls that overlap on one GPU (other
   streams or threads) race. */
void vis_fourier_volume_rendering_noalloc(MeshlessDataset* meshless_dataset, VisConfig* vis_config);

/* Picks number_of_interpolation_nodes, h_bin_ratio and tile for the MFMA form on
   this (registered) dataset: every candidate is timed on the config's stream and
   its spectrum compared with the plain FP32 kernel's; the fastest candidate whose
   relative L2 difference is at most `tolerance` is kept (the most accurate one
   when none is), with number_of_h_bins set to 0 (bins from the ratio).  The
   (nodes, ratio) candidates are timed first with the config's tile, then the
   kept setting with the other VisTile variants of the kernel it runs (split
   forms).  Leaves the config's other fields as they were and overwrites
   _d_freq_image.  Returns the kept candidate's spectrum time in milliseconds, or
   a negative value on error. */
double vis_autotune(MeshlessDataset* meshless_dataset, VisConfig* vis_config, double tolerance);

/* Kernel timing with HIP events on the config's stream around the direct-sum
   kernel (the dominant kernel of a render; for an MFMA render split into passes by
   vis_set_partial_budget, from its first pass's kernels to its last pass's, with the
   reductions between them).  Off by default. */
void vis_set_kernel_timing(VisConfig* vis_config, int enable);
/* Returns the number of timed launches since the last reset and their summed
   duration in milliseconds (synchronises the config's stream). */
int vis_get_kernel_timing(VisConfig* vis_config, double* total_ms);
void vis_reset_kernel_timing(VisConfig* vis_config);

/* Test support (no counterpart in the reference): after `after` (>= 0) more device
   allocations succeed, every later one fails as VIS_ERROR_OUT_OF_MEMORY without calling
   hipMalloc, until this is called with a negative value (the default).  For exercising
   the library's out-of-memory paths; process-wide. */
void vis_debug_fail_device_allocations(int after);

/* Test support (no counterpart in the reference): how many batch graphs the batched
   renders led by this config (the first config of each batch) have captured, and how many
   times they were launched, the capturing call included.  A batch's first call runs
   without a graph, its second captures and launches the graph, later identical calls
   replay it. */
void vis_debug_batch_graph_counts(const VisConfig* vis_config, int* captures, long long* launches);

/* Last error of the calling thread (VIS_SUCCESS / "" when none). */
int vis_get_last_error_code(void);
const char* vis_get_last_error(void);
void vis_clear_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
