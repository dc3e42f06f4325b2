/*
 * xrp.h - C ABI of libxrp, the MI355X (gfx950) X-ray projection renderer.
 *
 * Drop-in boundary for the reference renderer's per-pixel line-integral
 * path (igrega348/xray_projection_render).  Plain C types only; every
 * pointer is caller-owned unless a function says otherwise.  Every entry
 * point returns an xrp_status (0 = success); xrp_last_error() gives the
 * message of the last failure on the calling thread.
 *
 * Replaces:
 *   - RenderVolumeProjectionsCUDA   cuda_backend.h:101-112  -> xrp_render_volume_projections
 *   - the CPU render loop           main.go:429-471        -> xrp_render / xrp_render_device
 *   - computeCameraFromAngles       main.go:226-239        -> xrp_camera_from_angles
 *   - RenderProjections / FreeString api.go:82-180          -> same names (JSON in/out)
 */
#ifndef XRP_H
#define XRP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* 2: every output stats struct opens with struct_size (set by the caller);
 *    XRP_ERR_INTERNAL added. */
#define XRP_ABI_VERSION 2

/* Status codes.  1 and 2 keep the meaning of cuda_backend.cu:95-100. */
typedef enum {
  XRP_OK = 0,
  XRP_ERR_NULL = 1,        /* a required pointer is NULL */
  XRP_ERR_DIMS = 2,        /* non-positive size / resolution / view count */
  XRP_ERR_DEVICE = 3,      /* HIP runtime failure (message has the HIP error) */
  XRP_ERR_ALLOC = 4,       /* device or host allocation failed */
  XRP_ERR_SCENE = 10,      /* object description rejected (reference FromMap rules) */
  XRP_ERR_DEFORMATION = 11,/* deformation description rejected */
  XRP_ERR_ARG = 12,        /* invalid parameter value (ds <= 0, unknown enum, ...) */
  XRP_ERR_UNSUPPORTED = 13,/* valid for the reference, not supported by this library */
  XRP_ERR_IO = 14,         /* file could not be read or written */
  XRP_ERR_INTERNAL = 15    /* unexpected library failure (message says which) */
} xrp_status;

/* Integrators: main.go:144-154 and main.go:159-199. */
typedef enum { XRP_INTEGRATE_SIMPLE = 0, XRP_INTEGRATE_HIERARCHICAL = 1 } xrp_integrator;

/* Precision of the density evaluation.
 *  AUTO    - a voxel-grid root under linear deformations with the simple
 *            integrator takes the FP32 voxel kernel (exact FP64 sample
 *            range per ray, FP32 trilinear, Kahan-compensated sum); other
 *            scenes holding a voxel grid FP64; everything else GUARDED.
 *  GUARDED - FP32 membership tests with guard bands; samples inside a band
 *            are re-evaluated in FP64.  Densities and the integral are FP64.
 *  FP64    - every sample evaluated in FP64 in the reference's order. */
typedef enum {
  XRP_PRECISION_AUTO = 0,
  XRP_PRECISION_GUARDED = 1,
  XRP_PRECISION_FP64 = 2
} xrp_precision;

/* Number semantics of scene text: Go's encoding/json decodes every number
 * as float64; yaml.v3 keeps integer literals as int (so `rho: 1` is
 * rejected where the reference type-asserts float64). */
typedef enum { XRP_NUMBERS_JSON = 0, XRP_NUMBERS_YAML = 1 } xrp_number_mode;

typedef struct xrp_context xrp_context; /* one device + stream + caches */
typedef struct xrp_scene xrp_scene;     /* loaded object tree + deformation */

/* Camera position on the sphere of radius R (main.go:282-285). */
typedef struct {
  double azimuthal_deg;
  double polar_deg;
} xrp_view_angle;

/* Render parameters (main.go:310-330 globals and flags).  Set
 * struct_size = sizeof(xrp_render_params); fields added later are
 * appended so older callers keep working. */
typedef struct {
  uint32_t struct_size;
  int32_t resolution;        /* square image side, pixels */
  double ds;                 /* step (simple) or coarse step DS (hierarchical); > 0 */
  double R;                  /* camera distance, default 4 */
  double fov_deg;            /* field of view, default 40 */
  double flat_field;         /* optical depth added to every ray */
  double density_multiplier; /* must be non-zero (api.go:101-103) */
  int32_t integrator;        /* xrp_integrator */
  int32_t precision;         /* xrp_precision */
  int32_t audit;             /* 1: also evaluate every sample in FP64 and count
                                FP32 decisions that disagree (slow; testing) */
  int32_t output_format;     /* xrp_output_format */
  int32_t transparency;      /* XRP_OUTPUT_GRAY8_ALPHA8: alpha 0 where T == 1 */
  int32_t reserved;
} xrp_render_params;

/* Output pixel formats, layout [view][i][j] (i = image x, main.go:375-378).
 *  F32            - float32 transmittance exp(-T).
 *  GRAY8_ALPHA8   - two bytes per pixel as the reference's PNG stores them
 *                   (main.go:481-498): gray = uint16(val*0xffff) >> 8 from
 *                   the FP64 transmittance, alpha = 0 under --transparency
 *                   where val >= 1, else 255. */
typedef enum { XRP_OUTPUT_F32 = 0, XRP_OUTPUT_GRAY8_ALPHA8 = 1 } xrp_output_format;

/* Fills p with the reference's CLI defaults (main.go:651-761):
 * resolution 512, ds -1 (caller must resolve; see xrp_scene_auto_ds),
 * R 4, fov 40, flat_field 0, density_multiplier 1, hierarchical. */
void xrp_default_params(xrp_render_params *p);

/* Output structs (xrp_render_stats, xrp_batch_stats, xrp_train_stats,
 * xrp_surrogate_stats) open with struct_size: the caller sets it to
 * sizeof(the struct it was compiled with) and the library writes no byte
 * beyond it, so fields appended later never overrun an older caller.  A
 * struct_size below 8 is rejected with XRP_ERR_ARG; a NULL pointer asks for
 * no statistics. */
typedef struct {
  uint32_t struct_size;
  int32_t reserved0;
  uint64_t rays;
  uint64_t samples;        /* samples marched after bounding-box clipping (evaluated or
                              settled by a safe radius, see skipped_samples) */
  uint64_t reevaluated;    /* guard-band samples re-evaluated in FP64 */
  uint64_t audit_mismatch; /* audit mode: FP32-clear samples that differ in FP64 */
  uint64_t fine_samples;   /* hierarchical fine sub-step evaluations */
  double kernel_ms;        /* device time of the render kernels */
  double position_error;   /* FP32 object-space position error bound used */
  int32_t precision_used;  /* xrp_precision actually run */
  int32_t block;           /* samples per node visit of the guarded simple march
                              (1, 2 or 4), or coarse samples per node visit of the
                              two-pass hierarchical march's first pass; 0 for the
                              other kernels */
  uint64_t staged_chunks;   /* voxel kernel: sample chunks served from LDS */
  uint64_t staged_voxels;   /* voxel kernel: voxels copied into LDS */
  uint64_t fallback_chunks; /* voxel kernel: chunks gathered from global memory */
  uint64_t skipped_samples; /* guarded marches: (coarse) samples whose density was settled
                               without evaluation, from the safe radius (distance to
                               the nearest boundary, less every FP32 error) of an
                               evaluated sample before them; under audit each is
                               also evaluated in FP64 and compared */
  uint64_t redo_rays;       /* two-pass march: rays marched again (more guard-band
                               samples than a deferral slot holds, no slot free, or,
                               hierarchical, an FP32 guess that was not exact) */
  uint64_t deferred_rays;   /* two-pass march: rays whose guard-band samples were
                               re-evaluated (hierarchical: checked) in FP64 after
                               the march */
  int32_t two_pass;         /* 1 when the guarded simple march ran in two passes: the
                               march without FP64, guard-band samples re-evaluated after */
  int32_t reserved1;
} xrp_render_stats;

int xrp_abi_version(void);
const char *xrp_last_error(void);

/* ---- context -------------------------------------------------------------- */
/* Batches and surrogates live on a context: destroy them before it. */
int xrp_context_create(int device, xrp_context **out);
void xrp_context_destroy(xrp_context *ctx);
int xrp_context_synchronize(xrp_context *ctx);

/* ---- scenes (host only; no device work) ----------------------------------- */
/* object_text: an object description (objects.NewObject, objects.go:658-689).
 * deformation_text: NULL/"" or one deformation (deformations.go:308-343).
 * Text is JSON; number_mode chooses the decoder semantics.  base_dir
 * resolves relative voxel_grid raw paths (NULL = current directory). */
int xrp_scene_create(const char *object_text, const char *deformation_text,
                     int number_mode, const char *base_dir, xrp_scene **out);
/* Scene from files as load_object / load_deformation read them
 * (main.go:50-120): the last four characters pick YAML ("yaml", integer
 * literals stay int) or JSON ("json", every number float64); raw voxel
 * paths resolve against the working directory. */
int xrp_scene_load_files(const char *object_path, const char *deformation_path,
                         xrp_scene **out);
/* Voxel-grid scene from caller memory: rho[z*nx*ny + x*ny + y] in [0,1]. */
int xrp_scene_create_voxels(const float *rho, int nx, int ny, int nz,
                            const char *deformation_text, int number_mode,
                            xrp_scene **out);
void xrp_scene_destroy(xrp_scene *scene);
/* MinFeatureSize of the root object (objects.go MinFeatureSize methods). */
int xrp_scene_min_feature_size(const xrp_scene *scene, double *out);
/* ds rule of main.go:350-353: MinFeatureSize / 5. */
int xrp_scene_auto_ds(const xrp_scene *scene, double *out);
/* World-space box outside which the scene density is zero (after the
 * deformation preimage); infinite entries mean unbounded.  *empty = 1 when
 * the density is zero everywhere. */
int xrp_scene_clip_box(const xrp_scene *scene, double lo[3], double hi[3],
                       int *empty);
/* Number of flattened device nodes and deformation stages. */
int xrp_scene_layout(const xrp_scene *scene, int *nodes, int *deformations,
                     int *has_voxels);

/* ---- cameras and sample grids (host only) --------------------------------- */
/* eye and row-major cam2world = inverse(LookAt(eye, 0, +z)). */
int xrp_camera_from_angles(double azimuthal_deg, double polar_deg, double R,
                           double eye[3], double cam2world[16]);
/* Count of s = smin, smin+ds, ... with s < smax (accumulated in FP64). */
int xrp_sample_count(double smin, double smax, double ds, int64_t *count);

/* ---- rendering ------------------------------------------------------------ */
/* out: n_views * res * res pixels in p->output_format (host memory). */
int xrp_render(xrp_context *ctx, xrp_scene *scene, const xrp_render_params *p,
               const xrp_view_angle *views, int n_views, void *out,
               xrp_render_stats *stats);
/* Same, out is device memory on the context's device; returns when done. */
int xrp_render_device(xrp_context *ctx, xrp_scene *scene,
                      const xrp_render_params *p, const xrp_view_angle *views,
                      int n_views, void *out_device, xrp_render_stats *stats);

/* No counterpart in the reference.  xrp_render_device for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc): this call allocates no device or pinned host
 * memory.  Its per-call data (stat counters, node table, deformations,
 * views and sample table) goes into a fixed 2 MiB region of device memory,
 * one per GPU, held until the process ends; larger requests (for example
 * more than about 12,000 views in one call, or a sample table above
 * 250,000 entries) fail with XRP_ERR_DIMS.  A scene holding a voxel grid
 * must first be rendered once on the same context with xrp_render or
 * xrp_render_device (which upload the grid), else XRP_ERR_UNSUPPORTED.
 * Every call on one GPU shares that GPU's region: calls that overlap on one
 * GPU (from different contexts) race; calls on one context are serialized.
 * The output is the same as xrp_render_device's. */
int xrp_render_device_noalloc(xrp_context *ctx, xrp_scene *scene,
                              const xrp_render_params *p, const xrp_view_angle *views,
                              int n_views, void *out_device, xrp_render_stats *stats);

/* Explicit-camera voxel render replacing RenderVolumeProjectionsCUDA
 * (cuda_backend.h:70-112) with the CPU renderer's semantics: software
 * trilinear with the (N-1) convention, FP64 accumulated sample positions.
 * volume[k*nx*ny + i*ny + j]; cameras row-major cam2world (FP64). */
typedef struct {
  double eye[3];
  double cam2world[16];
  double fov_deg;
  double R;
} xrp_camera;
int xrp_render_volume_projections(xrp_context *ctx, const float *volume, int nx,
                                  int ny, int nz, const xrp_camera *cameras,
                                  int n_cameras, int image_res, double ds,
                                  double flat_field, float *out_images);

/* Volume export of main.go:549-635 (--export_volume): density (deformation
 * and density_multiplier applied) at x = i/res*2-1, the i/N convention of
 * main.go:208-214; divided by the maximum, times 255, converted to a byte
 * as Go does.  out[k*res*res + i*res + j], res^3 bytes, host memory. */
int xrp_export_volume(xrp_context *ctx, xrp_scene *scene, int res,
                      double density_multiplier, uint8_t *out);

/* ---- batched rendering ------------------------------------------------------
 * Many scenes, each with its own views and parameters, rendered by one
 * persistent launch per (precision, integrator) group over an atomic queue
 * of (scene, view, tile) items, the sequence captured once in a hipGraph
 * and replayed by xrp_batch_run.  It batches the per-view loop of render()
 * (main.go:427-523) across scenes; the reference renders one scene per
 * process.  Images are float32 transmittance, [view][i][j] per item, kept
 * in device memory owned by the batch. */
typedef struct {
  const xrp_scene *scene;
  const xrp_view_angle *views;
  int32_t n_views;
  int32_t reserved;
  xrp_render_params params; /* output_format must be XRP_OUTPUT_F32 */
} xrp_batch_item;

/* Work item of the queue: wave_rows x wave_cols blocks of 8x8 pixels that
 * one wave marches in sequence (at most 16 blocks); groups_per_cu persistent
 * 256-thread workgroups are launched per compute unit. */
typedef struct {
  int32_t wave_rows;
  int32_t wave_cols;
  int32_t groups_per_cu;
  int32_t reserved;
} xrp_tile_shape;

typedef struct {
  uint32_t struct_size; /* sizeof(xrp_batch_stats), set by the caller */
  int32_t reserved0;
  double kernel_ms;     /* graph replay, HIP events on the context stream */
  uint64_t rays;
  uint64_t samples;
  uint64_t reevaluated;
  uint64_t tiles;
  int32_t launches;     /* persistent launches in the graph */
  int32_t reserved;
} xrp_batch_stats;

typedef struct xrp_batch xrp_batch;

/* Plans every item (precision, guard bands, sample tables), uploads the
 * scenes, views and tables, allocates the images and captures the graph
 * with the default tile shape {1, 1, 2}.  The scenes must outlive the batch. */
int xrp_batch_create(xrp_context *ctx, const xrp_batch_item *items, int n_items,
                     xrp_batch **out);
/* Re-tiles the queue and re-captures the graph. */
int xrp_batch_set_shape(xrp_batch *batch, const xrp_tile_shape *shape);
/* Autotuner: times each candidate shape over `repeats` replays, keeps the
 * fastest and records it in a process-wide table keyed by the batch's
 * signature (and in the file named by $XRP_TUNE_CACHE when set), so a batch
 * with the same signature starts from the recorded shape. */
int xrp_batch_tune(xrp_batch *batch, int repeats, xrp_tile_shape *chosen);
/* Replays the graph and waits for it. */
int xrp_batch_run(xrp_batch *batch, xrp_batch_stats *stats);
/* Device pointer of item's images (valid until xrp_batch_destroy). */
int xrp_batch_output(const xrp_batch *batch, int item, const float **device_images);
/* Copies item's images to host memory (n_views * res * res floats). */
int xrp_batch_read(xrp_batch *batch, int item, float *host_images);
/* Per item, milliseconds from the batch's first tile start to the item's
 * last tile end in the last run (n_items values). */
int xrp_batch_latencies(const xrp_batch *batch, double *ms);
void xrp_batch_destroy(xrp_batch *batch);

/* ---- learned density surrogate ----------------------------------------------
 * No counterpart in the reference.  A multiresolution hash encoding (16
 * levels x 2 features, 2^log2_table entries per level) feeding a 32-64-64-1
 * MLP learns the scene's object-space density from exact samples; the
 * march then evaluates the network (BF16 MFMA, FP32 accumulation) at a
 * coarser step, and a sampled exact check re-renders any 16x16 tile whose
 * error exceeds a tolerance with the exact march. */
typedef struct xrp_surrogate xrp_surrogate;

typedef struct {
  uint32_t struct_size;
  int32_t log2_table; /* entries per level: default 19 */
  int32_t n_min;      /* coarsest level resolution: default 16 */
  int32_t n_max;      /* finest level resolution: default 512 */
  int32_t wrap;       /* -1: learn one unit cell when the root is a tessellation */
  int32_t reserved;
  uint64_t seed;      /* initialisation and sampling seed */
} xrp_surrogate_config;

typedef struct {
  uint32_t struct_size;
  int32_t batch;      /* samples per optimiser step: default 65536 */
  int64_t samples;    /* exact training samples: default 2^24 */
  int32_t epochs;     /* passes over the samples: default 4 */
  int32_t reserved;
  double lr_table;    /* Adam step for the tables: default 1e-2 */
  double lr_mlp;      /* Adam step for the MLP: default 1e-3 */
} xrp_train_config;

typedef struct {
  uint32_t struct_size; /* sizeof(xrp_train_stats), set by the caller */
  int32_t reserved0;
  double first_loss;  /* mean squared error of the first step */
  double final_loss;  /* mean over the last epoch */
  double label_ms;    /* exact sample generation */
  double train_ms;    /* forward/backward/Adam kernels */
  int32_t steps;
  int32_t reserved;
} xrp_train_stats;

typedef struct {
  uint32_t struct_size;  /* sizeof(xrp_surrogate_stats), set by the caller */
  int32_t reserved0;
  double surrogate_ms;   /* surrogate march */
  double check_ms;       /* exact march of the check pixels */
  double fallback_ms;    /* exact re-render of failing tiles */
  double max_sampled_error; /* largest |surrogate - exact| at check pixels */
  uint64_t tiles;
  uint64_t tiles_failed;
  uint64_t check_pixels;
  uint64_t network_samples; /* network evaluations issued by the surrogate march */
  uint64_t exact_samples;   /* samples the exact check and fallback marches evaluated
                               (in the clip box, less those settled by a safe radius) */
} xrp_surrogate_stats;

void xrp_surrogate_default_config(xrp_surrogate_config *c);
void xrp_train_default_config(xrp_train_config *c);
int xrp_surrogate_create(xrp_context *ctx, const xrp_scene *scene,
                         const xrp_surrogate_config *config, xrp_surrogate **out);
int xrp_surrogate_train(xrp_surrogate *s, const xrp_train_config *config,
                        xrp_train_stats *stats);
/* Renders with the network at step ds_surrogate into device memory (float32
 * transmittance [view][i][j]).  The exact march (params->ds, integrator,
 * precision) judges the result:
 *  tile_tolerance > 0  sampled: every 16x16 tile is checked at 16 pixels (a
 *                      stride-4 grid) and re-rendered exactly when any
 *                      differs by more than tile_tolerance / 2; pixels
 *                      between check pixels are not compared;
 *  tile_tolerance < 0  certified: the exact march renders every pixel and
 *                      every tile whose largest difference exceeds
 *                      |tile_tolerance| takes the exact values, so no pixel
 *                      is further than |tile_tolerance| from the exact render;
 *  tile_tolerance = 0  the network's image as it is. */
int xrp_surrogate_render_device(xrp_surrogate *s, const xrp_render_params *params,
                                const xrp_view_angle *views, int n_views,
                                double ds_surrogate, double tile_tolerance,
                                void *out_device, xrp_surrogate_stats *stats);
/* The same with the images copied to host memory (n_views * res * res). */
int xrp_surrogate_render(xrp_surrogate *s, const xrp_render_params *params,
                         const xrp_view_angle *views, int n_views, double ds_surrogate,
                         double tile_tolerance, float *out_host, xrp_surrogate_stats *stats);
/* Parameters (FP32, layout: tables[16][T][2], W1[32][64], b1[64], W2[64][64],
 * b2[64], w3[64], b3) and their count. */
int64_t xrp_surrogate_param_count(const xrp_surrogate *s);
int xrp_surrogate_get_params(const xrp_surrogate *s, float *host);
int xrp_surrogate_set_params(xrp_surrogate *s, const float *host);
/* Network output at n points u in [0,1]^3 (host arrays; BF16 MFMA path). */
int xrp_surrogate_predict(xrp_surrogate *s, const float *u_host, int64_t n, float *out_host);
/* Mean squared error and its gradients (parameter layout) for n points and
 * targets, computed by the training kernels (host arrays). */
int xrp_surrogate_gradients(xrp_surrogate *s, const float *u_host, const float *targets,
                            int64_t n, float *grads_host, double *loss);
void xrp_surrogate_destroy(xrp_surrogate *s);
/* Diagnostic: D = A B for 16x16 row-major float matrices through one
 * v_mfma_f32_16x16x16_bf16 with the operand layout the surrogate uses. */
int xrp_diag_mfma_bf16(xrp_context *ctx, const float *a, const float *b, float *d);
/* Diagnostic, no counterpart in the reference: measured L2 read bandwidth
 * (GB/s, best of 3 timed runs) of a kernel that streams one 2 MiB region
 * per XCD through its L2 with 16-byte loads; the second ceiling the voxel
 * kernel's gathers are compared with. */
int xrp_diag_l2_bandwidth(xrp_context *ctx, double *gbps);
/* Diagnostic, no counterpart in the reference: the largest absolute errors
 * of the hardware sine and cosine (argument in revolutions) over every FP32
 * argument in [-0.5, 0.5], against FP64; the FP32 gyroid evaluators' guard
 * bands assume at most 1.3e-7. */
int xrp_diag_sincos_error(xrp_context *ctx, double *max_sin, double *max_cos);
/* Diagnostic, no counterpart in the reference: the largest absolute errors
 * against FP64 of the FP32 deformations' hardware tanh (every FP32 argument
 * in [-20, 20]), logistic function ([-100, 100]) and e^-x ([0, 110]); the
 * deformations' position error bounds assume at most 3e-7, 2e-7, 2e-7. */
int xrp_diag_deform_error(xrp_context *ctx, double *max_tanh, double *max_sigmoid,
                          double *max_expneg);

/* ---- reference JSON API (api.go:82-180) ----------------------------------- */
/* Renders as the reference's RenderProjections does (PNG frames,
 * transforms.json, object.json) on the GPU.  Returns a malloc'd JSON
 * RenderResult; release it with FreeString. */
char *RenderProjections(const char *json_params);
void FreeString(char *str);

#ifdef __cplusplus
}
#endif
#endif /* XRP_H */
