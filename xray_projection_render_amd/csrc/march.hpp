// march.hpp - launch interface of the projection kernels (host side).
#pragma once
#include <stdint.h>

#include "dev_types.hpp"

namespace xrp {

// One camera: eye and the row-major cam2world matrix (main.go:445-453).
struct DevView {
  double eye[3];
  double cam[16];
  double pad_;
};

enum Integrator : int32_t { INT_SIMPLE = 0, INT_HIERARCHICAL = 1 };
enum Precision : int32_t { PREC_GUARDED = 0, PREC_FP64 = 1 };

// Stat counters accumulated by the kernels (device, uint64 each).
enum StatSlot {
  ST_SAMPLES = 0,      // density evaluations made (after clipping)
  ST_REEVAL = 1,       // guard-band samples re-evaluated in FP64
  ST_AUDIT_MISS = 2,   // audit: FP32-clear samples whose FP64 value differs
  ST_FINE = 3,         // hierarchical fine-step evaluations
  ST_STAGED_CHUNKS = 4,   // LDS voxel kernel: chunks staged in LDS
  ST_STAGED_VOXELS = 5,   //   voxels copied into LDS
  ST_FALLBACK_CHUNKS = 6, //   chunks too large for LDS (gathered)
  ST_SKIPPED = 7,      // samples settled by a safe radius without evaluation
  ST_DEFERRED = 8,     // two-pass march: rays whose guard-band samples were
                       // resolved by march_resolve_kernel
  ST_COUNT = 9
};

struct MarchParams {
  const DevNode* nodes;
  const DevDeform* defs;
  const DevView* views;
  const double* s_table;   // simple: s_k; hierarchical: coarse right_n
  float* out;              // [view][i][j] f32 transmittance
  unsigned long long* stats;
  int32_t n_defs;
  int32_t n_views;
  int32_t res;
  int32_t K;               // entries in s_table
  double focal;            // 1/tan(fov/2)
  double ds;               // simple: ds; hierarchical: DS (coarse)
  double ds_fine;          // hierarchical: DS/10
  double smin, smax;
  double flat_field;
  double dm;               // density multiplier
  double lo[3], hi[3];     // world clip box (inflated); infinite = no clip
  int32_t clip;
  int32_t audit;           // evaluate every sample in FP64 as well
  int32_t view_offset;     // first view index of this launch
  int32_t out_view_stride; // floats between consecutive views in out
  // Voxel-root kernel (and culled collections: M, t): world -> object
  // affine map (the composed linear deformation chain), grid dims, brick
  // layout and the grid-space guard.
  const float* vgrid;      // linear [z][x][y], or 4^3 bricks of 1, 2 or 4 floats
  int32_t vn[3];           // nx ny nz
  int32_t vbricks[3];      // bricks per axis (x, y, z)
  double M[9];             // row-major
  double t[3];
  float vguard;            // grid units
  uint32_t vbytes;         // size of vgrid in bytes (buffer descriptor range)
  const float* vlin;       // linear [z][x][y] copy (LDS staging source)
  uint32_t vlin_bytes;
  // Output: 0 = float32 transmittance in out; 1 = the reference PNG's
  // 8-bit (gray, alpha) pair per pixel in out8 (main.go:481-498).
  int32_t out_mode;
  unsigned char* out8;
  int32_t transparency;
  // Guarded marches: 1 / (lipschitz * ds), slightly reduced (ds: the coarse
  // step of the hierarchical integrator); a block's last sample with safe
  // radius r settles the next floor(r * skip_inv) samples without
  // evaluating them.  0 disables.
  float skip_inv;
  // Guarded simple march, per-scene kernel: samples per node visit (2 or 4).
  int32_t block;
  // Guarded marches: cull the root collection's children per wave
  // (march_ray.hpp wave_cull); M and t then hold the composed deformation.
  int32_t cull;
  // Guarded marches: carry the last sample's velocity through the
  // deformations for the gyroids' second-order reach (eval_block.hpp).
  int32_t vel;
  // Guarded simple march at block 2: queue the FP64 re-evaluations per wave
  // (march_ray.hpp march_T_guarded) instead of making them inline.
  int32_t queue;
  // Per-scene march kernel: rays of a wave as 2 x 32 pixels (32 along the
  // image's contiguous axis j) instead of 32 x 2.
  int32_t tile_j32;
  // Guarded simple march at block 1 or 2, and the guarded hierarchical
  // march, in two passes (march.hip) when redo_list is set.  The first pass
  // (march_try_kernel) marches every ray without the FP64 evaluator and
  // defers its guard-band samples (at most defer_max, <= kDeferMax; their
  // indices wait in LDS).  The hierarchical first pass instead goes on with
  // their FP32 values, stores the pixel and lists the samples, at most
  // kHierDeferMax, as (coarse index << 4) | position in the fine walk (15
  // for the coarse sample itself), for a pass that checks them in FP64.  At the end of a
  // wave its rays with deferred samples are appended (one atomic per wave)
  // as entries e < defer_cap: the ray (v * res^2 + i * res + j, v counted
  // from view_offset) to defer_ray[e], its T without those samples to
  // defer_T[e], their number to defer_cnt[e] and the sample indices to
  // defer_k from defer_kfirst[e] on (k_cap entries).  counts[1] and
  // counts[2] count entries and indices taken.  The resolve pass
  // (march_resolve_kernel) evaluates them in FP64, adds their terms in
  // sample order and stores the pixel.  A ray with more than defer_max
  // such samples, or finding the lists full, stops and is appended to
  // redo_list (counts[0], redo_cap entries); march_redo_kernel marches
  // those rays again.  A ray finding redo_list full sets counts[3]: the
  // caller renders the launch again in one pass.  counts starts zeroed.
  uint32_t* redo_list;
  uint32_t* counts;
  double* defer_T;
  uint32_t* defer_ray;
  uint32_t* defer_cnt;
  uint32_t* defer_kfirst;
  uint32_t* defer_k;
  uint32_t redo_cap, defer_cap, k_cap;
  uint32_t defer_max;
  // Culling: radius added to every child's bounding sphere, twice the
  // object-space position error, so that a culled child is outside every
  // sample's FP32 position by more than its guard band (the FP32 density
  // of a sample is then the same with or without the child).
  float cull_pad;
};

// Guard-band samples a ray may defer to the resolve pass (the first pass
// keeps them in LDS: 6 KiB per wave, so that 24 one-wave workgroups, 6 per
// SIMD, fit a CU).  A ray with more is marched again, by one wave with few
// rays, which takes long: on the sheared gyroid lattice 32 slots left 78
// rays of 33.5 M to the redo pass (0.9 ms).
constexpr uint32_t kDeferMax = 24;
// Guard-band samples a ray the first pass of the two-pass hierarchical
// march may list (LDS: 3.5 KiB per wave beside the toggle queue and the
// cull list, so that 12 one-wave workgroups fit a CU's 160 KiB).
constexpr uint32_t kHierDeferMax = 14;
// Two-pass list sizes for a launch of `rays` rays: room for every ray (two
// sample indices each) up to 2^22 rays; beyond, a ray entry per 8 rays, a
// sample index per 4 and a redo entry per 64, but never less than for 2^22.
struct TwoPassCaps {
  uint32_t redo, defer, k;
};
// The hierarchical march's guard-band samples include fine samples, which
// lie at surfaces, so more rays have some: beyond 2^22 rays a ray entry per
// 2 rays, two sample entries per ray and a redo entry per 16.
inline TwoPassCaps two_pass_caps(size_t rays, bool hierarchical = false) {
  const size_t small = rays < ((size_t)1 << 22) ? rays : ((size_t)1 << 22);
  auto cap = [&](size_t all, size_t share) {
    const size_t c = all > share ? all : share;
    return (uint32_t)(c < 64 ? 64 : c);
  };
  if (hierarchical) return {cap(small, rays / 16), cap(small, rays / 2), cap(2 * small, 2 * rays)};
  return {cap(small, rays / 64), cap(small, rays / 8), cap(2 * small, rays / 4)};
}
// Bytes of the two-pass buffers for those sizes (each array 256-byte aligned).
inline size_t two_pass_bytes(const TwoPassCaps& c) {
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  return al((size_t)c.redo * 4) + al((size_t)c.defer * 8) + 3 * al((size_t)c.defer * 4) +
         al((size_t)c.k * 4);
}

// VOX_LDS: per-wave sub-volumes staged in LDS from the linear copy, with
// the 4^3-brick copy as the gather fallback for oversized sub-volumes.
// *_CACHED: per-lane register cache of the last cell's 8 corners.
// VOX_BRICK4_PAIR / VOX_BRICK4_QUAD (cached): brick elements of 2 floats
// (v[y], v[y+1]) or 4 floats (v[z][y], v[z+1][y], v[z][y+1], v[z+1][y+1]),
// upper indices clamped, so a cell's corners come in 4 or 2 loads.
// VOX_BRICK4_QUAD_SLIDE: the quad copy with a cell cache that, when a cell
// moves by one along x, keeps the shared x plane and loads the other.
enum VoxelLayout : int32_t {
  VOX_LINEAR = 0,
  VOX_BRICK4 = 1,
  VOX_LDS = 2,
  VOX_BRICK4_CACHED = 3,
  VOX_LINEAR_CACHED = 4,
  VOX_BRICK4_PAIR = 5,
  VOX_BRICK4_QUAD = 6,
  VOX_BRICK4_QUAD_SLIDE = 7,
  // the quad copy stepped incrementally over each lane's own samples
  // (voxel.hip step_march)
  VOX_BRICK4_QUAD_STEP = 8
};
inline bool voxel_layout_linear(int layout) {
  return layout == VOX_LINEAR || layout == VOX_LINEAR_CACHED;
}
int launch_voxel(const MarchParams& p, int layout, void* stream);
// L2 read-bandwidth probe: workgroups x 256 lanes each read iters 16-byte
// words from a buffer of l2_probe_buffer_bytes(); out gets workgroups x 256
// words.
size_t l2_probe_buffer_bytes();
int launch_l2_probe(const void* buf, int workgroups, int iters, unsigned* out, void* stream);
// Hardware sine/cosine error over all FP32 arguments in [-0.5, 0.5]
// revolutions: out[0], out[1] (two uints, zeroed by the caller) receive the
// bits of the largest absolute errors of sin and cos.
int launch_sincos_error(unsigned* out, void* stream);
// Hardware tanh, logistic function and e^-x (eval_device.hpp) against FP64
// over every FP32 argument of their ranges (dev_types.hpp kHwTanhErr ...):
// out[0..2] (zeroed by the caller) receive the bits of the largest absolute
// errors, as non-negative doubles.
int launch_deform_error(unsigned long long* out, void* stream);
// Exact march of a list: with vals, items are (view, i, j) pixels and
// vals[k] receives the transmittance; without, items are (view, ti, tj)
// 16x16 tiles rendered into p.out.
int launch_march_list(const MarchParams& p, int precision, int integrator, const int* items,
                      int n, float* vals, void* stream);

// Batched persistent launch: every (scene, view, tile) item of one
// (precision, integrator) group is taken from an atomic work counter by the
// waves of a persistent grid until the counter passes n_tiles.
struct BatchGroup {
  const MarchParams* jobs;      // one per scene of the group, ordered by cost
  const uint32_t* tile_start;   // n_jobs + 1 prefix sums of tiles per job
  uint32_t* counter;            // work counter, zeroed by launch_batch_reset
  unsigned long long* t_first;  // per job: wall clock at its first tile start
  unsigned long long* t_done;   // per job: wall clock at its last tile end
  unsigned long long* sink;     // kBatchSinkBytes: one 128-byte line per lane
  int32_t n_jobs;
  uint32_t n_tiles;
  int32_t wave_rows, wave_cols; // item: (8 wave_rows) x (8 wave_cols) pixels, one wave
  int32_t pad_;
};
// Lines that take the bookkeeping atomics of a wave's lanes other than its
// first (batch.hip lane_slot).
constexpr int kBatchSinkBytes = 64 * 128;
int launch_batch(const BatchGroup& g, int precision, int integrator, int workgroups,
                 void* stream);
// Zero the work counters and per-job stats, and reset the timestamps,
// with a kernel (no memset nodes, so the sequence can be graph-captured).
int launch_batch_reset(uint32_t* counters, int n_counters, unsigned long long* t_first,
                       unsigned long long* t_done, int n_jobs, unsigned long long* stats,
                       int n_stats, void* stream);

// Volume export: densities (FP64) into dens, max into *maxbits (must start
// at 0), bytes into out; res^3 entries each.
int launch_volume(const MarchParams& p, int res, double* dens,
                  unsigned long long* maxbits, unsigned char* out, void* stream);

// Launch on the given stream; returns hipError_t as int.
int launch_march(const MarchParams& p, int precision, int integrator,
                 void* stream);

// Fixed per-GPU scratch region of xrp_render_device_noalloc (noalloc.hip):
// its size, and its device address on `device`, which must be current (the
// address is looked up once per device and kept).  Returns hipError_t.
constexpr size_t kNoAllocScratchBytes = (size_t)2 << 20;
int noalloc_scratch(int device, void** out);

}  // namespace xrp
