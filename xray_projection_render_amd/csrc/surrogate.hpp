// surrogate.hpp - learned density surrogate: a multiresolution hash
// encoding (16 levels x 2 features) feeding a 32-64-64-1 MLP, trained on
// exact density samples and evaluated with BF16 MFMA in the march.
//
// Parameter layout (FP32 master copy, one contiguous buffer):
//   tables [L][T][2] | W1 [32][64] | b1 [64] | W2 [64][64] | b2 [64] |
//   w3 [64] | b3 [1]       (W[in][out], row-major)
// Gradients and the Adam moments use the same layout.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "march.hpp"

namespace xrp {

constexpr int kSurLevels = 16;
constexpr int kSurHidden = 64;
constexpr int kSurIn = 32;
constexpr int kSurMlpParams = kSurIn * kSurHidden + kSurHidden + kSurHidden * kSurHidden +
                              kSurHidden + kSurHidden + 1;

struct SurNet {
  float* params;        // FP32 master parameters (layout above)
  uint32_t* table_bf16; // inference copy of the tables: 2 x bf16 per entry
  int32_t log2_table;
  int32_t res[kSurLevels];
  // Input map: u = (p - lo) * inv_ext.  With wrap (tessellated scenes) p
  // is first tested against the tessellation bounds (outside: density 0)
  // and wrapped into the unit cell [lo, lo + ext) as objects.go:568-582.
  int32_t wrap;
  float lo[3], ext[3], inv_ext[3];
  float bounds[6];      // wrap: tessellation bounds (xmin xmax ymin ymax zmin zmax)
  float rho_scale;      // density = rho_scale * sigmoid(z)
};

__host__ __device__ inline size_t sur_table_floats(int log2_table) {
  return (size_t)kSurLevels * ((size_t)1 << log2_table) * 2;
}
__host__ __device__ inline size_t sur_param_count(int log2_table) {
  return sur_table_floats(log2_table) + (size_t)kSurMlpParams;
}

// Exact training data: n points uniform in the network's input cube and
// the exact FP64 density there (the scene's tree without deformations).
int launch_sur_labels(const SurNet& net, const MarchParams& scene, uint64_t seed,
                      int64_t n, float* pts, float* labels, void* stream);
// Loss and gradients of a batch (grads accumulated, loss_sum += sum of
// squared errors).  Points are u in [0,1]^3.
int launch_sur_train(const SurNet& net, const float* pts, const float* labels, int64_t n,
                     float inv_n, float* grads, double* loss_sum, void* stream);
int launch_sur_adam(float* params, const float* grads, float* m, float* v, int64_t count,
                    int64_t table_count, float lr_table, float lr_mlp, float beta1,
                    float beta2, float eps, int step, void* stream);
int launch_sur_pack_tables(const SurNet& net, void* stream);
// Network output at points u (BF16 MFMA path, FP32 tables).
int launch_sur_predict(const SurNet& net, const float* pts, int64_t n, float* out,
                       void* stream);
// Surrogate march of every pixel: p carries the cameras, clip box, output
// and deformations; ds_sur is the surrogate's step (midpoint rule).
int launch_sur_render(const SurNet& net, const MarchParams& p, double ds_sur, void* stream);
// out[k] = img[items[3k]][items[3k+1]][items[3k+2]] of a res x res image stack.
int launch_gather_pixels(const float* img, const int* items, int n, int res, float* out,
                         void* stream);
// Certified tile check: one 256-lane workgroup per 16x16 tile of a
// res x res image stack (grid: tiles, views).  A tile whose largest
// |img - exact| exceeds tol is replaced by the exact values; counters[0]
// counts replaced tiles and counters[1] receives the bits of the largest
// difference seen (both zeroed by the caller).
int launch_sur_certify(float* img, const float* exact, int n_views, int res, float tol,
                       unsigned* counters, void* stream);
// MFMA operand layout probe: D = A (16x16) * B (16x16), bf16 inputs.
int launch_mfma_probe(const float* a, const float* b, float* d, void* stream);

}  // namespace xrp
