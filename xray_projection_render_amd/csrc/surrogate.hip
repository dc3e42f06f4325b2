// surrogate.hip - training and inference kernels of the learned density
// surrogate (hash encoding + 32-64-64-1 MLP, BF16 MFMA, FP32 accumulation).
//
// Wave layout.  One wave evaluates 16 samples at a time: lane l serves
// sample s = l % 16 and lane group g = l / 16.  The MLP layers are computed
// transposed, H^T = W^T X^T, with v_mfma_f32_16x16x16_bf16: the weights are
// the A operand (A[m][k] in lane m%16... see below) and the activations the
// B operand.  For that instruction lane l holds
//   A[m = l%16][k = 4(l/16) + i],  B[k = 4(l/16) + i][n = l%16],
//   D[m = 4(l/16) + i][n = l%16],  i = 0..3,
// so a layer's output tile (rows = units 4g+i, column = sample s) is, after
// bias, ReLU and BF16 rounding, exactly the B operand of the next layer's
// K block.  The 32 encoding features therefore split over the lane groups
// as features 4g..4g+3 and 16+4g..16+4g+3, i.e. hash levels 2g, 2g+1,
// 8+2g and 9+2g: every lane encodes four levels of its sample.
#include <hip/hip_runtime.h>

#include <cmath>

#include "eval_device.hpp"
#include "march.hpp"
#include "surrogate.hpp"

namespace xrp {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ short f2bf(float f) {  // round to nearest even
  unsigned u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned h) { return __uint_as_float(h << 16); }

__device__ __forceinline__ f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

struct SurOffsets {
  size_t W1, b1, W2, b2, w3, b3;
};
__host__ __device__ inline SurOffsets sur_offsets(int log2_table) {
  const size_t t = sur_table_floats(log2_table);
  SurOffsets o;
  o.W1 = t;
  o.b1 = o.W1 + kSurIn * kSurHidden;
  o.W2 = o.b1 + kSurHidden;
  o.b2 = o.W2 + kSurHidden * kSurHidden;
  o.w3 = o.b2 + kSurHidden;
  o.b3 = o.w3 + kSurHidden;
  return o;
}

// ---------------------------------------------------------------------------
// Hash encoding of one level: corner indices and trilinear weights.
// ---------------------------------------------------------------------------
struct LevelCell {
  uint32_t idx[8];
  float w[8];
};

__device__ __forceinline__ LevelCell level_cell(int N, uint32_t tmask, bool dense, float ux,
                                                float uy, float uz) {
  // inputs outside [0, 1] (and NaN) are clamped, so every index stays in the table
  ux = fminf(fmaxf(ux, 0.0f), 1.0f);
  uy = fminf(fmaxf(uy, 0.0f), 1.0f);
  uz = fminf(fmaxf(uz, 0.0f), 1.0f);
  const float gx = ux * (float)N, gy = uy * (float)N, gz = uz * (float)N;
  const float fx = fminf(floorf(gx), (float)(N - 1)), fy = fminf(floorf(gy), (float)(N - 1)),
              fz = fminf(floorf(gz), (float)(N - 1));
  const float wx = gx - fx, wy = gy - fy, wz = gz - fz;
  const uint32_t bx = (uint32_t)fx, by = (uint32_t)fy, bz = (uint32_t)fz;
  const uint32_t n1 = (uint32_t)N + 1u;
  LevelCell c;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const uint32_t ox = (k >> 2) & 1, oy = (k >> 1) & 1, oz = k & 1;
    const uint32_t cx = bx + ox, cy = by + oy, cz = bz + oz;
    c.idx[k] = dense ? cx + n1 * (cy + n1 * cz)
                     : ((cx * 1u) ^ (cy * 2654435761u) ^ (cz * 805459861u)) & tmask;
    c.w[k] = (ox ? wx : 1.0f - wx) * (oy ? wy : 1.0f - wy) * (oz ? wz : 1.0f - wz);
  }
  return c;
}

__device__ __forceinline__ bool level_dense(int N, int log2_table) {
  const double n1 = (double)N + 1.0;
  return n1 * n1 * n1 <= (double)(1u << log2_table);
}

// Levels of lane group g, in feature order: features 4g..4g+3 are levels
// 2g, 2g+1; features 16+4g..16+4g+3 are levels 8+2g, 9+2g.
__device__ __forceinline__ int lane_level(int g, int q) {
  return (q < 2 ? 2 * g : 8 + 2 * g) + (q & 1);
}

// Features of the lane's four levels from the FP32 tables.
__device__ __forceinline__ void encode_f32(const SurNet& net, int g, float ux, float uy,
                                           float uz, float (&feat)[8]) {
  const uint32_t T = 1u << net.log2_table, tmask = T - 1u;
  const float2* tab = reinterpret_cast<const float2*>(net.params);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int l = lane_level(g, q);
    const int N = net.res[l];
    const LevelCell c = level_cell(N, tmask, level_dense(N, net.log2_table), ux, uy, uz);
    float f0 = 0.0f, f1 = 0.0f;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const float2 e = tab[(size_t)l * T + c.idx[k]];
      f0 = fmaf(c.w[k], e.x, f0);
      f1 = fmaf(c.w[k], e.y, f1);
    }
    feat[2 * q] = f0;
    feat[2 * q + 1] = f1;
  }
}

// The same from the packed BF16 inference tables (one dword per entry).
__device__ __forceinline__ void encode_bf16(const SurNet& net, int g, float ux, float uy,
                                            float uz, float (&feat)[8]) {
  const uint32_t T = 1u << net.log2_table, tmask = T - 1u;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int l = lane_level(g, q);
    const int N = net.res[l];
    const LevelCell c = level_cell(N, tmask, level_dense(N, net.log2_table), ux, uy, uz);
    const uint32_t* tab = net.table_bf16 + (size_t)l * T;
    float f0 = 0.0f, f1 = 0.0f;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const uint32_t e = tab[c.idx[k]];
      f0 = fmaf(c.w[k], bf2f(e & 0xffffu), f0);
      f1 = fmaf(c.w[k], bf2f(e >> 16), f1);
    }
    feat[2 * q] = f0;
    feat[2 * q + 1] = f1;
  }
}

// ---------------------------------------------------------------------------
// Weight operands staged in LDS, pre-arranged per lane (one 8-byte read).
// ---------------------------------------------------------------------------
struct SurFwdLds {
  bf16x4 a1[4][2][64];  // layer 1: A = W1^T  [mt][k step][lane]
  bf16x4 a2[4][4][64];  // layer 2: A = W2^T  [mt][k block][lane]
};
struct SurBwdLds {
  bf16x4 bw2[4][4][64];  // dH1^T = W2 dH2^T:  A = W2  [in tile][out block][lane]
  bf16x4 bw1[2][4][64];  // dX^T  = W1 dH1^T:  A = W1  [feat tile][hidden block][lane]
};

__device__ __forceinline__ void stage_fwd(SurFwdLds& L, const float* W1, const float* W2) {
  for (int e = threadIdx.x; e < 4 * 2 * 64; e += blockDim.x) {
    const int lane = e & 63, ks = (e >> 6) & 1, mt = e >> 7;
    const int s = lane & 15, g = lane >> 4;
    bf16x4 v;
#pragma unroll
    for (int i = 0; i < 4; ++i)
      v[i] = f2bf(W1[(size_t)(16 * ks + 4 * g + i) * kSurHidden + 16 * mt + s]);
    L.a1[mt][ks][lane] = v;
  }
  for (int e = threadIdx.x; e < 4 * 4 * 64; e += blockDim.x) {
    const int lane = e & 63, kb = (e >> 6) & 3, mt = e >> 8;
    const int s = lane & 15, g = lane >> 4;
    bf16x4 v;
#pragma unroll
    for (int i = 0; i < 4; ++i)
      v[i] = f2bf(W2[(size_t)(16 * kb + 4 * g + i) * kSurHidden + 16 * mt + s]);
    L.a2[mt][kb][lane] = v;
  }
}

__device__ __forceinline__ void stage_bwd(SurBwdLds& L, const float* W1, const float* W2) {
  for (int e = threadIdx.x; e < 4 * 4 * 64; e += blockDim.x) {
    const int lane = e & 63, kb = (e >> 6) & 3, mt = e >> 8;
    const int s = lane & 15, g = lane >> 4;
    bf16x4 v;
#pragma unroll
    for (int i = 0; i < 4; ++i)
      v[i] = f2bf(W2[(size_t)(16 * mt + s) * kSurHidden + 16 * kb + 4 * g + i]);
    L.bw2[mt][kb][lane] = v;
  }
  for (int e = threadIdx.x; e < 2 * 4 * 64; e += blockDim.x) {
    const int lane = e & 63, kb = (e >> 6) & 3, mt = e >> 8;
    const int s = lane & 15, g = lane >> 4;
    bf16x4 v;
#pragma unroll
    for (int i = 0; i < 4; ++i)
      v[i] = f2bf(W1[(size_t)(16 * mt + s) * kSurHidden + 16 * kb + 4 * g + i]);
    L.bw1[mt][kb][lane] = v;
  }
}

// Forward pass of one 16-sample group.  Returns sigmoid(z) on every lane of
// the sample; h1/h2 receive the post-bias pre-activations.
struct SurAct {
  f32x4 h1[4], h2[4];
};

__device__ __forceinline__ float mlp_forward(const SurFwdLds& L, const float* b1,
                                             const float* b2, const float* w3, float b3,
                                             int lane, const float (&feat)[8], SurAct& a) {
  const int g = lane >> 4;
  bf16x4 x0, x1;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    x0[i] = f2bf(feat[i]);
    x1[i] = f2bf(feat[4 + i]);
  }
  bf16x4 h1b[4];
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) {
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    acc = mfma16(L.a1[mt][0][lane], x0, acc);
    acc = mfma16(L.a1[mt][1][lane], x1, acc);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      acc[i] += b1[16 * mt + 4 * g + i];
      h1b[mt][i] = f2bf(fmaxf(acc[i], 0.0f));
    }
    a.h1[mt] = acc;
  }
  float part = 0.0f;
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) {
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) acc = mfma16(L.a2[mt][kb][lane], h1b[kb], acc);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      acc[i] += b2[16 * mt + 4 * g + i];
      part = fmaf(w3[16 * mt + 4 * g + i], fmaxf(acc[i], 0.0f), part);
    }
    a.h2[mt] = acc;
  }
  part += __shfl_xor(part, 16, 64);
  part += __shfl_xor(part, 32, 64);
  const float z = part + b3;
  return 1.0f / (1.0f + expf(-z));
}

// ---------------------------------------------------------------------------
// Training data.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t mix64(uint64_t z) {  // splitmix64 finalizer
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__device__ __forceinline__ float unit_float(uint64_t h) {
  return (float)(h >> 40) * (1.0f / 16777216.0f);  // 24 bits -> [0, 1)
}

__global__ void sur_labels_kernel(SurNet net, MarchParams scene, uint64_t seed, int64_t n,
                                  float* pts, float* labels) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const uint64_t base = seed * 0x9E3779B97F4A7C15ull + (uint64_t)k * 3ull;
  const float u0 = unit_float(mix64(base + 1)), u1 = unit_float(mix64(base + 2)),
              u2 = unit_float(mix64(base + 3));
  const double x = (double)net.lo[0] + (double)u0 * (double)net.ext[0];
  const double y = (double)net.lo[1] + (double)u1 * (double)net.ext[1];
  const double z = (double)net.lo[2] + (double)u2 * (double)net.ext[2];
  const double rho = node64<0>(scene.nodes, 0, x, y, z);
  pts[3 * k] = u0;
  pts[3 * k + 1] = u1;
  pts[3 * k + 2] = u2;
  labels[k] = (float)(rho / (double)net.rho_scale);
}

// ---------------------------------------------------------------------------
// Training step: forward, backward, gradient accumulation.
// ---------------------------------------------------------------------------
constexpr int kSurTrainWaves = 4;
constexpr int kSurGroupsPerWave = 8;  // 128 samples per wave between flushes

struct SurWaveLds {
  short t1[64][16];  // h1 (post ReLU, bf16) [unit][sample]
  short t2[64][16];  // dH2 pre-activation gradient
  short tx[32][16];  // encoding features
  short t3[64][16];  // dH1 pre-activation gradient
};

__device__ __forceinline__ bf16x4 lds_row4(const short (*t)[16], int unit, int s0) {
  bf16x4 v;
#pragma unroll
  for (int i = 0; i < 4; ++i) v[i] = t[unit][s0 + i];
  return v;
}

__global__ __launch_bounds__(256) void sur_train_kernel(SurNet net, const float* pts,
                                                        const float* labels, int64_t n,
                                                        float inv_n, float* grads,
                                                        double* loss_sum) {
  __shared__ SurFwdLds F;
  __shared__ SurBwdLds B;
  __shared__ SurWaveLds WL[kSurTrainWaves];
  const SurOffsets off = sur_offsets(net.log2_table);
  const float* P = net.params;
  stage_fwd(F, P + off.W1, P + off.W2);
  stage_bwd(B, P + off.W1, P + off.W2);
  __syncthreads();

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int s = lane & 15, g = lane >> 4;
  SurWaveLds& W = WL[wave];
  const float* b1 = P + off.b1;
  const float* b2 = P + off.b2;
  const float* w3 = P + off.w3;
  const float b3 = P[off.b3];
  const uint32_t T = 1u << net.log2_table, tmask = T - 1u;
  float* gtab = grads;

  f32x4 accW2[4][4], accW1[2][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) accW2[a][b] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) accW1[a][b] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  float db1[16], db2[16], dw3[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) db1[q] = db2[q] = dw3[q] = 0.0f;
  float db3 = 0.0f, sq = 0.0f;

  const int64_t wave_id = (int64_t)blockIdx.x * kSurTrainWaves + wave;
  const int64_t first = wave_id * 16 * kSurGroupsPerWave;
  for (int grp = 0; grp < kSurGroupsPerWave; ++grp) {
    const int64_t k = first + (int64_t)grp * 16 + s;
    if (first + (int64_t)grp * 16 >= n) break;  // wave-uniform
    const bool ok = k < n;
    const float ux = ok ? pts[3 * k] : 0.5f, uy = ok ? pts[3 * k + 1] : 0.5f,
                uz = ok ? pts[3 * k + 2] : 0.5f;
    const float t = ok ? labels[k] : 0.0f;
    float feat[8];
    encode_f32(net, g, ux, uy, uz, feat);
    SurAct a;
    const float y = mlp_forward(F, b1, b2, w3, b3, lane, feat, a);
    const float err = ok ? y - t : 0.0f;
    if (g == 0) sq = fmaf(err, err, sq);
    const float dz = 2.0f * err * inv_n * y * (1.0f - y);
    if (g == 0) db3 += dz;
    // layer 3 and the H2 ReLU
    bf16x4 dh2b[4];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int u = 16 * mt + 4 * g + i;
        const float h2 = fmaxf(a.h2[mt][i], 0.0f);
        dw3[4 * mt + i] = fmaf(dz, h2, dw3[4 * mt + i]);
        const float d = a.h2[mt][i] > 0.0f ? dz * w3[u] : 0.0f;
        db2[4 * mt + i] += d;
        dh2b[mt][i] = f2bf(d);
        W.t2[u][s] = dh2b[mt][i];
        W.t1[u][s] = f2bf(fmaxf(a.h1[mt][i], 0.0f));
      }
    // dH1^T = W2 dH2^T, then the H1 ReLU
    bf16x4 dh1b[4];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int kb = 0; kb < 4; ++kb) acc = mfma16(B.bw2[mt][kb][lane], dh2b[kb], acc);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const float d = a.h1[mt][i] > 0.0f ? acc[i] : 0.0f;
        db1[4 * mt + i] += d;
        dh1b[mt][i] = f2bf(d);
        W.t3[16 * mt + 4 * g + i][s] = dh1b[mt][i];
      }
    }
    // dX^T = W1 dH1^T: the lane's own features
    float dx[8];
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
      f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int kb = 0; kb < 4; ++kb) acc = mfma16(B.bw1[mt][kb][lane], dh1b[kb], acc);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        dx[4 * mt + i] = acc[i];
        W.tx[16 * mt + 4 * g + i][s] = f2bf(feat[4 * mt + i]);
      }
    }
    // hash table gradients (only real samples)
    if (ok) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int l = lane_level(g, q);
        const int N = net.res[l];
        const LevelCell c = level_cell(N, tmask, level_dense(N, net.log2_table), ux, uy, uz);
        float* gl = gtab + (size_t)l * T * 2;
#pragma unroll
        for (int kk = 0; kk < 8; ++kk) {
          atomicAdd(gl + 2 * (size_t)c.idx[kk], c.w[kk] * dx[2 * q]);
          atomicAdd(gl + 2 * (size_t)c.idx[kk] + 1, c.w[kk] * dx[2 * q + 1]);
        }
      }
    }
    // weight gradients over the group's 16 samples, from the transposes in
    // LDS (this wave's own region; DS operations of one wave complete in
    // order, the fence keeps the compiler from reordering them)
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      const bf16x4 av = lds_row4(W.t1, 16 * mt + s, 4 * g);
#pragma unroll
      for (int nt = 0; nt < 4; ++nt)
        accW2[mt][nt] = mfma16(av, lds_row4(W.t2, 16 * nt + s, 4 * g), accW2[mt][nt]);
    }
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
      const bf16x4 av = lds_row4(W.tx, 16 * mt + s, 4 * g);
#pragma unroll
      for (int nt = 0; nt < 4; ++nt)
        accW1[mt][nt] = mfma16(av, lds_row4(W.t3, 16 * nt + s, 4 * g), accW1[mt][nt]);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }

  // flush: weight tiles hold D[m = 4g+i][n = s] of each (mt, nt) tile
  float* gW1 = grads + off.W1;
  float* gW2 = grads + off.W2;
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int nt = 0; nt < 4; ++nt)
#pragma unroll
      for (int i = 0; i < 4; ++i)
        atomicAdd(gW2 + (size_t)(16 * mt + 4 * g + i) * kSurHidden + 16 * nt + s,
                  accW2[mt][nt][i]);
#pragma unroll
  for (int mt = 0; mt < 2; ++mt)
#pragma unroll
    for (int nt = 0; nt < 4; ++nt)
#pragma unroll
      for (int i = 0; i < 4; ++i)
        atomicAdd(gW1 + (size_t)(16 * mt + 4 * g + i) * kSurHidden + 16 * nt + s,
                  accW1[mt][nt][i]);
  // per-unit sums over the 16 samples (lanes of one group)
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    float a1 = db1[q], a2 = db2[q], a3 = dw3[q];
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) {
      a1 += __shfl_xor(a1, o, 64);
      a2 += __shfl_xor(a2, o, 64);
      a3 += __shfl_xor(a3, o, 64);
    }
    if (s == 0) {
      const int u = 16 * (q >> 2) + 4 * g + (q & 3);
      atomicAdd(grads + off.b1 + u, a1);
      atomicAdd(grads + off.b2 + u, a2);
      atomicAdd(grads + off.w3 + u, a3);
    }
  }
  float r3 = db3, rs = sq;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    r3 += __shfl_xor(r3, o, 64);
    rs += __shfl_xor(rs, o, 64);
  }
  if (lane == 0) {
    atomicAdd(grads + off.b3, r3);
    atomicAdd(loss_sum, (double)rs);
  }
}

__global__ void sur_adam_kernel(float* params, const float* grads, float* m, float* v,
                                int64_t count, int64_t table_count, float lr_table,
                                float lr_mlp, float beta1, float beta2, float eps, float c1,
                                float c2) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= count) return;
  const float gk = grads[k];
  const float mk = beta1 * m[k] + (1.0f - beta1) * gk;
  const float vk = beta2 * v[k] + (1.0f - beta2) * gk * gk;
  m[k] = mk;
  v[k] = vk;
  const float lr = k < table_count ? lr_table : lr_mlp;
  params[k] -= lr * (mk * c1) / (sqrtf(vk * c2) + eps);
}

__global__ void sur_pack_kernel(const float* tables, uint32_t* out, int64_t entries) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= entries) return;
  const uint32_t lo = (uint16_t)f2bf(tables[2 * k]), hi = (uint16_t)f2bf(tables[2 * k + 1]);
  out[k] = lo | (hi << 16);
}

__global__ __launch_bounds__(256) void sur_predict_kernel(SurNet net, const float* pts,
                                                          int64_t n, float* out) {
  __shared__ SurFwdLds F;
  const SurOffsets off = sur_offsets(net.log2_table);
  const float* P = net.params;
  stage_fwd(F, P + off.W1, P + off.W2);
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int s = lane & 15, g = lane >> 4;
  const int64_t k = ((int64_t)blockIdx.x * 4 + wave) * 16 + s;
  const bool ok = k < n;
  const float ux = ok ? pts[3 * k] : 0.5f, uy = ok ? pts[3 * k + 1] : 0.5f,
              uz = ok ? pts[3 * k + 2] : 0.5f;
  float feat[8];
  encode_f32(net, g, ux, uy, uz, feat);
  SurAct a;
  const float y = mlp_forward(F, P + off.b1, P + off.b2, P + off.w3, P[off.b3], lane, feat, a);
  if (ok && g == 0) out[k] = y;
}

// ---------------------------------------------------------------------------
// Surrogate march.  One wave = 16 rays (a 4x4 pixel block), one sample of
// each per step, midpoint rule with step ds_sur; the four lanes of a ray
// encode four levels each.  FP64 ray set-up and clipping, FP32 positions
// re-anchored in FP64 every kChunk steps, Kahan-compensated FP32 sum.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool sur_domain(const SurNet& net, float x, float y, float z,
                                           float& ux, float& uy, float& uz) {
  if (net.wrap) {
    if (x < net.bounds[0] || x > net.bounds[1] || y < net.bounds[2] || y > net.bounds[3] ||
        z < net.bounds[4] || z > net.bounds[5])
      return false;
    ux = (x - net.lo[0]) * net.inv_ext[0];
    uy = (y - net.lo[1]) * net.inv_ext[1];
    uz = (z - net.lo[2]) * net.inv_ext[2];
    ux -= floorf(ux);
    uy -= floorf(uy);
    uz -= floorf(uz);
    return true;
  }
  ux = (x - net.lo[0]) * net.inv_ext[0];
  uy = (y - net.lo[1]) * net.inv_ext[1];
  uz = (z - net.lo[2]) * net.inv_ext[2];
  return ux >= 0.0f && ux <= 1.0f && uy >= 0.0f && uy <= 1.0f && uz >= 0.0f && uz <= 1.0f;
}

// One wave per workgroup (16 rays, a 4x4 pixel tile, each with its own LDS
// copy of the weights): waves whose rays end early free their slot at once
// (5 % faster than 8x8 tiles of four waves).
__global__ __launch_bounds__(64) void sur_render_kernel(SurNet net, MarchParams p,
                                                         double ds_sur, long K) {
  __shared__ SurFwdLds F;
  const SurOffsets off = sur_offsets(net.log2_table);
  const float* P = net.params;
  stage_fwd(F, P + off.W1, P + off.W2);
  __syncthreads();
  const float* b1 = P + off.b1;
  const float* b2 = P + off.b2;
  const float* w3 = P + off.w3;
  const float b3 = P[off.b3];

  const int lane = threadIdx.x & 63;
  const int s = lane & 15, g = lane >> 4;
  const int tiles_j = (p.res + 3) >> 2;
  const int ti = blockIdx.x / tiles_j, tj = blockIdx.x - ti * tiles_j;
  const int i = ti * 4 + (s >> 2);
  const int j = tj * 4 + (s & 3);
  const int v = blockIdx.y;
  const bool valid = i < p.res && j < p.res;

  long k_lo = 1, k_hi = 0;
  Ray r;
  r.o[0] = r.o[1] = r.o[2] = 0.0;
  r.d[0] = r.d[1] = r.d[2] = 0.0;
  if (valid) {
    r = make_ray(p, p.views[p.view_offset + v], i, j);
    double t_in, t_out;
    if (clip_ray(p, r, t_in, t_out)) {
      double a = floor((t_in - p.smin) / ds_sur - 0.5) - 1.0;
      double b = ceil((t_out - p.smin) / ds_sur - 0.5) + 1.0;
      k_lo = a > 0.0 ? (long)a : 0;
      k_hi = b < (double)(K - 1) ? (long)b : K - 1;
    }
  }
  long wlo = k_lo, whi = k_hi;
  // the union of the wave's sample ranges (empty ranges have lo > hi)
  if (k_lo > k_hi) {
    wlo = 0x7fffffffffffL;
    whi = -1;
  }
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    wlo = min(wlo, (long)__shfl_xor(wlo, o, 64));
    whi = max(whi, (long)__shfl_xor(whi, o, 64));
  }
  if (p.stats && lane == 0 && whi >= wlo)
    atomicAdd(p.stats + ST_SAMPLES, (unsigned long long)(whi - wlo + 1) * 16ull);
  float acc = 0.0f, comp = 0.0f;
  const float d0 = (float)r.d[0], d1 = (float)r.d[1], d2 = (float)r.d[2];
  const float dsf = (float)ds_sur;
  for (long k0 = wlo; k0 <= whi; k0 += kChunk) {
    const double sa = p.smin + ((double)k0 + 0.5) * ds_sur;
    const float a0 = (float)(r.o[0] + r.d[0] * sa), a1 = (float)(r.o[1] + r.d[1] * sa),
                a2 = (float)(r.o[2] + r.d[2] * sa);
    const int n = (int)min((long)kChunk, whi - k0 + 1);
    for (int jj = 0; jj < n; ++jj) {
      const float sj = (float)jj * dsf;
      float x = fmaf(d0, sj, a0), y = fmaf(d1, sj, a1), z = fmaf(d2, sj, a2);
      deform32(p.defs, p.n_defs, x, y, z);
      float ux = 0.5f, uy = 0.5f, uz = 0.5f;
      const long k = k0 + jj;
      const bool in = sur_domain(net, x, y, z, ux, uy, uz) && k >= k_lo && k <= k_hi;
      float feat[8];
      encode_bf16(net, g, in ? ux : 0.5f, in ? uy : 0.5f, in ? uz : 0.5f, feat);
      SurAct act;
      const float dens = mlp_forward(F, b1, b2, w3, b3, lane, feat, act);
      const float rho = in ? dens * net.rho_scale : 0.0f;
      const float yk = rho - comp;
      const float tt = acc + yk;
      comp = (tt - acc) - yk;
      acc = tt;
    }
  }
  if (valid && g == 0) {
    const double T = p.flat_field + ((double)acc * p.dm) * ds_sur;
    store_pixel(p, v, i, j, T);
  }
}

// ---------------------------------------------------------------------------
// MFMA operand layout probe: a, b, d are 16x16 row-major float matrices.
// ---------------------------------------------------------------------------
__global__ void mfma_probe_kernel(const float* a, const float* b, float* d) {
  const int lane = threadIdx.x & 63;
  const int r = lane & 15, g = lane >> 4;
  bf16x4 av, bv;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    av[i] = f2bf(a[r * 16 + 4 * g + i]);   // A[m = lane%16][k = 4g+i]
    bv[i] = f2bf(b[(4 * g + i) * 16 + r]);  // B[k = 4g+i][n = lane%16]
  }
  f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
  acc = mfma16(av, bv, acc);
#pragma unroll
  for (int i = 0; i < 4; ++i) d[(4 * g + i) * 16 + r] = acc[i];  // D[m = 4g+i][n = lane%16]
}

// Image values at listed pixels (view, i, j) of a [view][i][j] image.
__global__ void gather_pixels_kernel(const float* img, const int* items, int n, int res,
                                     float* out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const size_t off =
      ((size_t)items[3 * k] * res + (size_t)items[3 * k + 1]) * res + (size_t)items[3 * k + 2];
  out[k] = img[off];
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void sur_certify_kernel(float* img, const float* exact,
                                                          int res, float tol,
                                                          unsigned* counters) {
  __shared__ float wave_max[4];
  const int tiles_side = (res + 15) / 16;
  const int ti = blockIdx.x / tiles_side, tj = blockIdx.x - ti * tiles_side;
  const int i = ti * 16 + (threadIdx.x >> 4), j = tj * 16 + (threadIdx.x & 15);
  const bool in = i < res && j < res;
  const size_t at = ((size_t)blockIdx.y * res + (in ? i : 0)) * res + (in ? j : 0);
  const float e = in ? exact[at] : 0.0f;
  float d = in ? fabsf(img[at] - e) : 0.0f;
  if (!(d == d)) d = INFINITY;  // a NaN never passes
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) d = fmaxf(d, __shfl_xor(d, off, 64));
  if ((threadIdx.x & 63) == 0) wave_max[threadIdx.x >> 6] = d;
  __syncthreads();
  const float m = fmaxf(fmaxf(wave_max[0], wave_max[1]), fmaxf(wave_max[2], wave_max[3]));
  if (m > tol && in) img[at] = e;
  if (threadIdx.x == 0) {
    if (m > tol) atomicAdd(counters, 1u);
    atomicMax(counters + 1, __float_as_uint(m));  // non-negative: bits order like values
  }
}

int launch_sur_certify(float* img, const float* exact, int n_views, int res, float tol,
                       unsigned* counters, void* stream) {
  if (n_views <= 0 || res <= 0) return 0;
  const int tiles_side = (res + 15) / 16;
  hipLaunchKernelGGL(sur_certify_kernel, dim3(tiles_side * tiles_side, n_views), dim3(256), 0,
                     (hipStream_t)stream, img, exact, res, tol, counters);
  return (int)hipGetLastError();
}

int launch_gather_pixels(const float* img, const int* items, int n, int res, float* out,
                         void* stream) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(gather_pixels_kernel, dim3((n + 255) / 256), dim3(256), 0,
                     (hipStream_t)stream, img, items, n, res, out);
  return (int)hipGetLastError();
}

int launch_sur_labels(const SurNet& net, const MarchParams& scene, uint64_t seed, int64_t n,
                      float* pts, float* labels, void* stream) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(sur_labels_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, net, scene, seed, n, pts, labels);
  return (int)hipGetLastError();
}

int launch_sur_train(const SurNet& net, const float* pts, const float* labels, int64_t n,
                     float inv_n, float* grads, double* loss_sum, void* stream) {
  if (n <= 0) return 0;
  const int64_t per_block = (int64_t)kSurTrainWaves * 16 * kSurGroupsPerWave;
  hipLaunchKernelGGL(sur_train_kernel, dim3((unsigned)((n + per_block - 1) / per_block)),
                     dim3(64 * kSurTrainWaves), 0, (hipStream_t)stream, net, pts, labels, n,
                     inv_n, grads, loss_sum);
  return (int)hipGetLastError();
}

int launch_sur_adam(float* params, const float* grads, float* m, float* v, int64_t count,
                    int64_t table_count, float lr_table, float lr_mlp, float beta1,
                    float beta2, float eps, int step, void* stream) {
  const float c1 = 1.0f / (1.0f - powf(beta1, (float)step));
  const float c2 = 1.0f / (1.0f - powf(beta2, (float)step));
  hipLaunchKernelGGL(sur_adam_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, params, grads, m, v, count, table_count, lr_table,
                     lr_mlp, beta1, beta2, eps, c1, c2);
  return (int)hipGetLastError();
}

int launch_sur_pack_tables(const SurNet& net, void* stream) {
  const int64_t entries = (int64_t)kSurLevels << net.log2_table;
  hipLaunchKernelGGL(sur_pack_kernel, dim3((unsigned)((entries + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, net.params, net.table_bf16, entries);
  return (int)hipGetLastError();
}

int launch_sur_predict(const SurNet& net, const float* pts, int64_t n, float* out,
                       void* stream) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(sur_predict_kernel, dim3((unsigned)((n + 63) / 64)), dim3(256), 0,
                     (hipStream_t)stream, net, pts, n, out);
  return (int)hipGetLastError();
}

int launch_sur_render(const SurNet& net, const MarchParams& p, double ds_sur, void* stream) {
  const long K = (long)std::ceil((p.smax - p.smin) / ds_sur);
  const int tiles = (p.res + 3) / 4;
  hipLaunchKernelGGL(sur_render_kernel, dim3(tiles * tiles, p.n_views), dim3(64), 0,
                     (hipStream_t)stream, net, p, ds_sur, K);
  return (int)hipGetLastError();
}

int launch_mfma_probe(const float* a, const float* b, float* d, void* stream) {
  hipLaunchKernelGGL(mfma_probe_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream, a, b, d);
  return (int)hipGetLastError();
}

}  // namespace xrp
