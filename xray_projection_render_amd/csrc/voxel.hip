// voxel.hip - march kernel for scenes whose root is a voxel grid seen
// through a linear deformation (VoxelGrid.Density, objects.go:789-855,
// composed with affine/linear/rigid maps, deformations.go:87-175).
//
// Each ray is mapped once, in FP64, to a line in grid coordinates
// G(s) = A + B s with G = ((M p + t) + 1) / 2 * (N - 1).  Sample positions
// are FP32 offsets from an FP64 anchor refreshed every kChunk samples, so
// the grid coordinate carries at most a few 1e-5 voxel of error; trilinear
// weights and the interpolation are FP32 and the integral is FP64.  The
// closed [-1, 1] bounds test is the only discontinuity; it is settled per
// ray in FP64 (exact_range), so the loop runs over exactly the samples the
// reference finds inside and needs no per-sample bounds test.
#include <hip/hip_runtime.h>

#include "eval_device.hpp"
#include "march.hpp"

namespace xrp {

// Raw buffer descriptor (dword3 0x00020000 on gfx9 parts): 32-bit byte
// offsets from an SGPR base, and reads past num_records return 0.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t grid_rsrc(const float* g,
                                                            unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc((void*)g, (short)0, (int)bytes,
                                           0x00020000);
}

template <int LAYOUT>
__device__ __forceinline__ float trilinear(const MarchParams& p,
                                           __amdgpu_buffer_rsrc_t rs, float gx,
                                           float gy, float gz) {
  const int nx = p.vn[0], ny = p.vn[1], nz = p.vn[2];
  int x0 = min((int)gx, nx - 1), y0 = min((int)gy, ny - 1),
      z0 = min((int)gz, nz - 1);
  int x1 = min(x0 + 1, nx - 1), y1 = min(y0 + 1, ny - 1),
      z1 = min(z0 + 1, nz - 1);
  float wx = gx - (float)x0, wy = gy - (float)y0, wz = gz - (float)z0;
  unsigned ox0, ox1, oy0, oy1, oz0, oz1;
  if constexpr (LAYOUT == VOX_LINEAR) {
    oz0 = (unsigned)z0 * (unsigned)(nx * ny);
    oz1 = (unsigned)z1 * (unsigned)(nx * ny);
    ox0 = (unsigned)x0 * (unsigned)ny;
    ox1 = (unsigned)x1 * (unsigned)ny;
    oy0 = (unsigned)y0;
    oy1 = (unsigned)y1;
  } else {
    // 4x4x4 bricks ordered (bz, bx, by); inside a brick (lz, lx, ly).
    const unsigned by_n = (unsigned)p.vbricks[1];
    const unsigned sx = by_n * 64u, sz = (unsigned)p.vbricks[0] * by_n * 64u;
    oz0 = (unsigned)(z0 >> 2) * sz + (unsigned)(z0 & 3) * 16u;
    oz1 = (unsigned)(z1 >> 2) * sz + (unsigned)(z1 & 3) * 16u;
    ox0 = (unsigned)(x0 >> 2) * sx + (unsigned)(x0 & 3) * 4u;
    ox1 = (unsigned)(x1 >> 2) * sx + (unsigned)(x1 & 3) * 4u;
    oy0 = (unsigned)(y0 >> 2) * 64u + (unsigned)(y0 & 3);
    oy1 = (unsigned)(y1 >> 2) * 64u + (unsigned)(y1 & 3);
  }
  // the builtin returns the 32 raw bits as an unsigned int
#define XRP_LD(o) \
  __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs, (int)((o) * 4u), 0, 0))
  float v000 = XRP_LD(oz0 + ox0 + oy0), v001 = XRP_LD(oz1 + ox0 + oy0);
  float v010 = XRP_LD(oz0 + ox0 + oy1), v011 = XRP_LD(oz1 + ox0 + oy1);
  float v100 = XRP_LD(oz0 + ox1 + oy0), v101 = XRP_LD(oz1 + ox1 + oy0);
  float v110 = XRP_LD(oz0 + ox1 + oy1), v111 = XRP_LD(oz1 + ox1 + oy1);
#undef XRP_LD
  float v00 = fmaf(wz, v001 - v000, v000);
  float v01 = fmaf(wz, v011 - v010, v010);
  float v10 = fmaf(wz, v101 - v100, v100);
  float v11 = fmaf(wz, v111 - v110, v110);
  float v0 = fmaf(wy, v01 - v00, v00);
  float v1 = fmaf(wy, v11 - v10, v10);
  return fmaf(wx, v1 - v0, v0);
}

// Exact FP64 membership of sample k in the grid's closed [-1, 1]^3 support
// (objects.go:791), with the position and the linear deformations in the
// reference's operation order (main.go:148-150, deformations.go:87-175).
__device__ __forceinline__ bool inside64(const MarchParams& p, const Ray& r,
                                         long k) {
  double s = p.s_table[k];
  double x = r.o[0] + r.d[0] * s;
  double y = r.o[1] + r.d[1] * s;
  double z = r.o[2] + r.d[2] * s;
  for (int q = 0; q < p.n_defs; ++q) {
    const DevDeform& f = p.defs[q];
    if (f.kind == D_AFFINE) {
      const double* m = f.d;
      double nx = m[0] * x + m[1] * y + m[2] * z;
      double ny = m[3] * x + m[4] * y + m[5] * z;
      double nz = m[6] * x + m[7] * y + m[8] * z;
      x = nx;
      y = ny;
      z = nz;
    } else if (f.kind == D_LINEAR) {
      const double* e = f.d;
      double nx = x + e[0] * x + e[5] * y + e[4] * z;
      double ny = y + e[5] * x + e[1] * y + e[3] * z;
      double nz = z + e[4] * x + e[3] * y + e[2] * z;
      x = nx;
      y = ny;
      z = nz;
    } else {
      x = x + f.d[0];
      y = y + f.d[1];
      z = z + f.d[2];
    }
  }
  return x >= -1 && x <= 1 && y >= -1 && y <= 1 && z >= -1 && z <= 1;
}

// The samples of a ray inside a convex box form one index range.  Start
// from the analytic FP64 interval and settle each end with inside64, so
// the range holds exactly the samples the reference finds inside.
__device__ __forceinline__ bool exact_range(const MarchParams& p, const Ray& r,
                                            double ta, double tb, long& ka,
                                            long& kb) {
  const long last = (long)p.K - 1;
  if (last < 0) return false;
  double a = floor((ta - p.smin) / p.ds), b = floor((tb - p.smin) / p.ds);
  ka = a < 0.0 ? 0 : (a > (double)last ? last : (long)a);
  kb = b < 0.0 ? 0 : (b > (double)last ? last : (long)b);
  // A few steps settle each end: the analytic interval and the rounded
  // sample positions differ by ~1e-15, far less than one step.
  for (int it = 0; it < 4 && ka > 0 && inside64(p, r, ka - 1); ++it) --ka;
  for (int it = 0; it < 4 && ka <= kb + 1 && ka <= last && !inside64(p, r, ka); ++it) ++ka;
  for (int it = 0; it < 4 && kb < last && inside64(p, r, kb + 1); ++it) ++kb;
  for (int it = 0; it < 4 && kb >= ka && !inside64(p, r, kb); ++it) --kb;
  return ka <= kb && ka <= last;
}

template <int LAYOUT>
__global__ __launch_bounds__(256) void voxel_kernel(MarchParams p) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int tiles_j = (p.res + 15) >> 4;
  const int ti = blockIdx.x / tiles_j, tj = blockIdx.x - ti * tiles_j;
  const int i = ti * 16 + (wave >> 1) * 8 + (lane >> 3);
  const int j = tj * 16 + (wave & 1) * 8 + (lane & 7);
  const int view = p.view_offset + blockIdx.y;
  const bool valid = i < p.res && j < p.res;

  unsigned long long n_samples = 0;
  double T = p.flat_field;
  if (valid) {
    const DevView& V = p.views[view];
    Ray r = make_ray(p, V, i, j);
    // Object-space line Ao + Bo s and its slab interval in [-1, 1]^3.
    double Ao[3], Bo[3], ta = -INFINITY, tb = INFINITY;
    bool hit = r.d[0] == r.d[0] && r.d[1] == r.d[1] && r.d[2] == r.d[2];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      Ao[c] = p.M[3 * c] * r.o[0] + p.M[3 * c + 1] * r.o[1] +
              p.M[3 * c + 2] * r.o[2] + p.t[c];
      Bo[c] = p.M[3 * c] * r.d[0] + p.M[3 * c + 1] * r.d[1] +
              p.M[3 * c + 2] * r.d[2];
      if (Bo[c] != 0.0) {
        double t1 = (-1.0 - Ao[c]) / Bo[c], t2 = (1.0 - Ao[c]) / Bo[c];
        ta = fmax(ta, fmin(t1, t2));
        tb = fmin(tb, fmax(t1, t2));
      } else if (!(Ao[c] >= -1.0 && Ao[c] <= 1.0)) {
        hit = false;
      }
    }
    long ka, kb;
    // pad the analytic interval by one step on each side; exact_range
    // then trims it to the samples the reference counts as inside.
    if (hit && ta <= tb && exact_range(p, r, ta - p.ds, tb + p.ds, ka, kb)) {
      double A[3], B[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        double h = 0.5 * (double)(p.vn[c] - 1);
        A[c] = (Ao[c] + 1.0) * h;
        B[c] = Bo[c] * h;
      }
      const float b0 = (float)B[0], b1 = (float)B[1], b2 = (float)B[2];
      const float ds32 = (float)p.ds;
      const double ds = p.ds, dm = p.dm;
      const __amdgpu_buffer_rsrc_t rs = grid_rsrc(p.vgrid, p.vbytes);
      for (long k0 = ka; k0 <= kb; k0 += kChunk) {
        double sa = p.s_table[k0];
        float a0 = (float)(A[0] + B[0] * sa), a1 = (float)(A[1] + B[1] * sa),
              a2 = (float)(A[2] + B[2] * sa);
        int n = (int)min((long)kChunk, kb - k0 + 1);
        float acc = 0.0f, comp = 0.0f;  // Kahan-compensated chunk sum
#pragma unroll 4
        for (int jj = 0; jj < n; ++jj) {
          float sj = (float)jj * ds32;
          float gx = fmaf(b0, sj, a0), gy = fmaf(b1, sj, a1),
                gz = fmaf(b2, sj, a2);
          float rho = trilinear<LAYOUT>(p, rs, fmaxf(gx, 0.0f), fmaxf(gy, 0.0f),
                                        fmaxf(gz, 0.0f));
          float y = rho - comp;
          float tt = acc + y;
          comp = (tt - acc) - y;
          acc = tt;
        }
        T += ((double)acc * dm) * ds;
        n_samples += (unsigned long long)n;
      }
    }
    p.out[(size_t)blockIdx.y * p.out_view_stride + (size_t)i * p.res + j] =
        (float)exp(-T);
  }
  if (p.stats) {
    unsigned long long a = wave_sum(n_samples);
    if (lane == 0 && a) atomicAdd(p.stats + ST_SAMPLES, a);
  }
}

int launch_voxel(const MarchParams& p, int layout, void* stream) {
  dim3 block(256);
  int tiles = (p.res + 15) / 16;
  dim3 grid(tiles * tiles, p.n_views);
  hipStream_t s = (hipStream_t)stream;
  if (layout == VOX_BRICK4)
    hipLaunchKernelGGL((voxel_kernel<VOX_BRICK4>), grid, block, 0, s, p);
  else
    hipLaunchKernelGGL((voxel_kernel<VOX_LINEAR>), grid, block, 0, s, p);
  return (int)hipGetLastError();
}

}  // namespace xrp
