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

// Element offsets of the two x, y and z planes of cell (x0, y0, z0), upper
// indices clamped to N-1 as objects.go:811-830 does.
struct CellOffsets {
  unsigned x[2], y[2], z[2];
};

template <int LAYOUT>
__device__ __forceinline__ CellOffsets cell_offsets(const MarchParams& p, int x0, int y0,
                                                    int z0) {
  const int nx = p.vn[0], ny = p.vn[1], nz = p.vn[2];
  const int x1 = min(x0 + 1, nx - 1), y1 = min(y0 + 1, ny - 1), z1 = min(z0 + 1, nz - 1);
  CellOffsets o;
  if constexpr (LAYOUT == VOX_LINEAR) {
    o.z[0] = (unsigned)z0 * (unsigned)(nx * ny);
    o.z[1] = (unsigned)z1 * (unsigned)(nx * ny);
    o.x[0] = (unsigned)x0 * (unsigned)ny;
    o.x[1] = (unsigned)x1 * (unsigned)ny;
    o.y[0] = (unsigned)y0;
    o.y[1] = (unsigned)y1;
  } else {
    // 4x4x4 bricks ordered (bz, bx, by); inside a brick (lz, lx, ly).
    const unsigned by_n = (unsigned)p.vbricks[1];
    const unsigned sx = by_n * 64u, sz = (unsigned)p.vbricks[0] * by_n * 64u;
    o.z[0] = (unsigned)(z0 >> 2) * sz + (unsigned)(z0 & 3) * 16u;
    o.z[1] = (unsigned)(z1 >> 2) * sz + (unsigned)(z1 & 3) * 16u;
    o.x[0] = (unsigned)(x0 >> 2) * sx + (unsigned)(x0 & 3) * 4u;
    o.x[1] = (unsigned)(x1 >> 2) * sx + (unsigned)(x1 & 3) * 4u;
    o.y[0] = (unsigned)(y0 >> 2) * 64u + (unsigned)(y0 & 3);
    o.y[1] = (unsigned)(y1 >> 2) * 64u + (unsigned)(y1 & 3);
  }
  return o;
}

// the builtin returns the 32 raw bits as an unsigned int
__device__ __forceinline__ float grid_load(__amdgpu_buffer_rsrc_t rs, unsigned elem) {
  return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs, (int)(elem * 4u), 0, 0));
}

// c[4*dx + 2*dy + dz] is the voxel at (x0+dx, y0+dy, z0+dz).
template <int LAYOUT>
__device__ __forceinline__ void load_cell(const MarchParams& p, __amdgpu_buffer_rsrc_t rs,
                                          int x0, int y0, int z0, float (&c)[8]) {
  if constexpr (LAYOUT == VOX_BRICK4_PAIR) {
    const CellOffsets o = cell_offsets<VOX_BRICK4>(p, x0, y0, z0);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int dz = q & 1, dx = q >> 1;
      const auto v = __builtin_amdgcn_raw_buffer_load_b64(
          rs, (int)((o.z[dz] + o.x[dx] + o.y[0]) * 8u), 0, 0);
      c[4 * dx + dz] = __uint_as_float(v[0]);
      c[4 * dx + 2 + dz] = __uint_as_float(v[1]);
    }
  } else if constexpr (LAYOUT == VOX_BRICK4_QUAD) {
    const CellOffsets o = cell_offsets<VOX_BRICK4>(p, x0, y0, z0);
#pragma unroll
    for (int dx = 0; dx < 2; ++dx) {
      const auto v = __builtin_amdgcn_raw_buffer_load_b128(
          rs, (int)((o.z[0] + o.x[dx] + o.y[0]) * 16u), 0, 0);
      c[4 * dx + 0] = __uint_as_float(v[0]);
      c[4 * dx + 1] = __uint_as_float(v[1]);
      c[4 * dx + 2] = __uint_as_float(v[2]);
      c[4 * dx + 3] = __uint_as_float(v[3]);
    }
  } else {
    const CellOffsets o = cell_offsets<LAYOUT>(p, x0, y0, z0);
#pragma unroll
    for (int q = 0; q < 8; ++q)
      c[q] = grid_load(rs, o.z[q & 1] + o.x[q >> 2] + o.y[(q >> 1) & 1]);
  }
}

// Lerp order z, then y, then x (objects.go:845-851).
__device__ __forceinline__ float lerp_cell(const float (&c)[8], float wx, float wy,
                                           float wz) {
  float v00 = fmaf(wz, c[1] - c[0], c[0]);
  float v01 = fmaf(wz, c[3] - c[2], c[2]);
  float v10 = fmaf(wz, c[5] - c[4], c[4]);
  float v11 = fmaf(wz, c[7] - c[6], c[6]);
  float v0 = fmaf(wy, v01 - v00, v00);
  float v1 = fmaf(wy, v11 - v10, v10);
  return fmaf(wx, v1 - v0, v0);
}

template <int LAYOUT>
__device__ __forceinline__ float trilinear(const MarchParams& p,
                                           __amdgpu_buffer_rsrc_t rs, float gx,
                                           float gy, float gz) {
  int x0 = min((int)gx, p.vn[0] - 1), y0 = min((int)gy, p.vn[1] - 1),
      z0 = min((int)gz, p.vn[2] - 1);
  float c[8];
  load_cell<LAYOUT>(p, rs, x0, y0, z0, c);
  return lerp_cell(c, gx - (float)x0, gy - (float)y0, gz - (float)z0);
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

// CACHE_CELL keeps the last cell's 8 corners in registers and reloads them
// only when the cell changes (a step of ds moves a ray through the 256^3
// benchmark grid ~0.4 voxel, so most samples reuse the cell).  The values,
// and so the result, are the same as without the cache.  (Reusing the
// shared face when the cell moves by one along one axis, reading only four
// corners, measured 3x slower: the four divergent paths serialize.)
// CACHE_SLIDE (quad copy only, slide_march below): a cell that moves by one
// along x keeps the x plane it shares with the last cell and loads only the
// other one.
enum VoxelCache { CACHE_NONE = 0, CACHE_CELL = 1, CACHE_SLIDE = 2, CACHE_STEP = 3 };
// Sliding-cache march of one ray over the wave's sample range, on the quad
// copy: a quad holds (z, y) corners (v[z][y], v[z+1][y], v[z][y+1],
// v[z+1][y+1]), so a cell is two x planes, one 16-byte load each, and a
// step to the next cell along x reloads one plane.  Corners, weights and
// the lerp order are those of lerp_cell, so the result has the same bits as
// every other layout.  (A second quad copy of the x/y transpose, letting
// each wave or view slide along y when its rays cross y faster, measured
// 17 % slower on the 256^3 workload: the two 268 MB copies no longer share
// the 256 MB last-level cache the single copy is re-read from across views.)
__device__ __forceinline__ double slide_march(const MarchParams& p, __amdgpu_buffer_rsrc_t rs,
                                              const double (&A)[3], const double (&B)[3],
                                              int ka, int kb, int wk0, int wk1) {
  const int nx = p.vn[0], ny = p.vn[1], nz = p.vn[2];
  const unsigned sbx = (unsigned)p.vbricks[1] * 64u,
                 sbz = (unsigned)p.vbricks[0] * (unsigned)p.vbricks[1] * 64u;
  const float b0 = (float)B[0], b1 = (float)B[1], b2 = (float)B[2];
  const float ds32 = (float)p.ds;
  const double ds = p.ds, dm = p.dm;
  double T = p.flat_field;
  int cx = -1, cy = -1, cz = -1;
  float cell[8];  // cell[4*dx + 2*dy + dz]
#pragma unroll
  for (int q = 0; q < 8; ++q) cell[q] = 0.0f;
  for (int k0 = wk0; k0 <= wk1; k0 += kChunk) {
    double sa = p.s_table[k0];
    float a0 = (float)(A[0] + B[0] * sa), a1 = (float)(A[1] + B[1] * sa),
          a2 = (float)(A[2] + B[2] * sa);
    const int n = min(kChunk, wk1 - k0 + 1);
    const int jlo = ka - k0, jhi = kb - k0;
    float acc = 0.0f, comp = 0.0f;  // Kahan-compensated chunk sum
    for (int jj = 0; jj < n; ++jj) {
      if (jj >= jlo && jj <= jhi) {
        float sj = (float)jj * ds32;
        float gx = fmaxf(fmaf(b0, sj, a0), 0.0f), gy = fmaxf(fmaf(b1, sj, a1), 0.0f),
              gz = fmaxf(fmaf(b2, sj, a2), 0.0f);
        const int x0 = min((int)gx, nx - 1), y0 = min((int)gy, ny - 1),
                  z0 = min((int)gz, nz - 1);
        if (x0 != cx || y0 != cy || z0 != cz) {
          const bool same_yz = y0 == cy && z0 == cz;
          const bool up = same_yz && x0 == cx + 1, down = same_yz && x0 == cx - 1;
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const float lo = cell[q], hi = cell[4 + q];
            cell[q] = up ? hi : lo;
            cell[4 + q] = down ? lo : hi;
          }
          const int x1 = min(x0 + 1, nx - 1);
          const unsigned ezy = (unsigned)(z0 >> 2) * sbz + (unsigned)(z0 & 3) * 16u +
                               (unsigned)(y0 >> 2) * 64u + (unsigned)(y0 & 3);
          if (!up) {
            const unsigned e = ezy + (unsigned)(x0 >> 2) * sbx + (unsigned)(x0 & 3) * 4u;
            const auto v = __builtin_amdgcn_raw_buffer_load_b128(rs, (int)(e * 16u), 0, 0);
#pragma unroll
            for (int q = 0; q < 4; ++q) cell[q] = __uint_as_float(v[q]);
          }
          if (!down) {
            const unsigned e = ezy + (unsigned)(x1 >> 2) * sbx + (unsigned)(x1 & 3) * 4u;
            const auto v = __builtin_amdgcn_raw_buffer_load_b128(rs, (int)(e * 16u), 0, 0);
#pragma unroll
            for (int q = 0; q < 4; ++q) cell[4 + q] = __uint_as_float(v[q]);
          }
          cx = x0;
          cy = y0;
          cz = z0;
        }
        const float rho = lerp_cell(cell, gx - (float)x0, gy - (float)y0, gz - (float)z0);
        float y = rho - comp;
        float tt = acc + y;
        comp = (tt - acc) - y;
        acc = tt;
      }
    }
    T += ((double)acc * dm) * ds;
  }
  return T;
}

// Stepping march of one ray over its own sample range [ka, kb] (not the
// wave's union of ranges, so no lane idles past its own ends), on the quad
// copy.  Per chunk of kChunk samples the grid position is anchored in FP64
// and cell and weights are taken from it as the other layouts take them;
// after that the position advances by the constant FP32 increment of one
// step, kept as cell indices plus weights: a weight that leaves [0, 1)
// moves the cell by its floor.  A cell's four z-differences are formed when
// it is loaded, so a sample costs the seven lerps' FMAs and three
// differences, the increments and the Kahan sum.  The weights differ from
// the other layouts' by the increments' rounding (a few FP32 ulps of the
// grid coordinate per chunk), so the bits differ from theirs while the
// error against the FP64 oracle stays of the same size.  Indices are held
// in [0, N-1] as the reference clamps them (objects.go:811-830): a
// coordinate that rounds below 0 takes cell 0 with weight 0, one past N-1
// the last cell, whose two planes are then the same.
__device__ __forceinline__ double step_march(const MarchParams& p, __amdgpu_buffer_rsrc_t rs,
                                             const double (&A)[3], const double (&B)[3],
                                             int ka, int kb) {
  const int nx = p.vn[0], ny = p.vn[1], nz = p.vn[2];
  const unsigned sbx = (unsigned)p.vbricks[1] * 64u,
                 sbz = (unsigned)p.vbricks[0] * (unsigned)p.vbricks[1] * 64u;
  const double ds = p.ds, dm = p.dm;
  const float ix = (float)(B[0] * ds), iy = (float)(B[1] * ds), iz = (float)(B[2] * ds);
  double T = p.flat_field;
  int cx = -1, cy = -1, cz = -1;
  // corners (lerp_cell order) of the loaded cell: lo = x plane x0, hi =
  // x plane x0+1, each (v[z][y], v[z+1][y], v[z][y+1], v[z+1][y+1])
  float lo[4], hi[4], dlo[2], dhi[2];
#pragma unroll
  for (int q = 0; q < 4; ++q) lo[q] = hi[q] = 0.0f;
  dlo[0] = dlo[1] = dhi[0] = dhi[1] = 0.0f;
  for (int k0 = ka; k0 <= kb; k0 += kChunk) {
    const double sa = p.s_table[k0];
    const float gx = fmaxf((float)(A[0] + B[0] * sa), 0.0f),
                gy = fmaxf((float)(A[1] + B[1] * sa), 0.0f),
                gz = fmaxf((float)(A[2] + B[2] * sa), 0.0f);
    int x0 = min((int)gx, nx - 1), y0 = min((int)gy, ny - 1), z0 = min((int)gz, nz - 1);
    float wx = gx - (float)x0, wy = gy - (float)y0, wz = gz - (float)z0;
    const int n = min(kChunk, kb - k0 + 1);
    float acc = 0.0f, comp = 0.0f;  // Kahan-compensated chunk sum
    for (int j = 0; j < n; ++j) {
      if (x0 != cx || y0 != cy || z0 != cz) {
        const bool same_yz = y0 == cy && z0 == cz;
        const bool up = same_yz && x0 == cx + 1, down = same_yz && x0 == cx - 1;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const float a = lo[q], b = hi[q];
          lo[q] = up ? b : a;
          hi[q] = down ? a : b;
        }
        const int x1 = min(x0 + 1, nx - 1);
        const unsigned ezy = (unsigned)(z0 >> 2) * sbz + (unsigned)(z0 & 3) * 16u +
                             (unsigned)(y0 >> 2) * 64u + (unsigned)(y0 & 3);
        if (!up) {
          const unsigned e = ezy + (unsigned)(x0 >> 2) * sbx + (unsigned)(x0 & 3) * 4u;
          const auto v = __builtin_amdgcn_raw_buffer_load_b128(rs, (int)(e * 16u), 0, 0);
#pragma unroll
          for (int q = 0; q < 4; ++q) lo[q] = __uint_as_float(v[q]);
        }
        if (!down) {
          const unsigned e = ezy + (unsigned)(x1 >> 2) * sbx + (unsigned)(x1 & 3) * 4u;
          const auto v = __builtin_amdgcn_raw_buffer_load_b128(rs, (int)(e * 16u), 0, 0);
#pragma unroll
          for (int q = 0; q < 4; ++q) hi[q] = __uint_as_float(v[q]);
        }
        dlo[0] = lo[1] - lo[0];
        dlo[1] = lo[3] - lo[2];
        dhi[0] = hi[1] - hi[0];
        dhi[1] = hi[3] - hi[2];
        cx = x0;
        cy = y0;
        cz = z0;
      }
      // lerp_cell's order: z, then y, then x
      const float v00 = fmaf(wz, dlo[0], lo[0]), v01 = fmaf(wz, dlo[1], lo[2]);
      const float v10 = fmaf(wz, dhi[0], hi[0]), v11 = fmaf(wz, dhi[1], hi[2]);
      const float v0 = fmaf(wy, v01 - v00, v00), v1 = fmaf(wy, v11 - v10, v10);
      const float rho = fmaf(wx, v1 - v0, v0);
      const float yk = rho - comp;
      const float tt = acc + yk;
      comp = (tt - acc) - yk;
      acc = tt;
      // next sample: advance, move the cell by the weights' floors, clamp
      wx += ix;
      wy += iy;
      wz += iz;
      const float fx = floorf(wx), fy = floorf(wy), fz = floorf(wz);
      x0 += (int)fx;
      y0 += (int)fy;
      z0 += (int)fz;
      wx -= fx;
      wy -= fy;
      wz -= fz;
      if (x0 < 0) { x0 = 0; wx = 0.0f; }
      if (y0 < 0) { y0 = 0; wy = 0.0f; }
      if (z0 < 0) { z0 = 0; wz = 0.0f; }
      x0 = min(x0, nx - 1);
      y0 = min(y0, ny - 1);
      z0 = min(z0, nz - 1);
    }
    T += ((double)acc * dm) * ds;
  }
  return T;
}

// Ray tile of one wave: 32 pixels along i (the image's horizontal axis) by
// 2 along j.  At a given sample index the rays of a wave then lie along one
// horizontal line through the volume; on the 256^3 workload (views around
// the z axis) 32x2 and 64x1 tiles ran in 4.11 ms, 16x4 in 5.19 ms, 8x8 in
// 5.42 ms and 4x16 in 10.4 ms.
constexpr int kVoxTileI = 32, kVoxTileJ = 2;

template <int LAYOUT, int CACHE>
__device__ __forceinline__ void voxel_body(const MarchParams& p) {
  const int lane = threadIdx.x & 63;
  // a wave's 64 rays: kVoxTileI pixels along i by kVoxTileJ along j
  const int tiles_j = (p.res + kVoxTileJ - 1) / kVoxTileJ;
  const int ti = blockIdx.x / tiles_j, tj = blockIdx.x - ti * tiles_j;
  const int i = ti * kVoxTileI + lane / kVoxTileJ;
  const int j = tj * kVoxTileJ + lane % kVoxTileJ;
  const int view = p.view_offset + blockIdx.y;
  const bool valid = i < p.res && j < p.res;

  // Per lane: exact inside-sample range [ka, kb] and the grid-space line.
  int ka = 0x7fffffff, kb = -1;
  double A[3] = {0, 0, 0}, B[3] = {0, 0, 0};
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
    long la, lb;
    // pad the analytic interval by one step on each side; exact_range
    // then trims it to the samples the reference counts as inside.
    if (hit && ta <= tb && exact_range(p, r, ta - p.ds, tb + p.ds, la, lb)) {
      ka = (int)la;
      kb = (int)lb;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        double h = 0.5 * (double)(p.vn[c] - 1);
        A[c] = (Ao[c] + 1.0) * h;
        B[c] = Bo[c] * h;
      }
    }
  }
  // The wave walks one sample index at a time, so its 64 lanes gather from
  // one thin slab of the grid (neighbouring rays at the same depth).
  int wk0 = ka, wk1 = kb;
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    wk0 = min(wk0, __shfl_xor(wk0, off, 64));
    wk1 = max(wk1, __shfl_xor(wk1, off, 64));
  }
  double T = p.flat_field;
  const float b0 = (float)B[0], b1 = (float)B[1], b2 = (float)B[2];
  if constexpr (CACHE == CACHE_SLIDE)
    T = slide_march(p, grid_rsrc(p.vgrid, p.vbytes), A, B, ka, kb, wk0, wk1);
  if constexpr (CACHE == CACHE_STEP)
    T = step_march(p, grid_rsrc(p.vgrid, p.vbytes), A, B, ka, kb);
  const float ds32 = (float)p.ds;
  const double ds = p.ds, dm = p.dm;
  const __amdgpu_buffer_rsrc_t rs = grid_rsrc(p.vgrid, p.vbytes);
  int cx = -1, cy = -1, cz = -1;  // cached cell (CACHED only)
  float cell[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) cell[q] = 0.0f;
  for (int k0 = wk0; CACHE != CACHE_SLIDE && CACHE != CACHE_STEP && k0 <= wk1; k0 += kChunk) {
    double sa = p.s_table[k0];
    float a0 = (float)(A[0] + B[0] * sa), a1 = (float)(A[1] + B[1] * sa),
          a2 = (float)(A[2] + B[2] * sa);
    const int n = min(kChunk, wk1 - k0 + 1);
    const int jlo = ka - k0, jhi = kb - k0;
    float acc = 0.0f, comp = 0.0f;  // Kahan-compensated chunk sum
    for (int jj = 0; jj < n; ++jj) {
      if (jj >= jlo && jj <= jhi) {
        float sj = (float)jj * ds32;
        float gx = fmaxf(fmaf(b0, sj, a0), 0.0f), gy = fmaxf(fmaf(b1, sj, a1), 0.0f),
              gz = fmaxf(fmaf(b2, sj, a2), 0.0f);
        float rho;
        if constexpr (CACHE != CACHE_NONE) {
          int x0 = min((int)gx, p.vn[0] - 1), y0 = min((int)gy, p.vn[1] - 1),
              z0 = min((int)gz, p.vn[2] - 1);
          if (x0 != cx || y0 != cy || z0 != cz) {
            load_cell<LAYOUT>(p, rs, x0, y0, z0, cell);
            cx = x0;
            cy = y0;
            cz = z0;
          }
          rho = lerp_cell(cell, gx - (float)x0, gy - (float)y0, gz - (float)z0);
        } else {
          rho = trilinear<LAYOUT>(p, rs, gx, gy, gz);
        }
        float y = rho - comp;
        float tt = acc + y;
        comp = (tt - acc) - y;
        acc = tt;
      }
    }
    T += ((double)acc * dm) * ds;
  }
  if (valid)
    store_pixel(p, blockIdx.y, i, j, T);
  if (p.stats) {
    unsigned long long a = wave_sum(kb >= ka ? (unsigned long long)(kb - ka + 1) : 0ull);
    if (lane == 0 && a) atomicAdd(p.stats + ST_SAMPLES, a);
  }
}

// One wave per workgroup, one ray tile (kVoxTileI x kVoxTileJ): a wave
// whose rays are done frees its slot without waiting for the other waves of
// a larger workgroup (1.93x faster than 16x16 tiles of 256 lanes on the
// 256^3 workload).
// 8 waves per SIMD (<= 64 VGPRs) except for the sliding cache, which
// spills at 64 and gets 7.
template <int LAYOUT, int CACHE>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(CACHE == CACHE_SLIDE ? 7 : 8))) void voxel_kernel(
    MarchParams p) {
  voxel_body<LAYOUT, CACHE>(p);
}

// ---------------------------------------------------------------------------
// LDS-staged variant.  One 64-lane workgroup = one 8x8 ray tile.  For each
// chunk of kLdsChunk sample indices the wave takes the box of voxels its
// rays can touch (union of every active lane's segment, +1 for the upper
// trilinear corner), copies it from the linear [z][x][y] grid into LDS with
// row-contiguous loads, and interpolates from LDS.  A box larger than the
// LDS slot falls back to brick-layout gathers for that chunk.
// ---------------------------------------------------------------------------
constexpr int kLdsFloats = 4096;  // 16 KB per wave
constexpr int kLdsChunk = 16;

__device__ __forceinline__ float wave_minf(float v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = fminf(v, __shfl_xor(v, off, 64));
  return v;
}
__device__ __forceinline__ float wave_maxf(float v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
  return v;
}

__global__ __launch_bounds__(64) void voxel_lds_kernel(MarchParams p) {
  __shared__ float tile[kLdsFloats];
  const int lane = threadIdx.x;
  // Workgroup b runs on XCD b % 8.  Give each XCD a contiguous run of
  // tiles so its L2 holds one band of the image's footprint.
  const int tiles_j = (p.res + 7) >> 3;
  const int ntiles = tiles_j * tiles_j;
  int logical = blockIdx.x;
  if ((ntiles & 7) == 0) logical = (blockIdx.x & 7) * (ntiles >> 3) + (blockIdx.x >> 3);
  const int ti = logical / tiles_j, tj = logical - ti * tiles_j;
  const int i = ti * 8 + (lane >> 3);
  const int j = tj * 8 + (lane & 7);
  const int view = p.view_offset + blockIdx.y;
  const bool valid = i < p.res && j < p.res;

  int ka = 0x7fffffff, kb = -1;
  double A[3] = {0, 0, 0}, B[3] = {0, 0, 0};
  if (valid) {
    const DevView& V = p.views[view];
    Ray r = make_ray(p, V, i, j);
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
    long la, lb;
    if (hit && ta <= tb && exact_range(p, r, ta - p.ds, tb + p.ds, la, lb)) {
      ka = (int)la;
      kb = (int)lb;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        double h = 0.5 * (double)(p.vn[c] - 1);
        A[c] = (Ao[c] + 1.0) * h;
        B[c] = Bo[c] * h;
      }
    }
  }
  int wk0 = ka, wk1 = kb;
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    wk0 = min(wk0, __shfl_xor(wk0, off, 64));
    wk1 = max(wk1, __shfl_xor(wk1, off, 64));
  }
  const int nx = p.vn[0], ny = p.vn[1], nz = p.vn[2];
  double T = p.flat_field;
  const float b0 = (float)B[0], b1 = (float)B[1], b2 = (float)B[2];
  const float ds32 = (float)p.ds;
  const double ds = p.ds, dm = p.dm;
  const __amdgpu_buffer_rsrc_t rs_brick = grid_rsrc(p.vgrid, p.vbytes);
  const __amdgpu_buffer_rsrc_t rs_lin = grid_rsrc(p.vlin, p.vlin_bytes);
  unsigned long long n_staged = 0, n_chunks = 0, n_fallback = 0;
  for (int k0 = wk0; k0 <= wk1; k0 += kLdsChunk) {
    const double sa = p.s_table[k0];
    const float a0 = (float)(A[0] + B[0] * sa), a1 = (float)(A[1] + B[1] * sa),
                a2 = (float)(A[2] + B[2] * sa);
    const int n = min(kLdsChunk, wk1 - k0 + 1);
    const int jlo = ka - k0, jhi = kb - k0;
    const int j0 = max(jlo, 0), j1 = min(jhi, n - 1);
    const bool act = j0 <= j1;
    // This lane's footprint: the segment between its first and last
    // sample of the chunk (positions are monotone in the sample index).
    float lo0 = 3e38f, lo1 = 3e38f, lo2 = 3e38f, hi0 = -3e38f, hi1 = -3e38f,
          hi2 = -3e38f;
    if (act) {
      const float s0 = (float)j0 * ds32, s1 = (float)j1 * ds32;
      float x0 = fmaf(b0, s0, a0), x1 = fmaf(b0, s1, a0);
      float y0 = fmaf(b1, s0, a1), y1 = fmaf(b1, s1, a1);
      float z0 = fmaf(b2, s0, a2), z1 = fmaf(b2, s1, a2);
      lo0 = fminf(x0, x1);
      hi0 = fmaxf(x0, x1);
      lo1 = fminf(y0, y1);
      hi1 = fmaxf(y0, y1);
      lo2 = fminf(z0, z1);
      hi2 = fmaxf(z0, z1);
    }
    lo0 = wave_minf(lo0);
    lo1 = wave_minf(lo1);
    lo2 = wave_minf(lo2);
    hi0 = wave_maxf(hi0);
    hi1 = wave_maxf(hi1);
    hi2 = wave_maxf(hi2);
    if (!(lo0 <= hi0)) continue;  // no lane samples in this chunk
    const int X0 = max(0, (int)floorf(fmaxf(lo0, 0.0f))),
              X1 = min(nx - 1, (int)floorf(fminf(hi0, (float)nx)) + 1);
    const int Y0 = max(0, (int)floorf(fmaxf(lo1, 0.0f))),
              Y1 = min(ny - 1, (int)floorf(fminf(hi1, (float)ny)) + 1);
    const int Z0 = max(0, (int)floorf(fmaxf(lo2, 0.0f))),
              Z1 = min(nz - 1, (int)floorf(fminf(hi2, (float)nz)) + 1);
    // Samples lie inside the grid (exact_range), so X0 <= X1; the min()
    // only keeps every loop below finite whatever the inputs.
    const int bx = X1 - min(X0, X1) + 1, by = Y1 - min(Y0, Y1) + 1,
              bz = Z1 - min(Z0, Z1) + 1;
    float acc = 0.0f, comp = 0.0f;  // Kahan-compensated chunk sum
    if (by <= 64 && bx * by * bz <= kLdsFloats) {
      n_staged += (unsigned long long)(bx * by * bz);
      ++n_chunks;
      // Stage: lane = (row offset, y); rpi rows per pass.
      const int rpi = 64 / by;
      const int ro = lane / by, ly = lane - ro * by;
      const int R = bz * bx;
      if (ro < rpi) {
        int z = 0, x = ro;
        while (x >= bx) {
          x -= bx;
          ++z;
        }
        for (int rr = ro; rr < R; rr += rpi) {
          unsigned g = ((unsigned)(Z0 + z) * (unsigned)nx + (unsigned)(X0 + x)) *
                           (unsigned)ny + (unsigned)(Y0 + ly);
          tile[rr * by + ly] = __uint_as_float(
              __builtin_amdgcn_raw_buffer_load_b32(rs_lin, (int)(g * 4u), 0, 0));
          x += rpi;
          while (x >= bx) {
            x -= bx;
            ++z;
          }
        }
      }
      __syncthreads();
      const int sxy = bx * by;
      for (int jj = 0; jj < n; ++jj) {
        if (jj >= jlo && jj <= jhi) {
          float sj = (float)jj * ds32;
          float gx = fmaxf(fmaf(b0, sj, a0), 0.0f), gy = fmaxf(fmaf(b1, sj, a1), 0.0f),
                gz = fmaxf(fmaf(b2, sj, a2), 0.0f);
          int x0 = min((int)gx, nx - 1), y0 = min((int)gy, ny - 1),
              z0 = min((int)gz, nz - 1);
          float wx = gx - (float)x0, wy = gy - (float)y0, wz = gz - (float)z0;
          int dx = min(x0 + 1, nx - 1) - x0, dy = min(y0 + 1, ny - 1) - y0,
              dz = min(z0 + 1, nz - 1) - z0;
          int base = ((z0 - Z0) * bx + (x0 - X0)) * by + (y0 - Y0);
          int ox = dx * by, oz = dz * sxy;
          float v000 = tile[base], v010 = tile[base + dy];
          float v100 = tile[base + ox], v110 = tile[base + ox + dy];
          float v001 = tile[base + oz], v011 = tile[base + oz + dy];
          float v101 = tile[base + oz + ox], v111 = tile[base + oz + ox + dy];
          float v00 = fmaf(wz, v001 - v000, v000);
          float v01 = fmaf(wz, v011 - v010, v010);
          float v10 = fmaf(wz, v101 - v100, v100);
          float v11 = fmaf(wz, v111 - v110, v110);
          float v0 = fmaf(wy, v01 - v00, v00);
          float v1 = fmaf(wy, v11 - v10, v10);
          float rho = fmaf(wx, v1 - v0, v0);
          float y = rho - comp;
          float tt = acc + y;
          comp = (tt - acc) - y;
          acc = tt;
        }
      }
      __syncthreads();
    } else {
      ++n_fallback;
      for (int jj = 0; jj < n; ++jj) {
        if (jj >= jlo && jj <= jhi) {
          float sj = (float)jj * ds32;
          float gx = fmaf(b0, sj, a0), gy = fmaf(b1, sj, a1), gz = fmaf(b2, sj, a2);
          float rho = trilinear<VOX_BRICK4>(p, rs_brick, fmaxf(gx, 0.0f),
                                            fmaxf(gy, 0.0f), fmaxf(gz, 0.0f));
          float y = rho - comp;
          float tt = acc + y;
          comp = (tt - acc) - y;
          acc = tt;
        }
      }
    }
    T += ((double)acc * dm) * ds;
  }
  if (valid)
    store_pixel(p, blockIdx.y, i, j, T);
  if (p.stats) {
    unsigned long long a = wave_sum(kb >= ka ? (unsigned long long)(kb - ka + 1) : 0ull);
    if (lane == 0) {
      if (a) atomicAdd(p.stats + ST_SAMPLES, a);
      // wave-uniform staging counters
      if (n_chunks) atomicAdd(p.stats + ST_STAGED_CHUNKS, n_chunks);
      if (n_staged) atomicAdd(p.stats + ST_STAGED_VOXELS, n_staged);
      if (n_fallback) atomicAdd(p.stats + ST_FALLBACK_CHUNKS, n_fallback);
    }
  }
}

// ---------------------------------------------------------------------------
// L2 read-bandwidth probe, the second ceiling the voxel kernel is compared
// with.  Workgroup b reads 4 KiB chunks of region b % 8 (one 2 MiB region per
// XCD, workgroups being dealt to the 8 XCDs in turn), moving to another chunk
// of the region every iteration: 512 chunks per region defeat the 32 KiB L1,
// and the region fits the XCD's 4 MiB L2.  Each lane folds what it reads into
// one word so the loads stay live.
// ---------------------------------------------------------------------------
constexpr int kL2Regions = 8;
constexpr int kL2ChunkVec = 256;                         // uint4 per chunk (4 KiB)
constexpr int kL2ChunksPerRegion = 512;                  // 2 MiB per region
__global__ __launch_bounds__(256) void l2_probe_kernel(const uint4* __restrict__ buf,
                                                       int iters, unsigned* __restrict__ out) {
  const int region = blockIdx.x % kL2Regions;
  const int slot = blockIdx.x / kL2Regions;
  const uint4* base = buf + (size_t)region * kL2ChunksPerRegion * kL2ChunkVec + threadIdx.x;
  unsigned acc = 0;
  for (int it = 0; it < iters; ++it) {
    const int chunk = (slot + it * 7) & (kL2ChunksPerRegion - 1);
    const uint4 v = base[(size_t)chunk * kL2ChunkVec];
    acc ^= v.x + v.y + v.z + v.w;
  }
  out[(size_t)blockIdx.x * 256 + threadIdx.x] = acc;
}

size_t l2_probe_buffer_bytes() {
  return (size_t)kL2Regions * kL2ChunksPerRegion * kL2ChunkVec * sizeof(uint4);
}

int launch_l2_probe(const void* buf, int workgroups, int iters, unsigned* out, void* stream) {
  if (workgroups <= 0 || iters <= 0) return (int)hipErrorInvalidValue;
  hipLaunchKernelGGL(l2_probe_kernel, dim3(workgroups), dim3(256), 0, (hipStream_t)stream,
                     (const uint4*)buf, iters, out);
  return (int)hipGetLastError();
}

int launch_voxel(const MarchParams& p, int layout, void* stream) {
  if (layout == VOX_LDS) {
    int tiles = (p.res + 7) / 8;
    hipLaunchKernelGGL(voxel_lds_kernel, dim3(tiles * tiles, p.n_views), dim3(64), 0,
                       (hipStream_t)stream, p);
    return (int)hipGetLastError();
  }
  dim3 block(64);
  const int tiles = ((p.res + kVoxTileI - 1) / kVoxTileI) * ((p.res + kVoxTileJ - 1) / kVoxTileJ);
  dim3 grid(tiles, p.n_views);
  hipStream_t s = (hipStream_t)stream;
  switch (layout) {
    case VOX_BRICK4:
      hipLaunchKernelGGL((voxel_kernel<VOX_BRICK4, CACHE_NONE>), grid, block, 0, s, p);
      break;
    case VOX_BRICK4_CACHED:
      hipLaunchKernelGGL((voxel_kernel<VOX_BRICK4, CACHE_CELL>), grid, block, 0, s, p);
      break;
    case VOX_BRICK4_PAIR:
      hipLaunchKernelGGL((voxel_kernel<VOX_BRICK4_PAIR, CACHE_CELL>), grid, block, 0, s, p);
      break;
    case VOX_BRICK4_QUAD:
      hipLaunchKernelGGL((voxel_kernel<VOX_BRICK4_QUAD, CACHE_CELL>), grid, block, 0, s, p);
      break;
    case VOX_BRICK4_QUAD_SLIDE:
      hipLaunchKernelGGL((voxel_kernel<VOX_BRICK4_QUAD, CACHE_SLIDE>), grid, block, 0, s, p);
      break;
    case VOX_BRICK4_QUAD_STEP:
      hipLaunchKernelGGL((voxel_kernel<VOX_BRICK4_QUAD, CACHE_STEP>), grid, block, 0, s, p);
      break;
    case VOX_LINEAR_CACHED:
      hipLaunchKernelGGL((voxel_kernel<VOX_LINEAR, CACHE_CELL>), grid, block, 0, s, p);
      break;
    default:
      hipLaunchKernelGGL((voxel_kernel<VOX_LINEAR, CACHE_NONE>), grid, block, 0, s, p);
  }
  return (int)hipGetLastError();
}

}  // namespace xrp
