// eval_device.hpp - device-side density evaluators, ray set-up and
// clipping shared by the projection kernels.
//
// FP64 evaluators follow the reference's operation order (the including
// translation unit is built with -ffp-contract=off); the guarded FP32
// evaluators return the exact FP64 density when every membership decision
// is outside its guard band and flag the sample otherwise.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

#include "march.hpp"

namespace xrp {

constexpr int kChunk = 32;  // FP32 positions re-anchored from FP64 every kChunk samples

// ---------------------------------------------------------------------------
// FP64 evaluator: the reference's operation order.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double sphere64(const DevNode& n, double x, double y,
                                           double z) {  // objects.go:63-72
  x = x - n.d[0];
  y = y - n.d[1];
  z = z - n.d[2];
  double r_2 = x * x + y * y + z * z;
  return r_2 < n.d[3] * n.d[3] ? n.d[4] : 0.0;
}

__device__ __forceinline__ double box64(const DevNode& n, double x, double y,
                                        double z) {  // objects.go:171-179
  x = fabs(x - n.d[0]);
  y = fabs(y - n.d[1]);
  z = fabs(z - n.d[2]);
  return (x < n.d[3] && y < n.d[4] && z < n.d[5]) ? n.d[6] : 0.0;
}

__device__ __forceinline__ double pped64(const DevNode& n, double x, double y,
                                         double z) {  // objects.go:247-255
  double w0 = x - n.d[0], w1 = y - n.d[1], w2 = z - n.d[2];
  const double* m = n.d + 3;
  double u = m[0] * w0 + m[3] * w1 + m[6] * w2;
  double v = m[1] * w0 + m[4] * w1 + m[7] * w2;
  double t = m[2] * w0 + m[5] * w1 + m[8] * w2;
  return (u > 0.0 && u < 1.0 && v > 0.0 && v < 1.0 && t > 0.0 && t < 1.0)
             ? n.d[12]
             : 0.0;
}

__device__ __forceinline__ double cyl64(const DevNode& n, double x, double y,
                                        double z) {  // objects.go:334-350
  double v0 = n.d[3], v1 = n.d[4], v2 = n.d[5];
  double w0 = x - n.d[0], w1 = y - n.d[1], w2 = z - n.d[2];
  double c = (w0 * v0 + w1 * v1 + w2 * v2) / n.d[6];
  if (c < 0.0 || c > 1.0) return 0.0;
  double d0 = w0 - v0 * c, d1 = w1 - v1 * c, d2 = w2 - v2 * c;
  double d = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
  return d < n.d[7] ? n.d[8] : 0.0;
}

// out of line: FP64 sin/cos would otherwise set the register budget of
// every kernel that can re-evaluate a sample
__device__ __noinline__ double gyroid64(const DevNode& n, double x, double y,
                                           double z) {  // objects.go:1014-1032
  x = (x - n.d[0]) / n.d[3];
  y = (y - n.d[1]) / n.d[3];
  z = (z - n.d[2]) / n.d[3];
  double g = sin(x) * cos(y) + sin(y) * cos(z) + sin(z) * cos(x);
  return fabs(g) < n.d[4] ? n.d[5] : 0.0;
}

__device__ __noinline__ double voxel64(const DevNode& n, double x, double y,
                                          double z) {  // objects.go:789-855
  // Same test as the reference for finite points; NaN (which the
  // reference would turn into an out-of-range index) reads nothing.
  if (!(x >= -1 && x <= 1 && y >= -1 && y <= 1 && z >= -1 && z <= 1))
    return 0.0;
  x = (x + 1) / 2;
  y = (y + 1) / 2;
  z = (z + 1) / 2;
  x = x * (double)(n.nx - 1);
  y = y * (double)(n.ny - 1);
  z = z * (double)(n.nz - 1);
  int x0 = (int)floor(x), y0 = (int)floor(y), z0 = (int)floor(z);
  int x1 = x0 + 1, y1 = y0 + 1, z1 = z0 + 1;
  x0 = x0 < 0 ? 0 : x0;
  y0 = y0 < 0 ? 0 : y0;
  z0 = z0 < 0 ? 0 : z0;
  x1 = x1 >= n.nx ? n.nx - 1 : x1;
  y1 = y1 >= n.ny ? n.ny - 1 : y1;
  z1 = z1 >= n.nz ? n.nz - 1 : z1;
  double wx = x - (double)x0, wy = y - (double)y0, wz = z - (double)z0;
  const float* r = n.vox;
  size_t NX = (size_t)n.nx, NY = (size_t)n.ny;
  size_t zs0 = (size_t)z0 * NX * NY, zs1 = (size_t)z1 * NX * NY;
  size_t xs0 = (size_t)x0 * NY, xs1 = (size_t)x1 * NY;
  double v000 = r[zs0 + xs0 + y0], v001 = r[zs1 + xs0 + y0];
  double v010 = r[zs0 + xs0 + y1], v011 = r[zs1 + xs0 + y1];
  double v100 = r[zs0 + xs1 + y0], v101 = r[zs1 + xs1 + y0];
  double v110 = r[zs0 + xs1 + y1], v111 = r[zs1 + xs1 + y1];
  double v00 = v000 * (1 - wz) + v001 * wz;
  double v01 = v010 * (1 - wz) + v011 * wz;
  double v10 = v100 * (1 - wz) + v101 * wz;
  double v11 = v110 * (1 - wz) + v111 * wz;
  double v0 = v00 * (1 - wy) + v01 * wy;
  double v1 = v10 * (1 - wy) + v11 * wy;
  return v0 * (1 - wx) + v1 * wx;
}

template <int D>
__device__ __forceinline__ double coll64(const DevNode* __restrict__ N, int i, double x,
                         double y, double z);

template <int D>
__device__ __forceinline__ double tess64(const DevNode* __restrict__ N, int i,
                                         double x, double y, double z) {
  // objects.go:568-582 then UnitCell.Density 458-464
  const DevNode& n = N[i];
  const double* b = n.d;
  if (x < b[0] || x > b[1] || y < b[2] || y > b[3] || z < b[4] || z > b[5])
    return 0.0;
  const double* lo = n.d + 6;
  const double* hi = n.d + 9;
  double dx = hi[0] - lo[0];
  x = x - dx * floor((x - lo[0]) / dx);
  double dy = hi[1] - lo[1];
  y = y - dy * floor((y - lo[1]) / dy);
  double dz = hi[2] - lo[2];
  z = z - dz * floor((z - lo[2]) / dz);
  if (x < lo[0] || x > hi[0] || y < lo[1] || y > hi[1] || z < lo[2] ||
      z > hi[2])
    return 0.0;
  if constexpr (D + 1 <= kMaxDepth) {
    return coll64<D + 1>(N, n.first, x, y, z);
  } else {
    return 0.0;  // rejected by the loader
  }
}

template <int D>
__device__ __forceinline__ double node64(const DevNode* __restrict__ N, int i,
                                         double x, double y, double z) {
  const DevNode& n = N[i];
  switch (n.kind) {
    case K_SPHERE: return sphere64(n, x, y, z);
    case K_BOX: return box64(n, x, y, z);
    case K_PPED: return pped64(n, x, y, z);
    case K_CYL: return cyl64(n, x, y, z);
    case K_GYROID: return gyroid64(n, x, y, z);
    case K_VOXEL: return voxel64(n, x, y, z);
    case K_TESS: return tess64<D>(N, i, x, y, z);
    case K_COLL: return coll64<D>(N, i, x, y, z);
  }
  return 0.0;
}

template <int D>
__device__ __forceinline__ double coll64(const DevNode* __restrict__ N, int i, double x,
                         double y, double z) {  // objects.go:422-438
  const DevNode& n = N[i];
  double density = 0.0;
  for (int c = n.first; c < n.next; c = N[c].next) {
    double rho;
    const DevNode& k = N[c];
    switch (k.kind) {
      case K_SPHERE: rho = sphere64(k, x, y, z); break;
      case K_BOX: rho = box64(k, x, y, z); break;
      case K_PPED: rho = pped64(k, x, y, z); break;
      case K_CYL: rho = cyl64(k, x, y, z); break;
      case K_GYROID: rho = gyroid64(k, x, y, z); break;
      case K_VOXEL: rho = voxel64(k, x, y, z); break;
      case K_TESS: rho = tess64<D>(N, c, x, y, z); break;
      default: rho = 0.0;
    }
    if (n.greedy && rho > 0.0) return rho;
    density += rho;
  }
  if (density < 0.0)
    density = 0.0;
  else if (density > 1.0)
    density = 1.0;
  return density;
}

__device__ __forceinline__ void deform64(const DevDeform* __restrict__ D,
                                         int nd, double& x, double& y,
                                         double& z) {
  for (int k = 0; k < nd; ++k) {
    const DevDeform& f = D[k];
    switch (f.kind) {
      case D_GAUSS: {  // deformations.go:29-38
        double x0 = x - f.d[6], y0 = y - f.d[7], z0 = z - f.d[8];
        double r2 = x0 * x0 + y0 * y0 + z0 * z0;
        double ddx = f.d[0] * exp(-r2 / (2 * f.d[3] * f.d[3]));
        double ddy = f.d[1] * exp(-r2 / (2 * f.d[4] * f.d[4]));
        double ddz = f.d[2] * exp(-r2 / (2 * f.d[5] * f.d[5]));
        x = x + ddx;
        y = y + ddy;
        z = z + ddz;
        break;
      }
      case D_AFFINE: {  // deformations.go:87-92
        const double* m = f.d;
        double nx = m[0] * x + m[1] * y + m[2] * z;
        double ny = m[3] * x + m[4] * y + m[5] * z;
        double nz = m[6] * x + m[7] * y + m[8] * z;
        x = nx;
        y = ny;
        z = nz;
        break;
      }
      case D_LINEAR: {  // deformations.go:136-141
        const double* s = f.d;
        double nx = x + s[0] * x + s[5] * y + s[4] * z;
        double ny = y + s[5] * x + s[1] * y + s[3] * z;
        double nz = z + s[4] * x + s[3] * y + s[2] * z;
        x = nx;
        y = ny;
        z = nz;
        break;
      }
      case D_RIGID:  // deformations.go:173-175
        x = x + f.d[0];
        y = y + f.d[1];
        z = z + f.d[2];
        break;
      case D_SIGMOID: {  // deformations.go:210-222
        double q = f.axis == 0 ? x : (f.axis == 1 ? y : z);
        q = q + f.d[0] / (1 + exp(-(q - f.d[1]) / f.d[2]));
        if (f.axis == 0)
          x = q;
        else if (f.axis == 1)
          y = q;
        else
          z = q;
        break;
      }
      case D_TANH: {  // extension: no deformations.go counterpart
        double s = f.src == 0 ? x : (f.src == 1 ? y : z);
        double q = f.axis == 0 ? x : (f.axis == 1 ? y : z);
        q = q + f.d[0] * tanh(f.d[1] * s);
        if (f.axis == 0)
          x = q;
        else if (f.axis == 1)
          y = q;
        else
          z = q;
        break;
      }
    }
  }
}

// density() of main.go:137-140 without the multiplier.
__device__ __forceinline__ double density64(const MarchParams& p, double x,
                                            double y, double z) {
  deform64(p.defs, p.n_defs, x, y, z);
  return node64<0>(p.nodes, 0, x, y, z);
}

// density64 with the root collection's children limited to idx[0 .. n)
// (increasing node indices, the same for every lane of the wave; n < 0:
// all of them).  The list is the wave's cull list: every child left out
// has density 0 along the wave's rays, so the sum over the list (in the
// same order), and the result, are those of the full walk
// (objects.go:422-438).
__device__ __forceinline__ double density64(const MarchParams& p, double x, double y,
                                            double z, const int16_t* idx, int n) {
  deform64(p.defs, p.n_defs, x, y, z);
  if (n < 0) return node64<0>(p.nodes, 0, x, y, z);
  const DevNode* __restrict__ N = p.nodes;
  const bool greedy = N[0].greedy;
  double density = 0.0;
  for (int t = 0; t < n; ++t) {
    const int c = __builtin_amdgcn_readfirstlane((int)idx[t]);
    const DevNode& k = N[c];
    double rho;
    switch (k.kind) {
      case K_SPHERE: rho = sphere64(k, x, y, z); break;
      case K_BOX: rho = box64(k, x, y, z); break;
      case K_PPED: rho = pped64(k, x, y, z); break;
      case K_CYL: rho = cyl64(k, x, y, z); break;
      case K_GYROID: rho = gyroid64(k, x, y, z); break;
      case K_VOXEL: rho = voxel64(k, x, y, z); break;
      case K_TESS: rho = tess64<0>(N, c, x, y, z); break;
      default: rho = 0.0;
    }
    if (greedy && rho > 0.0) return rho;
    density += rho;
  }
  if (density < 0.0)
    density = 0.0;
  else if (density > 1.0)
    density = 1.0;
  return density;
}


// ---------------------------------------------------------------------------
// Guarded FP32 evaluator.  Returns the exact FP64 density when every
// membership decision is outside its guard band, else amb = true.
// ---------------------------------------------------------------------------
struct GR {
  double v;
  bool amb;
};

__device__ __forceinline__ GR sphere32(const DevNode& n, float x, float y,
                                       float z) {
  float dx = x - n.f[0], dy = y - n.f[1], dz = z - n.f[2];
  float q = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, -n.f[3])));
  bool amb = fabsf(q) < n.g[0];
  return {q < 0.0f ? n.d[4] : 0.0, amb};
}

__device__ __forceinline__ GR box32(const DevNode& n, float x, float y,
                                    float z) {
  float ax = fabsf(x - n.f[0]) - n.f[3];
  float ay = fabsf(y - n.f[1]) - n.f[4];
  float az = fabsf(z - n.f[2]) - n.f[5];
  bool out = ax >= n.g[0] || ay >= n.g[1] || az >= n.g[2];
  bool in = ax < -n.g[0] && ay < -n.g[1] && az < -n.g[2];
  return {in ? n.d[6] : 0.0, !(in || out)};
}

__device__ __forceinline__ GR pped32(const DevNode& n, float x, float y,
                                     float z) {
  float w0 = x - n.f[0], w1 = y - n.f[1], w2 = z - n.f[2];
  const float* m = n.f + 3;
  float u = fmaf(m[0], w0, fmaf(m[3], w1, m[6] * w2));
  float v = fmaf(m[1], w0, fmaf(m[4], w1, m[7] * w2));
  float t = fmaf(m[2], w0, fmaf(m[5], w1, m[8] * w2));
  bool out = u <= -n.g[0] || u >= 1.0f + n.g[0] || v <= -n.g[1] ||
             v >= 1.0f + n.g[1] || t <= -n.g[2] || t >= 1.0f + n.g[2];
  bool in = u > n.g[0] && u < 1.0f - n.g[0] && v > n.g[1] &&
            v < 1.0f - n.g[1] && t > n.g[2] && t < 1.0f - n.g[2];
  return {in ? n.d[12] : 0.0, !(in || out)};
}

__device__ __forceinline__ GR cyl32(const DevNode& n, float x, float y,
                                    float z) {
  float w0 = x - n.f[0], w1 = y - n.f[1], w2 = z - n.f[2];
  float v0 = n.f[3], v1 = n.f[4], v2 = n.f[5];
  float t = fmaf(w0, v0, fmaf(w1, v1, w2 * v2));
  float c = t * n.f[7];
  float e0 = fmaf(-c, v0, w0), e1 = fmaf(-c, v1, w1), e2 = fmaf(-c, v2, w2);
  float q = fmaf(e0, e0, fmaf(e1, e1, fmaf(e2, e2, -n.f[8])));
  float gt = n.g[1], gq = n.g[0], vv = n.f[6];
  bool cap_out = t < -gt || t > vv + gt;
  bool cap_in = t > gt && t < vv - gt;
  bool rad_out = q >= gq;
  bool rad_in = q < -gq;
  bool in = cap_in && rad_in;
  bool out = cap_out || rad_out;
  return {in ? n.d[8] : 0.0, !(in || out)};
}

// sin(2 pi t) and cos(2 pi t) on the hardware unit: t is reduced exactly to
// [-0.5, 0.5] (t - rint(t) has no rounding error), where the error of
// v_sin_f32 / v_cos_f32 is at most kHwSinCosErr.
__device__ __forceinline__ void sincos_rev(float t, float& s, float& c) {
  const float r = t - rintf(t);
  s = __builtin_amdgcn_sinf(r);
  c = __builtin_amdgcn_cosf(r);
}

// objects.go:1014-1032 in FP32: the arguments are taken in revolutions,
// t = (p - c) / (2 pi scale), the product rounding like that of the
// radians q = (p - c) / scale it replaces.
__device__ __forceinline__ GR gyroid32(const DevNode& n, float x, float y,
                                       float z) {
  float sx, cx, sy, cy, sz, cz;
  sincos_rev((x - n.f[0]) * n.f[6], sx, cx);
  sincos_rev((y - n.f[1]) * n.f[6], sy, cy);
  sincos_rev((z - n.f[2]) * n.f[6], sz, cz);
  float g = fmaf(sx, cy, fmaf(sy, cz, sz * cx));
  float a = fabsf(g) - n.f[4];
  return {a < 0.0f ? n.d[5] : 0.0, fabsf(a) < n.g[0]};
}

__device__ __forceinline__ GR voxel32(const DevNode& n, float x, float y,
                                      float z) {
  float gb = n.g[0];
  float ax = fabsf(x) - 1.0f, ay = fabsf(y) - 1.0f, az = fabsf(z) - 1.0f;
  if (ax > gb || ay > gb || az > gb) return {0.0, false};
  if (!(ax < -gb && ay < -gb && az < -gb)) return {0.0, true};
  float gx = fmaf(x, n.f[0], n.f[0]);
  float gy = fmaf(y, n.f[1], n.f[1]);
  float gz = fmaf(z, n.f[2], n.f[2]);
  int x0 = (int)floorf(gx), y0 = (int)floorf(gy), z0 = (int)floorf(gz);
  x0 = max(0, min(x0, n.nx - 1));
  y0 = max(0, min(y0, n.ny - 1));
  z0 = max(0, min(z0, n.nz - 1));
  int x1 = min(x0 + 1, n.nx - 1), y1 = min(y0 + 1, n.ny - 1),
      z1 = min(z0 + 1, n.nz - 1);
  float wx = gx - (float)x0, wy = gy - (float)y0, wz = gz - (float)z0;
  const float* r = n.vox;
  size_t NX = (size_t)n.nx, NY = (size_t)n.ny;
  size_t zs0 = (size_t)z0 * NX * NY, zs1 = (size_t)z1 * NX * NY;
  size_t xs0 = (size_t)x0 * NY, xs1 = (size_t)x1 * NY;
  float v00 = fmaf(wz, r[zs1 + xs0 + y0] - r[zs0 + xs0 + y0], r[zs0 + xs0 + y0]);
  float v01 = fmaf(wz, r[zs1 + xs0 + y1] - r[zs0 + xs0 + y1], r[zs0 + xs0 + y1]);
  float v10 = fmaf(wz, r[zs1 + xs1 + y0] - r[zs0 + xs1 + y0], r[zs0 + xs1 + y0]);
  float v11 = fmaf(wz, r[zs1 + xs1 + y1] - r[zs0 + xs1 + y1], r[zs0 + xs1 + y1]);
  float v0 = fmaf(wy, v01 - v00, v00);
  float v1 = fmaf(wy, v11 - v10, v10);
  return {(double)fmaf(wx, v1 - v0, v0), false};
}

template <int D>
__device__ __forceinline__ GR coll32(const DevNode* __restrict__ N, int i, float x, float y,
                     float z);

template <int D>
__device__ __forceinline__ GR tess32(const DevNode* __restrict__ N, int i,
                                     float x, float y, float z) {
  const DevNode& n = N[i];
  const float* b = n.f;
  float gb = n.g[1];
  bool out = x < b[0] - gb || x > b[1] + gb || y < b[2] - gb || y > b[3] + gb ||
             z < b[4] - gb || z > b[5] + gb;
  if (out) return {0.0, false};
  bool in = x > b[0] + gb && x < b[1] - gb && y > b[2] + gb && y < b[3] - gb &&
            z > b[4] + gb && z < b[5] - gb;
  if (!in) return {0.0, true};
  float fx = (x - n.f[6]) * n.f[12];
  float fy = (y - n.f[7]) * n.f[13];
  float fz = (z - n.f[8]) * n.f[14];
  float nx = floorf(fx), ny = floorf(fy), nz = floorf(fz);
  float rx = fx - nx, ry = fy - ny, rz = fz - nz;
  float gw = n.g[0];
  if (rx < gw || rx > 1.0f - gw || ry < gw || ry > 1.0f - gw || rz < gw ||
      rz > 1.0f - gw)
    return {0.0, true};
  x = fmaf(-n.f[9], nx, x);
  y = fmaf(-n.f[10], ny, y);
  z = fmaf(-n.f[11], nz, z);
  if constexpr (D + 1 <= kMaxDepth) {
    return coll32<D + 1>(N, n.first, x, y, z);
  } else {
    return {0.0, true};
  }
}

template <int D>
__device__ __forceinline__ GR leaf32(const DevNode* __restrict__ N, int c,
                                     float x, float y, float z) {
  const DevNode& k = N[c];
  switch (k.kind) {
    case K_SPHERE: return sphere32(k, x, y, z);
    case K_BOX: return box32(k, x, y, z);
    case K_PPED: return pped32(k, x, y, z);
    case K_CYL: return cyl32(k, x, y, z);
    case K_GYROID: return gyroid32(k, x, y, z);
    case K_VOXEL: return voxel32(k, x, y, z);
    case K_TESS: return tess32<D>(N, c, x, y, z);
  }
  return {0.0, false};
}

template <int D>
__device__ __forceinline__ GR coll32(const DevNode* __restrict__ N, int i, float x, float y,
                     float z) {
  const DevNode& n = N[i];
  double density = 0.0;
  for (int c = n.first; c < n.next; c = N[c].next) {
    GR r = leaf32<D>(N, c, x, y, z);
    if (r.amb) return r;
    if (n.greedy && r.v > 0.0) return r;
    density += r.v;
  }
  if (density < 0.0)
    density = 0.0;
  else if (density > 1.0)
    density = 1.0;
  return {density, false};
}

__device__ __forceinline__ GR root32(const DevNode* __restrict__ N, float x,
                                     float y, float z) {
  if (N[0].kind == K_COLL) return coll32<0>(N, 0, x, y, z);
  return leaf32<0>(N, 0, x, y, z);
}

// tanh, the logistic function and e^-x from the hardware base-2
// exponential (v_exp_f32) and reciprocal (v_rcp_f32), a handful of
// instructions each where the library functions branch on the argument.
// Absolute errors: kHwTanhErr, kHwSigmoidErr, kHwExpNegErr (dev_types.hpp).
// tanh x = sign(x) (1 - 2 / (e^(2|x|) + 1)); past |x| = 44 the exponential
// is infinite and the result exactly +-1.
__device__ __forceinline__ float tanh_hw(float x) {
  const float e = __builtin_amdgcn_exp2f(fabsf(x) * 2.88539008f);  // 2 log2(e)
  return copysignf(fmaf(-2.0f, __builtin_amdgcn_rcpf(e + 1.0f), 1.0f), x);
}
// 1 / (1 + e^-t)
__device__ __forceinline__ float sigmoid_hw(float t) {
  return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(t * -1.44269504f));
}
// e^-x for x >= 0
__device__ __forceinline__ float expneg_hw(float x) {
  return __builtin_amdgcn_exp2f(x * -1.44269504f);
}

__device__ __forceinline__ void deform32(const DevDeform* __restrict__ D,
                                         int nd, float& x, float& y,
                                         float& z) {
  for (int k = 0; k < nd; ++k) {
    const DevDeform& f = D[k];
    switch (f.kind) {
      case D_GAUSS: {
        float x0 = x - f.f[6], y0 = y - f.f[7], z0 = z - f.f[8];
        float r2 = fmaf(x0, x0, fmaf(y0, y0, z0 * z0));
        float ddx = f.f[0] * expneg_hw(r2 * f.f[3]);
        float ddy = f.f[1] * expneg_hw(r2 * f.f[4]);
        float ddz = f.f[2] * expneg_hw(r2 * f.f[5]);
        x = x + ddx;
        y = y + ddy;
        z = z + ddz;
        break;
      }
      case D_AFFINE: {
        const float* m = f.f;
        float nx = fmaf(m[0], x, fmaf(m[1], y, m[2] * z));
        float ny = fmaf(m[3], x, fmaf(m[4], y, m[5] * z));
        float nz = fmaf(m[6], x, fmaf(m[7], y, m[8] * z));
        x = nx;
        y = ny;
        z = nz;
        break;
      }
      case D_LINEAR: {
        const float* s = f.f;
        float nx = x + fmaf(s[0], x, fmaf(s[5], y, s[4] * z));
        float ny = y + fmaf(s[5], x, fmaf(s[1], y, s[3] * z));
        float nz = z + fmaf(s[4], x, fmaf(s[3], y, s[2] * z));
        x = nx;
        y = ny;
        z = nz;
        break;
      }
      case D_RIGID:
        x = x + f.f[0];
        y = y + f.f[1];
        z = z + f.f[2];
        break;
      case D_SIGMOID: {
        float q = f.axis == 0 ? x : (f.axis == 1 ? y : z);
        q = q + f.f[0] * sigmoid_hw((q - f.f[1]) * f.f[2]);
        if (f.axis == 0)
          x = q;
        else if (f.axis == 1)
          y = q;
        else
          z = q;
        break;
      }
      case D_TANH: {
        float s = f.src == 0 ? x : (f.src == 1 ? y : z);
        float q = f.axis == 0 ? x : (f.axis == 1 ? y : z);
        q = q + f.f[0] * tanh_hw(f.f[1] * s);
        if (f.axis == 0)
          x = q;
        else if (f.axis == 1)
          y = q;
        else
          z = q;
        break;
      }
    }
  }
}

// deform32 of one point together with the image of a tangent vector v: on
// return v holds J v, J the chain's Jacobian at the point (FP32).  The
// guarded marches take v = the ray direction at the block's last sample,
// so that v is the object-space velocity of the sample path per unit of
// ray length there.
__device__ __forceinline__ void deform32_jvp(const DevDeform* __restrict__ D, int nd,
                                             float& x, float& y, float& z, float& vx,
                                             float& vy, float& vz) {
  for (int k = 0; k < nd; ++k) {
    const DevDeform& f = D[k];
    switch (f.kind) {
      case D_GAUSS: {
        // x_i += A_i e^(-r^2 / (2 s_i^2)): d/dv = -2 A_i e_i / (2 s_i^2) (r . v)
        float x0 = x - f.f[6], y0 = y - f.f[7], z0 = z - f.f[8];
        float r2 = fmaf(x0, x0, fmaf(y0, y0, z0 * z0));
        float rv = fmaf(x0, vx, fmaf(y0, vy, z0 * vz));
        float ex = expneg_hw(r2 * f.f[3]), ey = expneg_hw(r2 * f.f[4]),
              ez = expneg_hw(r2 * f.f[5]);
        float ddx = f.f[0] * ex, ddy = f.f[1] * ey, ddz = f.f[2] * ez;
        vx = fmaf(-2.0f * f.f[3] * ddx, rv, vx);
        vy = fmaf(-2.0f * f.f[4] * ddy, rv, vy);
        vz = fmaf(-2.0f * f.f[5] * ddz, rv, vz);
        x = x + ddx;
        y = y + ddy;
        z = z + ddz;
        break;
      }
      case D_AFFINE: {
        const float* m = f.f;
        float nx = fmaf(m[0], x, fmaf(m[1], y, m[2] * z));
        float ny = fmaf(m[3], x, fmaf(m[4], y, m[5] * z));
        float nz = fmaf(m[6], x, fmaf(m[7], y, m[8] * z));
        float wx = fmaf(m[0], vx, fmaf(m[1], vy, m[2] * vz));
        float wy = fmaf(m[3], vx, fmaf(m[4], vy, m[5] * vz));
        float wz = fmaf(m[6], vx, fmaf(m[7], vy, m[8] * vz));
        x = nx;
        y = ny;
        z = nz;
        vx = wx;
        vy = wy;
        vz = wz;
        break;
      }
      case D_LINEAR: {
        const float* s = f.f;
        float nx = x + fmaf(s[0], x, fmaf(s[5], y, s[4] * z));
        float ny = y + fmaf(s[5], x, fmaf(s[1], y, s[3] * z));
        float nz = z + fmaf(s[4], x, fmaf(s[3], y, s[2] * z));
        float wx = vx + fmaf(s[0], vx, fmaf(s[5], vy, s[4] * vz));
        float wy = vy + fmaf(s[5], vx, fmaf(s[1], vy, s[3] * vz));
        float wz = vz + fmaf(s[4], vx, fmaf(s[3], vy, s[2] * vz));
        x = nx;
        y = ny;
        z = nz;
        vx = wx;
        vy = wy;
        vz = wz;
        break;
      }
      case D_RIGID:
        x = x + f.f[0];
        y = y + f.f[1];
        z = z + f.f[2];
        break;
      case D_SIGMOID: {
        // q += A sigma((q - c) / l): dq'/dq = 1 + (A / l) sigma (1 - sigma)
        // (the position rounds as deform32's)
        float q = f.axis == 0 ? x : (f.axis == 1 ? y : z);
        float sg = sigmoid_hw((q - f.f[1]) * f.f[2]);
        float gain = fmaf(f.f[0] * f.f[2] * sg, 1.0f - sg, 1.0f);
        q = q + f.f[0] * sg;
        if (f.axis == 0) {
          x = q;
          vx *= gain;
        } else if (f.axis == 1) {
          y = q;
          vy *= gain;
        } else {
          z = q;
          vz *= gain;
        }
        break;
      }
      case D_TANH: {
        // q += A tanh(k s): dq'/ds = A k (1 - tanh^2)
        float s = f.src == 0 ? x : (f.src == 1 ? y : z);
        float vs = f.src == 0 ? vx : (f.src == 1 ? vy : vz);
        float q = f.axis == 0 ? x : (f.axis == 1 ? y : z);
        float th = tanh_hw(f.f[1] * s);
        float dv = f.f[0] * f.f[1] * fmaf(-th, th, 1.0f) * vs;
        q = q + f.f[0] * th;
        if (f.axis == 0) {
          x = q;
          vx += dv;
        } else if (f.axis == 1) {
          y = q;
          vy += dv;
        } else {
          z = q;
          vz += dv;
        }
        break;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Ray set-up (main.go:457-465 and the Normalize of main.go:145/160).
// ---------------------------------------------------------------------------
struct Ray {
  double o[3], d[3];
};

__device__ __forceinline__ Ray make_ray(const MarchParams& p,
                                        const DevView& V, int i, int j) {
  double res_f = (double)p.res;
  double v0 = (double)i / (res_f / 2) - 1, v1 = (double)j / (res_f / 2) - 1,
         v2 = -p.focal;
  double t[4];
#pragma unroll
  for (int r = 0; r < 4; ++r)
    t[r] = V.cam[r * 4 + 0] * v0 + V.cam[r * 4 + 1] * v1 +
           V.cam[r * 4 + 2] * v2 + V.cam[r * 4 + 3] * 1.0;
  double iw = 1 / t[3];
  double w[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) w[k] = t[k] * iw - V.eye[k];
  double l = 1.0 / sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
  Ray r;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    r.o[k] = V.eye[k];
    r.d[k] = w[k] * l;
  }
  return r;
}

// Parameter interval [t_in, t_out] of the ray inside the clip box;
// false when it misses (NaN rays, produced by polar angles 0/180, miss).
__device__ __forceinline__ bool clip_ray(const MarchParams& p, const Ray& r,
                                         double& t_in, double& t_out) {
  t_in = -INFINITY;
  t_out = INFINITY;
  // A NaN ray (degenerate camera) has zero density at every sample in the
  // reference: every membership comparison is false.
  if (!(r.d[0] == r.d[0] && r.d[1] == r.d[1] && r.d[2] == r.d[2] &&
        r.o[0] == r.o[0] && r.o[1] == r.o[1] && r.o[2] == r.o[2]))
    return false;
  if (!p.clip) return true;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    double o = r.o[a], d = r.d[a];
    if (d != 0.0) {
      double t1 = (p.lo[a] - o) / d, t2 = (p.hi[a] - o) / d;
      double tn = fmin(t1, t2), tf = fmax(t1, t2);
      t_in = fmax(t_in, tn);
      t_out = fmin(t_out, tf);
    } else if (!(o >= p.lo[a] && o <= p.hi[a])) {
      return false;
    }
  }
  return t_in <= t_out;  // false for NaN
}

// Stores pixel (view slot v, i, j) with transmittance exp(-T).  Mode 1
// quantizes like main.go:485-498: uint16(val*0xffff) (Go converts through
// int64 on amd64, keeping the low 16 bits) and the RGBA image keeps the
// high byte; alpha is 0 where val >= 1 under transparency.
__device__ __forceinline__ void store_pixel(const MarchParams& p, int v, int i,
                                            int j, double T) {
  const double val = exp(-T);
  const size_t idx = (size_t)v * p.out_view_stride + (size_t)i * p.res + j;
  if (p.out_mode == 0) {
    p.out[idx] = (float)val;
  } else {
    unsigned v16 = (unsigned)((long long)(val * 65535.0) & 0xffff);
    unsigned a16 = (p.transparency && !(val < 1.0)) ? 0u : 0xffffu;
    p.out8[2 * idx] = (unsigned char)(v16 >> 8);
    p.out8[2 * idx + 1] = (unsigned char)(a16 >> 8);
  }
}

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

}  // namespace xrp
