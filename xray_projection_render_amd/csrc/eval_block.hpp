// eval_block.hpp - the guarded FP32 evaluator of eval_device.hpp applied to
// a block of S consecutive samples of one ray per node visit.
//
// Every lane of a wave visits the same node at the same time, so the node's
// record is read (scalar loads) and its kind branched on once per block
// instead of once per sample, and the S independent membership tests give
// the wave instruction-level parallelism to cover the load latency.  The
// FP32 arithmetic per sample is the one of the single-sample functions
// (sphere32, cyl32, ... in eval_device.hpp), so membership, guard flags and
// densities are identical; a collection stops walking its children for a
// sample exactly where the single-sample walk would (first ambiguous child,
// or first positive child of a greedy collection).
//
// With DIST the walk also returns, for the block's last sample, a lower
// bound on the distance from its object-space position to the nearest
// boundary that could change the density there (primitive surfaces, the
// tessellation bounds and cell seams), with every FP32 error the bound can
// carry already subtracted (slack terms from compute_guards, g[4..6]).  Inside
// that radius the exact density is the sample's own, which lets the march
// settle the following samples without evaluating them.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

#include "eval_device.hpp"

namespace xrp {

// Positions of a block of S samples, and the object-space velocity of the
// sample path at the last one (per unit of ray length; the DIST walks read
// it, translations such as the tessellation wrap leave it unchanged).  Zero
// when the march carries none: no second-order reach is taken then.
template <int S>
struct PBlock {
  float x[S], y[S], z[S];
  float vx = 0.0f, vy = 0.0f, vz = 0.0f;
};

// Safe-radius bounds of single points, for the neighbouring cell that
// tess_blk looks into: the same bounds and slacks as the blocked
// evaluators apply to their block's last sample.
__device__ __forceinline__ float sphere_dist(const DevNode& n, float x, float y, float z) {
  const float dx = x - n.f[0], dy = y - n.f[1], dz = z - n.f[2];
  const float l = __builtin_amdgcn_sqrtf(fmaf(dx, dx, fmaf(dy, dy, dz * dz)));
  return fabsf(l - n.g[4]) - n.g[5];
}
__device__ __forceinline__ float box_dist(const DevNode& n, float x, float y, float z) {
  const float ax = fabsf(x - n.f[0]) - n.f[3];
  const float ay = fabsf(y - n.f[1]) - n.f[4];
  const float az = fabsf(z - n.f[2]) - n.f[5];
  return fabsf(fmaxf(ax, fmaxf(ay, az))) - n.g[5];
}
__device__ __forceinline__ float pped_dist(const DevNode& n, float x, float y, float z) {
  const float w0 = x - n.f[0], w1 = y - n.f[1], w2 = z - n.f[2];
  const float u = fmaf(n.f[3], w0, fmaf(n.f[6], w1, n.f[9] * w2));
  const float v = fmaf(n.f[4], w0, fmaf(n.f[7], w1, n.f[10] * w2));
  const float t = fmaf(n.f[5], w0, fmaf(n.f[8], w1, n.f[11] * w2));
  const float e = fmaxf(fmaxf(fmaxf(-u, u - 1.0f), fmaxf(-v, v - 1.0f)), fmaxf(-t, t - 1.0f));
  return (fabsf(e) - n.g[5]) * n.g[4];
}
__device__ __forceinline__ float cyl_dist(const DevNode& n, float x, float y, float z) {
  const float w0 = x - n.f[0], w1 = y - n.f[1], w2 = z - n.f[2];
  const float t = fmaf(w0, n.f[3], fmaf(w1, n.f[4], w2 * n.f[5]));
  const float c = t * n.f[7];
  const float e0 = fmaf(-c, n.f[3], w0), e1 = fmaf(-c, n.f[4], w1), e2 = fmaf(-c, n.f[5], w2);
  const float ra = __builtin_amdgcn_sqrtf(fmaf(e0, e0, fmaf(e1, e1, e2 * e2))) - n.g[4];
  const float aa = fmaxf(-t, t - n.f[6]) * n.g[6];
  return fabsf(fmaxf(ra, aa)) - n.g[5];
}
__device__ __forceinline__ float gyroid_reach(const DevNode& n, float gg, float g1);
// With a velocity (vx, vy, vz) of the path at the point (zero when the
// march carries none), also the second-order reach along it, as an
// object-space radius (gyroid_blk).
__device__ __forceinline__ float gyroid_dist(const DevNode& n, float x, float y, float z,
                                             float vx, float vy, float vz) {
  float sx, cxs, sy, cys, sz, czs;
  sincos_rev((x - n.f[0]) * n.f[6], sx, cxs);
  sincos_rev((y - n.f[1]) * n.f[6], sy, cys);
  sincos_rev((z - n.f[2]) * n.f[6], sz, czs);
  const float gg = fmaf(sx, cys, fmaf(sy, czs, sz * cxs));
  const float a = fabsf(gg) - n.f[4];
  float r = (fabsf(a) - n.g[0]) * n.g[4];
  if (n.g[6] > 0.0f && (vx != 0.0f || vy != 0.0f || vz != 0.0f)) {
    const float gx = fmaf(cxs, cys, -sz * sx);
    const float gy = fmaf(cys, czs, -sx * sy);
    const float gz = fmaf(czs, cxs, -sy * sz);
    const float g1 = fmaf(gx, vx, fmaf(gy, vy, gz * vz)) * n.g[3];
    r = fmaxf(r, gyroid_reach(n, gg, g1) * n.g[5]);
  }
  return r;
}
// Smallest bound over a collection's children (no skipping past a child
// without one: nested tessellations and voxel grids give 0).
__device__ __forceinline__ float coll_dist(const DevNode* __restrict__ N, int i, float x,
                                           float y, float z, float vx, float vy, float vz) {
  const DevNode& n = N[i];
  float d = INFINITY;
  for (int c = n.first; c < n.next; c = N[c].next) {
    const DevNode& k = N[c];
    switch (k.kind) {
      case K_SPHERE: d = fminf(d, sphere_dist(k, x, y, z)); break;
      case K_BOX: d = fminf(d, box_dist(k, x, y, z)); break;
      case K_PPED: d = fminf(d, pped_dist(k, x, y, z)); break;
      case K_CYL: d = fminf(d, cyl_dist(k, x, y, z)); break;
      case K_GYROID: d = fminf(d, gyroid_dist(k, x, y, z, vx, vy, vz)); break;
      default: return 0.0f;
    }
  }
  return d;
}

template <int S, bool DIST>
__device__ __forceinline__ void sphere_blk(const DevNode& n, const PBlock<S>& p,
                                           double (&v)[S], unsigned& amb, float& dist) {
  const float cx = n.f[0], cy = n.f[1], cz = n.f[2], r2 = n.f[3], g = n.g[0];
  const double rho = n.d[4];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    float dx = p.x[s] - cx, dy = p.y[s] - cy, dz = p.z[s] - cz;
    float q = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, -r2)));
    v[s] = q < 0.0f ? rho : 0.0;
    amb |= (fabsf(q) < g ? 1u : 0u) << s;
    if constexpr (DIST) {
      if (s == S - 1) {
        // | |p - c| - r |; the hardware square root (within an ulp, which
        // the slack allows for) instead of the correctly rounded sequence
        const float l = __builtin_amdgcn_sqrtf(fmaf(dx, dx, fmaf(dy, dy, dz * dz)));
        dist = fminf(dist, fabsf(l - n.g[4]) - n.g[5]);
      }
    }
  }
}

template <int S, bool DIST, bool G = false>
__device__ __forceinline__ void box_blk(const DevNode& n, const PBlock<S>& p,
                                        double (&v)[S], unsigned& amb, float& dist) {
  const float cx = n.f[0], cy = n.f[1], cz = n.f[2], hx = n.f[3], hy = n.f[4], hz = n.f[5];
  const float g0 = n.g[0], g1 = n.g[1], g2 = n.g[2];
  const double rho = n.d[6];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    float ax = fabsf(p.x[s] - cx) - hx;
    float ay = fabsf(p.y[s] - cy) - hy;
    float az = fabsf(p.z[s] - cz) - hz;
    bool out = ax >= g0 || ay >= g1 || az >= g2;
    bool in = ax < -g0 && ay < -g1 && az < -g2;
    v[s] = (G ? ax < 0.0f && ay < 0.0f && az < 0.0f : in) ? rho : 0.0;
    amb |= (!(in || out) ? 1u : 0u) << s;
    if constexpr (DIST) {
      // outside: the largest face excess; inside: the nearest face
      if (s == S - 1) dist = fminf(dist, fabsf(fmaxf(ax, fmaxf(ay, az))) - n.g[5]);
    }
  }
}

template <int S, bool DIST, bool G = false>
__device__ __forceinline__ void pped_blk(const DevNode& n, const PBlock<S>& p,
                                         double (&v)[S], unsigned& amb, float& dist) {
  const float ox = n.f[0], oy = n.f[1], oz = n.f[2];
  const float m0 = n.f[3], m1 = n.f[4], m2 = n.f[5], m3 = n.f[6], m4 = n.f[7], m5 = n.f[8],
              m6 = n.f[9], m7 = n.f[10], m8 = n.f[11];
  const float g0 = n.g[0], g1 = n.g[1], g2 = n.g[2];
  const double rho = n.d[12];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    float w0 = p.x[s] - ox, w1 = p.y[s] - oy, w2 = p.z[s] - oz;
    float u = fmaf(m0, w0, fmaf(m3, w1, m6 * w2));
    float vv = fmaf(m1, w0, fmaf(m4, w1, m7 * w2));
    float t = fmaf(m2, w0, fmaf(m5, w1, m8 * w2));
    bool out = u <= -g0 || u >= 1.0f + g0 || vv <= -g1 || vv >= 1.0f + g1 || t <= -g2 ||
               t >= 1.0f + g2;
    bool in = u > g0 && u < 1.0f - g0 && vv > g1 && vv < 1.0f - g1 && t > g2 &&
              t < 1.0f - g2;
    const bool guess = u > 0.0f && u < 1.0f && vv > 0.0f && vv < 1.0f && t > 0.0f && t < 1.0f;
    v[s] = (G ? guess : in) ? rho : 0.0;
    amb |= (!(in || out) ? 1u : 0u) << s;
    if constexpr (DIST) {
      if (s == S - 1) {
        // distance to the unit cube's faces in (u, v, t), less the bands,
        // over the largest stretch of M^-1 (its Frobenius norm)
        float e = fmaxf(fmaxf(fmaxf(-u, u - 1.0f), fmaxf(-vv, vv - 1.0f)),
                        fmaxf(-t, t - 1.0f));
        dist = fminf(dist, (fabsf(e) - n.g[5]) * n.g[4]);
      }
    }
  }
}

template <int S, bool DIST, bool G = false>
__device__ __forceinline__ void cyl_blk(const DevNode& n, const PBlock<S>& p,
                                        double (&v)[S], unsigned& amb, float& dist) {
  const float px = n.f[0], py = n.f[1], pz = n.f[2];
  const float v0 = n.f[3], v1 = n.f[4], v2 = n.f[5], vv = n.f[6], ivv = n.f[7], r2 = n.f[8];
  const float gq = n.g[0], gt = n.g[1];
  const double rho = n.d[8];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    float w0 = p.x[s] - px, w1 = p.y[s] - py, w2 = p.z[s] - pz;
    float t = fmaf(w0, v0, fmaf(w1, v1, w2 * v2));
    float c = t * ivv;
    float e0 = fmaf(-c, v0, w0), e1 = fmaf(-c, v1, w1), e2 = fmaf(-c, v2, w2);
    float q = fmaf(e0, e0, fmaf(e1, e1, fmaf(e2, e2, -r2)));
    bool cap_out = t < -gt || t > vv + gt;
    bool cap_in = t > gt && t < vv - gt;
    bool in = cap_in && q < -gq;
    bool out = cap_out || q >= gq;
    v[s] = (G ? t >= 0.0f && t <= vv && q < 0.0f : in) ? rho : 0.0;
    amb |= (!(in || out) ? 1u : 0u) << s;
    if constexpr (DIST) {
      if (s == S - 1) {
        // radial |e| - r and axial excess past the caps, both in length
        // units; the solid's boundary is at least |max(.,.)| away
        const float ra = __builtin_amdgcn_sqrtf(fmaf(e0, e0, fmaf(e1, e1, e2 * e2))) - n.g[4];
        const float aa = fmaxf(-t, t - vv) * n.g[6];
        dist = fminf(dist, fabsf(fmaxf(ra, aa)) - n.g[5]);
      }
    }
  }
}

// Length of ray, from a point where g = g0 and dg/ds = g1 (FP32, within
// the band and e1 of the exact values), over which the exact g keeps its
// side of the level |g| = th: with |g''| <= M along the path,
// sigma g(s) - th >= c + (sigma g1 - e1) s - M s^2 / 2 outside (c = |g0| -
// th - band, sigma = sign g0), and both th - g and g + th are bounded the
// same way inside.  The first root of c + b s - M s^2 / 2, written without
// cancellation: (b + sqrt(b^2 + 2 M c)) / M for b > 0, else
// 2c / (sqrt(b^2 + 2 M c) - b).
// The division by M is a product with 1/M (n.g[1], rounded to nearest) and
// the other one the hardware reciprocal (within an ulp): a few ulps in all,
// which the final 1e-4 shrink covers.
__device__ __forceinline__ float quad_reach(float c, float b, float M, float inv_M) {
  if (!(c > 0.0f)) return 0.0f;
  const float q = __builtin_amdgcn_sqrtf(fmaf(b, b, 2.0f * M * c));
  return b > 0.0f ? (b + q) * inv_M : 2.0f * c * __builtin_amdgcn_rcpf(q - b);
}
__device__ __forceinline__ float gyroid_reach(const DevNode& n, float gg, float g1) {
  const float th = n.f[4], band = n.g[0], M = n.g[6], e1 = n.g[7], inv_M = n.g[1];
  const float ag = fabsf(gg);
  float s;
  if (ag >= th) {
    s = quad_reach(ag - th - band, (gg < 0.0f ? -g1 : g1) - e1, M, inv_M);
  } else {
    s = fminf(quad_reach(th - band - gg, -g1 - e1, M, inv_M),
              quad_reach(gg + th - band, g1 - e1, M, inv_M));
  }
  return s * (1.0f - 1e-4f);
}

template <int S, bool DIST>
__device__ __forceinline__ void gyroid_blk(const DevNode& n, const PBlock<S>& p,
                                           double (&v)[S], unsigned& amb, float& dist) {
  const float cx = n.f[0], cy = n.f[1], cz = n.f[2], rv = n.f[6], th = n.f[4], g = n.g[0];
  const double rho = n.d[5];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    float sx, cxs, sy, cys, sz, czs;
    sincos_rev((p.x[s] - cx) * rv, sx, cxs);
    sincos_rev((p.y[s] - cy) * rv, sy, cys);
    sincos_rev((p.z[s] - cz) * rv, sz, czs);
    float gg = fmaf(sx, cys, fmaf(sy, czs, sz * cxs));
    float a = fabsf(gg) - th;
    v[s] = a < 0.0f ? rho : 0.0;
    amb |= (fabsf(a) < g ? 1u : 0u) << s;
    if constexpr (DIST) {
      if (s == S - 1) {
        // |grad g| <= 1.77 / scale: the level |g| = th is at least
        // (|a| - band) * scale / 1.77 away
        float r = (fabsf(a) - g) * n.g[4];
        if (n.g[6] > 0.0f && (p.vx != 0.0f || p.vy != 0.0f || p.vz != 0.0f)) {
          // along the ray: dg/ds = grad g . v / scale, and the second-order
          // reach, as an object-space radius (reach times the chain's
          // Lipschitz bound n.g[5]: the march divides by it again)
          const float gx = fmaf(cxs, cys, -sz * sx);
          const float gy = fmaf(cys, czs, -sx * sy);
          const float gz = fmaf(czs, cxs, -sy * sz);
          const float g1 = fmaf(gx, p.vx, fmaf(gy, p.vy, gz * p.vz)) * n.g[3];
          r = fmaxf(r, gyroid_reach(n, gg, g1) * n.g[5]);
        }
        dist = fminf(dist, r);
      }
    }
  }
}

// The children of the root collection that can be non-zero somewhere along
// the rays of the calling wave, as increasing node indices in LDS; n < 0
// walks every child.  A child left out has density 0 at every sample of
// those rays, so its terms (+0) and its safe-radius bound (the radius is
// only used along the same rays) change nothing.
struct Cull {
  const int16_t* idx = nullptr;
  int n = -1;
};

template <int D, int S, bool DIST, bool G = false>
__device__ __forceinline__ void coll_blk(const DevNode* __restrict__ N, int i,
                                         const PBlock<S>& p, unsigned live, double (&out)[S],
                                         unsigned& amb_out, float& dist,
                                         const Cull& cl = Cull());

// Tessellation wrap (objects.go:568-582) per sample, then the unit cell's
// collection for the samples that passed the bounds and seam guards.
// G (a best guess for samples inside a guard band): those inside the
// bounds by the FP32 test are evaluated in the cell the FP32 wrap picks.
template <int D, int S, bool DIST, bool G = false>
__device__ __forceinline__ void tess_blk(const DevNode* __restrict__ N, int i,
                                         const PBlock<S>& p, unsigned live, double (&out)[S],
                                         unsigned& amb_out, float& dist) {
  const DevNode& n = N[i];
  const float b0 = n.f[0], b1 = n.f[1], b2 = n.f[2], b3 = n.f[3], b4 = n.f[4], b5 = n.f[5];
  const float gb = n.g[1], gw = n.g[0];
  // Seamless content (scene.cpp: every child of the unit cell a gyroid whose
  // period divides the cell along each axis): either side of a seam wraps to
  // the same density function, so seams need no guard and the safe radius
  // may cross cell faces.
  const bool seamless = n.g[2] != 0.0f;
  PBlock<S> q;
  unsigned go = 0, amb = 0;
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const float x = p.x[s], y = p.y[s], z = p.z[s];
    out[s] = 0.0;
    q.x[s] = x;
    q.y[s] = y;
    q.z[s] = z;
    bool o = x < b0 - gb || x > b1 + gb || y < b2 - gb || y > b3 + gb || z < b4 - gb ||
             z > b5 + gb;
    bool in = x > b0 + gb && x < b1 - gb && y > b2 + gb && y < b3 - gb && z > b4 + gb &&
              z < b5 - gb;
    float fx = (x - n.f[6]) * n.f[12];
    float fy = (y - n.f[7]) * n.f[13];
    float fz = (z - n.f[8]) * n.f[14];
    float nx = floorf(fx), ny = floorf(fy), nz = floorf(fz);
    float rx = fx - nx, ry = fy - ny, rz = fz - nz;
    bool seam = rx < gw || rx > 1.0f - gw || ry < gw || ry > 1.0f - gw || rz < gw ||
                rz > 1.0f - gw;
    q.x[s] = fmaf(-n.f[9], nx, x);
    q.y[s] = fmaf(-n.f[10], ny, y);
    q.z[s] = fmaf(-n.f[11], nz, z);
    if (s == S - 1) {
      q.vx = p.vx;
      q.vy = p.vy;
      q.vz = p.vz;
    }
    if (!o) {
      if (!in || (seam && !seamless)) {
        amb |= 1u << s;
        if (G && x >= b0 && x <= b1 && y >= b2 && y <= b3 && z >= b4 && z <= b5)
          go |= 1u << s;
      } else {
        go |= 1u << s;
      }
    }
    if constexpr (DIST) {
      if (s == S - 1) {
        // the bounds box, either side (the cell faces: after the walk)
        const float eb = fmaxf(fmaxf(fmaxf(b0 - x, x - b1), fmaxf(b2 - y, y - b3)),
                               fmaxf(b4 - z, z - b5));
        dist = fminf(dist, fabsf(eb) - gb);
      }
    }
  }
  amb &= live;
  go &= live;
  amb_out |= amb;
  if constexpr (D + 1 <= kMaxDepth) {
    if (go) coll_blk<D + 1, S, DIST, G>(N, n.first, q, go, out, amb_out, dist);
  } else {
    amb_out |= go;
  }
  if (DIST && !seamless) {
    // Last sample inside the bounds: a radius up to the nearest cell face
    // stays in this cell, whose content the walk above bounded.  Beyond it,
    // the neighbour across that face holds the same content translated, and
    // up to the second-nearest face a ball meets only these two cells.  The
    // wrap is recomputed here (the same arithmetic as above) rather than
    // kept live through the walk.  Face distances are lower bounds (the
    // fractions are within the band gw).
    const float x = p.x[S - 1], y = p.y[S - 1], z = p.z[S - 1];
    const bool o = x < b0 - gb || x > b1 + gb || y < b2 - gb || y > b3 + gb ||
                   z < b4 - gb || z > b5 + gb;
    const float sx = n.f[9], sy = n.f[10], sz = n.f[11];
    const float fx = (x - n.f[6]) * n.f[12];
    const float fy = (y - n.f[7]) * n.f[13];
    const float fz = (z - n.f[8]) * n.f[14];
    const float nx = floorf(fx), ny = floorf(fy), nz = floorf(fz);
    const float rx = fx - nx, ry = fy - ny, rz = fz - nz;
    const float fc[6] = {(rx - gw) * sx, (1.0f - rx - gw) * sx, (ry - gw) * sy,
                         (1.0f - ry - gw) * sy, (rz - gw) * sz, (1.0f - rz - gw) * sz};
    int a = 0;
    float f1 = fc[0], f2 = INFINITY;
#pragma unroll
    for (int t = 1; t < 6; ++t) {
      if (fc[t] < f1) {
        f2 = f1;
        f1 = fc[t];
        a = t;
      } else {
        f2 = fminf(f2, fc[t]);
      }
    }
    if (!o && dist > f1) {
      // this cell's point, moved one cell side up (a even: the neighbour
      // below) or down (a odd: the neighbour above)
      const float sh = (a & 1) ? -1.0f : 1.0f;
      float qx = fmaf(-sx, nx, x), qy = fmaf(-sy, ny, y), qz = fmaf(-sz, nz, z);
      if (a < 2)
        qx = fmaf(sh, sx, qx);
      else if (a < 4)
        qy = fmaf(sh, sy, qy);
      else
        qz = fmaf(sh, sz, qz);
      // Past the face the density is the neighbour's, whose content is
      // this cell's translated: constant within its bound dn around this
      // point (the path keeps its velocity there, zero when the march
      // carries none), but equal to this sample's only where the two cells
      // agree.  The content need not be continuous across faces (a
      // gyroid of 1.5 periods per cell is not), so the neighbour's density
      // at this point is evaluated; unless it is clear and equal, the
      // radius stops at the face.
      float cap = f1;
      if constexpr (D + 1 <= kMaxDepth) {
        // one walk of the content at the neighbour's point: its density,
        // whether that is clear, and its bound
        PBlock<1> nq;
        nq.x[0] = qx;
        nq.y[0] = qy;
        nq.z[0] = qz;
        nq.vx = p.vx;
        nq.vy = p.vy;
        nq.vz = p.vz;
        double nv[1];
        unsigned namb = 0;
        float dn = INFINITY;
        coll_blk<D + 1, 1, true>(N, n.first, nq, 1u, nv, namb, dn);
        if (!namb && nv[0] == out[S - 1]) cap = fminf(f2, dn);
      }
      dist = fminf(dist, cap);
    }
  }
}

template <int D, int S, bool DIST, bool G = false>
__device__ __forceinline__ void leaf_blk(const DevNode* __restrict__ N, int c,
                                         const PBlock<S>& p, unsigned live, double (&v)[S],
                                         unsigned& amb, float& dist) {
  const DevNode& k = N[c];
  switch (k.kind) {
    case K_SPHERE: sphere_blk<S, DIST>(k, p, v, amb, dist); return;
    case K_BOX: box_blk<S, DIST, G>(k, p, v, amb, dist); return;
    case K_PPED: pped_blk<S, DIST, G>(k, p, v, amb, dist); return;
    case K_CYL: cyl_blk<S, DIST, G>(k, p, v, amb, dist); return;
    case K_GYROID: gyroid_blk<S, DIST>(k, p, v, amb, dist); return;
    case K_TESS:
      // Tessellations below the first level (a tessellation inside a
      // tessellated cell) are left to the FP64 evaluator: the blocked
      // walk keeps one level of cell state in registers.
      if constexpr (D == 0) {
        tess_blk<D, S, DIST, G>(N, c, p, live, v, amb, dist);
      } else {
#pragma unroll
        for (int s = 0; s < S; ++s) v[s] = 0.0;
        amb |= live;
      }
      return;
    case K_VOXEL:
      // trilinear density varies continuously: no radius
      if constexpr (DIST) dist = 0.0f;
#pragma unroll
      for (int s = 0; s < S; ++s) {
        GR r = voxel32(k, p.x[s], p.y[s], p.z[s]);
        v[s] = r.v;
        amb |= (r.amb ? 1u : 0u) << s;
      }
      return;
  }
  if constexpr (DIST) dist = 0.0f;
#pragma unroll
  for (int s = 0; s < S; ++s) v[s] = 0.0;
}

// objects.go:422-438 per sample: sum then clamp to [0, 1]; greedy returns
// the first positive child unclamped.  A sample whose child is inside a
// guard band is flagged and left (its value is not used), or with G (a
// best guess) summed on with the child's FP32 value.
template <int D, int S, bool DIST, bool G>
__device__ __forceinline__ void coll_blk(const DevNode* __restrict__ N, int i,
                                         const PBlock<S>& p, unsigned live, double (&out)[S],
                                         unsigned& amb_out, float& dist, const Cull& cl) {
  const DevNode& n = N[i];
  const bool greedy = n.greedy;
  // out[s] holds the running sum until sample s finishes
#pragma unroll
  for (int s = 0; s < S; ++s) out[s] = 0.0;
  unsigned done = 0, amb = 0, hit = 0;
  const bool listed = cl.n >= 0;
  int t = 0;
  int c = !listed ? n.first
                  : (cl.n > 0 ? __builtin_amdgcn_readfirstlane((int)cl.idx[0]) : n.next);
  for (; c < n.next;
       c = !listed ? N[c].next
                   : (++t < cl.n ? __builtin_amdgcn_readfirstlane((int)cl.idx[t]) : n.next)) {
    const unsigned act = live & ~done;
    if (!act) break;
    double v[S];
    unsigned a = 0;
    // A greedy collection's result is fixed by its children up to the
    // first positive one, so the distances of the children walked bound it.
    leaf_blk<D, S, DIST, G>(N, c, p, act, v, a, dist);
    a &= act;
    amb |= a;
    if (!G) done |= a;
#pragma unroll
    for (int s = 0; s < S; ++s) {
      if (((act & (G ? ~0u : ~a)) >> s) & 1u) {
        if (greedy && v[s] > 0.0) {
          out[s] = v[s];
          hit |= 1u << s;
          done |= 1u << s;
        } else {
          out[s] += v[s];
        }
      }
    }
  }
#pragma unroll
  for (int s = 0; s < S; ++s) {
    if (((live & (G ? ~0u : ~amb) & ~hit) >> s) & 1u) {
      double d = out[s];
      out[s] = d < 0.0 ? 0.0 : (d > 1.0 ? 1.0 : d);
    }
  }
  amb_out |= amb;
}

// dist (DIST only) starts at +inf in the caller and is lowered by every
// node the last sample visits.
// cl: the wave's culled child list when the root is a collection.
// G: samples inside a guard band get a best guess, their FP32 value.
template <int S, bool DIST = false, bool G = false>
__device__ __forceinline__ void root_blk(const DevNode* __restrict__ N, const PBlock<S>& p,
                                         unsigned live, double (&out)[S], unsigned& amb,
                                         float& dist, const Cull& cl = Cull()) {
  if (N[0].kind == K_COLL)
    coll_blk<0, S, DIST, G>(N, 0, p, live, out, amb, dist, cl);
  else
    leaf_blk<0, S, DIST, G>(N, 0, p, live, out, amb, dist);
  amb &= live;
}
template <int S, bool G = false>
__device__ __forceinline__ void root_blk(const DevNode* __restrict__ N, const PBlock<S>& p,
                                         unsigned live, double (&out)[S], unsigned& amb,
                                         const Cull& cl = Cull()) {
  float unused = 0.0f;
  root_blk<S, false, G>(N, p, live, out, amb, unused, cl);
}


// With vel (wave-uniform) the last sample's velocity (p.vx..vz, the ray
// direction on entry) is carried through the chain as well.
template <int S>
__device__ __forceinline__ void deform_blk(const DevDeform* __restrict__ D, int nd,
                                           PBlock<S>& p, bool vel = false) {
  for (int k = 0; k < nd; ++k) {
#pragma unroll
    for (int s = 0; s < S - 1; ++s) deform32(D + k, 1, p.x[s], p.y[s], p.z[s]);
    if (vel)
      deform32_jvp(D + k, 1, p.x[S - 1], p.y[S - 1], p.z[S - 1], p.vx, p.vy, p.vz);
    else
      deform32(D + k, 1, p.x[S - 1], p.y[S - 1], p.z[S - 1]);
  }
}

}  // namespace xrp
