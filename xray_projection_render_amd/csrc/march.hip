// march.hip - gfx950 kernels for the per-pixel X-ray line integral.
//
// One lane per camera ray; a 64-lane wave covers an 8x8 pixel tile and a
// 256-thread workgroup a 16x16 tile.  Every lane of a wave walks the same
// flattened scene node at the same time, so node records are read through
// the scalar data cache and branch on node kind is wave-uniform.
//
// Numerics (reference: main.go:137-199, objects.go, deformations.go):
//  * Ray set-up, sample positions s_k and the integral T are FP64 and
//    follow the reference's operation order (this file is built with
//    -ffp-contract=off; FP32 multiply-adds are explicit fmaf calls).
//  * PREC_FP64 evaluates the density tree in FP64 exactly as the oracle.
//  * PREC_GUARDED decides primitive membership in FP32.  Each test also
//    checks a guard band, sized on the host from a rigorous bound on the
//    FP32 position error; a sample inside any band is re-evaluated with
//    the FP64 evaluator.  Membership outside the bands provably agrees
//    with FP64, and the density value returned is the FP64 constant, so T
//    is accumulated from exactly the oracle's per-sample densities.
#include <hip/hip_runtime.h>

#include <cmath>

#include "eval_device.hpp"
#include "march.hpp"

namespace xrp {

template <int PREC, int INTEG>
__global__ __launch_bounds__(256) void march_kernel(MarchParams p) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int tiles_j = (p.res + 15) >> 4;
  const int ti = blockIdx.x / tiles_j, tj = blockIdx.x - ti * tiles_j;
  const int i = ti * 16 + (wave >> 1) * 8 + (lane >> 3);
  const int j = tj * 16 + (wave & 1) * 8 + (lane & 7);
  const int view = p.view_offset + blockIdx.y;
  const bool valid = i < p.res && j < p.res;
  const DevView& V = p.views[view];

  unsigned long long n_samples = 0, n_reeval = 0, n_miss = 0, n_fine = 0;
  double T = p.flat_field;

  if (valid) {
    Ray r = make_ray(p, V, i, j);
    double t_in, t_out;
    if (clip_ray(p, r, t_in, t_out)) {
      if constexpr (INTEG == INT_SIMPLE) {
        // main.go:144-154; s_k = s_table[k], s_{k+1} = s_k + ds.
        long k_lo = 0, k_hi = p.K - 1;
        if (p.clip) {
          double a = floor((t_in - p.smin) / p.ds) - 2.0;
          double b = ceil((t_out - p.smin) / p.ds) + 2.0;
          k_lo = a > 0.0 ? (long)a : 0;
          k_hi = b < (double)(p.K - 1) ? (long)b : (long)(p.K - 1);
        }
        const double ds = p.ds;
        if constexpr (PREC == PREC_FP64) {
          if (k_lo <= k_hi) {
            double s = p.s_table[k_lo];
            for (long k = k_lo; k <= k_hi; ++k) {
              double x = r.o[0] + r.d[0] * s;
              double y = r.o[1] + r.d[1] * s;
              double z = r.o[2] + r.d[2] * s;
              T += (density64(p, x, y, z) * p.dm) * ds;
              s += ds;
            }
            n_samples += (unsigned long long)(k_hi - k_lo + 1);
          }
        } else {
          const float ds32 = (float)ds;
          const float d0 = (float)r.d[0], d1 = (float)r.d[1], d2 = (float)r.d[2];
          for (long k0 = k_lo; k0 <= k_hi; k0 += kChunk) {
            double sa = p.s_table[k0];
            float a0 = (float)(r.o[0] + r.d[0] * sa);
            float a1 = (float)(r.o[1] + r.d[1] * sa);
            float a2 = (float)(r.o[2] + r.d[2] * sa);
            int n = (int)min((long)kChunk, k_hi - k0 + 1);
            for (int jj = 0; jj < n; ++jj) {
              float sj = (float)jj * ds32;
              float x = fmaf(d0, sj, a0), y = fmaf(d1, sj, a1),
                    z = fmaf(d2, sj, a2);
              deform32(p.defs, p.n_defs, x, y, z);
              GR g = root32(p.nodes, x, y, z);
              double rho = g.v;
              if (g.amb || p.audit) {
                double s = p.s_table[k0 + jj];
                double e = density64(p, r.o[0] + r.d[0] * s,
                                     r.o[1] + r.d[1] * s, r.o[2] + r.d[2] * s);
                if (g.amb) {
                  ++n_reeval;
                } else if (e != rho) {
                  ++n_miss;
                }
                rho = e;
              }
              T += (rho * p.dm) * ds;
            }
            n_samples += (unsigned long long)n;
          }
        }
      } else {
        // main.go:159-199: coarse right-endpoint samples R[n] (s_table),
        // fine sub-steps of DS/10 from the previous coarse point where
        // (rho == 0) toggles.
        const double DS = p.ds, dsf = p.ds_fine;
        long n_lo = 0, n_hi = p.K - 1;
        if (p.clip) {
          double a = floor((t_in - p.smin) / DS) - 3.0;
          double b = ceil((t_out - p.smin) / DS) + 3.0;
          n_lo = a > 0.0 ? (long)a : 0;
          n_hi = b < (double)(p.K - 1) ? (long)b : (long)(p.K - 1);
        }
        bool prev_zero = true;
        for (long n = n_lo; n <= n_hi; ++n) {
          double right = p.s_table[n];
          double rho;
          {
            double x = r.o[0] + r.d[0] * right;
            double y = r.o[1] + r.d[1] * right;
            double z = r.o[2] + r.d[2] * right;
            if constexpr (PREC == PREC_FP64) {
              rho = density64(p, x, y, z) * p.dm;
            } else {
              float fx = (float)x, fy = (float)y, fz = (float)z;
              deform32(p.defs, p.n_defs, fx, fy, fz);
              GR g = root32(p.nodes, fx, fy, fz);
              rho = g.v;
              if (g.amb || p.audit) {
                double e = density64(p, x, y, z);
                if (g.amb)
                  ++n_reeval;
                else if (e != rho)
                  ++n_miss;
                rho = e;
              }
              rho = rho * p.dm;
            }
            ++n_samples;
          }
          if ((rho == 0) != prev_zero) {
            double left = n == 0 ? p.smin : p.s_table[n - 1];
            left += dsf;
            while (left < right) {
              double x = r.o[0] + r.d[0] * left;
              double y = r.o[1] + r.d[1] * left;
              double z = r.o[2] + r.d[2] * left;
              double rf;
              if constexpr (PREC == PREC_FP64) {
                rf = density64(p, x, y, z);
              } else {
                float fx = (float)x, fy = (float)y, fz = (float)z;
                deform32(p.defs, p.n_defs, fx, fy, fz);
                GR g = root32(p.nodes, fx, fy, fz);
                rf = g.v;
                if (g.amb || p.audit) {
                  double e = density64(p, x, y, z);
                  if (g.amb)
                    ++n_reeval;
                  else if (e != rf)
                    ++n_miss;
                  rf = e;
                }
              }
              T += (rf * p.dm) * dsf;
              left += dsf;
              ++n_fine;
            }
            T += rho * dsf;
          } else {
            T += rho * DS;
          }
          prev_zero = rho == 0;
        }
      }
    }
    store_pixel(p, blockIdx.y, i, j, T);
  }

  if (p.stats) {
    unsigned long long a = wave_sum(n_samples), b = wave_sum(n_reeval),
                       c = wave_sum(n_miss), d = wave_sum(n_fine);
    if (lane == 0) {
      if (a) atomicAdd(p.stats + ST_SAMPLES, a);
      if (b) atomicAdd(p.stats + ST_REEVAL, b);
      if (c) atomicAdd(p.stats + ST_AUDIT_MISS, c);
      if (d) atomicAdd(p.stats + ST_FINE, d);
    }
  }
}

// Volume export (main.go:549-635): density() at x = i/res*2-1 (the i/N
// convention of computeVoxel, main.go:208-214), out[k*res*res + i*res + j].
__global__ __launch_bounds__(256) void volume_density_kernel(MarchParams p, int res,
                                                             double* out,
                                                             unsigned long long* maxbits) {
  const size_t n = (size_t)res * res * res;
  const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
  double v = 0.0;
  if (idx < n) {
    size_t k = idx / ((size_t)res * res), i = (idx / res) % res, j = idx % res;
    double x = (double)i / (double)res * 2.0 - 1.0;
    double y = (double)j / (double)res * 2.0 - 1.0;
    double z = (double)k / (double)res * 2.0 - 1.0;
    v = density64(p, x, y, z) * p.dm;
    out[idx] = v;
  }
  // max_val starts at 0 (main.go:619-624), so only positive values compete;
  // their IEEE bit patterns order like unsigned integers.
  unsigned long long b = v > 0.0 ? (unsigned long long)__double_as_longlong(v) : 0ull;
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    unsigned long long o = __shfl_xor(b, off, 64);
    b = o > b ? o : b;
  }
  if ((threadIdx.x & 63) == 0 && b) atomicMax(maxbits, b);
}

// byte(v / max_val * 255): Go converts through int64 (CVTTSD2SQ on amd64)
// and keeps the low byte; NaN and out-of-range values give 0.
__global__ __launch_bounds__(256) void volume_quantize_kernel(const double* dens, size_t n,
                                                              const unsigned long long* maxbits,
                                                              unsigned char* out) {
  const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= n) return;
  double mx = __longlong_as_double((long long)*maxbits);
  double q = dens[idx] / mx * 255;
  long long iq = (q == q && q > -9.2e18 && q < 9.2e18) ? (long long)q : 0;
  out[idx] = (unsigned char)(iq & 0xff);
}

int launch_volume(const MarchParams& p, int res, double* dens, unsigned long long* maxbits,
                  unsigned char* out, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  size_t n = (size_t)res * res * res;
  unsigned blocks = (unsigned)((n + 255) / 256);
  hipLaunchKernelGGL(volume_density_kernel, dim3(blocks), dim3(256), 0, s, p, res, dens,
                     maxbits);
  hipLaunchKernelGGL(volume_quantize_kernel, dim3(blocks), dim3(256), 0, s, dens, n, maxbits,
                     out);
  return (int)hipGetLastError();
}

int launch_march(const MarchParams& p, int precision, int integrator,
                 void* stream) {
  dim3 block(256);
  int tiles = (p.res + 15) / 16;
  dim3 grid(tiles * tiles, p.n_views);
  hipStream_t s = (hipStream_t)stream;
  if (precision == PREC_FP64) {
    if (integrator == INT_HIERARCHICAL)
      hipLaunchKernelGGL((march_kernel<PREC_FP64, INT_HIERARCHICAL>), grid,
                         block, 0, s, p);
    else
      hipLaunchKernelGGL((march_kernel<PREC_FP64, INT_SIMPLE>), grid, block, 0,
                         s, p);
  } else {
    if (integrator == INT_HIERARCHICAL)
      hipLaunchKernelGGL((march_kernel<PREC_GUARDED, INT_HIERARCHICAL>), grid,
                         block, 0, s, p);
    else
      hipLaunchKernelGGL((march_kernel<PREC_GUARDED, INT_SIMPLE>), grid, block,
                         0, s, p);
  }
  return (int)hipGetLastError();
}

}  // namespace xrp
