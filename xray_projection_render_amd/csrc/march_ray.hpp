// march_ray.hpp - the per-ray march shared by the per-scene kernel
// (march.hip) and the batched persistent kernel (batch.hip).
//
// Numerics (reference: main.go:137-199, objects.go, deformations.go):
//  * Ray set-up, sample positions s_k and the integral T are FP64 and
//    follow the reference's operation order (built with -ffp-contract=off;
//    FP32 multiply-adds are explicit fmaf calls).
//  * PREC_FP64 evaluates the density tree in FP64 exactly as the oracle.
//  * PREC_GUARDED decides primitive membership in FP32 with guard bands and
//    re-evaluates in FP64 any sample inside a band, so the densities summed
//    are the oracle's.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

#include "eval_block.hpp"
#include "eval_device.hpp"
#include "march.hpp"

namespace xrp {

// Samples per node visit of the guarded FP32 simple march.
constexpr int kBlock = 4;

// Counters one ray adds to the launch's stats.
struct RayCounts {
  unsigned long long samples = 0, reeval = 0, miss = 0, fine = 0;
};

// Optical depth T (flat field included) of pixel (i, j) of view v (v counts
// from p.view_offset).
template <int PREC, int INTEG>
__device__ __forceinline__ double march_T(const MarchParams& p, int v, int i, int j,
                                          RayCounts& rc) {
  const DevView& V = p.views[p.view_offset + v];
  unsigned long long n_samples = 0, n_reeval = 0, n_miss = 0, n_fine = 0;
  double T = p.flat_field;
  {
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
            // kBlock consecutive samples per node visit (eval_block.hpp);
            // T still adds the samples in order.
            for (int jb = 0; jb < n; jb += kBlock) {
              PBlock<kBlock> P;
#pragma unroll
              for (int q = 0; q < kBlock; ++q) {
                float sj = (float)(jb + q) * ds32;
                P.x[q] = fmaf(d0, sj, a0);
                P.y[q] = fmaf(d1, sj, a1);
                P.z[q] = fmaf(d2, sj, a2);
              }
              deform_blk<kBlock>(p.defs, p.n_defs, P);
              const unsigned live =
                  n - jb >= kBlock ? (1u << kBlock) - 1u : (1u << (n - jb)) - 1u;
              double v[kBlock];
              unsigned amb = 0;
              root_blk<kBlock>(p.nodes, P, live, v, amb);
#pragma unroll
              for (int q = 0; q < kBlock; ++q) {
                if (!((live >> q) & 1u)) break;
                double rho = v[q];
                const bool a = (amb >> q) & 1u;
                if (a || p.audit) {
                  double s = p.s_table[k0 + jb + q];
                  double e = density64(p, r.o[0] + r.d[0] * s, r.o[1] + r.d[1] * s,
                                       r.o[2] + r.d[2] * s);
                  if (a) {
                    ++n_reeval;
                  } else if (e != rho) {
                    ++n_miss;
                  }
                  rho = e;
                }
                T += (rho * p.dm) * ds;
              }
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
  }
  rc.samples += n_samples;
  rc.reeval += n_reeval;
  rc.miss += n_miss;
  rc.fine += n_fine;
  return T;
}

// March pixel (i, j) of view v and store its transmittance (the image is
// written at out + v * out_view_stride).
template <int PREC, int INTEG>
__device__ __forceinline__ void march_ray(const MarchParams& p, int v, int i, int j,
                                          RayCounts& rc) {
  store_pixel(p, v, i, j, march_T<PREC, INTEG>(p, v, i, j, rc));
}

// Wave-reduced counters into a launch's stats slots (lane 0 writes).
__device__ __forceinline__ void flush_counts(unsigned long long* stats, const RayCounts& rc) {
  if (!stats) return;
  unsigned long long a = wave_sum(rc.samples), b = wave_sum(rc.reeval),
                     c = wave_sum(rc.miss), d = wave_sum(rc.fine);
  if ((threadIdx.x & 63) == 0) {
    if (a) atomicAdd(stats + ST_SAMPLES, a);
    if (b) atomicAdd(stats + ST_REEVAL, b);
    if (c) atomicAdd(stats + ST_AUDIT_MISS, c);
    if (d) atomicAdd(stats + ST_FINE, d);
  }
}

}  // namespace xrp
