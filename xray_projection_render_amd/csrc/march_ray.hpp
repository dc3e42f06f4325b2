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
//    are the oracle's.  Samples inside an evaluated sample's safe radius
//    (eval_block.hpp) are settled without evaluation: their exact density is
//    that sample's, and their terms are added to T in order all the same.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

#include "eval_block.hpp"
#include "eval_device.hpp"
#include "march.hpp"

namespace xrp {

// Samples per node visit of the guarded FP32 simple march (the per-scene
// kernel picks 2 or 4 by scene size, the batched kernel uses 4), and of the
// hierarchical march's coarse and fine blocks.
constexpr int kBlock = 4;
constexpr int kHierBlock = 4;

// Counters one ray adds to the launch's stats.
struct RayCounts {
  unsigned long long samples = 0, reeval = 0, miss = 0, fine = 0, skipped = 0;
};

// First and last sample index of a ray's clipped segment (main.go:144-154;
// s_k = s_table[k], s_{k+1} = s_k + ds).  Empty when k_hi < k_lo.
__device__ __forceinline__ void clip_range(const MarchParams& p, double t_in, double t_out,
                                           int& k_lo, int& k_hi) {
  k_lo = 0;
  k_hi = p.K - 1;
  if (p.clip) {
    double a = floor((t_in - p.smin) / p.ds) - 2.0;
    double b = ceil((t_out - p.smin) / p.ds) + 2.0;
    k_lo = a > 0.0 ? (a < (double)p.K ? (int)a : p.K) : 0;
    k_hi = b < (double)(p.K - 1) ? (b > -1.0 ? (int)b : -1) : p.K - 1;
  }
}

// Most children a culled root collection may have (p.cull is off above).
constexpr int kMaxCull = 256;

// The wave's culled child list (Cull, eval_block.hpp) when p.cull is set:
// the root is a collection of at most kMaxCull children and the
// deformations compose to x -> M x + t (p.M, p.t), so each ray is a line in
// object space.  A bounded child (sphere, box, parallelepiped, cylinder) is
// kept when its bounding sphere, which the host inflates past rounding,
// meets the line of any lane's ray; every other child is kept.  The test is
// FP64 on the infinite line, a superset of the samples marched.  Called by
// every lane of the wave.
template <int QW>
__device__ __forceinline__ Cull wave_cull(const MarchParams& p, int v, int i, int j,
                                          bool valid) {
  __shared__ int16_t lists[QW][kMaxCull];
  Cull cl;
  if (!p.cull) return cl;
  int16_t* L = lists[QW == 1 ? 0 : __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6))];
  bool has = false;
  double o[3] = {0.0, 0.0, 0.0}, d[3] = {0.0, 0.0, 0.0};
  if (valid) {
    const Ray r = make_ray(p, p.views[p.view_offset + v], i, j);
    double t_in, t_out;
    if (clip_ray(p, r, t_in, t_out)) {
      has = true;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        o[c] = p.M[3 * c] * r.o[0] + p.M[3 * c + 1] * r.o[1] + p.M[3 * c + 2] * r.o[2] + p.t[c];
        d[c] = p.M[3 * c] * r.d[0] + p.M[3 * c + 1] * r.d[1] + p.M[3 * c + 2] * r.d[2];
      }
    }
  }
  const double dd = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  const double inv = dd > 0.0 ? 1.0 / dd : 0.0;
  const DevNode* __restrict__ N = p.nodes;
  int n = 0;
  for (int c = N[0].first; c < N[0].next; c = N[c].next) {
    const DevNode& ch = N[c];
    bool keep = has;
    if (has && (ch.kind == K_SPHERE || ch.kind == K_BOX || ch.kind == K_PPED ||
                ch.kind == K_CYL)) {
      const double w0 = (double)ch.bsph[0] - o[0], w1 = (double)ch.bsph[1] - o[1],
                   w2 = (double)ch.bsph[2] - o[2];
      const double t = (w0 * d[0] + w1 * d[1] + w2 * d[2]) * inv;
      const double q0 = w0 - t * d[0], q1 = w1 - t * d[1], q2 = w2 - t * d[2];
      const double rr = (double)ch.bsph[3] + (double)p.cull_pad;
      keep = q0 * q0 + q1 * q1 + q2 * q2 <= rr * rr;
    }
    if (__ballot(keep)) {
      if ((threadIdx.x & 63) == 0) L[n] = (int16_t)c;
      ++n;
    }
  }
  cl.idx = L;
  cl.n = n;
  return cl;
}

// Optical depth T (flat field included) of pixel (i, j) of view v (v counts
// from p.view_offset), simple integrator, FP64 evaluator.
__device__ __forceinline__ double march_T_simple64(const MarchParams& p, int v, int i, int j,
                                                   RayCounts& rc) {
  double T = p.flat_field;
  Ray r = make_ray(p, p.views[p.view_offset + v], i, j);
  double t_in, t_out;
  if (clip_ray(p, r, t_in, t_out)) {
    int k_lo, k_hi;
    clip_range(p, t_in, t_out, k_lo, k_hi);
    if (k_lo <= k_hi) {
      const double ds = p.ds;
      double s = p.s_table[k_lo];
      for (int k = k_lo; k <= k_hi; ++k) {
        double x = r.o[0] + r.d[0] * s;
        double y = r.o[1] + r.d[1] * s;
        double z = r.o[2] + r.d[2] * s;
        T += (density64(p, x, y, z) * p.dm) * ds;
        s += ds;
      }
      rc.samples += (unsigned long long)(k_hi - k_lo + 1);
    }
  }
  return T;
}

// Optical depth T of pixel (i, j) of view v, simple integrator, guarded FP32
// membership with B samples per node visit and inline FP64 re-evaluation of
// the samples inside a guard band.  Those terms are summed, in sample order,
// into Ta, the others into T, and the result is T + Ta: the same two sums,
// and so the same bits, as the queued form below.
// State of one ray in the first pass of the two-pass march.
struct TryState {
  uint32_t* ks = nullptr;  // the lane's kDeferMax sample indices (LDS)
  unsigned cnt = 0;        // guard-band samples deferred (hierarchical: listed)
  unsigned samples = 0, skipped = 0;  // the ray's counts, when it was not stopped
  unsigned fine = 0;                  // hierarchical: fine samples of its own steps
  bool aborted = false;
};

// TRY: the first pass of the two-pass march, without the FP64 evaluator
// (MarchParams::redo_list).  Guard-band samples are deferred: their indices
// go to ts->ks and the march goes on; the return value is T without their
// terms (Ta = 0), which the resolve pass adds.  A ray with more than
// p.defer_max of them stops and sets `aborted`; the counts of a ray that
// ends go to ts, not rc.  Not for audit.
template <int B, bool TRY = false>
__device__ __forceinline__ double march_T_guarded_inline(const MarchParams& p, int v, int i,
                                                         int j, RayCounts& rc, const Cull& cl,
                                                         TryState* ts = nullptr) {
  // per-ray counts fit 32 bits (a ray has fewer than 2^31 samples)
  unsigned n_samples = 0, n_reeval = 0, n_miss = 0, n_skip = 0;
  const double ds = p.ds, dm = p.dm;
  double T = p.flat_field, Ta = 0.0;
  Ray r = make_ray(p, p.views[p.view_offset + v], i, j);
  double t_in, t_out;
  if (clip_ray(p, r, t_in, t_out)) {
    int k, k_hi;
    clip_range(p, t_in, t_out, k, k_hi);
    const float ds32 = (float)ds;
    const float d0 = (float)r.d[0], d1 = (float)r.d[1], d2 = (float)r.d[2];
    // One block of B samples per iteration.  Positions are FP32 strides of
    // at most kChunk - 1 steps from an FP64 anchor at sample ka, s_ka =
    // smin + ka ds (within ~K ulp of the accumulated table value; the error
    // model allows for it).
    int ka = -kChunk;
    float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f;
    while (k <= k_hi) {
      if (k + (B - 1) - ka >= kChunk) {
        ka = k;
        const double sa = p.smin + (double)ka * ds;
        a0 = (float)(r.o[0] + r.d[0] * sa);
        a1 = (float)(r.o[1] + r.d[1] * sa);
        a2 = (float)(r.o[2] + r.d[2] * sa);
      }
      const int j0 = k - ka;
      PBlock<B> P;
#pragma unroll
      for (int q = 0; q < B; ++q) {
        float sj = (float)(j0 + q) * ds32;
        P.x[q] = fmaf(d0, sj, a0);
        P.y[q] = fmaf(d1, sj, a1);
        P.z[q] = fmaf(d2, sj, a2);
      }
      if (p.vel) {  // velocity through the chain only when the reach uses it
        P.vx = d0;
        P.vy = d1;
        P.vz = d2;
      }
      deform_blk<B>(p.defs, p.n_defs, P, p.vel != 0);
      const int left = k_hi - k + 1;
      const unsigned live = left >= B ? (1u << B) - 1u : (1u << left) - 1u;
      double vv[B];
      unsigned amb = 0;
      float dist = INFINITY;
      root_blk<B, true>(p.nodes, P, live, vv, amb, dist, cl);
      if constexpr (TRY) {
        if (amb) {
          unsigned m = amb;
          while (m) {
            const int q = __builtin_ctz(m);
            m &= m - 1u;
            if (ts->cnt < p.defer_max) ts->ks[ts->cnt] = (uint32_t)(k + q);
            ++ts->cnt;
          }
          if (ts->cnt > p.defer_max) {
            ts->aborted = true;
            return 0.0;
          }
        }
      }
      const bool last_amb = (amb >> (B - 1)) & 1u;
      // FP64 re-evaluation (under audit of every sample; the FP32-clear
      // ones must agree) through one call site
      unsigned redo = TRY ? 0u : p.audit ? live : amb;
      while (!TRY && redo) {
        const int q = __builtin_ctz(redo);
        redo &= redo - 1u;
        double vq = vv[0];
#pragma unroll
        for (int t = 1; t < B; ++t)
          if (q == t) vq = vv[t];
        const double s = p.s_table[k + q];
        const double e = density64(p, r.o[0] + r.d[0] * s, r.o[1] + r.d[1] * s,
                                   r.o[2] + r.d[2] * s, cl.idx, cl.n);
        if ((amb >> q) & 1u) {
          ++n_reeval;
          Ta += (e * dm) * ds;
        } else if (e != vq) {
          ++n_miss;
        }
      }
#pragma unroll
      for (int q = 0; q < B; ++q)
        if (((live & ~amb) >> q) & 1u) T += (vv[q] * dm) * ds;
      const int n_live = __builtin_popcount(live);
      n_samples += (unsigned)n_live;
      k += n_live;
      // Safe radius of the block's last sample: within dist of its
      // object-space position the exact density is vv[B-1], and sample
      // k+j lies at most lipschitz * (j+1) * ds from it.
      if (live == (1u << B) - 1u && !last_amb) {
        const float f = dist * p.skip_inv;
        if (f >= 1.0f && k <= k_hi) {
          const int m = min(f < 65536.0f ? (int)f : 65536, k_hi - k + 1);
          const double rho_m = vv[B - 1];
          if (!TRY && p.audit) {
            for (int jj = 0; jj < m; ++jj) {
              double s = p.s_table[k + jj];
              double e = density64(p, r.o[0] + r.d[0] * s, r.o[1] + r.d[1] * s,
                                   r.o[2] + r.d[2] * s, cl.idx, cl.n);
              if (e != rho_m) ++n_miss;
            }
          }
          // the terms the evaluated samples would add, in order (a zero
          // term leaves T as it is)
          const double c = (rho_m * dm) * ds;
          if (c != 0.0)
            for (int jj = 0; jj < m; ++jj) T += c;
          n_samples += (unsigned)m;
          n_skip += (unsigned)m;
          k += m;
        }
      }
    }
  }
  if constexpr (TRY) {
    // counted by the caller once the ray is known not to be marched again
    ts->samples = n_samples;
    ts->skipped = n_skip;
  } else {
    rc.samples += n_samples;
    rc.reeval += n_reeval;
    rc.miss += n_miss;
    rc.skipped += n_skip;
  }
  return T + Ta;
}

// The marches that keep a per-wave queue in LDS take the number of waves per
// workgroup, QW, as a template argument (the kernel's launch bound / 64):
// the LDS a workgroup reserves is QW queues.

// Per-wave queue of ambiguous samples awaiting FP64 re-evaluation.
constexpr int kAmbCap = 512;
struct AmbQueue {
  double val[kAmbCap];     // FP64 density of the sample (dm not applied)
  int32_t k[kAmbCap];      // its sample index
  int16_t link[kAmbCap];   // the owning lane's next entry, -1 for none
  uint8_t owner[kAmbCap];  // lane whose ray it belongs to
};

// Optical depth T of pixel (i, j) of view v, simple integrator, guarded FP32
// membership with B samples per node visit.  Every lane of the wave calls it
// together (lanes without a pixel pass valid = false and still help with the
// FP64 re-evaluations); T is meaningful for valid lanes.
//
// Queued form of march_T_guarded_inline, for kernels with registers to
// spare (the drain's FP64 evaluator and the queue state spill at 4 waves per
// SIMD).  A sample inside a guard band waits in a per-wave queue in LDS;
// once the queue is half full, or every lane has finished, the wave
// re-evaluates the queued samples 64 at a time, one per lane, instead of
// stalling 63 lanes behind each one.  Each lane adds those terms in sample
// order into Ta, so the result has the same bits as the inline form.  Under
// audit the re-evaluations are made inline.
template <int B, int QW>
__device__ __forceinline__ double march_T_guarded(const MarchParams& p, int v, int i, int j,
                                                  bool valid, RayCounts& rc, const Cull& cl) {
  __shared__ AmbQueue queues[QW];
  AmbQueue& Q = queues[QW == 1 ? 0 : (threadIdx.x >> 6)];
  const int lane = threadIdx.x & 63;
  // per-ray counts fit 32 bits (a ray has fewer than 2^31 samples)
  unsigned n_samples = 0, n_reeval = 0, n_miss = 0, n_skip = 0;
  const double ds = p.ds, dm = p.dm;
  double T = p.flat_field;  // terms decided in FP32 or settled, in sample order
  double Ta = 0.0;          // FP64 re-evaluated terms, in sample order
  Ray r;
  r.o[0] = r.o[1] = r.o[2] = 0.0;
  r.d[0] = r.d[1] = 0.0;
  r.d[2] = 1.0;
  int k = 0, k_hi = -1;
  if (valid) {
    r = make_ray(p, p.views[p.view_offset + v], i, j);
    double t_in, t_out;
    if (clip_ray(p, r, t_in, t_out)) clip_range(p, t_in, t_out, k, k_hi);
  }
  const float ds32 = (float)ds;
  const float d0 = (float)r.d[0], d1 = (float)r.d[1], d2 = (float)r.d[2];
  // One block of B samples per iteration.  Positions are FP32 strides of at
  // most kChunk - 1 steps from an FP64 anchor at sample ka, s_ka = smin +
  // ka ds (within ~K ulp of the accumulated table value; the error model
  // allows for it).
  int ka = -kChunk;
  float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f;
  int head = -1, tail = -1;  // this lane's queue entries, in sample order
  int queued = 0;            // entries in the wave's queue (same on every lane)
  for (;;) {
    const int kb = k;
    unsigned defer = 0;  // this block's samples to re-evaluate later
    if (k <= k_hi) {
      if (k + (B - 1) - ka >= kChunk) {
        ka = k;
        const double sa = p.smin + (double)ka * ds;
        a0 = (float)(r.o[0] + r.d[0] * sa);
        a1 = (float)(r.o[1] + r.d[1] * sa);
        a2 = (float)(r.o[2] + r.d[2] * sa);
      }
      const int j0 = k - ka;
      PBlock<B> P;
#pragma unroll
      for (int q = 0; q < B; ++q) {
        float sj = (float)(j0 + q) * ds32;
        P.x[q] = fmaf(d0, sj, a0);
        P.y[q] = fmaf(d1, sj, a1);
        P.z[q] = fmaf(d2, sj, a2);
      }
      if (p.vel) {  // velocity through the chain only when the reach uses it
        P.vx = d0;
        P.vy = d1;
        P.vz = d2;
      }
      deform_blk<B>(p.defs, p.n_defs, P, p.vel != 0);
      const int left = k_hi - k + 1;
      const unsigned live = left >= B ? (1u << B) - 1u : (1u << left) - 1u;
      double vv[B];
      unsigned amb = 0;
      float dist = INFINITY;
      root_blk<B, true>(p.nodes, P, live, vv, amb, dist, cl);
      n_reeval += (unsigned)__builtin_popcount(amb);
      if (p.audit) {
        // under audit every sample in FP64, inline (the FP32-clear ones
        // must agree)
        unsigned redo = live;
        while (redo) {
          const int q = __builtin_ctz(redo);
          redo &= redo - 1u;
          double vq = vv[0];
#pragma unroll
          for (int t = 1; t < B; ++t)
            if (q == t) vq = vv[t];
          const double s = p.s_table[k + q];
          const double e = density64(p, r.o[0] + r.d[0] * s, r.o[1] + r.d[1] * s,
                                     r.o[2] + r.d[2] * s, cl.idx, cl.n);
          if ((amb >> q) & 1u)
            Ta += (e * dm) * ds;
          else if (e != vq)
            ++n_miss;
        }
      } else {
        defer = amb;
      }
#pragma unroll
      for (int q = 0; q < B; ++q)
        if (((live & ~amb) >> q) & 1u) T += (vv[q] * dm) * ds;
      const int n_live = __builtin_popcount(live);
      n_samples += (unsigned)n_live;
      k += n_live;
      // Safe radius of the block's last sample: within dist of its
      // object-space position the exact density is vv[B-1], and sample
      // k+j lies at most lipschitz * (j+1) * ds from it.
      if (live == (1u << B) - 1u && !((amb >> (B - 1)) & 1u)) {
        const float f = dist * p.skip_inv;
        if (f >= 1.0f && k <= k_hi) {
          const int m = min(f < 65536.0f ? (int)f : 65536, k_hi - k + 1);
          const double rho_m = vv[B - 1];
          if (p.audit) {
            for (int jj = 0; jj < m; ++jj) {
              double s = p.s_table[k + jj];
              double e = density64(p, r.o[0] + r.d[0] * s, r.o[1] + r.d[1] * s,
                                   r.o[2] + r.d[2] * s, cl.idx, cl.n);
              if (e != rho_m) ++n_miss;
            }
          }
          // the terms the evaluated samples would add, in order (a zero
          // term leaves T as it is)
          const double c = (rho_m * dm) * ds;
          if (c != 0.0)
            for (int jj = 0; jj < m; ++jj) T += c;
          n_samples += (unsigned)m;
          n_skip += (unsigned)m;
          k += m;
        }
      }
    }
    // Queue this block's deferred samples in (sample, lane) order.
#pragma unroll
    for (int q = 0; q < B; ++q) {
      const bool t = (defer >> q) & 1u;
      const unsigned long long mask = __ballot(t);
      if (t) {
        const int pos = queued + (int)__popcll(mask & __lanemask_lt());
        Q.k[pos] = kb + q;
        Q.owner[pos] = (uint8_t)lane;
        Q.link[pos] = -1;
        if (tail >= 0)
          Q.link[tail] = (int16_t)pos;
        else
          head = pos;
        tail = pos;
      }
      queued += (int)__popcll(mask);
    }
    const bool more = __ballot(k <= k_hi) != 0ull;
    if (queued > kAmbCap - 64 * B || (!more && queued > 0)) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      for (int base = 0; base < queued; base += 64) {
        const int e = base + lane;
        const bool has = e < queued;
        const int src = has ? (int)Q.owner[e] : lane;
        const int kk = has ? Q.k[e] : 0;
        double o[3], d[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          o[c] = __shfl(r.o[c], src, 64);
          d[c] = __shfl(r.d[c], src, 64);
        }
        if (has) {
          const double s = p.s_table[kk];
          Q.val[e] = density64(p, o[0] + d[0] * s, o[1] + d[1] * s, o[2] + d[2] * s, cl.idx, cl.n);
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      while (head >= 0) {
        Ta += (Q.val[head] * dm) * ds;
        head = Q.link[head];
      }
      tail = -1;
      queued = 0;
      __builtin_amdgcn_wave_barrier();
    }
    if (!more) break;
  }
  rc.samples += n_samples;
  rc.reeval += n_reeval;
  rc.miss += n_miss;
  rc.skipped += n_skip;
  return T + Ta;
}

// ---------------------------------------------------------------------------
// Hierarchical integrator (main.go:159-199), wave-cooperative.
//
// Coarse samples sit at the right endpoints R[n] = s_table[n] of steps of
// DS; where (rho == 0) toggles between R[n-1] and R[n] the reference walks
// that step again at DS/10 (9 or 10 fine samples).  Done inside each lane,
// a fine walk stalls the other 63 lanes of the wave for its ten
// evaluations, and on lattices some lane toggles at most coarse samples.
// Here a toggle is appended to a per-wave queue in LDS; once the queue is
// half full, or every lane has finished its coarse samples, the whole wave
// drains it, one queued toggle per lane and kHierBlock fine samples per node
// visit.  Each lane links its own toggles and adds their fine sums, in
// coarse order, to a separate sum, so the result does not depend on when
// the queue is drained; it differs from the reference's only in the FP64
// summation order.

constexpr int kHierCap = 512;   // queued toggles per wave

// Per-wave toggle queue of CAP entries.  TRY (the first pass of the
// two-pass march) also keeps each walk's guard-band positions and counts
// for the owning lane.
template <bool TRY, int CAP = kHierCap>
struct HierQueue {
  double sum[CAP];               // fine sum of the step: sum of (rho * dm) * DS/10
  int32_t n[CAP];                // coarse index of the step's right endpoint
  int16_t link[CAP];             // the owning lane's next entry, -1 for none
  uint16_t ambm[TRY ? CAP : 1];  // bit m: the walk's m-th sample was in a guard band
  uint8_t cnt[TRY ? CAP : 1];    // the walk's fine samples | settled ones << 4
  uint8_t owner[CAP];            // lane whose ray the step belongs to
};

// Densities (dm not applied) at s[q] along the ray o + s d, for the samples
// in `live`.  Guarded: FP32 membership with guard bands over the block and
// FP64 re-evaluation of the ambiguous samples (all of them under audit),
// through one FP64 evaluator call site; amb is the mask of the ambiguous
// ones.  TRY: no FP64 evaluator; the ambiguous samples get a best guess,
// their FP32 value (root_blk with G).  With DIST (guarded only), dist is the last sample's safe radius
// (eval_block.hpp) and last_amb whether that sample was ambiguous.  A
// sample's value depends on its position alone (and not on the block, the
// cull list, whose children are further than the guard band from the ray,
// or DIST).
template <int PREC, int B, bool DIST = false, bool TRY = false>
__device__ __forceinline__ void eval_along(const MarchParams& p, const double (&o)[3],
                                           const double (&d)[3], const double (&s)[B],
                                           unsigned live, double (&rho)[B],
                                           unsigned long long& n_reeval,
                                           unsigned long long& n_miss, const Cull& cl,
                                           unsigned& amb, float* dist = nullptr,
                                           bool* last_amb = nullptr) {
  amb = 0;
  if constexpr (PREC == PREC_FP64) {
#pragma unroll
    for (int q = 0; q < B; ++q)
      rho[q] = ((live >> q) & 1u)
                   ? density64(p, o[0] + d[0] * s[q], o[1] + d[1] * s[q], o[2] + d[2] * s[q], cl.idx, cl.n)
                   : 0.0;
  } else {
    PBlock<B> P;
#pragma unroll
    for (int q = 0; q < B; ++q) {
      P.x[q] = (float)(o[0] + d[0] * s[q]);
      P.y[q] = (float)(o[1] + d[1] * s[q]);
      P.z[q] = (float)(o[2] + d[2] * s[q]);
    }
    if (DIST && p.vel) {  // velocity through the chain only when the reach uses it
      P.vx = (float)d[0];
      P.vy = (float)d[1];
      P.vz = (float)d[2];
    }
    deform_blk<B>(p.defs, p.n_defs, P, DIST && p.vel != 0);
    if constexpr (DIST) {
      root_blk<B, true, TRY>(p.nodes, P, live, rho, amb, *dist, cl);
      *last_amb = (amb >> (B - 1)) & 1u;
    } else {
      root_blk<B, TRY>(p.nodes, P, live, rho, amb, cl);
    }
    if constexpr (!TRY) {
      unsigned redo = p.audit ? live : amb;
      while (redo) {
        const int q = __builtin_ctz(redo);
        redo &= redo - 1u;
        double sq = s[0], rq = rho[0];
#pragma unroll
        for (int t = 1; t < B; ++t)
          if (q == t) {
            sq = s[t];
            rq = rho[t];
          }
        const double e = density64(p, o[0] + d[0] * sq, o[1] + d[1] * sq, o[2] + d[2] * sq, cl.idx, cl.n);
        if ((amb >> q) & 1u)
          ++n_reeval;
        else if (e != rq)
          ++n_miss;
#pragma unroll
        for (int t = 0; t < B; ++t)
          if (q == t) rho[t] = e;
      }
    }
  }
}

// The fine walk of the step whose right endpoint is coarse sample nn
// (main.go:180-188): left = R[nn-1] + DS/10, then left += DS/10 while left <
// right, each sample adding (rho * dm) * DS/10.  Returns that sum, in
// order; ambm gets the positions of the samples inside a guard band (bit m:
// the walk's m-th sample), nf the walk's samples and nsk those a safe
// radius settled.  TRY: no FP64 evaluator, those samples add their FP32
// values.  BF: fine samples per node visit; the sum does not depend on it.
template <int PREC, bool TRY = false, int BF = kHierBlock>
__device__ __forceinline__ double fine_walk(const MarchParams& p, const double (&o)[3],
                                            const double (&d)[3], int nn, const Cull& cl,
                                            unsigned& nf, unsigned& nsk, unsigned& ambm,
                                            unsigned long long& n_reeval,
                                            unsigned long long& n_miss) {
  constexpr int B = PREC == PREC_FP64 ? 1 : BF;
  constexpr bool kDist = PREC == PREC_GUARDED;
  const double dsf = p.ds_fine, dm = p.dm;
  const double right = p.s_table[nn];
  double left = (nn == 0 ? p.smin : p.s_table[nn - 1]) + dsf;
  double S = 0.0;
  nf = 0;
  nsk = 0;
  ambm = 0;
  while (left < right) {
    double s[B], rf[B];
    unsigned live = 0;
#pragma unroll
    for (int q = 0; q < B; ++q) {
      s[q] = left;
      if (left < right) {
        live |= 1u << q;
        left += dsf;
      }
    }
    float dist = INFINITY;
    bool last_amb = true;
    unsigned a = 0;
    eval_along<PREC, B, kDist, TRY>(p, o, d, s, live, rf, n_reeval, n_miss, cl, a, &dist,
                                    &last_amb);
    ambm |= a << nf;
#pragma unroll
    for (int q = 0; q < B; ++q)
      if ((live >> q) & 1u) S += (rf[q] * dm) * dsf;
    nf += (unsigned)__builtin_popcount(live);
    // Fine samples inside the last one's safe radius: same terms, same
    // left += ds walk, no evaluation.
    if constexpr (kDist) {
      float f = dist * p.skip_inv * 10.0f;  // skip_inv is per DS = 10 ds
      if (live == (1u << B) - 1u && !last_amb && f >= 1.0f) {
        const double c = (rf[B - 1] * dm) * dsf;
        for (int m = f < 64.0f ? (int)f : 64; m > 0 && left < right; --m) {
          if constexpr (!TRY) {
            if (p.audit) {
              const double e2 = density64(p, o[0] + d[0] * left, o[1] + d[1] * left,
                                          o[2] + d[2] * left, cl.idx, cl.n);
              if (e2 != rf[B - 1]) ++n_miss;
            }
          }
          S += c;
          left += dsf;
          ++nf;
          ++nsk;
        }
      }
    }
  }
  return S;
}

// Optical depth T of pixel (i, j) of view v, hierarchical integrator.  Every
// lane of the wave calls it together (lanes without a pixel pass valid =
// false and still help drain the queue); T is meaningful for valid lanes.
// The coarse terms go into T and the fine sums, in coarse order, into Tf:
// the result, T + Tf, does not depend on when the queue is drained (nor,
// therefore, on which other rays share the wave or on how many samples are
// settled).
//
// TRY: the first pass of the two-pass hierarchical march, without the FP64
// evaluator.  Guard-band samples, coarse and fine, keep their FP32 values
// (almost always the exact ones: the bands are worst-case bounds) and are
// listed in ts->ks as (nn << 4) | m: the m-th sample of the fine walk of
// step nn, m = 15 for coarse sample nn itself.  A later pass evaluates them
// in FP64; where every listed value was exact, T + Tf is the one-pass
// march's, bit for bit.  A ray with more than p.defer_max of them, or a walk
// with 15 samples or more, stops (ts->aborted).  The ray's counts go to ts,
// not rc.
//
// BC, BF: coarse and fine samples per node visit.  Every term is the exact
// density of its sample added in order, so the sums do not depend on them.
// CAP: toggle queue entries (drained once more than CAP - 64 BC wait).
template <int PREC, int QW, bool TRY = false, int BC = kHierBlock, int BF = kHierBlock,
          int CAP = kHierCap>
__device__ __forceinline__ double march_T_hier(const MarchParams& p, int v, int i, int j,
                                               bool valid, RayCounts& rc, const Cull& cl,
                                               TryState* ts = nullptr) {
  constexpr int B = PREC == PREC_FP64 ? 1 : BC;
  static_assert(CAP >= 64 * (PREC == PREC_FP64 ? 1 : BC) + 64, "queue too small");
  __shared__ HierQueue<TRY, CAP> queues[QW];
  HierQueue<TRY, CAP>& Q = queues[QW == 1 ? 0 : (threadIdx.x >> 6)];
  const int lane = threadIdx.x & 63;
  const double DS = p.ds, dsf = p.ds_fine, dm = p.dm;
  unsigned long long n_samples = 0, n_reeval = 0, n_miss = 0, n_fine = 0, n_skip = 0;
  unsigned own_fine = 0, own_skip = 0;  // TRY: fine counts of the ray's own steps
  double T = p.flat_field;
  Ray r;
  r.o[0] = r.o[1] = r.o[2] = 0.0;
  r.d[0] = r.d[1] = 0.0;
  r.d[2] = 1.0;
  int n = 0, n_hi = -1;
  if (valid) {
    r = make_ray(p, p.views[p.view_offset + v], i, j);
    double t_in, t_out;
    if (clip_ray(p, r, t_in, t_out)) {
      n_hi = p.K - 1;
      if (p.clip) {
        const double a = floor((t_in - p.smin) / DS) - 3.0;
        const double b = ceil((t_out - p.smin) / DS) + 3.0;
        n = a > 0.0 ? (int)a : 0;
        n_hi = b < (double)(p.K - 1) ? (int)b : p.K - 1;
      }
    }
  }
  bool prev_zero = true;  // main.go: prev_rho starts at 0
  int head = -1, tail = -1;  // this lane's queue entries, in coarse order
  int queued = 0;            // entries in the wave's queue (same on every lane)
  double Tf = 0.0;
  // TRY: lists a guard-band sample; a ray with too many stops
  auto list = [&](uint32_t item) {
    if (ts->cnt < p.defer_max) ts->ks[ts->cnt] = item;
    if (++ts->cnt > p.defer_max) {
      ts->aborted = true;
      n = n_hi + 1;
    }
  };
  for (;;) {
    const int nb = n;
    unsigned tog = 0;
    if (n <= n_hi) {
      double s[B], rho[B];
      unsigned live = 0;
#pragma unroll
      for (int q = 0; q < B; ++q) {
        s[q] = p.s_table[min(nb + q, n_hi)];
        if (nb + q <= n_hi) live |= 1u << q;
      }
      constexpr bool kDist = PREC == PREC_GUARDED;
      float dist = INFINITY;
      bool last_amb = true;
      unsigned amb = 0;
      eval_along<PREC, B, kDist, TRY>(p, r.o, r.d, s, live, rho, n_reeval, n_miss, cl, amb,
                                      &dist, &last_amb);
#pragma unroll
      for (int q = 0; q < B; ++q) {
        if (!((live >> q) & 1u)) break;
        const double rr = rho[q] * dm;
        const bool z = rr == 0;
        if (z != prev_zero) {
          tog |= 1u << q;
          T += rr * dsf;  // the reference adds rho * DS/10 after the fine walk
        } else {
          T += rr * DS;
        }
        prev_zero = z;
      }
      n_samples += (unsigned long long)__builtin_popcount(live);
      n = nb + B;
      if constexpr (TRY) {
        for (unsigned m = amb; m; m &= m - 1u)
          list(((uint32_t)(nb + __builtin_ctz(m)) << 4) | 15u);
      }
      // Coarse samples inside the last one's safe radius have its density:
      // no toggle, T += rho * DS each (main.go:189-191).
      if constexpr (kDist) {
        const float f = dist * p.skip_inv;
        if (live == (1u << B) - 1u && !last_amb && f >= 1.0f && n <= n_hi) {
          const int m = min(f < 65536.0f ? (int)f : 65536, n_hi - n + 1);
          const double rho_m = rho[B - 1];
          if constexpr (!TRY) {
            if (p.audit) {
              for (int j = 0; j < m; ++j) {
                const double sj = p.s_table[n + j];
                const double e = density64(p, r.o[0] + r.d[0] * sj, r.o[1] + r.d[1] * sj,
                                           r.o[2] + r.d[2] * sj, cl.idx, cl.n);
                if (e != rho_m) ++n_miss;
              }
            }
          }
          const double c = (rho_m * dm) * DS;
          if (c != 0.0)
            for (int j = 0; j < m; ++j) T += c;
          n_samples += (unsigned long long)m;
          n_skip += (unsigned long long)m;
          n += m;
        }
      }
    }
    // Append this block's toggles in (sample, lane) order.
#pragma unroll
    for (int q = 0; q < B; ++q) {
      const bool t = (tog >> q) & 1u;
      const unsigned long long m = __ballot(t);
      if (t) {
        const int pos = queued + (int)__popcll(m & __lanemask_lt());
        Q.n[pos] = nb + q;
        Q.owner[pos] = (uint8_t)lane;
        Q.link[pos] = -1;
        if (tail >= 0)
          Q.link[tail] = (int16_t)pos;
        else
          head = pos;
        tail = pos;
      }
      queued += (int)__popcll(m);
    }
    const bool more = __ballot(n <= n_hi) != 0ull;
    if (queued > CAP - 64 * B || (!more && queued > 0)) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      for (int base = 0; base < queued; base += 64) {
        const int e = base + lane;
        const bool has = e < queued;
        const int src = has ? (int)Q.owner[e] : lane;
        const int nn = has ? Q.n[e] : 0;
        double o[3], d[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          o[c] = __shfl(r.o[c], src, 64);
          d[c] = __shfl(r.d[c], src, 64);
        }
        if (has) {
          unsigned nf, nsk, ambm;
          Q.sum[e] = fine_walk<PREC, TRY, BF>(p, o, d, nn, cl, nf, nsk, ambm, n_reeval, n_miss);
          if constexpr (TRY) {
            // positions 0..14 fit the listing; 0xff (nsk < nf otherwise)
            // marks a longer walk, whose ray stops
            Q.ambm[e] = (uint16_t)ambm;
            Q.cnt[e] = nf < 15u ? (uint8_t)(nf | (nsk << 4)) : (uint8_t)0xff;
          } else {
            n_fine += nf;
            n_skip += nsk;
          }
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      while (head >= 0) {
        Tf += Q.sum[head];
        if constexpr (TRY) {
          const unsigned c = Q.cnt[head];
          if (c == 0xffu) {
            ts->aborted = true;
            n = n_hi + 1;
          } else {
            own_fine += c & 15u;
            own_skip += c >> 4;
            for (unsigned m = Q.ambm[head]; m; m &= m - 1u)
              list(((uint32_t)Q.n[head] << 4) | (uint32_t)__builtin_ctz(m));
          }
        }
        head = Q.link[head];
      }
      tail = -1;
      queued = 0;
      __builtin_amdgcn_wave_barrier();
    }
    if (!more) break;
  }
  if constexpr (TRY) {
    ts->samples = (unsigned)n_samples;
    ts->fine = own_fine;
    ts->skipped = (unsigned)n_skip + own_skip;
  } else {
    rc.samples += n_samples;
    rc.reeval += n_reeval;
    rc.miss += n_miss;
    rc.fine += n_fine;
    rc.skipped += n_skip;
  }
  return T + Tf;
}

// T of pixel (i, j) of view v.  Every lane of the wave calls it together;
// lanes with valid = false have no pixel and their return value is unused.
// B: samples per node visit of the guarded simple march; DEFER: queue its
// FP64 re-evaluations per wave (same result); QW: waves per workgroup.
template <int PREC, int INTEG, int QW, int B = kBlock, bool DEFER = false>
__device__ __forceinline__ double march_T(const MarchParams& p, int v, int i, int j, bool valid,
                                          RayCounts& rc) {
  Cull cl;
  if constexpr (PREC == PREC_GUARDED) cl = wave_cull<QW>(p, v, i, j, valid);
  if constexpr (INTEG == INT_HIERARCHICAL) {
    return march_T_hier<PREC, QW>(p, v, i, j, valid, rc, cl);
  } else if constexpr (PREC == PREC_FP64) {
    return valid ? march_T_simple64(p, v, i, j, rc) : 0.0;
  } else if constexpr (DEFER) {
    return march_T_guarded<B, QW>(p, v, i, j, valid, rc, cl);
  } else {
    return valid ? march_T_guarded_inline<B>(p, v, i, j, rc, cl) : 0.0;
  }
}

// March pixel (i, j) of view v and store its transmittance (the image is
// written at out + v * out_view_stride).  Called by every lane of the wave.
template <int PREC, int INTEG, int QW, int B = kBlock, bool DEFER = false>
__device__ __forceinline__ void march_ray(const MarchParams& p, int v, int i, int j, bool valid,
                                          RayCounts& rc) {
  const double T = march_T<PREC, INTEG, QW, B, DEFER>(p, v, i, j, valid, rc);
  if (valid) store_pixel(p, v, i, j, T);
}

// Wave-reduced counters into a launch's stats slots (lane 0 writes).
__device__ __forceinline__ void flush_counts(unsigned long long* stats, const RayCounts& rc) {
  if (!stats) return;
  unsigned long long a = wave_sum(rc.samples), b = wave_sum(rc.reeval),
                     c = wave_sum(rc.miss), d = wave_sum(rc.fine),
                     e = wave_sum(rc.skipped);
  if ((threadIdx.x & 63) == 0) {
    if (a) atomicAdd(stats + ST_SAMPLES, a);
    if (b) atomicAdd(stats + ST_REEVAL, b);
    if (c) atomicAdd(stats + ST_AUDIT_MISS, c);
    if (d) atomicAdd(stats + ST_FINE, d);
    if (e) atomicAdd(stats + ST_SKIPPED, e);
  }
}

}  // namespace xrp
