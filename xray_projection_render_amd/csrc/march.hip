// march.hip - gfx950 kernels for the per-pixel X-ray line integral.
//
// One lane per camera ray; a 64-lane wave covers a 64-pixel tile (one
// wave per workgroup in the per-scene kernel).  Every lane of a wave walks the same
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

#include <algorithm>
#include <cmath>

#include "eval_device.hpp"
#include "march.hpp"
#include "march_ray.hpp"

namespace xrp {

// One wave per workgroup, a 64-pixel tile: rays end at very different
// sample counts (the safe radius settles empty space in a few blocks), and a
// wave that finishes frees its slot at once instead of waiting for the
// other waves of a larger workgroup (lattice 1.86x, gyroid lattice 1.73x
// faster than 16x16 tiles of 256 lanes).
// Simple integrator: 4 waves per SIMD (<= 128 VGPRs), measured fastest on
// the lattice and CSG workloads (+62 % / +12 % over the unconstrained 2
// waves; 3 waves: -4 % / -8 %).  Large collections (4 samples per node
// visit, 24 nodes and up): the FP64 re-evaluations wait in the per-wave LDS
// queue (march_ray.hpp) and the kernel runs 3 waves per SIMD (168 VGPRs;
// at 4 the queue spills 352 B): 1.8-2.1x faster on 124-179-primitive
// scenes than inline re-evaluation at 4 waves, the same bits.
// Hierarchical: 3 waves (168 VGPRs, the toggle queue's state stays in
// registers), 7 % faster than 4 on the lattice.
// Ray tile of one wave: 32 pixels along i (the image's horizontal axis) by
// 2 along j, so a wave's samples of one index lie along a horizontal line.
// Against 8 x 8 tiles: simple lattice 39.3 -> 33.2 ms, gyroid lattice 2.57
// -> 2.25 ms, CSG (launch-bound) 0.091 -> 0.094 ms; hierarchical lattice
// 104.8 -> 87.2 ms, hierarchical CSG 0.305 -> 0.229 ms.
constexpr int march_tile_i(int) { return 32; }
constexpr int march_tile_j(int) { return 2; }

// Q: queue the FP64 re-evaluations per wave (always at B = 4).
template <int PREC, int INTEG, int B, bool Q = false>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(
    INTEG == INT_HIERARCHICAL ? 3 : (PREC == PREC_GUARDED && (B == 4 || Q) ? 3 : 4)))) void
march_kernel(MarchParams p) {
  const int lane = threadIdx.x;
  // p.tile_j32: the transposed tile, 2 along i by 32 along j (the image's
  // contiguous axis), so a wave writes two whole 128-byte row segments
  const int TI = p.tile_j32 ? 2 : march_tile_i(INTEG);
  const int TJ = p.tile_j32 ? 32 : march_tile_j(INTEG);
  const int tiles_j = (p.res + TJ - 1) / TJ;
  const int ti = blockIdx.x / tiles_j, tj = blockIdx.x - ti * tiles_j;
  const int i = ti * TI + lane / TJ;
  const int j = tj * TJ + lane % TJ;
  RayCounts rc;
  march_ray<PREC, INTEG, 1, B, PREC == PREC_GUARDED && INTEG == INT_SIMPLE && (B == 4 || Q)>(
      p, blockIdx.y, i, j, i < p.res && j < p.res, rc);
  flush_counts(p.stats, rc);
}

// End of a first-pass wave of the two-pass marches; every lane calls it.
// A ray with nothing deferred stores its pixel (SPEC: every ray not stopped
// does, its listed samples to be checked, and its counts follow its items
// in defer_k: samples, fine samples, settled samples).  The rays with
// deferred items are appended, with T and the items (ts.ks), to the
// resolve pass's lists in lane order (an exclusive scan of the counts over
// the wave, one atomic per wave and list).  The stopped ones, and a wave's deferring rays when
// the lists are full, go to p.redo_list.  The counts of a ray go to rc only
// when it is not marched again.
template <bool SPEC = false>
__device__ __forceinline__ void try_finish(const MarchParams& p, int v, int i, int j,
                                           bool valid, double T, TryState& ts,
                                           RayCounts& rc) {
  const int lane = threadIdx.x & 63;
  const unsigned id = ((unsigned)v * (unsigned)p.res + (unsigned)i) * (unsigned)p.res +
                      (unsigned)j;
  if (valid && !ts.aborted && (SPEC || ts.cnt == 0)) store_pixel(p, v, i, j, T);
  bool defer = valid && !ts.aborted && ts.cnt > 0;
  const unsigned kc = ts.cnt + (SPEC ? 3u : 0u);  // words in defer_k
  const unsigned long long dm = __ballot(defer);
  if (dm) {
    unsigned incl = defer ? kc : 0u;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const unsigned t = (unsigned)__shfl_up((int)incl, off, 64);
      if (lane >= off) incl += t;
    }
    const unsigned total = (unsigned)__shfl((int)incl, 63, 64);
    unsigned eb = 0, kb = 0;
    if (lane == 0) {
      eb = atomicAdd(p.counts + 1, (unsigned)__popcll(dm));
      kb = atomicAdd(p.counts + 2, total);
    }
    eb = (unsigned)__shfl((int)eb, 0, 64);
    kb = (unsigned)__shfl((int)kb, 0, 64);
    const bool fits = eb + (unsigned)__popcll(dm) <= p.defer_cap && kb + total <= p.k_cap;
    if (defer && fits) {
      const unsigned e = eb + (unsigned)__popcll(dm & __lanemask_lt());
      const unsigned k0 = kb + incl - kc;
      p.defer_T[e] = T;
      p.defer_ray[e] = id;
      p.defer_cnt[e] = ts.cnt;
      p.defer_kfirst[e] = k0;
      for (unsigned c = 0; c < ts.cnt; ++c) p.defer_k[k0 + c] = ts.ks[c];
      if (SPEC) {
        p.defer_k[k0 + ts.cnt] = ts.samples;
        p.defer_k[k0 + ts.cnt + 1] = ts.fine;
        p.defer_k[k0 + ts.cnt + 2] = ts.skipped;
      }
    } else if (defer) {
      // the lists are full: the wave's rays go to the redo list, and its
      // entries below the cap are marked empty for the resolve pass
      const unsigned e = eb + (unsigned)__popcll(dm & __lanemask_lt());
      if (e < p.defer_cap) p.defer_cnt[e] = 0u;
      ts.aborted = true;
    }
  }
  if (valid && !ts.aborted) {
    rc.samples += ts.samples;
    rc.skipped += ts.skipped;
    rc.fine += ts.fine;
  }
  const unsigned long long m = __ballot(valid && ts.aborted);
  if (m) {
    unsigned base = 0;
    if (lane == 0) base = atomicAdd(p.counts, (unsigned)__popcll(m));
    base = (unsigned)__shfl((int)base, 0, 64);
    const unsigned e = base + (unsigned)__popcll(m & __lanemask_lt());
    if (valid && ts.aborted) {
      if (e < p.redo_cap)
        p.redo_list[e] = id;
      else
        p.counts[3] = 1u;  // no room: the caller renders again in one pass
    }
  }
  flush_counts(p.stats, rc);
}

// Two-pass guarded simple march, first pass: every ray marched without the
// FP64 evaluator (march_T_guarded_inline<B, true>), one launch over all
// views like the one-pass kernel.  A ray with no guard-band sample stores
// its pixel; one with a few is appended, with its T and the deferred
// sample indices, to the resolve pass's lists; one with more, or finding
// the lists full, to p.redo_list for march_redo_kernel.  Without the FP64
// path in the loop no lane waits on another's FP64 re-evaluation.
// 6 waves per SIMD at one sample per visit (80 VGPRs, some spilled; 7 and 8
// waves spill more and ran 21 % and 18 % slower), 5 at two: on C5 5 -> 6
// waves took 57.5 to 54.5 ms, on the lattice 4 -> 5 took 39.3 to 35.4 ms.
template <int B>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(B == 1 ? 6 : 5))) void
march_try_kernel(MarchParams p) {
  __shared__ uint32_t kbuf[64 * kDeferMax];
  const int lane = threadIdx.x;
  const int TI = p.tile_j32 ? 2 : march_tile_i(INT_SIMPLE);
  const int TJ = p.tile_j32 ? 32 : march_tile_j(INT_SIMPLE);
  const int tiles_j = (p.res + TJ - 1) / TJ;
  const int ti = blockIdx.x / tiles_j, tj = blockIdx.x - ti * tiles_j;
  const int i = ti * TI + lane / TJ;
  const int j = tj * TJ + lane % TJ;
  const int v = blockIdx.y;
  const bool valid = i < p.res && j < p.res;
  RayCounts rc;
  const Cull cl = wave_cull<1>(p, v, i, j, valid);
  TryState ts;
  ts.ks = kbuf + lane * kDeferMax;
  const double T = valid ? march_T_guarded_inline<B, true>(p, v, i, j, rc, cl, &ts) : 0.0;
  try_finish(p, v, i, j, valid, T, ts, rc);
}

// Two-pass hierarchical march, first pass (march_T_hier<PREC_GUARDED, 1,
// true>): no FP64 evaluator in the kernel.  Guard-band samples keep their
// FP32 values; every ray not stopped stores its pixel, and the rays with
// such samples (at most p.defer_max <= kHierDeferMax) list them for
// march_hier_check_kernel.  Stopped rays go to march_hier_redo_kernel.
// At 2 coarse samples per visit 4 waves per SIMD, with a 320-entry toggle
// queue so that 16 workgroups fit a CU's LDS: lattice 84.0 -> 80.2 ms.  At
// 1, where the extra wave spills in the march loop, 3 waves and 512
// entries (C5: 176.5 against 191.6 ms at 4 waves).
template <int BC, int BF>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(BC == 2 ? 4 : 3))) void
march_hier_try_kernel(MarchParams p) {
  __shared__ uint32_t kbuf[64 * kHierDeferMax];
  const int lane = threadIdx.x;
  const int TI = p.tile_j32 ? 2 : march_tile_i(INT_HIERARCHICAL);
  const int TJ = p.tile_j32 ? 32 : march_tile_j(INT_HIERARCHICAL);
  const int tiles_j = (p.res + TJ - 1) / TJ;
  const int ti = blockIdx.x / tiles_j, tj = blockIdx.x - ti * tiles_j;
  const int i = ti * TI + lane / TJ;
  const int j = tj * TJ + lane % TJ;
  const int v = blockIdx.y;
  const bool valid = i < p.res && j < p.res;
  RayCounts rc;
  const Cull cl = wave_cull<1>(p, v, i, j, valid);
  TryState ts;
  ts.ks = kbuf + lane * kHierDeferMax;
  const double T = march_T_hier<PREC_GUARDED, 1, true, BC, BF, BC == 2 ? 320 : kHierCap>(
      p, v, i, j, valid, rc, cl, &ts);
  try_finish<true>(p, v, i, j, valid, T, ts, rc);
}

// Check pass of the two-pass hierarchical march: 64 entries per wave
// (grid-stride), their listed samples spread one per lane.  Each listed
// sample, (nn << 4) | m, is at R[nn] for m = 15 and otherwise at
// R[nn-1] + DS/10 plus m more additions of DS/10 (the
// m-th sample of step nn's fine walk, main.go:180-186).  Its FP32 value,
// which the first pass used, is computed again (it depends on the position
// alone) and compared with its FP64 density; a ray with any that differs
// goes to the redo list, whose pass overwrites its pixel, and its first
// pass counts (after its items) are taken back from the stats.  Each wave
// culls the root collection's children to its own 64 rays first.
// One listed sample per lane rather than one ray per lane: on C5 18.6 ->
// 11.3 ms (a wave ran as many rounds as its longest list, 1-14 samples).
// 4 waves per SIMD (128 VGPRs, the FP64 evaluator's spilled): 11.2 ms, 11.4
// at 3, 13.9 at 5 (one ray per lane: 24.5 ms at 2 waves, 18.5 at 4, 22.2
// at 5).
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) void
march_hier_check_kernel(MarchParams p) {
  const unsigned n = min(p.counts[1], p.defer_cap), nk = min(p.counts[2], p.k_cap);
  const unsigned rr = (unsigned)p.res * (unsigned)p.res;
  const double dsf = p.ds_fine;
  unsigned long long reeval = 0, rays = 0, back_s = 0, back_f = 0, back_k = 0;
  for (unsigned base = blockIdx.x * 64u; base < n; base += gridDim.x * 64u) {
    const unsigned e = base + threadIdx.x;
    unsigned cnt = 0u, k0 = 0u, id = 0u;
    if (e < n) {
      cnt = p.defer_cnt[e];
      k0 = p.defer_kfirst[e];
      if (k0 + cnt + 3u > nk) cnt = 0u;
      // an entry marked empty (cnt 0, lists full) has no ray or k0 written
      if (cnt != 0u)
        id = p.defer_ray[e];
      else
        k0 = 0u;
    }
    const unsigned v = id / rr, pix = id - v * rr;
    const int i = (int)(pix / (unsigned)p.res), j = (int)(pix - (unsigned)i * p.res);
    // this wave's own cull list over its rays (a child culled here is clear
    // of every one of them in FP32 and 0 in FP64, as in the first pass)
    const Cull cl = wave_cull<1>(p, (int)v, i, j, cnt != 0u);
    const Ray r = make_ray(p, p.views[p.view_offset + (int)v], i, j);
    // The wave's listed samples, one per lane: lane l of round fb takes
    // sample t = fb + l of the concatenated lists, owned by the ray (lane o)
    // whose inclusive prefix sum of counts is the first above t.
    unsigned incl = cnt;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const unsigned u = __shfl_up(incl, off, 64);
      if ((int)threadIdx.x >= off) incl += u;
    }
    const unsigned total = __builtin_amdgcn_readfirstlane(__shfl(incl, 63, 64));
    unsigned long long bad = 0ull;  // lanes whose ray holds a wrong guess
    for (unsigned fb = 0; fb < total; fb += 64u) {
      const unsigned t = fb + threadIdx.x;
      int o = 0;
#pragma unroll
      for (int step = 32; step >= 1; step >>= 1)
        if (__shfl(incl, o + step - 1, 64) <= t) o += step;
      const bool has = t < total;
      const unsigned c = t - (__shfl(incl, o, 64) - __shfl(cnt, o, 64));
      const unsigned ko = __shfl(k0, o, 64);
      double ro[3], rd[3];
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        ro[q] = __shfl(r.o[q], o, 64);
        rd[q] = __shfl(r.d[q], o, 64);
      }
      bool wrong = false;
      if (has) {
        const unsigned item = p.defer_k[ko + c], nn = item >> 4, m = item & 15u;
        double sq[1];
        if (m == 15u) {
          sq[0] = p.s_table[nn];
        } else {
          sq[0] = (nn == 0 ? p.smin : p.s_table[nn - 1]) + dsf;
          for (unsigned q = 0; q < m; ++q) sq[0] += dsf;
        }
        double f32[1];
        unsigned amb;
        unsigned long long unused = 0;
        eval_along<PREC_GUARDED, 1, false, true>(p, ro, rd, sq, 1u, f32, unused, unused, cl,
                                                 amb);
        const double ev = density64(p, ro[0] + rd[0] * sq[0], ro[1] + rd[1] * sq[0],
                                    ro[2] + rd[2] * sq[0], cl.idx, cl.n);
        wrong = ev != f32[0];
      }
      for (unsigned long long w = __ballot(wrong); w; w &= w - 1ull) {
        const int l = (int)__builtin_ctzll(w);
        bad |= 1ull << __builtin_amdgcn_readlane(o, l);
      }
    }
    if (cnt == 0u) continue;
    const bool same = ((bad >> threadIdx.x) & 1ull) == 0ull;
    if (same) {
      reeval += cnt;
      ++rays;
    } else {
      back_s += p.defer_k[k0 + cnt];
      back_f += p.defer_k[k0 + cnt + 1];
      back_k += p.defer_k[k0 + cnt + 2];
      const unsigned at = atomicAdd(p.counts, 1u);
      if (at < p.redo_cap)
        p.redo_list[at] = id;
      else
        p.counts[3] = 1u;  // no room: the caller renders again in one pass
    }
  }
  if (p.stats) {
    const unsigned long long a = wave_sum(reeval), b = wave_sum(rays), c = wave_sum(back_s),
                             d = wave_sum(back_f), f = wave_sum(back_k);
    if (threadIdx.x == 0) {
      if (a) atomicAdd(p.stats + ST_REEVAL, a);
      if (b) atomicAdd(p.stats + ST_DEFERRED, b);
      // the counts of rays marched again, taken back (mod 2^64)
      if (c) atomicAdd(p.stats + ST_SAMPLES, 0ull - c);
      if (d) atomicAdd(p.stats + ST_FINE, 0ull - d);
      if (f) atomicAdd(p.stats + ST_SKIPPED, 0ull - f);
    }
  }
}

// Second pass of the two-pass hierarchical march: the listed rays (stopped
// in the first pass, or with a listed sample whose FP32 value was not
// exact), 64 per wave in list order (grid-stride), marched again by the
// one-pass march.
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void
march_hier_redo_kernel(MarchParams p) {
  const unsigned n = min(p.counts[0], p.redo_cap);
  const unsigned rr = (unsigned)p.res * (unsigned)p.res;
  RayCounts rc;
  for (unsigned base = blockIdx.x * 64u; base < n; base += gridDim.x * 64u) {
    const unsigned e = base + threadIdx.x;
    const bool valid = e < n;
    const unsigned id = valid ? p.redo_list[e] : 0u;
    const unsigned v = id / rr, pix = id - v * rr;
    const int i = (int)(pix / (unsigned)p.res), j = (int)(pix - (unsigned)i * p.res);
    march_ray<PREC_GUARDED, INT_HIERARCHICAL, 1>(p, (int)v, i, j, valid, rc);
  }
  flush_counts(p.stats, rc);
}

// Resolve pass: 64 entries of the first pass's lists per wave (grid-stride
// over min(counts[1], defer_cap) entries), their deferred samples spread
// one per lane.  Each is evaluated in FP64, its term (rho * dm) * ds staged
// in LDS, and each entry's lane adds its own terms in sample order into Ta,
// as the one-pass march sums its re-evaluated terms; the pixel is T + Ta.
// The FP64 walks of a root collection visit the wave's own cull list.
// A wave whose entries did not all fit wrote none of them but their count
// 0 (its rays are in the redo list), and entries past the cap are not read.
// 3 waves per SIMD: on C5 1.57 ms, 1.65 ms at 4, 2.80 ms at 6 (one entry
// per lane, lists of 1-24 samples: 2.37 ms at 3 waves, 2.21 ms at 4).
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void
march_resolve_kernel(MarchParams p) {
  __shared__ double terms[64];
  const unsigned n = min(p.counts[1], p.defer_cap), nk = min(p.counts[2], p.k_cap);
  const unsigned rr = (unsigned)p.res * (unsigned)p.res;
  const double ds = p.ds, dm = p.dm;
  unsigned long long reeval = 0, rays = 0;
  for (unsigned base = blockIdx.x * 64u; base < n; base += gridDim.x * 64u) {
    const unsigned e = base + threadIdx.x;
    unsigned cnt = 0u, k0 = 0u, id = 0u;
    if (e < n) {
      cnt = p.defer_cnt[e];
      k0 = p.defer_kfirst[e];
      if (k0 + cnt > nk) cnt = 0u;
      // an entry marked empty (cnt 0, lists full) has no ray or k0 written
      if (cnt != 0u)
        id = p.defer_ray[e];
      else
        k0 = 0u;
    }
    const unsigned v = id / rr, pix = id - v * rr;
    const int i = (int)(pix / (unsigned)p.res), j = (int)(pix - (unsigned)i * p.res);
    // this wave's own cull list over its rays (a culled child is 0 along them)
    const Cull cl = wave_cull<1>(p, (int)v, i, j, cnt != 0u);
    const Ray r = make_ray(p, p.views[p.view_offset + (int)v], i, j);
    // sample t of the concatenated lists belongs to the lane o whose
    // inclusive prefix sum of counts is the first above t
    unsigned incl = cnt;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const unsigned u = __shfl_up(incl, off, 64);
      if ((int)threadIdx.x >= off) incl += u;
    }
    const unsigned excl = incl - cnt;
    const unsigned total = __builtin_amdgcn_readfirstlane(__shfl(incl, 63, 64));
    double Ta = 0.0;
    for (unsigned fb = 0; fb < total; fb += 64u) {
      const unsigned t = fb + threadIdx.x;
      int o = 0;
#pragma unroll
      for (int step = 32; step >= 1; step >>= 1)
        if (__shfl(incl, o + step - 1, 64) <= t) o += step;
      const unsigned c = t - __shfl(excl, o, 64);
      const unsigned ko = __shfl(k0, o, 64);
      double ro[3], rd[3];
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        ro[q] = __shfl(r.o[q], o, 64);
        rd[q] = __shfl(r.d[q], o, 64);
      }
      double term = 0.0;
      if (t < total) {
        const double sk = p.s_table[p.defer_k[ko + c]];
        const double ev = density64(p, ro[0] + rd[0] * sk, ro[1] + rd[1] * sk,
                                    ro[2] + rd[2] * sk, cl.idx, cl.n);
        term = (ev * dm) * ds;
      }
      terms[threadIdx.x] = term;
      __syncthreads();
      const unsigned lo = excl > fb ? excl : fb, hi = incl < fb + 64u ? incl : fb + 64u;
      for (unsigned x = lo; x < hi; ++x) Ta += terms[x - fb];
      __syncthreads();
    }
    if (cnt != 0u) {
      store_pixel(p, (int)v, i, j, p.defer_T[e] + Ta);
      reeval += cnt;
      ++rays;
    }
  }
  if (p.stats) {
    const unsigned long long a = wave_sum(reeval), b = wave_sum(rays);
    if (threadIdx.x == 0) {
      if (a) atomicAdd(p.stats + ST_REEVAL, a);
      if (b) atomicAdd(p.stats + ST_DEFERRED, b);
    }
  }
}

// Second pass: the rays the first pass listed, 64 per wave in list order
// (rays of one first-pass wave stay together), marched again.  Every lane
// here has FP64 re-evaluations, so they wait in the per-wave LDS queue and
// are drained 64 at a time (march_T_guarded, 3 waves per SIMD) rather than
// one lane at a time: the queued form gives each pixel exactly the value
// the one-pass march computes.  A grid-stride loop over the listed rays.
template <int B>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void
march_redo_kernel(MarchParams p) {
  const unsigned n = min(p.counts[0], p.redo_cap);
  const unsigned rr = (unsigned)p.res * (unsigned)p.res;
  RayCounts rc;
  for (unsigned base = blockIdx.x * 64u; base < n; base += gridDim.x * 64u) {
    const unsigned e = base + threadIdx.x;
    const bool valid = e < n;
    const unsigned id = valid ? p.redo_list[e] : 0u;
    const unsigned v = id / rr, pix = id - v * rr;
    const int i = (int)(pix / (unsigned)p.res), j = (int)(pix - (unsigned)i * p.res);
    march_ray<PREC_GUARDED, INT_SIMPLE, 1, B, true>(p, (int)v, i, j, valid, rc);
  }
  flush_counts(p.stats, rc);
}

// Exact transmittance of listed pixels: item k = (view, i, j) -> vals[k].
template <int PREC, int INTEG>
__global__ __launch_bounds__(64) void march_pixels_kernel(MarchParams p, const int* items,
                                                          int n, float* vals) {
  const int k = blockIdx.x * 64 + threadIdx.x;
  RayCounts rc;
  const bool valid = k < n;
  int v = 0, i = 0, j = 0;
  if (valid) {
    v = items[3 * k];
    i = items[3 * k + 1];
    j = items[3 * k + 2];
  }
  const double T = march_T<PREC, INTEG, 1>(p, v, i, j, valid, rc);
  if (valid) vals[k] = (float)exp(-T);
  flush_counts(p.stats, rc);
}

// Exact re-render of listed 16x16 tiles: item k = (view, ti, tj).
template <int PREC, int INTEG>
__global__ __launch_bounds__(256) void march_tiles_kernel(MarchParams p, const int* items) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int* it = items + 3 * blockIdx.x;
  const int i = it[1] * 16 + (wave >> 1) * 8 + (lane >> 3);
  const int j = it[2] * 16 + (wave & 1) * 8 + (lane & 7);
  RayCounts rc;
  march_ray<PREC, INTEG, 4>(p, it[0], i, j, i < p.res && j < p.res, rc);
  flush_counts(p.stats, rc);
}

int launch_march_list(const MarchParams& p, int precision, int integrator, const int* items,
                      int n, float* vals, void* stream) {
  if (n <= 0) return 0;
  hipStream_t s = (hipStream_t)stream;
  const bool f64 = precision == PREC_FP64, hier = integrator == INT_HIERARCHICAL;
  if (vals) {
    dim3 grid((n + 63) / 64), block(64);
    if (f64 && hier)
      hipLaunchKernelGGL((march_pixels_kernel<PREC_FP64, INT_HIERARCHICAL>), grid, block, 0, s,
                         p, items, n, vals);
    else if (f64)
      hipLaunchKernelGGL((march_pixels_kernel<PREC_FP64, INT_SIMPLE>), grid, block, 0, s, p,
                         items, n, vals);
    else if (hier)
      hipLaunchKernelGGL((march_pixels_kernel<PREC_GUARDED, INT_HIERARCHICAL>), grid, block, 0,
                         s, p, items, n, vals);
    else
      hipLaunchKernelGGL((march_pixels_kernel<PREC_GUARDED, INT_SIMPLE>), grid, block, 0, s, p,
                         items, n, vals);
  } else {
    dim3 grid(n), block(256);
    if (f64 && hier)
      hipLaunchKernelGGL((march_tiles_kernel<PREC_FP64, INT_HIERARCHICAL>), grid, block, 0, s, p,
                         items);
    else if (f64)
      hipLaunchKernelGGL((march_tiles_kernel<PREC_FP64, INT_SIMPLE>), grid, block, 0, s, p,
                         items);
    else if (hier)
      hipLaunchKernelGGL((march_tiles_kernel<PREC_GUARDED, INT_HIERARCHICAL>), grid, block, 0,
                         s, p, items);
    else
      hipLaunchKernelGGL((march_tiles_kernel<PREC_GUARDED, INT_SIMPLE>), grid, block, 0, s, p,
                         items);
  }
  return (int)hipGetLastError();
}

// Volume export (main.go:549-635): density() at x = i/res*2-1 (the i/N
// convention of computeVoxel, main.go:208-214), out[k*res*res + i*res + j].
__global__ __launch_bounds__(256) void volume_density_kernel(MarchParams p, int res,
                                                             double* out,
                                                             unsigned long long* maxbits) {
  // blockIdx.y is the slice k; blockIdx.x covers its res*res voxels
  const unsigned slice = (unsigned)res * (unsigned)res;
  const unsigned ij = blockIdx.x * 256u + threadIdx.x;
  const size_t k = blockIdx.y;
  double v = 0.0;
  if (ij < slice) {
    const size_t idx = k * slice + ij;
    const unsigned i = ij / (unsigned)res, j = ij % (unsigned)res;
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
__global__ __launch_bounds__(256) void volume_quantize_kernel(const double* dens,
                                                              unsigned slice,
                                                              const unsigned long long* maxbits,
                                                              unsigned char* out) {
  const unsigned ij = blockIdx.x * 256u + threadIdx.x;
  if (ij >= slice) return;
  const size_t idx = (size_t)blockIdx.y * slice + ij;
  double mx = __longlong_as_double((long long)*maxbits);
  double q = dens[idx] / mx * 255;
  long long iq = (q == q && q > -9.2e18 && q < 9.2e18) ? (long long)q : 0;
  out[idx] = (unsigned char)(iq & 0xff);
}

// Largest |v_sin_f32(t) - sin(2 pi t)| and |v_cos_f32(t) - cos(2 pi t)| over
// every FP32 t in [-0.5, 0.5] (both signs of each magnitude up to 0.5), the
// FP64 functions as reference.  Each lane walks bit patterns lo + k * stride
// up to hi; the maxima (non-negative floats, ordered like their bits) are
// combined with atomicMax into out[0..1].
__global__ __launch_bounds__(256) void sincos_error_kernel(unsigned hi, unsigned* out) {
  const unsigned stride = gridDim.x * blockDim.x;
  float es = 0.0f, ec = 0.0f;
  for (unsigned b = blockIdx.x * blockDim.x + threadIdx.x; b <= hi; b += stride) {
#pragma unroll
    for (int sgn = 0; sgn < 2; ++sgn) {
      const float t = __uint_as_float(b | (sgn ? 0x80000000u : 0u));
      const double a = 6.283185307179586476925286766559 * (double)t;
      es = fmaxf(es, (float)fabs((double)__builtin_amdgcn_sinf(t) - sin(a)));
      ec = fmaxf(ec, (float)fabs((double)__builtin_amdgcn_cosf(t) - cos(a)));
    }
    if (b > hi - stride) break;  // the next b would pass hi (or wrap)
  }
  atomicMax(out + 0, __float_as_uint(es));
  atomicMax(out + 1, __float_as_uint(ec));
}

int launch_sincos_error(unsigned* out, void* stream) {
  hipLaunchKernelGGL(sincos_error_kernel, dim3(8192), dim3(256), 0, (hipStream_t)stream,
                     0x3F000000u /* 0.5f */, out);
  return (int)hipGetLastError();
}

// Largest absolute error against FP64 of tanh_hw (which = 0), sigmoid_hw
// (1) or expneg_hw (2) over every FP32 argument of magnitude bits <= hi
// (both signs for tanh and the logistic function, x >= 0 for e^-x), as
// the bits of a non-negative double, maxed into *out.
__global__ __launch_bounds__(256) void deform_error_kernel(int which, unsigned hi,
                                                           unsigned long long* out) {
  const unsigned stride = gridDim.x * blockDim.x;
  double e = 0.0;
  for (unsigned b = blockIdx.x * blockDim.x + threadIdx.x; b <= hi; b += stride) {
#pragma unroll
    for (int sgn = 0; sgn < 2; ++sgn) {
      if (which == 2 && sgn) continue;
      const float t = __uint_as_float(b | (sgn ? 0x80000000u : 0u));
      double d;
      if (which == 0)
        d = (double)tanh_hw(t) - tanh((double)t);
      else if (which == 1)
        d = (double)sigmoid_hw(t) - 1.0 / (1.0 + exp(-(double)t));
      else
        d = (double)expneg_hw(t) - exp(-(double)t);
      e = fmax(e, fabs(d));
    }
    if (b > hi - stride) break;  // the next b would pass hi (or wrap)
  }
  atomicMax(out, (unsigned long long)__double_as_longlong(e));
}

int launch_deform_error(unsigned long long* out, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(deform_error_kernel, dim3(8192), dim3(256), 0, s, 0,
                     0x41A00000u /* 20.0f */, out);
  hipLaunchKernelGGL(deform_error_kernel, dim3(8192), dim3(256), 0, s, 1,
                     0x42C80000u /* 100.0f */, out + 1);
  hipLaunchKernelGGL(deform_error_kernel, dim3(8192), dim3(256), 0, s, 2,
                     0x42DC0000u /* 110.0f */, out + 2);
  return (int)hipGetLastError();
}

int launch_volume(const MarchParams& p, int res, double* dens, unsigned long long* maxbits,
                  unsigned char* out, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  // one grid row per z slice: res <= 2048 keeps both grid dimensions in
  // range (x: res^2/256 <= 16384 blocks, y: res <= 65535)
  const unsigned slice = (unsigned)res * (unsigned)res;
  const dim3 grid((slice + 255) / 256, (unsigned)res);
  hipLaunchKernelGGL(volume_density_kernel, grid, dim3(256), 0, s, p, res, dens, maxbits);
  hipLaunchKernelGGL(volume_quantize_kernel, grid, dim3(256), 0, s, dens, slice, maxbits,
                     out);
  return (int)hipGetLastError();
}

int launch_march(const MarchParams& p, int precision, int integrator,
                 void* stream) {
  dim3 block(64);
  const int ti = p.tile_j32 ? 2 : march_tile_i(integrator);
  const int tj = p.tile_j32 ? 32 : march_tile_j(integrator);
  const int tiles = ((p.res + ti - 1) / ti) * ((p.res + tj - 1) / tj);
  dim3 grid(tiles, p.n_views);
  hipStream_t s = (hipStream_t)stream;
  if (precision == PREC_FP64) {
    if (integrator == INT_HIERARCHICAL)
      hipLaunchKernelGGL((march_kernel<PREC_FP64, INT_HIERARCHICAL, kBlock>), grid,
                         block, 0, s, p);
    else
      hipLaunchKernelGGL((march_kernel<PREC_FP64, INT_SIMPLE, kBlock>), grid, block, 0,
                         s, p);
  } else {
    if (integrator == INT_HIERARCHICAL && p.redo_list && !p.audit) {
      const dim3 grid3(std::min((p.defer_cap + 63u) / 64u, 8192u));
      // 16384 one-wave groups for the redo pass's grid-stride: about one
      // batch of 64 rays each on C5 (891 k rays; 15.7 ms, 16.6 ms at 3072,
      // 24.5 ms at 1536), since a batch of such rays can take milliseconds
      const dim3 grid2(std::min((p.redo_cap + 63u) / 64u, 16384u));
      // coarse samples per node visit as the simple march's block (1 where
      // gyroids reach far, else 2), 2 per fine visit
      if (p.block == 1)
        hipLaunchKernelGGL((march_hier_try_kernel<1, 2>), grid, block, 0, s, p);
      else
        hipLaunchKernelGGL((march_hier_try_kernel<2, 2>), grid, block, 0, s, p);
      hipLaunchKernelGGL(march_hier_check_kernel, grid3, block, 0, s, p);
      hipLaunchKernelGGL(march_hier_redo_kernel, grid2, block, 0, s, p);
    } else if (integrator == INT_HIERARCHICAL)
      hipLaunchKernelGGL((march_kernel<PREC_GUARDED, INT_HIERARCHICAL, kBlock>), grid,
                         block, 0, s, p);
    else if (p.redo_list && !p.audit && !p.queue && (p.block == 1 || p.block == 2)) {
      // two passes and the resolve pass between them; the grids of the last
      // two stride over however many entries the first appended
      const dim3 grid3(std::min((p.defer_cap + 63u) / 64u, 8192u));
      const dim3 grid2(std::min((p.redo_cap + 63u) / 64u, 1024u));
      if (p.block == 1) {
        hipLaunchKernelGGL(march_try_kernel<1>, grid, block, 0, s, p);
        hipLaunchKernelGGL(march_resolve_kernel, grid3, block, 0, s, p);
        hipLaunchKernelGGL(march_redo_kernel<1>, grid2, block, 0, s, p);
      } else {
        hipLaunchKernelGGL(march_try_kernel<2>, grid, block, 0, s, p);
        hipLaunchKernelGGL(march_resolve_kernel, grid3, block, 0, s, p);
        hipLaunchKernelGGL(march_redo_kernel<2>, grid2, block, 0, s, p);
      }
    } else if (p.block == 2 && p.queue)
      hipLaunchKernelGGL((march_kernel<PREC_GUARDED, INT_SIMPLE, 2, true>), grid, block,
                         0, s, p);
    else if (p.block == 2)
      hipLaunchKernelGGL((march_kernel<PREC_GUARDED, INT_SIMPLE, 2>), grid, block,
                         0, s, p);
    else if (p.block == 1)
      hipLaunchKernelGGL((march_kernel<PREC_GUARDED, INT_SIMPLE, 1>), grid, block,
                         0, s, p);
    else
      hipLaunchKernelGGL((march_kernel<PREC_GUARDED, INT_SIMPLE, kBlock>), grid, block,
                         0, s, p);
  }
  return (int)hipGetLastError();
}

}  // namespace xrp
