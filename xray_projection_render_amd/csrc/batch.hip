// batch.hip - many scenes and views of uneven size in one persistent launch.
//
// The host flattens a batch into work items (scene, view, tile), grouped by
// (precision, integrator) and ordered within a group by estimated cost per
// tile, most expensive first.  Each wave of the persistent grid takes the
// next item from an atomic counter, marches its pixel tile (wave_rows x
// wave_cols blocks of 8x8 rays, one block after another) with the same
// per-ray code as the per-scene kernel (march_ray.hpp), and stamps the
// scene's first-start and last-end wall-clock times for per-scene latency.
// Every wave leaves the loop once the counter passes the item count, so the
// grid always drains.
#include <hip/hip_runtime.h>

#include "march.hpp"
#include "march_ray.hpp"

namespace xrp {

// Atomics issued by every lane of a wave: control flow stays wave-uniform
// inside the persistent loop (no lane-conditional blocks next to the loop's
// back edge, where a divergent branch could split one wave's lanes across
// loop iterations; a lane-0-only form faulted for that reason).  Only the
// first active lane targets the real address; every other lane hits a line
// of its own in the sink, so an item costs one same-address operation per
// counter instead of 64 (a one-scene batch of 8192 items: 3.2 us per item,
// serialized on its scene's counters).
__device__ __forceinline__ bool first_lane() {
  return __builtin_amdgcn_readfirstlane((int)(threadIdx.x & 63)) == (int)(threadIdx.x & 63);
}

template <typename T>
__device__ __forceinline__ T* lane_slot(T* real, unsigned long long* sink) {
  return first_lane() ? real : (T*)(sink + 16 * (threadIdx.x & 63));
}

__device__ __forceinline__ void wave_add(unsigned long long* p, unsigned long long v,
                                         unsigned long long* sink) {
  atomicAdd(lane_slot(p, sink), first_lane() ? v : 0ull);
}

// Items a wave takes per counter operation.  A wave adds its stats and its
// scene's last-end stamp once per run of consecutive items of one scene
// (items are queued scene by scene), so a one-scene batch costs a quarter
// of the same-address operations per item.
constexpr uint32_t kItemsPerGrab = 4;

// Close a wave's run of items of scene `job`: its stats, once, and its
// last-end stamp.  All conditions are wave-uniform (job and p.stats are).
__device__ __forceinline__ void close_run(const BatchGroup& g,
                                          const MarchParams* __restrict__ jobs, int job,
                                          const RayCounts& rc) {
  const MarchParams& p = jobs[job];
  if (p.stats) {
    wave_add(p.stats + ST_SAMPLES, wave_sum(rc.samples), g.sink);
    wave_add(p.stats + ST_REEVAL, wave_sum(rc.reeval), g.sink);
    wave_add(p.stats + ST_AUDIT_MISS, wave_sum(rc.miss), g.sink);
    wave_add(p.stats + ST_FINE, wave_sum(rc.fine), g.sink);
    wave_add(p.stats + ST_SKIPPED, wave_sum(rc.skipped), g.sink);
  }
  atomicMax(lane_slot(g.t_done + job, g.sink), (unsigned long long)wall_clock64());
}

// Two waves per SIMD: the batched scenes' hot path (gyroids, up to 200
// primitives) spills at 4 waves per SIMD and measured slower there, and
// without the cap the allocation passes 256 registers (1 wave per SIMD; the
// hierarchical form ran 16 scenes in 1,223 ms there, 751 ms at 2).  With
// registers to spare the guarded march queues its FP64 re-evaluations per
// wave (march_ray.hpp), 24 % faster on the 32-scene batch.
template <int PREC, int INTEG>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void batch_kernel(BatchGroup g,
                                                    const MarchParams* __restrict__ jobs) {
  const int lane = threadIdx.x & 63;
  const int tile_i = 8 * g.wave_rows, tile_j = 8 * g.wave_cols;
  // Each wave takes items on its own: no LDS, no barriers.
  for (;;) {
    const uint32_t got =
        atomicAdd(lane_slot(g.counter, g.sink), first_lane() ? kItemsPerGrab : 0u);
    const uint32_t t0 = __builtin_amdgcn_readfirstlane(got);
    if (t0 >= g.n_tiles) break;
    const uint32_t t1 = min(t0 + kItemsPerGrab, g.n_tiles);
    int run_job = -1;
    RayCounts rc;
    for (uint32_t t = t0; t < t1; ++t) {
      // job holding item t: last job whose tile_start <= t
      int lo = 0, hi = g.n_jobs - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (g.tile_start[mid] <= t)
          lo = mid;
        else
          hi = mid - 1;
      }
      const int job = __builtin_amdgcn_readfirstlane(lo);
      if (job != run_job) {
        if (run_job >= 0) {
          close_run(g, jobs, run_job, rc);
          rc = RayCounts();
        }
        atomicMin(lane_slot(g.t_first + job, g.sink), (unsigned long long)wall_clock64());
        run_job = job;
      }
      const MarchParams& p = jobs[job];
      const uint32_t local = t - g.tile_start[job];
      const int tiles_i = (p.res + tile_i - 1) / tile_i, tiles_j = (p.res + tile_j - 1) / tile_j;
      const uint32_t per_view = (uint32_t)tiles_i * (uint32_t)tiles_j;
      const int v = (int)(local / per_view);
      const uint32_t rem = local - (uint32_t)v * per_view;
      const int ti = (int)(rem / (uint32_t)tiles_j), tj = (int)(rem - (uint32_t)ti * tiles_j);
      for (int bi = 0; bi < g.wave_rows; ++bi)
        for (int bj = 0; bj < g.wave_cols; ++bj) {
          const int i = ti * tile_i + bi * 8 + (lane >> 3);
          const int j = tj * tile_j + bj * 8 + (lane & 7);
          march_ray<PREC, INTEG, 4, kBlock, true>(p, v, i, j, i < p.res && j < p.res, rc);
        }
    }
    close_run(g, jobs, run_job, rc);
  }
}

__global__ void batch_reset_kernel(uint32_t* counters, int n_counters,
                                   unsigned long long* t_first, unsigned long long* t_done,
                                   int n_jobs, unsigned long long* stats, int n_stats) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n_counters) counters[k] = 0u;
  if (k < n_jobs) {
    t_first[k] = ~0ull;
    t_done[k] = 0ull;
  }
  if (k < n_stats) stats[k] = 0ull;
}

int launch_batch(const BatchGroup& g, int precision, int integrator, int workgroups,
                 void* stream) {
  hipStream_t s = (hipStream_t)stream;
  dim3 block(256);
  dim3 grid(workgroups);
  if (precision == PREC_FP64) {
    if (integrator == INT_HIERARCHICAL)
      hipLaunchKernelGGL((batch_kernel<PREC_FP64, INT_HIERARCHICAL>), grid, block, 0, s, g,
                         g.jobs);
    else
      hipLaunchKernelGGL((batch_kernel<PREC_FP64, INT_SIMPLE>), grid, block, 0, s, g, g.jobs);
  } else {
    if (integrator == INT_HIERARCHICAL)
      hipLaunchKernelGGL((batch_kernel<PREC_GUARDED, INT_HIERARCHICAL>), grid, block, 0, s, g,
                         g.jobs);
    else
      hipLaunchKernelGGL((batch_kernel<PREC_GUARDED, INT_SIMPLE>), grid, block, 0, s, g,
                         g.jobs);
  }
  return (int)hipGetLastError();
}

int launch_batch_reset(uint32_t* counters, int n_counters, unsigned long long* t_first,
                       unsigned long long* t_done, int n_jobs, unsigned long long* stats,
                       int n_stats, void* stream) {
  int n = n_counters > n_jobs ? n_counters : n_jobs;
  if (n_stats > n) n = n_stats;
  if (n <= 0) return 0;
  hipLaunchKernelGGL(batch_reset_kernel, dim3((n + 255) / 256), dim3(256), 0,
                     (hipStream_t)stream, counters, n_counters, t_first, t_done, n_jobs, stats,
                     n_stats);
  return (int)hipGetLastError();
}

}  // namespace xrp
