// api.cpp - C ABI of libxrp (declared in include/xrp.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <new>
#include <set>
#include <string>
#include <vector>

#include "../../include/xrp.h"
#include "camera.hpp"
#include "json.hpp"
#include "march.hpp"
#include "scene.hpp"

namespace {
thread_local std::string g_error;

int fail(int code, const std::string& msg) {
  g_error = msg;
  return code;
}

// Every int-returning entry point runs its body through this: no C++
// exception crosses the C ABI (a host allocation that fails, for instance,
// becomes XRP_ERR_ALLOC).
template <class F>
int guarded(F&& body) noexcept {
  try {
    return body();
  } catch (const std::bad_alloc&) {
    return fail(XRP_ERR_ALLOC, "host memory allocation failed");
  } catch (const std::exception& e) {
    return fail(XRP_ERR_INTERNAL, std::string("internal error: ") + e.what());
  } catch (...) {
    return fail(XRP_ERR_INTERNAL, "internal error");
  }
}

// Output structs open with {uint32_t struct_size; int32_t reserved}: the
// caller states the size it allocated and the library writes no further,
// so fields appended later never overrun an older caller's struct.
constexpr size_t kOutHeader = 8;

template <class T>
int check_out(const T* st, const char* name) {
  if (st && st->struct_size < kOutHeader)
    return fail(XRP_ERR_ARG, std::string(name) + ".struct_size is not set");
  return XRP_OK;
}

template <class T>
void write_out(T* dst, const T& src) {
  if (!dst) return;
  const size_t n = std::min<size_t>(dst->struct_size, sizeof(T));
  std::memcpy((char*)dst + kOutHeader, (const char*)&src + kOutHeader, n - kOutHeader);
}

// Pinned host staging buffer (hipHostMalloc), grown on demand.
struct HostBuf {
  void* p = nullptr;
  size_t cap = 0;
  hipError_t ensure(size_t bytes) {
    if (bytes <= cap && p) return hipSuccess;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
    size_t want = std::max<size_t>(bytes, 4096);
    hipError_t e = hipHostMalloc(&p, want, hipHostMallocDefault);
    if (e == hipSuccess) cap = want;
    return e;
  }
  void release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
  }
};

// Device buffer owned by exactly one holder: released explicitly (with its
// device current) or, at the latest, by the destructor.
struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  hipError_t ensure(size_t bytes) {
    if (bytes <= cap && p) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    size_t want = std::max<size_t>(bytes, 256);
    hipError_t e = hipMalloc(&p, want);
    if (e == hipSuccess) cap = want;
    return e;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
};

std::mutex g_registry_mu;  // scene uploads <-> contexts

// Largest voxel grid accepted from caller memory (32 GiB of FP32).
constexpr double kMaxVoxels = 8589934592.0;
}  // namespace

struct xrp_scene;

struct xrp_context {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::mutex mu;
  DevBuf views, table, stats, out, nodes, defs;
  // per-render upload (nodes | deformations | views | zeroed stats), staged
  // through pinned memory and sent with one copy
  DevBuf upload;
  HostBuf staging;
  // sample table cache
  int table_mode = -1;
  double table_smin = NAN, table_smax = NAN, table_ds = NAN;
  std::vector<double> table_host;
  std::set<xrp_scene*> scenes;  // scenes holding device copies here
  // xrp_render_volume_projections: the last volume as a scene, reused while
  // the caller passes the same grid
  xrp_scene* vol_scene = nullptr;
  // the fixed scratch region of xrp_render_device_noalloc on this device
  void* scratch = nullptr;
  // ray list of the two-pass guarded march (march.hpp redo_list)
  DevBuf redo;
};

struct xrp_scene {
  std::unique_ptr<xrp::Scene> s;
  std::map<xrp_context*, std::vector<void*>> uploads;  // voxel grids per context
  // a two-pass render of this scene found its lists too short: the scene's
  // guard-band samples are too many for the two passes to pay, so its
  // renders stay in one pass
  std::atomic<bool> one_pass{false};
};

namespace xrp {
const HNode& scene_root(const xrp_scene* s) { return s->s->root; }
int set_error(int code, const std::string& msg) { return fail(code, msg); }
}  // namespace xrp

#define HIPCHK(expr)                                                        \
  do {                                                                      \
    hipError_t e_ = (expr);                                                 \
    if (e_ != hipSuccess)                                                   \
      return fail(XRP_ERR_DEVICE, std::string(#expr) + " failed: " +        \
                                      hipGetErrorString(e_));               \
  } while (0)

static void free_uploads(xrp_scene* sc, xrp_context* ctx) {
  auto it = sc->uploads.find(ctx);
  if (it == sc->uploads.end()) return;
  int prev = 0;
  (void)hipGetDevice(&prev);
  (void)hipSetDevice(ctx->device);
  for (void* p : it->second) (void)hipFree(p);
  (void)hipSetDevice(prev);
  sc->uploads.erase(it);
}

extern "C" {

int xrp_abi_version(void) { return XRP_ABI_VERSION; }

const char* xrp_last_error(void) { return g_error.c_str(); }

void xrp_default_params(xrp_render_params* p) {
  if (!p) return;
  std::memset(p, 0, sizeof *p);
  p->struct_size = sizeof *p;
  p->resolution = 512;
  p->ds = -1.0;
  p->R = 4.0;
  p->fov_deg = 40.0;
  p->flat_field = 0.0;
  p->density_multiplier = 1.0;
  p->integrator = XRP_INTEGRATE_HIERARCHICAL;
  p->precision = XRP_PRECISION_AUTO;
}

int xrp_context_create(int device, xrp_context** out) {
  return guarded([&]() -> int {
    if (!out) return fail(XRP_ERR_NULL, "out is NULL");
    *out = nullptr;
    int n = 0;
    HIPCHK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n)
      return fail(XRP_ERR_ARG, "device " + std::to_string(device) +
                                   " out of range (" + std::to_string(n) +
                                   " devices)");
    HIPCHK(hipSetDevice(device));
    auto* c = new xrp_context();
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&c->ev0);
    if (e == hipSuccess) e = hipEventCreate(&c->ev1);
    if (e == hipSuccess) e = (hipError_t)xrp::noalloc_scratch(device, &c->scratch);
    if (e != hipSuccess) {
      delete c;
      return fail(XRP_ERR_DEVICE, std::string("context set-up failed: ") +
                                      hipGetErrorString(e));
    }
    *out = c;
    return XRP_OK;
  });
}

void xrp_context_destroy(xrp_context* ctx) {
  if (!ctx) return;
  if (ctx->vol_scene) xrp_scene_destroy(ctx->vol_scene);
  ctx->vol_scene = nullptr;
  {
    std::lock_guard<std::mutex> g(g_registry_mu);
    for (xrp_scene* sc : ctx->scenes) free_uploads(sc, ctx);
    ctx->scenes.clear();
  }
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  for (DevBuf* b : {&ctx->views, &ctx->table, &ctx->stats, &ctx->out,
                    &ctx->nodes, &ctx->defs, &ctx->upload})
    b->release();
  ctx->staging.release();
  if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
  if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

int xrp_context_synchronize(xrp_context* ctx) {
  return guarded([&]() -> int {
    if (!ctx) return fail(XRP_ERR_NULL, "ctx is NULL");
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return XRP_OK;
  });
}

static int parse_text(const char* text, xrp::json::Value& out, bool& present) {
  present = text && text[0];
  if (!present) return XRP_OK;
  try {
    out = xrp::json::parse(text);
  } catch (const std::exception& e) {
    return fail(XRP_ERR_ARG, e.what());
  }
  return XRP_OK;
}

int xrp_scene_create(const char* object_text, const char* deformation_text,
                     int number_mode, const char* base_dir, xrp_scene** out) {
  return guarded([&]() -> int {
    if (!out) return fail(XRP_ERR_NULL, "out is NULL");
    *out = nullptr;
    if (!object_text) return fail(XRP_ERR_NULL, "object_text is NULL");
    if (number_mode != XRP_NUMBERS_JSON && number_mode != XRP_NUMBERS_YAML)
      return fail(XRP_ERR_ARG, "unknown number_mode");
    xrp::json::Value obj, def;
    bool has_obj, has_def;
    int rc = parse_text(object_text, obj, has_obj);
    if (rc) return rc;
    if (!has_obj) return fail(XRP_ERR_SCENE, "empty object description");
    rc = parse_text(deformation_text, def, has_def);
    if (rc) return rc;
    auto* sc = new xrp_scene();
    try {
      sc->s = xrp::load_scene(obj, nullptr, (xrp::NumberMode)number_mode,
                              base_dir ? base_dir : "");
    } catch (const std::exception& e) {
      delete sc;
      return fail(XRP_ERR_SCENE, e.what());
    }
    if (has_def) {
      try {
        xrp::append_deformation(def, (xrp::NumberMode)number_mode, sc->s->defs);
        xrp::finalize_scene(*sc->s);
      } catch (const std::exception& e) {
        delete sc;
        return fail(XRP_ERR_DEFORMATION, e.what());
      }
    }
    *out = sc;
    return XRP_OK;
  });
}

int xrp_scene_create_voxels(const float* rho, int nx, int ny, int nz,
                            const char* deformation_text, int number_mode,
                            xrp_scene** out) {
  return guarded([&]() -> int {
    if (!out) return fail(XRP_ERR_NULL, "out is NULL");
    *out = nullptr;
    if (!rho) return fail(XRP_ERR_NULL, "rho is NULL");
    if (nx <= 0 || ny <= 0 || nz <= 0)
      return fail(XRP_ERR_DIMS, "voxel grid dimensions must be positive");
    if ((double)nx * ny * nz > kMaxVoxels)
      return fail(XRP_ERR_DIMS, "voxel grid larger than 2^33 voxels");
    xrp::json::Value def;
    bool has_def;
    int rc = parse_text(deformation_text, def, has_def);
    if (rc) return rc;
    auto* sc = new xrp_scene();
    try {
      sc->s = xrp::voxel_scene(rho, nx, ny, nz, has_def ? &def : nullptr,
                               (xrp::NumberMode)number_mode);
    } catch (const std::exception& e) {
      delete sc;
      return fail(XRP_ERR_DEFORMATION, e.what());
    }
    *out = sc;
    return XRP_OK;
  });
}

void xrp_scene_destroy(xrp_scene* sc) {
  if (!sc) return;
  {
    std::lock_guard<std::mutex> g(g_registry_mu);
    std::vector<xrp_context*> ctxs;
    for (auto& kv : sc->uploads) ctxs.push_back(kv.first);
    for (xrp_context* c : ctxs) {
      free_uploads(sc, c);
      c->scenes.erase(sc);
    }
  }
  delete sc;
}

int xrp_scene_min_feature_size(const xrp_scene* sc, double* out) {
  return guarded([&]() -> int {
    if (!sc || !out) return fail(XRP_ERR_NULL, "NULL argument");
    *out = sc->s->min_feature;
    return XRP_OK;
  });
}

int xrp_scene_auto_ds(const xrp_scene* sc, double* out) {
  return guarded([&]() -> int {
    if (!sc || !out) return fail(XRP_ERR_NULL, "NULL argument");
    *out = sc->s->min_feature / 5.0;  // main.go:350-353
    return XRP_OK;
  });
}

int xrp_scene_clip_box(const xrp_scene* sc, double lo[3], double hi[3],
                       int* empty) {
  return guarded([&]() -> int {
    if (!sc || !lo || !hi || !empty) return fail(XRP_ERR_NULL, "NULL argument");
    const auto& b = sc->s->world_box;
    *empty = b.empty() ? 1 : 0;
    for (int k = 0; k < 3; ++k) {
      lo[k] = b.lo[k];
      hi[k] = b.hi[k];
    }
    return XRP_OK;
  });
}

int xrp_scene_layout(const xrp_scene* sc, int* nodes, int* deformations,
                     int* has_voxels) {
  return guarded([&]() -> int {
    if (!sc) return fail(XRP_ERR_NULL, "scene is NULL");
    if (nodes) *nodes = (int)sc->s->nodes.size();
    if (deformations) *deformations = (int)sc->s->ddefs.size();
    if (has_voxels) *has_voxels = sc->s->has_voxels ? 1 : 0;
    return XRP_OK;
  });
}

int xrp_camera_from_angles(double az, double polar, double R, double eye[3],
                           double cam[16]) {
  return guarded([&]() -> int {
    if (!eye || !cam) return fail(XRP_ERR_NULL, "NULL argument");
    xrp::camera_from_angles(az, polar, R, eye, cam);
    return XRP_OK;
  });
}

int xrp_sample_count(double smin, double smax, double ds, int64_t* count) {
  return guarded([&]() -> int {
    if (!count) return fail(XRP_ERR_NULL, "count is NULL");
    if (!(ds > 0.0) || !std::isfinite(ds) || !std::isfinite(smin) ||
        !std::isfinite(smax))
      return fail(XRP_ERR_ARG, "ds must be positive and bounds finite");
    int64_t n = 0;
    for (double s = smin; s < smax; s += ds) ++n;
    *count = n;
    return XRP_OK;
  });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Render core.
// ---------------------------------------------------------------------------
namespace {

struct RenderJob {
  std::vector<xrp::DevView> views;
  int res;
  double ds, R, fov, flat_field, dm;
  int integrator, precision, audit;
  int output = 0, transparency = 0;
};

size_t pixel_bytes(int output) { return output == XRP_OUTPUT_GRAY8_ALPHA8 ? 2 : 4; }

// $XRP_SKIP=0 turns safe-radius skipping off (read once per process).
bool skip_enabled() {
  static const bool on = [] {
    const char* e = std::getenv("XRP_SKIP");
    return !(e && std::strcmp(e, "0") == 0);
  }();
  return on;
}

// Second-order reach of gyroids in the guarded marches (XRP_REACH=0 turns
// it off, for comparison; the image is the same either way).
bool reach_enabled() {
  static const bool on = [] {
    const char* e = std::getenv("XRP_REACH");
    return !(e && std::strcmp(e, "0") == 0);
  }();
  return on;
}

// Two-pass guarded simple march (XRP_TWO_PASS=0 turns it off, for
// comparison; 1 takes it for small jobs too, for tests; read per render;
// the image is the same either way).
bool two_pass_enabled() {
  const char* e = std::getenv("XRP_TWO_PASS");
  return !(e && std::strcmp(e, "0") == 0);
}
bool two_pass_forced() {
  const char* e = std::getenv("XRP_TWO_PASS");
  return e && std::strcmp(e, "1") == 0;
}

// Per-wave culling of a root collection's children (XRP_CULL=0 turns it
// off, for comparison; the image is the same either way).
bool cull_enabled() {
  static const bool on = [] {
    const char* e = std::getenv("XRP_CULL");
    return !(e && std::strcmp(e, "0") == 0);
  }();
  return on;
}

// [z][x][y] grid -> 4x4x4 bricks ordered (bz, bx, by), voxels (lz, lx, ly)
// inside a brick; padding voxels are zero and never read.
std::vector<float> brick4(const std::vector<float>& v, int nx, int ny, int nz) {
  int bx = (nx + 3) / 4, by = (ny + 3) / 4, bz = (nz + 3) / 4;
  std::vector<float> out((size_t)bx * by * bz * 64, 0.0f);
  for (int z = 0; z < nz; ++z)
    for (int x = 0; x < nx; ++x)
      for (int y = 0; y < ny; ++y) {
        size_t b = ((size_t)(z >> 2) * bx + (x >> 2)) * by + (y >> 2);
        size_t l = (size_t)(z & 3) * 16 + (x & 3) * 4 + (y & 3);
        out[b * 64 + l] = v[((size_t)z * nx + x) * ny + y];
      }
  return out;
}

// Brick-order copies with each voxel's upper-y neighbour (PAIR: v[y], v[y1])
// or upper-y and upper-z neighbours (QUAD: v[z][y], v[z1][y], v[z][y1],
// v[z1][y1]) stored beside it, upper indices clamped to N-1 as
// objects.go:811-830 clamps them, so a cell's 8 corners come in 4 or 2 wide
// loads.
std::vector<float> brick4_wide(const std::vector<float>& v, int nx, int ny, int nz, int w) {
  int bx = (nx + 3) / 4, by = (ny + 3) / 4, bz = (nz + 3) / 4;
  std::vector<float> out((size_t)bx * by * bz * 64 * w, 0.0f);
  auto at = [&](int z, int x, int y) { return v[((size_t)z * nx + x) * ny + y]; };
  for (int z = 0; z < nz; ++z)
    for (int x = 0; x < nx; ++x)
      for (int y = 0; y < ny; ++y) {
        size_t b = ((size_t)(z >> 2) * bx + (x >> 2)) * by + (y >> 2);
        size_t l = (size_t)(z & 3) * 16 + (x & 3) * 4 + (y & 3);
        float* o = &out[(b * 64 + l) * w];
        const int y1 = std::min(y + 1, ny - 1), z1 = std::min(z + 1, nz - 1);
        if (w == 2) {
          o[0] = at(z, x, y);
          o[1] = at(z, x, y1);
        } else {
          o[0] = at(z, x, y);
          o[1] = at(z1, x, y);
          o[2] = at(z, x, y1);
          o[3] = at(z1, x, y1);
        }
      }
  return out;
}

// World -> object map of a chain of affine/linear/rigid deformations.
bool compose_linear(const std::vector<xrp::HDeform>& defs, double M[9], double t[3]) {
  double m[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, v[3] = {0, 0, 0};
  for (auto& f : defs) {
    double Mi[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, ti[3] = {0, 0, 0};
    if (f.kind == xrp::D_AFFINE) {
      std::memcpy(Mi, f.p, sizeof Mi);
    } else if (f.kind == xrp::D_LINEAR) {
      const double* e = f.p;
      double L[9] = {1 + e[0], e[5], e[4], e[5], 1 + e[1], e[3], e[4], e[3], 1 + e[2]};
      std::memcpy(Mi, L, sizeof Mi);
    } else if (f.kind == xrp::D_RIGID) {
      for (int k = 0; k < 3; ++k) ti[k] = f.p[k];
    } else {
      return false;
    }
    double nm[9], nv[3];
    for (int r = 0; r < 3; ++r) {
      for (int c = 0; c < 3; ++c)
        nm[3 * r + c] = Mi[3 * r] * m[c] + Mi[3 * r + 1] * m[3 + c] + Mi[3 * r + 2] * m[6 + c];
      nv[r] = Mi[3 * r] * v[0] + Mi[3 * r + 1] * v[1] + Mi[3 * r + 2] * v[2] + ti[r];
    }
    std::memcpy(m, nm, sizeof m);
    std::memcpy(v, nv, sizeof v);
  }
  std::memcpy(M, m, sizeof m);
  std::memcpy(t, v, sizeof v);
  return true;
}

// Brick-order copies of a voxel root that the voxel-root kernel reads.
enum BrickSlot { SLOT_PLAIN = 0, SLOT_PAIR = 1, SLOT_QUAD = 2, N_SLOTS = 3 };

// Device copies of a scene's voxel grids on one context: the linear copy of
// every voxel node (preorder, matching the flattened node order), then, for
// a voxel root, N_SLOTS brick-copy slots.  A brick slot is filled on first
// use (slot < 0 asks for none), so a scene holds only the copies its
// layouts need.
int upload_voxels(xrp_context* ctx, xrp_scene* sc, std::vector<void*>& ptrs,
                  int slot = -1, bool no_alloc = false) {
  std::lock_guard<std::mutex> g(g_registry_mu);
  const char* not_resident =
      "voxel grid not resident on this context: render the scene once with "
      "xrp_render or xrp_render_device first (this call allocates nothing)";
  const bool root_vox = sc->s->root.kind == xrp::K_VOXEL;
  auto put = [](const std::vector<float>& host, size_t bytes, void*& p) -> hipError_t {
    p = nullptr;
    hipError_t e = hipMalloc(&p, std::max<size_t>(bytes, 4));
    if (e == hipSuccess) e = hipMemcpy(p, host.data(), bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess && p) {
      (void)hipFree(p);
      p = nullptr;
    }
    return e;
  };
  auto it = sc->uploads.find(ctx);
  if (it == sc->uploads.end() && no_alloc) return fail(XRP_ERR_UNSUPPORTED, not_resident);
  if (it == sc->uploads.end()) {
    std::vector<const xrp::HNode*> vox;
    std::function<void(const xrp::HNode&)> walk = [&](const xrp::HNode& n) {
      if (n.kind == xrp::K_VOXEL) vox.push_back(&n);
      for (auto& k : n.kids) walk(k);
    };
    walk(sc->s->root);
    std::vector<void*> out;
    for (auto* n : vox) {
      void* p = nullptr;
      hipError_t e = put(*n->vox, n->vox->size() * sizeof(float), p);
      if (e != hipSuccess) {
        for (void* q : out) (void)hipFree(q);
        return fail(XRP_ERR_ALLOC, std::string("voxel upload failed: ") +
                                       hipGetErrorString(e));
      }
      out.push_back(p);
    }
    if (root_vox) out.insert(out.end(), N_SLOTS, nullptr);
    it = sc->uploads.emplace(ctx, std::move(out)).first;
    ctx->scenes.insert(sc);
  }
  std::vector<void*>& out = it->second;
  if (root_vox && slot >= 0) {
    void*& p = out[out.size() - N_SLOTS + slot];
    if (!p && no_alloc) return fail(XRP_ERR_UNSUPPORTED, not_resident);
    if (!p) {
      const xrp::HNode& n = sc->s->root;
      std::vector<float> br;
      if (slot == SLOT_PLAIN) {
        br = brick4(*n.vox, n.nx, n.ny, n.nz);
      } else {
        br = brick4_wide(*n.vox, n.nx, n.ny, n.nz, slot == SLOT_PAIR ? 2 : 4);
      }
      hipError_t e = put(br, br.size() * sizeof(float), p);
      if (e != hipSuccess)
        return fail(XRP_ERR_ALLOC, std::string("voxel upload failed: ") +
                                       hipGetErrorString(e));
    }
  }
  ptrs = out;
  return XRP_OK;
}

size_t mp_bricks(const xrp::HNode& n) {
  return (size_t)((n.nx + 3) / 4) * ((n.ny + 3) / 4) * ((n.nz + 3) / 4);
}

// Everything a launch of one scene needs, computed on the host: the
// resolved precision, the sample table, the node table with guard bands and
// voxel pointers, and the launch parameters.  Device pointers for nodes,
// deformations, views, the sample table, stats and the output are left for
// the caller to fill in.
struct LaunchPlan {
  xrp::MarchParams mp;
  std::vector<xrp::DevNode> nodes;
  std::vector<double> table;
  int precision = XRP_PRECISION_GUARDED;
  bool vox_kernel = false;
  int layout = xrp::VOX_BRICK4_QUAD_SLIDE;
  double delta = 0.0;
};

int plan_job(xrp_context* ctx, xrp_scene* sc, const RenderJob& job, LaunchPlan& pl,
             bool allow_voxel_kernel = true, bool no_alloc = false) {
  using namespace xrp;
  Scene& s = *sc->s;
  const int res = job.res;

  // Voxel-root scenes under linear deformations take the dedicated voxel
  // kernel (FP32 grid stepping, exact FP64 sample range per ray).
  double vM[9], vt[3];
  const bool vox_kernel =
      allow_voxel_kernel && s.root.kind == K_VOXEL && compose_linear(s.defs, vM, vt) &&
      job.integrator == XRP_INTEGRATE_SIMPLE && !job.audit &&
      job.precision != XRP_PRECISION_FP64 &&
      (size_t)mp_bricks(s.root) * 64 < ((size_t)1 << 30);
  int precision = job.precision;
  if (precision == XRP_PRECISION_AUTO)
    precision = (s.has_voxels && !vox_kernel) ? XRP_PRECISION_FP64 : XRP_PRECISION_GUARDED;

  const double smin = job.R - kCubeHalfDiagonal, smax = job.R + kCubeHalfDiagonal;
  pl.table = job.integrator == XRP_INTEGRATE_HIERARCHICAL ? coarse_samples(smin, smax, job.ds)
                                                          : simple_samples(smin, smax, job.ds);
  const int K = (int)pl.table.size();

  // Clip box and FP32 error model.
  const Box3& wb = s.world_box;
  int clip = 1;
  double lo[3], hi[3];
  for (int k = 0; k < 3; ++k) {
    lo[k] = wb.lo[k];
    hi[k] = wb.hi[k];
  }
  bool all_inf = true;
  for (int k = 0; k < 3; ++k)
    if (std::isfinite(lo[k]) || std::isfinite(hi[k])) all_inf = false;
  if (wb.empty())
    clip = 2;
  else if (all_inf)
    clip = 0;
  double Bw = 0;
  if (clip == 1 && wb.finite()) {
    for (int k = 0; k < 3; ++k)
      Bw = std::max({Bw, std::fabs(lo[k]), std::fabs(hi[k])});
    Bw += 3.0 * job.ds;
  } else {
    Bw = 2.0 * std::fabs(job.R) + kCubeHalfDiagonal + 1.0;
  }
  double delta = 0.0, lipschitz = 1.0;
  if (precision == XRP_PRECISION_GUARDED && !vox_kernel) {
    double step = job.integrator == XRP_INTEGRATE_SIMPLE ? 32.0 * job.ds : 0.0;
    // simple march anchors: smin + k ds against the table's K accumulated
    // additions (each off by at most half an ulp of |s|)
    double s_err = job.integrator == XRP_INTEGRATE_SIMPLE
                       ? (K + 4.0) * 2.3e-16 * std::max(std::fabs(smin), std::fabs(smax))
                       : 0.0;
    delta = compute_guards(s, Bw, step, &lipschitz, s_err);
    // The bands keep the result exact at any delta; they only cost time when
    // wide.  Fall back to FP64 when the FP32 position error reaches a tenth
    // of the smallest feature (MinFeatureSize is conservative: the gyroid's
    // is scale*thickness/10).  At 1e-3 the sheared gyroid lattice fell back
    // and ran 2.8x slower than guarded.
    if (!(delta < 0.1 * s.min_feature)) precision = XRP_PRECISION_FP64;
  }

  std::vector<void*> vptr;
  if (s.has_voxels) {
    int rc = upload_voxels(ctx, sc, vptr, -1, no_alloc);
    if (rc) return rc;
  }
  pl.nodes = s.nodes;
  size_t vi = 0;
  for (auto& n : pl.nodes)
    if (n.kind == K_VOXEL) n.vox = (const float*)vptr[vi++];

  MarchParams& mp = pl.mp;
  std::memset(&mp, 0, sizeof mp);
  mp.n_defs = (int)s.ddefs.size();
  mp.res = res;
  mp.K = K;
  mp.focal = focal_length(job.fov);
  mp.ds = job.ds;
  mp.ds_fine = job.ds / 10.0;
  mp.smin = smin;
  mp.smax = smax;
  mp.flat_field = job.flat_field;
  mp.dm = job.dm;
  for (int k = 0; k < 3; ++k) {
    mp.lo[k] = lo[k];
    mp.hi[k] = hi[k];
  }
  mp.clip = clip == 1 ? 1 : 0;
  if (clip == 2) mp.K = 0;  // zero density everywhere: T = flat_field
  mp.audit = job.audit;
  mp.out_view_stride = res * res;
  mp.out_mode = job.output;
  mp.transparency = job.transparency;
  // Safe-radius skipping in the guarded marches (XRP_SKIP=0 turns it off,
  // for comparison; the image is the same either way).  ds is the coarse
  // step DS for the hierarchical integrator.
  if (precision == XRP_PRECISION_GUARDED && !vox_kernel && skip_enabled()) {
    double inv = 1.0 / (lipschitz * job.ds) * (1.0 - 1e-6);
    mp.skip_inv = std::isfinite(inv) && inv > 0 ? (float)inv : 0.0f;
    // gyroids reach further along the ray with the second-order bound,
    // which needs the last sample's velocity (XRP_REACH=0 turns it off,
    // for comparison; the image is the same either way)
    if (mp.skip_inv > 0.0f && reach_enabled())
      for (const auto& n : pl.nodes)
        if (n.kind == K_GYROID && n.g[6] > 0.0f) mp.vel = 1;
  }
  // Samples per node visit of the per-scene guarded simple march: with the
  // safe radius, 2 is faster on small trees (lattice cell 10 nodes: 6 %
  // faster than 1), 4 on large collections, whose node walk it amortises
  // (1-200 primitives: 4 is 14 % faster than 2).  Gyroids settling by their
  // second-order reach skip far enough that the second sample of a block is
  // mostly wasted: 1 is 6 % faster there (6^3 sheared gyroid cells, 256 x
  // 1024^2: 90.1 against 95.9 ms; 512^2 gyroid lattice 1.135 against 1.212).
  mp.block = s.nodes.size() >= 24 ? 4 : mp.vel ? 1 : 2;
  if (const char* e = std::getenv("XRP_BLOCK")) {  // for comparison: 1, 2 or 4
    const int b = std::atoi(e);
    if (b == 1 || b == 2 || b == 4) mp.block = b;
  }
  if (const char* e = std::getenv("XRP_QUEUE"))  // for comparison: 0 or 1
    mp.queue = std::strcmp(e, "1") == 0 ? 1 : 0;
  if (const char* e = std::getenv("XRP_TILE"))  // for comparison: 32x2 or 2x32
    mp.tile_j32 = std::strcmp(e, "2x32") == 0 ? 1 : 0;
  // Per-wave culling of the root collection's children (guarded marches):
  // needs the deformations composed to one affine map, so that rays stay
  // lines in object space.  From 8 children up: on the 3-child CSG scene
  // the per-wave test cost 6 % (32-scene batch of 1-200 children: 1.9x
  // faster).
  if (precision == XRP_PRECISION_GUARDED && !vox_kernel && s.root.kind == K_COLL &&
      s.root.kids.size() >= 8 && s.root.kids.size() <= 256 && s.nodes.size() <= 32767 &&
      cull_enabled()) {
    double cM[9], ct[3];
    if (compose_linear(s.defs, cM, ct)) {
      mp.cull = 1;
      std::memcpy(mp.M, cM, sizeof cM);
      std::memcpy(mp.t, ct, sizeof ct);
      // a culled child stays further than the guard band from every FP32
      // sample position, so it changes no FP32 density (the check pass of
      // the two-pass hierarchical march culls again for its own waves);
      // its FP64 density along the wave's rays is 0, so the FP64
      // re-evaluations walk the same list
      const double pad = 2.0 * delta * (1.0 + 1e-6);
      mp.cull_pad = (float)pad;
      if ((double)mp.cull_pad < pad) mp.cull_pad = std::nextafter(mp.cull_pad, INFINITY);
    }
  }
  mp.n_views = (int)job.views.size();
  // 4^3 bricks of corner quads (two 16-byte loads per cell) with the
  // per-lane cell cache sliding along x measured fastest on the 256^3 voxel
  // workload (11.6k 512^2 projections/s; quads with a plain cell cache
  // 9.87k, plain bricks 7.94k, corner pairs 7.92k, uncached bricks 5.4k,
  // cached linear 4.8k, LDS-staged 2.2k, linear 1.8k); the others stay
  // selectable through XRP_VOXEL_LAYOUT for comparison.
  int layout = VOX_BRICK4_QUAD_SLIDE;
  if (vox_kernel) {
    const char* env = std::getenv("XRP_VOXEL_LAYOUT");
    if (env) {
      const std::string e(env);
      if (e == "brick4") layout = VOX_BRICK4;
      if (e == "linear") layout = VOX_LINEAR;
      if (e == "lds") layout = VOX_LDS;
      if (e == "cached") layout = VOX_BRICK4_CACHED;
      if (e == "linear_cached") layout = VOX_LINEAR_CACHED;
      if (e == "pair") layout = VOX_BRICK4_PAIR;
      if (e == "quad") layout = VOX_BRICK4_QUAD;
      if (e == "quad_slide") layout = VOX_BRICK4_QUAD_SLIDE;
      if (e == "quad_step") layout = VOX_BRICK4_QUAD_STEP;
    }
    // The kernels address a copy with 32-bit byte offsets: a grid whose
    // pair or quad copy would pass 4 GiB (2^28 voxels and up for quads)
    // reads the plain brick copy with the cell cache instead.
    const size_t belems = (size_t)mp_bricks(s.root) * 64;
    if ((layout == VOX_BRICK4_PAIR && belems >= ((size_t)1 << 29)) ||
        (layout >= VOX_BRICK4_QUAD && belems >= ((size_t)1 << 28)))
      layout = VOX_BRICK4_CACHED;
    const bool lin = voxel_layout_linear(layout);
    const size_t nlin = (size_t)s.root.nx * s.root.ny * s.root.nz;
    const int width = layout == VOX_BRICK4_PAIR ? 2 : layout >= VOX_BRICK4_QUAD ? 4 : 1;
    const int slot = width == 1 ? SLOT_PLAIN : width == 2 ? SLOT_PAIR : SLOT_QUAD;
    if (!lin) {
      int rc = upload_voxels(ctx, sc, vptr, slot, no_alloc);
      if (rc) return rc;
    }
    const size_t nb = vptr.size();
    mp.vgrid = (const float*)(lin ? vptr.front() : vptr[nb - N_SLOTS + slot]);
    mp.vbytes = (uint32_t)(4ull * (lin ? nlin : (size_t)mp_bricks(s.root) * 64 * width));
    mp.vlin = (const float*)vptr.front();
    mp.vlin_bytes = (uint32_t)(4ull * nlin);
    const HNode& vn = s.root;
    mp.vn[0] = vn.nx;
    mp.vn[1] = vn.ny;
    mp.vn[2] = vn.nz;
    mp.vbricks[0] = (vn.nx + 3) / 4;
    mp.vbricks[1] = (vn.ny + 3) / 4;
    mp.vbricks[2] = (vn.nz + 3) / 4;
    std::memcpy(mp.M, vM, sizeof vM);
    std::memcpy(mp.t, vt, sizeof vt);
    double L = 0;
    for (int r = 0; r < 3; ++r)
      L = std::max(L, std::fabs(vM[3 * r]) + std::fabs(vM[3 * r + 1]) + std::fabs(vM[3 * r + 2]));
    double N = std::max({vn.nx, vn.ny, vn.nz});
    const double u = 1.0 / 16777216.0;
    // FP32 anchor rounding (u*|G|) plus the FP32 step (4u * |B| * chunk*ds).
    mp.vguard = (float)(4.0 * (u * (N + 4.0 * L * 0.5 * N * 32.0 * job.ds) + 1e-9));
  }
  pl.precision = precision;
  pl.vox_kernel = vox_kernel;
  pl.layout = layout;
  pl.delta = delta;
  return XRP_OK;
}

// With `fixed` (the device address of a region of kNoAllocScratchBytes),
// the render allocates nothing: stat counters, node table, deformations,
// views and the sample table are staged from pageable host memory into that
// region instead of the context's buffers, and a voxel grid must already be
// resident on the context.
int render_core(xrp_context* ctx, xrp_scene* sc, const RenderJob& job,
                void* out_dev, xrp_render_stats* st, void* fixed = nullptr,
                bool one_pass = false) {
  using namespace xrp;
  Scene& s = *sc->s;
  const int nv = (int)job.views.size();
  const int res = job.res;
  HIPCHK(hipSetDevice(ctx->device));
  LaunchPlan pl;
  int rc = plan_job(ctx, sc, job, pl, true, fixed != nullptr);
  if (rc) return rc;
  const int precision = pl.precision;
  const bool vox_kernel = pl.vox_kernel;
  const int layout = pl.layout;
  const double delta = pl.delta;
  // Guarded simple and hierarchical marches at block 1 or 2 (small trees)
  // in two passes (march.hip march_try_kernel, march_hier_try_kernel), for
  // jobs of 2^21 rays and up.  On large collections a ray meets more
  // guard-band samples than it keeps (179 primitives, hierarchical: 76 % of
  // the rays marched again, 497 against 336 ms in one pass).  The
  // hierarchical form is not taken for root collections (XRP_TWO_PASS=1
  // takes it, for tests): boxes and other faces that views cross edge-on
  // put samples on the faces, whose FP32 guesses often fail (random
  // collections of 8-20 children, 16 x 512^2: 3.6-12.5 ms in one pass,
  // 26-42 % slower in two) (below, the two extra
  // launches cost more than they save: CSG 8 x 256^2 0.107 against 0.097
  // ms).  A launch covers at most 2^31 rays, which the lists name by 32-bit
  // index.  Not under audit, nor without allocation.
  const int kMaxGridY = 65535;
  const bool hier = job.integrator == XRP_INTEGRATE_HIERARCHICAL;
  const bool two_pass = !fixed && !one_pass && !sc->one_pass &&
                        precision == XRP_PRECISION_GUARDED && !vox_kernel && !job.audit &&
                        (hier ? s.root.kind != K_COLL || two_pass_forced() : !pl.mp.queue) &&
                        (pl.mp.block == 1 || pl.mp.block == 2) &&
                        ((size_t)nv * res * res >= ((size_t)1 << 21) || two_pass_forced()) &&
                        two_pass_enabled();
  int vpc = two_pass ? (int)std::min<size_t>(kMaxGridY, ((size_t)1 << 31) / ((size_t)res * res))
                     : kMaxGridY;
  if (const char* e = std::getenv("XRP_LAUNCH_VIEWS"))  // fewer views per launch, for tests
    vpc = std::max(1, std::min(vpc, std::atoi(e)));
  const int n_chunks = (nv + vpc - 1) / vpc;

  // Sample table (device copy cached per context).
  const double smin = job.R - kCubeHalfDiagonal, smax = job.R + kCubeHalfDiagonal;
  if (!fixed && (ctx->table_mode != job.integrator || ctx->table_smin != smin ||
                 ctx->table_smax != smax || ctx->table_ds != job.ds)) {
    ctx->table_host = pl.table;
    HIPCHK(ctx->table.ensure(std::max<size_t>(1, ctx->table_host.size()) *
                             sizeof(double)));
    if (!ctx->table_host.empty())
      HIPCHK(hipMemcpy(ctx->table.p, ctx->table_host.data(),
                       ctx->table_host.size() * sizeof(double),
                       hipMemcpyHostToDevice));
    ctx->table_mode = job.integrator;
    ctx->table_smin = smin;
    ctx->table_smax = smax;
    ctx->table_ds = job.ds;
  }

  // One pinned staging copy of nodes, deformations, views and zeroed
  // stats (256-byte aligned sections), sent with one copy: the per-call
  // host cost matters for small renders.
  auto align = [](size_t x) { return (x + 255) & ~(size_t)255; };
  const size_t b_nodes = pl.nodes.size() * sizeof(DevNode);
  const size_t b_defs = s.ddefs.size() * sizeof(DevDeform);
  const size_t b_views = (size_t)nv * sizeof(DevView);
  const size_t b_stats = ST_COUNT * sizeof(unsigned long long);
  const size_t o_defs = align(b_nodes), o_views = o_defs + align(std::max<size_t>(b_defs, 1));
  const size_t o_stats = o_views + align(b_views);
  // zeroed list counters of the two-pass march, one per launch
  const size_t o_ctr = o_stats + align(b_stats);
  const size_t b_ctr = two_pass ? (size_t)n_chunks * 4 * sizeof(uint32_t) : 0;
  // the no-allocation form also carries the sample table
  const size_t o_table = o_ctr + align(b_ctr);
  const size_t total = fixed ? o_table + pl.table.size() * sizeof(double)
                             : two_pass ? o_ctr + b_ctr : o_stats + b_stats;
  std::vector<unsigned char> pageable;
  unsigned char *hb, *db;
  if (fixed) {
    if (total > kNoAllocScratchBytes)
      return fail(XRP_ERR_DIMS,
                  "nodes, views and sample table need " + std::to_string(total) +
                      " bytes; the no-allocation scratch region holds " +
                      std::to_string(kNoAllocScratchBytes));
    pageable.assign(total, 0);
    hb = pageable.data();
    db = (unsigned char*)fixed;
    if (!pl.table.empty())
      std::memcpy(hb + o_table, pl.table.data(), pl.table.size() * sizeof(double));
  } else {
    HIPCHK(ctx->staging.ensure(total));
    HIPCHK(ctx->upload.ensure(total));
    hb = (unsigned char*)ctx->staging.p;
    db = (unsigned char*)ctx->upload.p;
  }
  std::memcpy(hb, pl.nodes.data(), b_nodes);
  if (b_defs) std::memcpy(hb + o_defs, s.ddefs.data(), b_defs);
  std::memcpy(hb + o_views, job.views.data(), b_views);
  std::memset(hb + o_stats, 0, b_stats);
  if (b_ctr) std::memset(hb + o_ctr, 0, b_ctr);
  // the lists, sized for the largest launch (XRP_DEFER_CAP and
  // XRP_REDO_CAP shorten them, for tests: 1 and up)
  TwoPassCaps caps = two_pass_caps(two_pass ? (size_t)std::min(vpc, nv) * res * res : 0, hier);
  if (const char* e = std::getenv("XRP_DEFER_CAP"))
    caps.defer = (uint32_t)std::min<long>(caps.defer, std::max(1L, std::atol(e)));
  if (const char* e = std::getenv("XRP_REDO_CAP"))
    caps.redo = (uint32_t)std::min<long>(caps.redo, std::max(1L, std::atol(e)));
  // guard-band samples a ray may defer (XRP_DEFER_MAX, fewer for tests: 0
  // and up)
  const uint32_t defer_lim = hier ? kHierDeferMax : kDeferMax;
  uint32_t defer_max = defer_lim;
  if (const char* e = std::getenv("XRP_DEFER_MAX"))
    defer_max = (uint32_t)std::min<long>(defer_lim, std::max(0L, std::atol(e)));
  if (two_pass) HIPCHK(ctx->redo.ensure(two_pass_bytes(caps)));
  HIPCHK(hipMemcpyAsync(db, hb, total, hipMemcpyHostToDevice, ctx->stream));
  // pageable staging: wait for the copy, so that no early return below can
  // free the vector while the copy still reads it
  if (fixed) HIPCHK(hipStreamSynchronize(ctx->stream));

  MarchParams mp = pl.mp;
  mp.nodes = (const DevNode*)db;
  mp.defs = (const DevDeform*)(db + o_defs);
  mp.views = (const DevView*)(db + o_views);
  mp.s_table = fixed ? (const double*)(db + o_table) : (const double*)ctx->table.p;
  mp.stats = (unsigned long long*)(db + o_stats);
  (void)res;

  HIPCHK(hipEventRecord(ctx->ev0, ctx->stream));
  for (int v0 = 0, c = 0; v0 < nv; v0 += vpc, ++c) {
    mp.view_offset = v0;
    mp.n_views = std::min(vpc, nv - v0);
    if (two_pass) {
      auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
      unsigned char* b = (unsigned char*)ctx->redo.p;
      mp.redo_list = (uint32_t*)b;
      b += al((size_t)caps.redo * 4);
      mp.defer_T = (double*)b;
      b += al((size_t)caps.defer * 8);
      mp.defer_ray = (uint32_t*)b;
      b += al((size_t)caps.defer * 4);
      mp.defer_cnt = (uint32_t*)b;
      b += al((size_t)caps.defer * 4);
      mp.defer_kfirst = (uint32_t*)b;
      b += al((size_t)caps.defer * 4);
      mp.defer_k = (uint32_t*)b;
      mp.redo_cap = caps.redo;
      mp.defer_cap = caps.defer;
      mp.k_cap = caps.k;
      mp.defer_max = defer_max;
      mp.counts = (uint32_t*)(db + o_ctr) + 4 * c;
    }
    if (job.output == XRP_OUTPUT_GRAY8_ALPHA8)
      mp.out8 = (unsigned char*)out_dev + (size_t)v0 * res * res * 2;
    else
      mp.out = (float*)out_dev + (size_t)v0 * res * res;
    int prec = precision == XRP_PRECISION_FP64 ? PREC_FP64 : PREC_GUARDED;
    int integ = job.integrator == XRP_INTEGRATE_HIERARCHICAL ? INT_HIERARCHICAL
                                                             : INT_SIMPLE;
    int e = vox_kernel ? launch_voxel(mp, layout, ctx->stream)
                       : launch_march(mp, prec, integ, ctx->stream);
    if (e != 0) {
      // the staging copy may still be reading the pinned buffer
      (void)hipStreamSynchronize(ctx->stream);
      return fail(XRP_ERR_DEVICE, std::string("kernel launch failed: ") +
                                      hipGetErrorString((hipError_t)e));
    }
  }
  HIPCHK(hipEventRecord(ctx->ev1, ctx->stream));
  // the staging copy above has completed in stream order: reuse its stats
  // section (and the two-pass list counters after it) for the read-back
  unsigned long long* hstp = (unsigned long long*)(hb + o_stats);
  const size_t b_back = two_pass ? o_ctr + b_ctr - o_stats : b_stats;
  HIPCHK(hipMemcpyAsync(hstp, mp.stats, b_back, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  unsigned long long hst[ST_COUNT];
  std::memcpy(hst, hstp, sizeof hst);
  uint64_t redo_rays = 0;
  bool overflow = false;
  for (int c = 0; two_pass && c < n_chunks; ++c) {
    uint32_t n[4];
    std::memcpy(n, hb + o_ctr + 4 * c * sizeof(uint32_t), sizeof n);
    redo_rays += n[0];
    overflow = overflow || n[3] != 0;
  }
  // a redo list too short for the rays it had to hold: the two-pass images
  // lack those rays, so the job is rendered again in one pass
  if (overflow) {
    sc->one_pass = true;
    return render_core(ctx, sc, job, out_dev, st, fixed, true);
  }
  if (st) {
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    xrp_render_stats o;
    std::memset(&o, 0, sizeof o);
    o.rays = (uint64_t)nv * res * res;
    o.samples = hst[ST_SAMPLES];
    o.reevaluated = hst[ST_REEVAL];
    o.audit_mismatch = hst[ST_AUDIT_MISS];
    o.fine_samples = hst[ST_FINE];
    o.staged_chunks = hst[ST_STAGED_CHUNKS];
    o.staged_voxels = hst[ST_STAGED_VOXELS];
    o.fallback_chunks = hst[ST_FALLBACK_CHUNKS];
    o.skipped_samples = hst[ST_SKIPPED];
    o.kernel_ms = ms;
    o.position_error = delta;
    o.precision_used = precision;
    o.two_pass = two_pass ? 1 : 0;
    o.redo_rays = redo_rays;
    o.deferred_rays = hst[ST_DEFERRED];
    o.block = precision == XRP_PRECISION_GUARDED && !vox_kernel &&
                      (job.integrator == XRP_INTEGRATE_SIMPLE || two_pass)
                  ? pl.mp.block
                  : 0;
    write_out(st, o);
  }
  return XRP_OK;
}

int check_params(const xrp_render_params* p, int n_views) {
  if (!p) return fail(XRP_ERR_NULL, "params is NULL");
  if (p->struct_size < sizeof(xrp_render_params))
    return fail(XRP_ERR_ARG, "params.struct_size too small");
  if (p->resolution <= 0 || n_views <= 0)
    return fail(XRP_ERR_DIMS, "resolution and view count must be positive");
  if (p->resolution > 32768)
    return fail(XRP_ERR_DIMS, "resolution above 32768");
  if (!(p->ds > 0.0) || !std::isfinite(p->ds))
    return fail(XRP_ERR_ARG,
                "ds must be positive (resolve ds < 0 with xrp_scene_auto_ds)");
  if (p->density_multiplier == 0.0)
    return fail(XRP_ERR_ARG,
                "density_multiplier is 0; all densities will be zero and the "
                "render will produce a blank image");
  if (!std::isfinite(p->R) || !std::isfinite(p->fov_deg) ||
      !std::isfinite(p->flat_field) || !std::isfinite(p->density_multiplier))
    return fail(XRP_ERR_ARG, "R, fov, flat_field and density_multiplier must be finite");
  if (p->integrator != XRP_INTEGRATE_SIMPLE &&
      p->integrator != XRP_INTEGRATE_HIERARCHICAL)
    return fail(XRP_ERR_ARG, "unknown integrator");
  if (p->precision < XRP_PRECISION_AUTO || p->precision > XRP_PRECISION_FP64)
    return fail(XRP_ERR_ARG, "unknown precision");
  if (p->output_format != XRP_OUTPUT_F32 && p->output_format != XRP_OUTPUT_GRAY8_ALPHA8)
    return fail(XRP_ERR_ARG, "unknown output_format");
  double K = 2 * xrp::kCubeHalfDiagonal / p->ds;
  if (K > 5e7) return fail(XRP_ERR_ARG, "ds too small: more than 5e7 samples per ray");
  return XRP_OK;
}

RenderJob make_job(const xrp_render_params* p, const xrp_view_angle* views,
                   int n_views) {
  RenderJob job;
  job.res = p->resolution;
  job.ds = p->ds;
  job.R = p->R;
  job.fov = p->fov_deg;
  job.flat_field = p->flat_field;
  job.dm = p->density_multiplier;
  job.integrator = p->integrator;
  job.precision = p->precision;
  job.audit = p->audit;
  job.output = p->output_format;
  job.transparency = p->transparency;
  job.views.resize(n_views);
  for (int v = 0; v < n_views; ++v) {
    xrp::DevView& dv = job.views[v];
    std::memset(&dv, 0, sizeof dv);
    xrp::camera_from_angles(views[v].azimuthal_deg, views[v].polar_deg, p->R,
                            dv.eye, dv.cam);
  }
  return job;
}

}  // namespace

extern "C" {

int xrp_render_device(xrp_context* ctx, xrp_scene* sc,
                      const xrp_render_params* p, const xrp_view_angle* views,
                      int n_views, void* out_device, xrp_render_stats* st) {
  return guarded([&]() -> int {
    if (!ctx || !sc || !views || !out_device)
      return fail(XRP_ERR_NULL, "NULL argument");
    int rc = check_params(p, n_views);
    if (!rc) rc = check_out(st, "stats");
    if (rc) return rc;
    std::lock_guard<std::mutex> g(ctx->mu);
    RenderJob job = make_job(p, views, n_views);
    return render_core(ctx, sc, job, out_device, st);
  });
}

int xrp_render_device_noalloc(xrp_context* ctx, xrp_scene* sc, const xrp_render_params* p,
                              const xrp_view_angle* views, int n_views, void* out_device,
                              xrp_render_stats* st) {
  return guarded([&]() -> int {
    if (!ctx || !sc || !views || !out_device) return fail(XRP_ERR_NULL, "NULL argument");
    int rc = check_params(p, n_views);
    if (!rc) rc = check_out(st, "stats");
    if (rc) return rc;
    if (!ctx->scratch) return fail(XRP_ERR_DEVICE, "no scratch region on this device");
    std::lock_guard<std::mutex> g(ctx->mu);
    RenderJob job = make_job(p, views, n_views);
    return render_core(ctx, sc, job, out_device, st, ctx->scratch);
  });
}

int xrp_render(xrp_context* ctx, xrp_scene* sc, const xrp_render_params* p,
               const xrp_view_angle* views, int n_views, void* out,
               xrp_render_stats* st) {
  return guarded([&]() -> int {
    if (!ctx || !sc || !views || !out) return fail(XRP_ERR_NULL, "NULL argument");
    int rc = check_params(p, n_views);
    if (!rc) rc = check_out(st, "stats");
    if (rc) return rc;
    std::lock_guard<std::mutex> g(ctx->mu);
    HIPCHK(hipSetDevice(ctx->device));
    size_t bytes = (size_t)n_views * p->resolution * p->resolution *
                   pixel_bytes(p->output_format);
    HIPCHK(ctx->out.ensure(bytes));
    RenderJob job = make_job(p, views, n_views);
    rc = render_core(ctx, sc, job, ctx->out.p, st);
    if (rc) return rc;
    HIPCHK(hipMemcpy(out, ctx->out.p, bytes, hipMemcpyDeviceToHost));
    return XRP_OK;
  });
}

int xrp_export_volume(xrp_context* ctx, xrp_scene* sc, int res,
                      double density_multiplier, uint8_t* out) {
  return guarded([&]() -> int {
    using namespace xrp;
    if (!ctx || !sc || !out) return fail(XRP_ERR_NULL, "NULL argument");
    if (res <= 0 || res > 2048) return fail(XRP_ERR_DIMS, "res must be in 1..2048");
    std::lock_guard<std::mutex> g(ctx->mu);
    HIPCHK(hipSetDevice(ctx->device));
    Scene& s = *sc->s;
    std::vector<void*> vptr;
    if (s.has_voxels) {
      int rc = upload_voxels(ctx, sc, vptr);
      if (rc) return rc;
    }
    size_t vi = 0;
    for (auto& n : s.nodes)
      if (n.kind == K_VOXEL) n.vox = (const float*)vptr[vi++];
    const size_t n = (size_t)res * res * res;
    HIPCHK(ctx->nodes.ensure(s.nodes.size() * sizeof(DevNode)));
    HIPCHK(ctx->defs.ensure(std::max<size_t>(1, s.ddefs.size()) * sizeof(DevDeform)));
    HIPCHK(ctx->stats.ensure(ST_COUNT * sizeof(unsigned long long)));
    DevBuf dens, bytes;
    HIPCHK(dens.ensure(n * sizeof(double)));
    hipError_t he = bytes.ensure(n);
    if (he != hipSuccess) {
      dens.release();
      return fail(XRP_ERR_ALLOC, hipGetErrorString(he));
    }
    MarchParams mp;
    std::memset(&mp, 0, sizeof mp);
    mp.nodes = (const DevNode*)ctx->nodes.p;
    mp.defs = (const DevDeform*)ctx->defs.p;
    mp.n_defs = (int)s.ddefs.size();
    mp.dm = density_multiplier;
    he = hipMemcpyAsync(ctx->nodes.p, s.nodes.data(), s.nodes.size() * sizeof(DevNode),
                        hipMemcpyHostToDevice, ctx->stream);
    if (he == hipSuccess && !s.ddefs.empty())
      he = hipMemcpyAsync(ctx->defs.p, s.ddefs.data(), s.ddefs.size() * sizeof(DevDeform),
                          hipMemcpyHostToDevice, ctx->stream);
    if (he == hipSuccess)
      he = hipMemsetAsync(ctx->stats.p, 0, sizeof(unsigned long long), ctx->stream);
    if (he == hipSuccess)
      he = (hipError_t)launch_volume(mp, res, (double*)dens.p, (unsigned long long*)ctx->stats.p,
                                     (unsigned char*)bytes.p, ctx->stream);
    if (he == hipSuccess)
      he = hipMemcpyAsync(out, bytes.p, n, hipMemcpyDeviceToHost, ctx->stream);
    if (he == hipSuccess) he = hipStreamSynchronize(ctx->stream);
    dens.release();
    bytes.release();
    if (he != hipSuccess)
      return fail(XRP_ERR_DEVICE, std::string("volume export failed: ") + hipGetErrorString(he));
    return XRP_OK;
  });
}

int xrp_render_volume_projections(xrp_context* ctx, const float* volume, int nx,
                                  int ny, int nz, const xrp_camera* cams,
                                  int n_cams, int image_res, double ds,
                                  double flat_field, float* out) {
  return guarded([&]() -> int {
    // cuda_backend.cu:95-100 argument checks keep their codes (1 NULL, 2
    // dimensions); the rest are the limits xrp_render applies.
    if (!ctx || !volume || !cams || !out) return fail(XRP_ERR_NULL, "NULL argument");
    if (nx <= 0 || ny <= 0 || nz <= 0 || image_res <= 0 || n_cams <= 0)
      return fail(XRP_ERR_DIMS, "dimensions must be positive");
    if (image_res > 32768) return fail(XRP_ERR_DIMS, "image_res above 32768");
    if ((double)nx * ny * nz > kMaxVoxels)
      return fail(XRP_ERR_DIMS, "volume larger than 2^33 voxels");
    if (!(ds > 0.0) || !std::isfinite(ds)) return fail(XRP_ERR_ARG, "ds must be positive");
    if (2 * xrp::kCubeHalfDiagonal / ds > 5e7)
      return fail(XRP_ERR_ARG, "ds too small: more than 5e7 samples per ray");
    if (!std::isfinite(flat_field)) return fail(XRP_ERR_ARG, "flat_field must be finite");
    for (int c = 0; c < n_cams; ++c) {
      bool ok = std::isfinite(cams[c].fov_deg) && std::isfinite(cams[c].R);
      for (int k = 0; k < 3; ++k) ok = ok && std::isfinite(cams[c].eye[k]);
      for (int k = 0; k < 16; ++k) ok = ok && std::isfinite(cams[c].cam2world[k]);
      if (!ok)
        return fail(XRP_ERR_ARG, "camera " + std::to_string(c) +
                                     ": eye, cam2world, fov_deg and R must be finite");
    }
    std::lock_guard<std::mutex> g(ctx->mu);
    HIPCHK(hipSetDevice(ctx->device));
    // The volume's scene, with its device copy and brick copies, stays on
    // the context while the caller passes the same dimensions and values.
    const size_t nvox = (size_t)nx * ny * nz;
    xrp_scene*& sc = ctx->vol_scene;
    if (sc) {
      const xrp::HNode& r = sc->s->root;
      if (r.nx != nx || r.ny != ny || r.nz != nz ||
          std::memcmp(r.vox->data(), volume, nvox * sizeof(float)) != 0) {
        xrp_scene_destroy(sc);
        sc = nullptr;
      }
    }
    if (!sc) {
      int rc = xrp_scene_create_voxels(volume, nx, ny, nz, nullptr, XRP_NUMBERS_JSON, &sc);
      if (rc) return rc;
    }
    const size_t img = (size_t)image_res * image_res;
    HIPCHK(ctx->out.ensure(img * n_cams * sizeof(float)));
    // Consecutive cameras sharing (fov, R) go into one launch.
    for (int c0 = 0; c0 < n_cams;) {
      int c1 = c0 + 1;
      while (c1 < n_cams && c1 - c0 < 65535 && cams[c1].fov_deg == cams[c0].fov_deg &&
             cams[c1].R == cams[c0].R)
        ++c1;
      RenderJob job;
      job.res = image_res;
      job.ds = ds;
      job.R = cams[c0].R;
      job.fov = cams[c0].fov_deg;
      job.flat_field = flat_field;
      job.dm = 1.0;
      job.integrator = XRP_INTEGRATE_SIMPLE;
      job.precision = XRP_PRECISION_AUTO;  // the voxel kernel
      job.audit = 0;
      for (int c = c0; c < c1; ++c) {
        xrp::DevView dv;
        std::memset(&dv, 0, sizeof dv);
        std::memcpy(dv.eye, cams[c].eye, sizeof dv.eye);
        std::memcpy(dv.cam, cams[c].cam2world, sizeof dv.cam);
        job.views.push_back(dv);
      }
      int rc = render_core(ctx, sc, job, (void*)((float*)ctx->out.p + img * c0), nullptr);
      if (rc) return rc;
      c0 = c1;
    }
    HIPCHK(hipMemcpy(out, ctx->out.p, img * n_cams * sizeof(float), hipMemcpyDeviceToHost));
    return XRP_OK;
  });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batched rendering.
// ---------------------------------------------------------------------------
struct xrp_batch {
  xrp_context* ctx = nullptr;
  int n_items = 0;
  std::vector<int> res, n_views;
  std::vector<xrp::MarchParams> mps;  // per item, device pointers filled in
  std::vector<int> prec, integ;
  std::vector<double> cost;           // per item: estimated work per tile
  struct Group {
    int prec = 0, integ = 0;
    std::vector<int> items;           // queue order (smallest total work first)
    int first = 0;                    // offset of the group's jobs in `jobs`
    uint32_t n_tiles = 0;
  };
  std::vector<Group> groups;
  DevBuf nodes, defs, views, tables, jobs, tile_start, counters, t_first, t_done, stats, out,
      sink;
  std::vector<size_t> out_off;        // floats
  xrp_tile_shape shape{1, 1, 2, 0};
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int n_cu = 0;
  double ticks_per_ms = 0.0;
  std::vector<double> latency;
  std::string signature;
  ~xrp_batch() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (DevBuf* b : {&nodes, &defs, &views, &tables, &jobs, &tile_start, &counters, &t_first,
                      &t_done, &stats, &out, &sink})
      b->release();
  }
};

namespace {

std::mutex g_tune_mu;
std::map<std::string, xrp_tile_shape> g_tune_table;

bool valid_shape(const xrp_tile_shape& t) {
  return t.wave_rows >= 1 && t.wave_cols >= 1 && t.wave_rows * t.wave_cols <= 16 &&
         t.groups_per_cu >= 1 && t.groups_per_cu <= 16;
}

bool lookup_tuned(const std::string& sig, xrp_tile_shape& out) {
  std::lock_guard<std::mutex> g(g_tune_mu);
  auto it = g_tune_table.find(sig);
  if (it != g_tune_table.end()) {
    out = it->second;
    return true;
  }
  const char* path = std::getenv("XRP_TUNE_CACHE");
  if (!path || !*path) return false;
  FILE* f = std::fopen(path, "r");
  if (!f) return false;
  char key[4096];
  xrp_tile_shape t{0, 0, 0, 0};
  bool found = false;
  while (std::fscanf(f, "%4095s %d %d %d", key, &t.wave_rows, &t.wave_cols,
                     &t.groups_per_cu) == 4) {
    if (sig == key && valid_shape(t)) {
      out = t;  // the last record wins
      found = true;
    }
  }
  std::fclose(f);
  if (found) g_tune_table[sig] = out;
  return found;
}

void record_tuned(const std::string& sig, const xrp_tile_shape& t) {
  std::lock_guard<std::mutex> g(g_tune_mu);
  g_tune_table[sig] = t;
  const char* path = std::getenv("XRP_TUNE_CACHE");
  if (!path || !*path || sig.size() >= 4096) return;
  if (FILE* f = std::fopen(path, "a")) {
    std::fprintf(f, "%s %d %d %d\n", sig.c_str(), t.wave_rows, t.wave_cols, t.groups_per_cu);
    std::fclose(f);
  }
}

// Tile the queue for b->shape and capture reset + persistent launches.
int apply_shape(xrp_batch* b) {
  using namespace xrp;
  xrp_context* ctx = b->ctx;
  const int ti = 8 * b->shape.wave_rows, tj = 8 * b->shape.wave_cols;
  std::vector<uint32_t> starts;
  std::vector<size_t> group_start_off;
  for (auto& g : b->groups) {
    group_start_off.push_back(starts.size());
    uint64_t acc = 0;
    for (int it : g.items) {
      starts.push_back((uint32_t)acc);
      const uint64_t tiles = (uint64_t)((b->res[it] + ti - 1) / ti) *
                             (uint64_t)((b->res[it] + tj - 1) / tj) * (uint64_t)b->n_views[it];
      acc += tiles;
    }
    starts.push_back((uint32_t)acc);
    if (acc >= (1ull << 31)) return fail(XRP_ERR_DIMS, "batch holds more than 2^31 tiles");
    g.n_tiles = (uint32_t)acc;
  }
  HIPCHK(b->tile_start.ensure(starts.size() * sizeof(uint32_t)));
  HIPCHK(hipMemcpy(b->tile_start.p, starts.data(), starts.size() * sizeof(uint32_t),
                   hipMemcpyHostToDevice));
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;

  HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  int e = launch_batch_reset((uint32_t*)b->counters.p, (int)b->groups.size(),
                             (unsigned long long*)b->t_first.p,
                             (unsigned long long*)b->t_done.p, b->n_items,
                             (unsigned long long*)b->stats.p, b->n_items * ST_COUNT,
                             ctx->stream);
  for (size_t gi = 0; gi < b->groups.size() && e == 0; ++gi) {
    const auto& g = b->groups[gi];
    BatchGroup bg;
    std::memset(&bg, 0, sizeof bg);
    bg.jobs = (const MarchParams*)b->jobs.p + g.first;
    bg.tile_start = (const uint32_t*)b->tile_start.p + group_start_off[gi];
    bg.counter = (uint32_t*)b->counters.p + gi;
    bg.t_first = (unsigned long long*)b->t_first.p + g.first;
    bg.t_done = (unsigned long long*)b->t_done.p + g.first;
    bg.sink = (unsigned long long*)b->sink.p;
    bg.n_jobs = (int)g.items.size();
    bg.n_tiles = g.n_tiles;
    bg.wave_rows = b->shape.wave_rows;
    bg.wave_cols = b->shape.wave_cols;
    const int wgs = std::max(1, b->n_cu * b->shape.groups_per_cu);
    e = launch_batch(bg, g.prec, g.integ, wgs, ctx->stream);
  }
  hipGraph_t graph = nullptr;
  hipError_t ce = hipStreamEndCapture(ctx->stream, &graph);
  if (e != 0) {
    if (graph) (void)hipGraphDestroy(graph);
    return fail(XRP_ERR_DEVICE, std::string("batch launch capture failed: ") +
                                    hipGetErrorString((hipError_t)e));
  }
  if (ce != hipSuccess)
    return fail(XRP_ERR_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(ce));
  b->graph = graph;
  HIPCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  return XRP_OK;
}

int batch_run_locked(xrp_batch* b, xrp_batch_stats* st) {
  using namespace xrp;
  xrp_context* ctx = b->ctx;
  HIPCHK(hipSetDevice(ctx->device));
  HIPCHK(hipEventRecord(ctx->ev0, ctx->stream));
  HIPCHK(hipGraphLaunch(b->exec, ctx->stream));
  HIPCHK(hipEventRecord(ctx->ev1, ctx->stream));
  std::vector<unsigned long long> hst((size_t)b->n_items * ST_COUNT), t0(b->n_items),
      t1(b->n_items);
  HIPCHK(hipMemcpyAsync(hst.data(), b->stats.p, hst.size() * sizeof(unsigned long long),
                        hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipMemcpyAsync(t0.data(), b->t_first.p, t0.size() * sizeof(unsigned long long),
                        hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipMemcpyAsync(t1.data(), b->t_done.p, t1.size() * sizeof(unsigned long long),
                        hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  // t_first / t_done are stored in job (group) order
  unsigned long long start = ~0ull;
  for (auto v : t0) start = std::min(start, v);
  b->latency.assign(b->n_items, 0.0);
  for (auto& g : b->groups)
    for (size_t k = 0; k < g.items.size(); ++k) {
      const unsigned long long done = t1[g.first + k];
      b->latency[g.items[k]] =
          done >= start && b->ticks_per_ms > 0 ? (double)(done - start) / b->ticks_per_ms : 0.0;
    }
  if (st) {
    xrp_batch_stats o;
    std::memset(&o, 0, sizeof o);
    o.kernel_ms = ms;
    for (int it = 0; it < b->n_items; ++it) {
      o.rays += (uint64_t)b->n_views[it] * b->res[it] * b->res[it];
      o.samples += hst[(size_t)it * ST_COUNT + ST_SAMPLES];
      o.reevaluated += hst[(size_t)it * ST_COUNT + ST_REEVAL];
    }
    for (auto& g : b->groups) o.tiles += g.n_tiles;
    o.launches = (int)b->groups.size();
    write_out(st, o);
  }
  return XRP_OK;
}

}  // namespace

extern "C" {

int xrp_batch_create(xrp_context* ctx, const xrp_batch_item* items, int n_items,
                     xrp_batch** out) {
  return guarded([&]() -> int {
    using namespace xrp;
    if (!ctx || !items || !out) return fail(XRP_ERR_NULL, "NULL argument");
    *out = nullptr;
    if (n_items <= 0) return fail(XRP_ERR_DIMS, "n_items must be positive");
    for (int it = 0; it < n_items; ++it) {
      if (!items[it].scene || !items[it].views)
        return fail(XRP_ERR_NULL, "item " + std::to_string(it) + ": NULL scene or views");
      int rc = check_params(&items[it].params, items[it].n_views);
      if (rc) return fail(rc, "item " + std::to_string(it) + ": " + g_error);
      if (items[it].params.output_format != XRP_OUTPUT_F32)
        return fail(XRP_ERR_ARG, "item " + std::to_string(it) + ": batches write float32 images");
    }
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIPCHK(hipSetDevice(ctx->device));
    std::unique_ptr<xrp_batch> b(new xrp_batch());
    b->ctx = ctx;
    b->n_items = n_items;
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, ctx->device));
    b->n_cu = prop.multiProcessorCount;
    int khz = 0;
    HIPCHK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, ctx->device));
    b->ticks_per_ms = (double)khz;

    // Plan every item; concatenate nodes, deformations, views and tables.
    std::vector<DevNode> all_nodes;
    std::vector<DevDeform> all_defs;
    std::vector<DevView> all_views;
    std::vector<double> all_tables;
    std::vector<size_t> node_off, def_off, view_off, table_off;
    size_t out_floats = 0;
    std::string sig = std::string(prop.gcnArchName) + "/" + std::to_string(b->n_cu);
    for (int it = 0; it < n_items; ++it) {
      xrp_scene* sc = const_cast<xrp_scene*>(items[it].scene);
      RenderJob job = make_job(&items[it].params, items[it].views, items[it].n_views);
      LaunchPlan pl;
      int rc = plan_job(ctx, sc, job, pl, /*allow_voxel_kernel=*/false);
      if (rc) return rc;
      const Scene& s = *sc->s;
      node_off.push_back(all_nodes.size());
      def_off.push_back(all_defs.size());
      view_off.push_back(all_views.size());
      table_off.push_back(all_tables.size());
      all_nodes.insert(all_nodes.end(), pl.nodes.begin(), pl.nodes.end());
      all_defs.insert(all_defs.end(), s.ddefs.begin(), s.ddefs.end());
      all_views.insert(all_views.end(), job.views.begin(), job.views.end());
      all_tables.insert(all_tables.end(), pl.table.begin(), pl.table.end());
      b->res.push_back(job.res);
      b->n_views.push_back((int)job.views.size());
      b->out_off.push_back(out_floats);
      out_floats += (size_t)job.views.size() * job.res * job.res;
      b->mps.push_back(pl.mp);
      b->prec.push_back(pl.precision == XRP_PRECISION_FP64 ? PREC_FP64 : PREC_GUARDED);
      b->integ.push_back(job.integrator == XRP_INTEGRATE_HIERARCHICAL ? INT_HIERARCHICAL
                                                                     : INT_SIMPLE);
      // work per tile ~ nodes x samples in the clip box (x2 for FP64)
      double frac = 1.0;
      if (pl.mp.clip) {
        double ext = 0;
        for (int k = 0; k < 3; ++k) ext += (pl.mp.hi[k] - pl.mp.lo[k]) * (pl.mp.hi[k] - pl.mp.lo[k]);
        frac = std::min(1.0, std::sqrt(ext) / (2 * kCubeHalfDiagonal));
      }
      b->cost.push_back((double)pl.nodes.size() * pl.mp.K * frac *
                        (b->prec.back() == PREC_FP64 ? 2.0 : 1.0));
      sig += ";" + std::to_string(pl.nodes.size()) + "," + std::to_string(job.res) + "," +
             std::to_string(job.views.size()) + "," + std::to_string(pl.mp.K) + "," +
             std::to_string(b->prec.back()) + "," + std::to_string(b->integ.back());
    }
    b->signature = sig;

    // Groups by (precision, integrator).  Within a group the scenes are
    // queued whole, in order of estimated total work (work per tile x
    // pixels), smallest first: each scene then finishes as early as its
    // place allows, which minimises the mean and median completion time,
    // while the tiles of the last, largest scenes still spread over every
    // compute unit.  $XRP_BATCH_ORDER=tile restores the order by work per
    // tile, largest first (for comparison).
    for (int it = 0; it < n_items; ++it) {
      auto g = std::find_if(b->groups.begin(), b->groups.end(), [&](const xrp_batch::Group& x) {
        return x.prec == b->prec[it] && x.integ == b->integ[it];
      });
      if (g == b->groups.end()) {
        b->groups.push_back(xrp_batch::Group());
        g = b->groups.end() - 1;
        g->prec = b->prec[it];
        g->integ = b->integ[it];
      }
      g->items.push_back(it);
    }
    const char* order_env = std::getenv("XRP_BATCH_ORDER");
    const bool by_tile = order_env && std::strcmp(order_env, "tile") == 0;
    auto total = [&](int it) {
      return b->cost[it] * (double)b->n_views[it] * b->res[it] * b->res[it];
    };
    int first = 0;
    for (auto& g : b->groups) {
      if (by_tile)
        std::stable_sort(g.items.begin(), g.items.end(),
                         [&](int a, int c) { return b->cost[a] > b->cost[c]; });
      else
        std::stable_sort(g.items.begin(), g.items.end(),
                         [&](int a, int c) { return total(a) < total(c); });
      g.first = first;
      first += (int)g.items.size();
    }

    // Device copies.
    HIPCHK(b->nodes.ensure(std::max<size_t>(1, all_nodes.size()) * sizeof(DevNode)));
    HIPCHK(b->defs.ensure(std::max<size_t>(1, all_defs.size()) * sizeof(DevDeform)));
    HIPCHK(b->views.ensure(all_views.size() * sizeof(DevView)));
    HIPCHK(b->tables.ensure(std::max<size_t>(1, all_tables.size()) * sizeof(double)));
    HIPCHK(b->jobs.ensure((size_t)n_items * sizeof(MarchParams)));
    HIPCHK(b->counters.ensure(b->groups.size() * sizeof(uint32_t)));
    HIPCHK(b->t_first.ensure((size_t)n_items * sizeof(unsigned long long)));
    HIPCHK(b->t_done.ensure((size_t)n_items * sizeof(unsigned long long)));
    HIPCHK(b->sink.ensure(kBatchSinkBytes));
    HIPCHK(b->stats.ensure((size_t)n_items * ST_COUNT * sizeof(unsigned long long)));
    HIPCHK(b->out.ensure(out_floats * sizeof(float)));
    if (!all_nodes.empty())
      HIPCHK(hipMemcpy(b->nodes.p, all_nodes.data(), all_nodes.size() * sizeof(DevNode),
                       hipMemcpyHostToDevice));
    if (!all_defs.empty())
      HIPCHK(hipMemcpy(b->defs.p, all_defs.data(), all_defs.size() * sizeof(DevDeform),
                       hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b->views.p, all_views.data(), all_views.size() * sizeof(DevView),
                     hipMemcpyHostToDevice));
    if (!all_tables.empty())
      HIPCHK(hipMemcpy(b->tables.p, all_tables.data(), all_tables.size() * sizeof(double),
                       hipMemcpyHostToDevice));
    std::vector<MarchParams> jobs(n_items);
    for (auto& g : b->groups)
      for (size_t k = 0; k < g.items.size(); ++k) {
        const int it = g.items[k];
        MarchParams mp = b->mps[it];
        mp.nodes = (const DevNode*)b->nodes.p + node_off[it];
        mp.defs = (const DevDeform*)b->defs.p + def_off[it];
        mp.views = (const DevView*)b->views.p + view_off[it];
        mp.s_table = (const double*)b->tables.p + table_off[it];
        mp.stats = (unsigned long long*)b->stats.p + (size_t)it * ST_COUNT;
        mp.out = (float*)b->out.p + b->out_off[it];
        mp.view_offset = 0;
        mp.n_views = b->n_views[it];
        b->mps[it] = mp;
        jobs[g.first + k] = mp;
      }
    HIPCHK(hipMemcpy(b->jobs.p, jobs.data(), jobs.size() * sizeof(MarchParams),
                     hipMemcpyHostToDevice));

    xrp_tile_shape tuned;
    if (lookup_tuned(b->signature, tuned)) b->shape = tuned;
    int rc = apply_shape(b.get());
    if (rc) return rc;
    *out = b.release();
    return XRP_OK;
  });
}

int xrp_batch_set_shape(xrp_batch* b, const xrp_tile_shape* shape) {
  return guarded([&]() -> int {
    if (!b || !shape) return fail(XRP_ERR_NULL, "NULL argument");
    if (!valid_shape(*shape))
      return fail(XRP_ERR_ARG, "tile shape: 1 <= rows, cols; rows*cols <= 16; "
                               "1 <= groups_per_cu <= 16");
    std::lock_guard<std::mutex> lk(b->ctx->mu);
    HIPCHK(hipSetDevice(b->ctx->device));
    b->shape = *shape;
    b->shape.reserved = 0;
    return apply_shape(b);
  });
}

int xrp_batch_run(xrp_batch* b, xrp_batch_stats* st) {
  return guarded([&]() -> int {
    if (!b) return fail(XRP_ERR_NULL, "NULL batch");
    int rc = check_out(st, "stats");
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(b->ctx->mu);
    return batch_run_locked(b, st);
  });
}

int xrp_batch_tune(xrp_batch* b, int repeats, xrp_tile_shape* chosen) {
  return guarded([&]() -> int {
    if (!b) return fail(XRP_ERR_NULL, "NULL batch");
    if (repeats < 1) repeats = 1;
    static const xrp_tile_shape cands[] = {
        {1, 1, 2, 0}, {1, 1, 4, 0}, {2, 1, 2, 0}, {2, 2, 2, 0}, {4, 1, 2, 0},
        {4, 2, 2, 0}, {8, 1, 2, 0}, {8, 2, 2, 0}, {4, 4, 2, 0}, {8, 1, 3, 0}};
    std::lock_guard<std::mutex> lk(b->ctx->mu);
    HIPCHK(hipSetDevice(b->ctx->device));
    double best_ms = INFINITY;
    xrp_tile_shape best = b->shape;
    for (const auto& c : cands) {
      b->shape = c;
      int rc = apply_shape(b);
      if (rc) return rc;
      double ms = INFINITY;
      for (int r = 0; r < repeats; ++r) {
        xrp_batch_stats st;
        st.struct_size = sizeof st;
        rc = batch_run_locked(b, &st);
        if (rc) return rc;
        ms = std::min(ms, st.kernel_ms);
      }
      if (ms < best_ms) {
        best_ms = ms;
        best = c;
      }
    }
    b->shape = best;
    int rc = apply_shape(b);
    if (rc) return rc;
    record_tuned(b->signature, best);
    if (chosen) *chosen = best;
    return XRP_OK;
  });
}

int xrp_batch_output(const xrp_batch* b, int item, const float** device_images) {
  return guarded([&]() -> int {
    if (!b || !device_images) return fail(XRP_ERR_NULL, "NULL argument");
    if (item < 0 || item >= b->n_items) return fail(XRP_ERR_ARG, "item out of range");
    *device_images = (const float*)b->out.p + b->out_off[item];
    return XRP_OK;
  });
}

int xrp_batch_read(xrp_batch* b, int item, float* host_images) {
  return guarded([&]() -> int {
    if (!b || !host_images) return fail(XRP_ERR_NULL, "NULL argument");
    if (item < 0 || item >= b->n_items) return fail(XRP_ERR_ARG, "item out of range");
    std::lock_guard<std::mutex> lk(b->ctx->mu);
    HIPCHK(hipSetDevice(b->ctx->device));
    const size_t n = (size_t)b->n_views[item] * b->res[item] * b->res[item];
    HIPCHK(hipMemcpy(host_images, (const float*)b->out.p + b->out_off[item], n * sizeof(float),
                     hipMemcpyDeviceToHost));
    return XRP_OK;
  });
}

int xrp_batch_latencies(const xrp_batch* b, double* ms) {
  return guarded([&]() -> int {
    if (!b || !ms) return fail(XRP_ERR_NULL, "NULL argument");
    for (int it = 0; it < b->n_items; ++it)
      ms[it] = it < (int)b->latency.size() ? b->latency[it] : 0.0;
    return XRP_OK;
  });
}

void xrp_batch_destroy(xrp_batch* b) {
  if (!b) return;
  {
    std::lock_guard<std::mutex> lk(b->ctx->mu);
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
  }
  delete b;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Learned density surrogate.
// ---------------------------------------------------------------------------
#include "surrogate.hpp"

struct xrp_surrogate {
  xrp_context* ctx = nullptr;
  xrp_scene* scene = nullptr;
  xrp::SurNet net;
  int log2_table = 19;
  int64_t count = 0;        // parameters
  int64_t table_count = 0;  // of which tables
  DevBuf params, grads, m, v, tbf16, loss, pts, labels, scene_nodes, items, vals;
  DevBuf judge, counters;  // certified check: exact images, tile counters
  int adam_step = 0;
  bool packed = false;
  ~xrp_surrogate() {
    for (DevBuf* b : {&params, &grads, &m, &v, &tbf16, &loss, &pts, &labels, &scene_nodes,
                      &items, &vals, &judge, &counters})
      b->release();
  }
};

namespace {

double rho_scale(const xrp::HNode& n) {
  using namespace xrp;
  if (n.kind == K_COLL) {
    double mx = 0, sum = 0;
    for (auto& k : n.kids) {
      const double r = rho_scale(k);
      mx = std::max(mx, r);
      sum += r;
    }
    return n.greedy ? mx : std::min(1.0, sum);
  }
  if (n.kind == K_TESS) return n.kids.empty() ? 0.0 : rho_scale(n.kids[0]);
  if (n.kind == K_VOXEL) return 1.0;
  return std::max(0.0, n.rho);
}

// Gaussian-free deterministic initial values: tables U(-1e-4, 1e-4),
// weights U(-sqrt(6/(in+out)), +) (Glorot), biases 0.
void init_params(std::vector<float>& p, int log2_table, uint64_t seed) {
  uint64_t st = seed * 0x9E3779B97F4A7C15ull + 12345;
  auto next = [&]() {
    st += 0x9E3779B97F4A7C15ull;
    uint64_t z = st;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
  };
  const size_t t = xrp::sur_table_floats(log2_table);
  for (size_t k = 0; k < t; ++k) p[k] = (float)((2 * next() - 1) * 1e-4);
  size_t o = t;
  auto dense = [&](int in, int out) {
    const double a = std::sqrt(6.0 / (in + out));
    for (int k = 0; k < in * out; ++k) p[o++] = (float)((2 * next() - 1) * a);
    for (int k = 0; k < out; ++k) p[o++] = 0.0f;  // bias
  };
  dense(xrp::kSurIn, xrp::kSurHidden);
  dense(xrp::kSurHidden, xrp::kSurHidden);
  dense(xrp::kSurHidden, 1);
}

}  // namespace

extern "C" {

void xrp_surrogate_default_config(xrp_surrogate_config* c) {
  if (!c) return;
  std::memset(c, 0, sizeof *c);
  c->struct_size = sizeof *c;
  c->log2_table = 19;
  c->n_min = 16;
  c->n_max = 512;
  c->wrap = -1;
  c->seed = 1;
}

void xrp_train_default_config(xrp_train_config* c) {
  if (!c) return;
  std::memset(c, 0, sizeof *c);
  c->struct_size = sizeof *c;
  c->batch = 65536;
  c->samples = (int64_t)1 << 24;
  c->epochs = 4;
  c->lr_table = 1e-2;
  c->lr_mlp = 1e-3;
}

int xrp_surrogate_create(xrp_context* ctx, const xrp_scene* scene,
                         const xrp_surrogate_config* cfg, xrp_surrogate** out) {
  return guarded([&]() -> int {
    using namespace xrp;
    if (!ctx || !scene || !cfg || !out) return fail(XRP_ERR_NULL, "NULL argument");
    *out = nullptr;
    if (cfg->struct_size < sizeof(xrp_surrogate_config))
      return fail(XRP_ERR_ARG, "config.struct_size too small");
    if (cfg->log2_table < 4 || cfg->log2_table > 24)
      return fail(XRP_ERR_ARG, "log2_table must be in 4..24");
    if (cfg->n_min < 1 || cfg->n_max < cfg->n_min || cfg->n_max > 65535)
      return fail(XRP_ERR_ARG, "need 1 <= n_min <= n_max <= 65535");
    const Scene& s = *scene->s;
    if (s.has_voxels) return fail(XRP_ERR_UNSUPPORTED, "surrogate of voxel scenes");
    std::unique_ptr<xrp_surrogate> sr(new xrp_surrogate());
    sr->ctx = ctx;
    sr->scene = const_cast<xrp_scene*>(scene);
    sr->log2_table = cfg->log2_table;
    SurNet& net = sr->net;
    std::memset(&net, 0, sizeof net);
    net.log2_table = cfg->log2_table;
    const double b = std::exp((std::log((double)cfg->n_max) - std::log((double)cfg->n_min)) /
                              (kSurLevels - 1));
    for (int l = 0; l < kSurLevels; ++l)
      net.res[l] = (int)std::floor((double)cfg->n_min * std::pow(b, (double)l));
    const bool tess = s.root.kind == K_TESS;
    net.wrap = cfg->wrap < 0 ? (tess ? 1 : 0) : cfg->wrap;
    if (net.wrap && !tess) return fail(XRP_ERR_ARG, "wrap needs a tessellated root");
    double lo[3], ext[3];
    if (net.wrap) {
      for (int k = 0; k < 3; ++k) {
        lo[k] = s.root.ucb[2 * k];
        ext[k] = s.root.ucb[2 * k + 1] - s.root.ucb[2 * k];
      }
      for (int k = 0; k < 6; ++k) net.bounds[k] = (float)s.root.bounds[k];
    } else {
      const Box3& ob = s.obj_box;
      if (ob.empty() || !ob.finite())
        return fail(XRP_ERR_UNSUPPORTED, "surrogate needs a bounded object");
      for (int k = 0; k < 3; ++k) {
        lo[k] = ob.lo[k];
        ext[k] = ob.hi[k] - ob.lo[k];
      }
    }
    for (int k = 0; k < 3; ++k) {
      if (!(ext[k] > 0)) return fail(XRP_ERR_ARG, "degenerate surrogate domain");
      net.lo[k] = (float)lo[k];
      net.ext[k] = (float)ext[k];
      net.inv_ext[k] = (float)(1.0 / ext[k]);
    }
    net.rho_scale = (float)std::max(1e-6, rho_scale(s.root));
    sr->count = (int64_t)sur_param_count(cfg->log2_table);
    sr->table_count = (int64_t)sur_table_floats(cfg->log2_table);
    std::vector<float> host((size_t)sr->count);
    init_params(host, cfg->log2_table, cfg->seed);
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIPCHK(hipSetDevice(ctx->device));
    const size_t bytes = (size_t)sr->count * sizeof(float);
    HIPCHK(sr->params.ensure(bytes));
    HIPCHK(sr->grads.ensure(bytes));
    HIPCHK(sr->m.ensure(bytes));
    HIPCHK(sr->v.ensure(bytes));
    HIPCHK(sr->tbf16.ensure(((size_t)kSurLevels << cfg->log2_table) * sizeof(uint32_t)));
    HIPCHK(sr->loss.ensure(sizeof(double)));
    HIPCHK(hipMemcpy(sr->params.p, host.data(), bytes, hipMemcpyHostToDevice));
    HIPCHK(hipMemset(sr->m.p, 0, bytes));
    HIPCHK(hipMemset(sr->v.p, 0, bytes));
    HIPCHK(sr->scene_nodes.ensure(s.nodes.size() * sizeof(DevNode)));
    HIPCHK(hipMemcpy(sr->scene_nodes.p, s.nodes.data(), s.nodes.size() * sizeof(DevNode),
                     hipMemcpyHostToDevice));
    net.params = (float*)sr->params.p;
    net.table_bf16 = (uint32_t*)sr->tbf16.p;
    *out = sr.release();
    return XRP_OK;
  });
}

int64_t xrp_surrogate_param_count(const xrp_surrogate* s) { return s ? s->count : 0; }

int xrp_surrogate_get_params(const xrp_surrogate* s, float* host) {
  return guarded([&]() -> int {
    if (!s || !host) return fail(XRP_ERR_NULL, "NULL argument");
    std::lock_guard<std::mutex> lk(s->ctx->mu);
    HIPCHK(hipSetDevice(s->ctx->device));
    HIPCHK(hipMemcpy(host, s->params.p, (size_t)s->count * sizeof(float), hipMemcpyDeviceToHost));
    return XRP_OK;
  });
}

int xrp_surrogate_set_params(xrp_surrogate* s, const float* host) {
  return guarded([&]() -> int {
    if (!s || !host) return fail(XRP_ERR_NULL, "NULL argument");
    std::lock_guard<std::mutex> lk(s->ctx->mu);
    HIPCHK(hipSetDevice(s->ctx->device));
    HIPCHK(hipMemcpy(s->params.p, host, (size_t)s->count * sizeof(float), hipMemcpyHostToDevice));
    s->packed = false;
    return XRP_OK;
  });
}

int xrp_surrogate_predict(xrp_surrogate* s, const float* u_host, int64_t n, float* out_host) {
  return guarded([&]() -> int {
    using namespace xrp;
    if (!s || !u_host || !out_host) return fail(XRP_ERR_NULL, "NULL argument");
    if (n <= 0) return fail(XRP_ERR_DIMS, "n must be positive");
    std::lock_guard<std::mutex> lk(s->ctx->mu);
    HIPCHK(hipSetDevice(s->ctx->device));
    HIPCHK(s->pts.ensure((size_t)n * 3 * sizeof(float)));
    HIPCHK(s->labels.ensure((size_t)n * sizeof(float)));
    HIPCHK(hipMemcpy(s->pts.p, u_host, (size_t)n * 3 * sizeof(float), hipMemcpyHostToDevice));
    int e = launch_sur_predict(s->net, (const float*)s->pts.p, n, (float*)s->labels.p,
                               s->ctx->stream);
    if (e) return fail(XRP_ERR_DEVICE, hipGetErrorString((hipError_t)e));
    HIPCHK(hipStreamSynchronize(s->ctx->stream));
    HIPCHK(hipMemcpy(out_host, s->labels.p, (size_t)n * sizeof(float), hipMemcpyDeviceToHost));
    return XRP_OK;
  });
}

int xrp_surrogate_gradients(xrp_surrogate* s, const float* u_host, const float* targets,
                            int64_t n, float* grads_host, double* loss) {
  return guarded([&]() -> int {
    using namespace xrp;
    if (!s || !u_host || !targets || !grads_host) return fail(XRP_ERR_NULL, "NULL argument");
    if (n <= 0) return fail(XRP_ERR_DIMS, "n must be positive");
    std::lock_guard<std::mutex> lk(s->ctx->mu);
    HIPCHK(hipSetDevice(s->ctx->device));
    HIPCHK(s->pts.ensure((size_t)n * 3 * sizeof(float)));
    HIPCHK(s->labels.ensure((size_t)n * sizeof(float)));
    HIPCHK(hipMemcpy(s->pts.p, u_host, (size_t)n * 3 * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(s->labels.p, targets, (size_t)n * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(hipMemsetAsync(s->grads.p, 0, (size_t)s->count * sizeof(float), s->ctx->stream));
    HIPCHK(hipMemsetAsync(s->loss.p, 0, sizeof(double), s->ctx->stream));
    int e = launch_sur_train(s->net, (const float*)s->pts.p, (const float*)s->labels.p, n,
                             (float)(1.0 / (double)n), (float*)s->grads.p, (double*)s->loss.p,
                             s->ctx->stream);
    if (e) return fail(XRP_ERR_DEVICE, hipGetErrorString((hipError_t)e));
    HIPCHK(hipStreamSynchronize(s->ctx->stream));
    HIPCHK(hipMemcpy(grads_host, s->grads.p, (size_t)s->count * sizeof(float),
                     hipMemcpyDeviceToHost));
    double l = 0;
    HIPCHK(hipMemcpy(&l, s->loss.p, sizeof l, hipMemcpyDeviceToHost));
    if (loss) *loss = l / (double)n;
    return XRP_OK;
  });
}

int xrp_surrogate_train(xrp_surrogate* s, const xrp_train_config* cfg, xrp_train_stats* st) {
  return guarded([&]() -> int {
    using namespace xrp;
    if (!s || !cfg) return fail(XRP_ERR_NULL, "NULL argument");
    if (cfg->struct_size < sizeof(xrp_train_config))
      return fail(XRP_ERR_ARG, "config.struct_size too small");
    if (cfg->batch <= 0 || cfg->samples <= 0 || cfg->epochs <= 0)
      return fail(XRP_ERR_ARG, "batch, samples and epochs must be positive");
    int rc0 = check_out(st, "stats");
    if (rc0) return rc0;
    xrp_context* ctx = s->ctx;
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIPCHK(hipSetDevice(ctx->device));
    const int64_t n = cfg->samples;
    HIPCHK(s->pts.ensure((size_t)n * 3 * sizeof(float)));
    HIPCHK(s->labels.ensure((size_t)n * sizeof(float)));
    MarchParams scene;
    std::memset(&scene, 0, sizeof scene);
    scene.nodes = (const DevNode*)s->scene_nodes.p;
    hipEvent_t e0 = ctx->ev0, e1 = ctx->ev1;
    HIPCHK(hipEventRecord(e0, ctx->stream));
    int e = launch_sur_labels(s->net, scene, (uint64_t)(s->adam_step + 1) * 7919ull, n,
                              (float*)s->pts.p, (float*)s->labels.p, ctx->stream);
    if (e) return fail(XRP_ERR_DEVICE, hipGetErrorString((hipError_t)e));
    HIPCHK(hipEventRecord(e1, ctx->stream));
    HIPCHK(hipEventSynchronize(e1));
    float label_ms = 0;
    HIPCHK(hipEventElapsedTime(&label_ms, e0, e1));

    double first = NAN, epoch_loss = 0;
    int steps = 0;
    HIPCHK(hipEventRecord(e0, ctx->stream));
    for (int ep = 0; ep < cfg->epochs; ++ep) {
      HIPCHK(hipMemsetAsync(s->loss.p, 0, sizeof(double), ctx->stream));
      for (int64_t b0 = 0; b0 < n; b0 += cfg->batch) {
        const int64_t nb = std::min<int64_t>(cfg->batch, n - b0);
        HIPCHK(hipMemsetAsync(s->grads.p, 0, (size_t)s->count * sizeof(float), ctx->stream));
        e = launch_sur_train(s->net, (const float*)s->pts.p + 3 * b0,
                             (const float*)s->labels.p + b0, nb, (float)(1.0 / (double)nb),
                             (float*)s->grads.p, (double*)s->loss.p, ctx->stream);
        if (!e)
          e = launch_sur_adam((float*)s->params.p, (const float*)s->grads.p, (float*)s->m.p,
                              (float*)s->v.p, s->count, s->table_count, (float)cfg->lr_table,
                              (float)cfg->lr_mlp, 0.9f, 0.99f, 1e-15f, ++s->adam_step,
                              ctx->stream);
        if (e) return fail(XRP_ERR_DEVICE, hipGetErrorString((hipError_t)e));
        ++steps;
        if (std::isnan(first)) {
          double l = 0;
          HIPCHK(hipStreamSynchronize(ctx->stream));
          HIPCHK(hipMemcpy(&l, s->loss.p, sizeof l, hipMemcpyDeviceToHost));
          first = l / (double)nb;
        }
      }
      double l = 0;
      HIPCHK(hipStreamSynchronize(ctx->stream));
      HIPCHK(hipMemcpy(&l, s->loss.p, sizeof l, hipMemcpyDeviceToHost));
      epoch_loss = l / (double)n;
    }
    HIPCHK(hipEventRecord(e1, ctx->stream));
    HIPCHK(hipEventSynchronize(e1));
    float train_ms = 0;
    HIPCHK(hipEventElapsedTime(&train_ms, e0, e1));
    s->packed = false;
    if (st) {
      xrp_train_stats o;
      std::memset(&o, 0, sizeof o);
      o.first_loss = first;
      o.final_loss = epoch_loss;
      o.label_ms = label_ms;
      o.train_ms = train_ms;
      o.steps = steps;
      write_out(st, o);
    }
    return XRP_OK;
  });
}

int xrp_surrogate_render_device(xrp_surrogate* s, const xrp_render_params* p,
                                const xrp_view_angle* views, int n_views, double ds_sur,
                                double tol, void* out_device, xrp_surrogate_stats* st) {
  return guarded([&]() -> int {
    using namespace xrp;
    if (!s || !views || !out_device) return fail(XRP_ERR_NULL, "NULL argument");
    int rc = check_params(p, n_views);
    if (!rc) rc = check_out(st, "stats");
    if (rc) return rc;
    if (p->output_format != XRP_OUTPUT_F32)
      return fail(XRP_ERR_ARG, "the surrogate writes float32 images");
    if (!(ds_sur > 0) || !std::isfinite(ds_sur)) return fail(XRP_ERR_ARG, "ds_surrogate > 0");
    if (n_views > 65535) return fail(XRP_ERR_DIMS, "at most 65535 views per call");
    xrp_context* ctx = s->ctx;
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIPCHK(hipSetDevice(ctx->device));
    RenderJob job = make_job(p, views, n_views);
    LaunchPlan pl;
    rc = plan_job(ctx, s->scene, job, pl, /*allow_voxel_kernel=*/false);
    if (rc) return rc;
    const Scene& sc = *s->scene->s;
    const int res = job.res;
    // exact plan on the context buffers
    HIPCHK(ctx->table.ensure(std::max<size_t>(1, pl.table.size()) * sizeof(double)));
    if (!pl.table.empty())
      HIPCHK(hipMemcpy(ctx->table.p, pl.table.data(), pl.table.size() * sizeof(double),
                       hipMemcpyHostToDevice));
    ctx->table_mode = -1;  // the cached table key no longer matches
    HIPCHK(ctx->nodes.ensure(pl.nodes.size() * sizeof(DevNode)));
    HIPCHK(ctx->defs.ensure(std::max<size_t>(1, sc.ddefs.size()) * sizeof(DevDeform)));
    HIPCHK(ctx->views.ensure(job.views.size() * sizeof(DevView)));
    HIPCHK(ctx->stats.ensure(ST_COUNT * sizeof(unsigned long long)));
    HIPCHK(hipMemcpy(ctx->nodes.p, pl.nodes.data(), pl.nodes.size() * sizeof(DevNode),
                     hipMemcpyHostToDevice));
    if (!sc.ddefs.empty())
      HIPCHK(hipMemcpy(ctx->defs.p, sc.ddefs.data(), sc.ddefs.size() * sizeof(DevDeform),
                       hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->views.p, job.views.data(), job.views.size() * sizeof(DevView),
                     hipMemcpyHostToDevice));
    MarchParams mp = pl.mp;
    mp.nodes = (const DevNode*)ctx->nodes.p;
    mp.defs = (const DevDeform*)ctx->defs.p;
    mp.views = (const DevView*)ctx->views.p;
    mp.s_table = (const double*)ctx->table.p;
    mp.stats = nullptr;
    mp.out = (float*)out_device;
    mp.view_offset = 0;
    mp.n_views = n_views;
    if (!s->packed) {
      int e = launch_sur_pack_tables(s->net, ctx->stream);
      if (e) return fail(XRP_ERR_DEVICE, hipGetErrorString((hipError_t)e));
      s->packed = true;
    }
    xrp_surrogate_stats out_st;
    std::memset(&out_st, 0, sizeof out_st);
    HIPCHK(hipMemsetAsync(ctx->stats.p, 0, ST_COUNT * sizeof(unsigned long long), ctx->stream));
    MarchParams smp = mp;
    smp.stats = (unsigned long long*)ctx->stats.p;
    HIPCHK(hipEventRecord(ctx->ev0, ctx->stream));
    int e = launch_sur_render(s->net, smp, ds_sur, ctx->stream);
    if (e) return fail(XRP_ERR_DEVICE, hipGetErrorString((hipError_t)e));
    HIPCHK(hipEventRecord(ctx->ev1, ctx->stream));
    HIPCHK(hipEventSynchronize(ctx->ev1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    out_st.surrogate_ms = ms;
    unsigned long long nsamp = 0;
    HIPCHK(hipMemcpy(&nsamp, (unsigned long long*)ctx->stats.p + ST_SAMPLES, sizeof nsamp,
                     hipMemcpyDeviceToHost));
    out_st.network_samples = nsamp;
    const int tiles_side = (res + 15) / 16;
    out_st.tiles = (uint64_t)tiles_side * tiles_side * n_views;
    const int prec_x = pl.precision == XRP_PRECISION_FP64 ? PREC_FP64 : PREC_GUARDED;
    const int integ_x = job.integrator == XRP_INTEGRATE_HIERARCHICAL ? INT_HIERARCHICAL
                                                                     : INT_SIMPLE;
    if (tol < 0) {
      // Certified: the exact march renders every pixel as the judge, and
      // every tile whose largest difference exceeds |tol| is replaced by
      // the exact values, so no pixel of the result is further than |tol|
      // from the exact render.
      HIPCHK(s->judge.ensure((size_t)n_views * res * res * sizeof(float)));
      HIPCHK(s->counters.ensure(2 * sizeof(unsigned)));
      HIPCHK(hipMemsetAsync(ctx->stats.p, 0, ST_COUNT * sizeof(unsigned long long),
                            ctx->stream));
      HIPCHK(hipMemsetAsync(s->counters.p, 0, 2 * sizeof(unsigned), ctx->stream));
      MarchParams ep = mp;
      ep.out = (float*)s->judge.p;
      ep.stats = (unsigned long long*)ctx->stats.p;
      HIPCHK(hipEventRecord(ctx->ev0, ctx->stream));
      e = launch_march(ep, prec_x, integ_x, ctx->stream);
      if (!e)
        e = launch_sur_certify((float*)out_device, (const float*)s->judge.p, n_views, res,
                               (float)-tol, (unsigned*)s->counters.p, ctx->stream);
      if (e) return fail(XRP_ERR_DEVICE, hipGetErrorString((hipError_t)e));
      HIPCHK(hipEventRecord(ctx->ev1, ctx->stream));
      HIPCHK(hipEventSynchronize(ctx->ev1));
      HIPCHK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
      unsigned cnt[2];
      unsigned long long st_x[ST_COUNT];
      HIPCHK(hipMemcpy(cnt, s->counters.p, sizeof cnt, hipMemcpyDeviceToHost));
      HIPCHK(hipMemcpy(st_x, ctx->stats.p, sizeof st_x, hipMemcpyDeviceToHost));
      float worst;
      std::memcpy(&worst, &cnt[1], sizeof worst);
      out_st.check_ms = ms;
      out_st.check_pixels = (uint64_t)n_views * res * res;
      out_st.tiles_failed = cnt[0];
      out_st.max_sampled_error = worst;
      out_st.exact_samples = st_x[ST_SAMPLES] - st_x[ST_SKIPPED];
    } else if (tol > 0) {
      // kCheck check pixels per 16x16 tile on a stride-4 grid (rows and
      // columns 1, 5, 9, 13 of the tile); a tile passes when every one is
      // within tol / 2, the other half of the tolerance being the margin
      // left for the pixels between them
      constexpr int kCheck = 16;
      std::vector<int> items;
      items.reserve(out_st.tiles * 3 * kCheck);
      for (int v = 0; v < n_views; ++v)
        for (int ti = 0; ti < tiles_side; ++ti)
          for (int tj = 0; tj < tiles_side; ++tj)
            for (int c = 0; c < kCheck; ++c) {
              items.push_back(v);
              items.push_back(std::min(res - 1, ti * 16 + 1 + 4 * (c >> 2)));
              items.push_back(std::min(res - 1, tj * 16 + 1 + 4 * (c & 3)));
            }
      const int n_check = (int)(items.size() / 3);
      HIPCHK(s->items.ensure(items.size() * sizeof(int)));
      HIPCHK(s->vals.ensure((size_t)n_check * sizeof(float)));
      HIPCHK(hipMemcpy(s->items.p, items.data(), items.size() * sizeof(int),
                       hipMemcpyHostToDevice));
      // the exact marches count their samples (evaluated + settled) here
      HIPCHK(hipMemsetAsync(ctx->stats.p, 0, ST_COUNT * sizeof(unsigned long long),
                            ctx->stream));
      mp.stats = (unsigned long long*)ctx->stats.p;
      const int prec = pl.precision == XRP_PRECISION_FP64 ? PREC_FP64 : PREC_GUARDED;
      const int integ = job.integrator == XRP_INTEGRATE_HIERARCHICAL ? INT_HIERARCHICAL
                                                                     : INT_SIMPLE;
      HIPCHK(hipEventRecord(ctx->ev0, ctx->stream));
      e = launch_march_list(mp, prec, integ, (const int*)s->items.p, n_check,
                            (float*)s->vals.p, ctx->stream);
      if (e) return fail(XRP_ERR_DEVICE, hipGetErrorString((hipError_t)e));
      HIPCHK(hipEventRecord(ctx->ev1, ctx->stream));
      HIPCHK(hipEventSynchronize(ctx->ev1));
      HIPCHK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
      out_st.check_ms = ms;
      out_st.check_pixels = (uint64_t)n_check;
      std::vector<float> exact((size_t)n_check);
      HIPCHK(hipMemcpy(exact.data(), s->vals.p, exact.size() * sizeof(float),
                       hipMemcpyDeviceToHost));
      // the surrogate's values at the same pixels (vals is reused)
      e = launch_gather_pixels((const float*)out_device, (const int*)s->items.p, n_check, res,
                               (float*)s->vals.p, ctx->stream);
      if (e) return fail(XRP_ERR_DEVICE, hipGetErrorString((hipError_t)e));
      std::vector<float> sur((size_t)n_check);
      HIPCHK(hipStreamSynchronize(ctx->stream));
      HIPCHK(hipMemcpy(sur.data(), s->vals.p, sur.size() * sizeof(float), hipMemcpyDeviceToHost));
      std::vector<int> fail_tiles;
      double worst = 0;
      for (int t = 0; t < n_check / kCheck; ++t) {
        double err = 0;
        for (int q = 0; q < kCheck; ++q)
          err = std::max(err, (double)std::fabs(sur[kCheck * t + q] - exact[kCheck * t + q]));
        worst = std::max(worst, err);
        if (err > 0.5 * tol) {
          fail_tiles.push_back(items[3 * kCheck * t]);
          fail_tiles.push_back(items[3 * kCheck * t + 1] / 16);
          fail_tiles.push_back(items[3 * kCheck * t + 2] / 16);
        }
      }
      out_st.max_sampled_error = worst;
      out_st.tiles_failed = fail_tiles.size() / 3;
      if (!fail_tiles.empty()) {
        HIPCHK(s->items.ensure(fail_tiles.size() * sizeof(int)));
        HIPCHK(hipMemcpy(s->items.p, fail_tiles.data(), fail_tiles.size() * sizeof(int),
                         hipMemcpyHostToDevice));
        HIPCHK(hipEventRecord(ctx->ev0, ctx->stream));
        e = launch_march_list(mp, prec, integ, (const int*)s->items.p,
                              (int)(fail_tiles.size() / 3), nullptr, ctx->stream);
        if (e) return fail(XRP_ERR_DEVICE, hipGetErrorString((hipError_t)e));
        HIPCHK(hipEventRecord(ctx->ev1, ctx->stream));
        HIPCHK(hipEventSynchronize(ctx->ev1));
        HIPCHK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        out_st.fallback_ms = ms;
      }
      unsigned long long cnt[ST_COUNT];
      HIPCHK(hipMemcpy(cnt, ctx->stats.p, sizeof cnt, hipMemcpyDeviceToHost));
      out_st.exact_samples = cnt[ST_SAMPLES] - cnt[ST_SKIPPED];
    }
    write_out(st, out_st);
    return XRP_OK;
  });
}

int xrp_surrogate_render(xrp_surrogate* s, const xrp_render_params* p,
                         const xrp_view_angle* views, int n_views, double ds_sur, double tol,
                         float* out_host, xrp_surrogate_stats* st) {
  return guarded([&]() -> int {
    if (!s || !p || !out_host) return fail(XRP_ERR_NULL, "NULL argument");
    int rc = check_params(p, n_views);
    if (rc) return rc;
    const size_t bytes = (size_t)n_views * p->resolution * p->resolution * sizeof(float);
    void* dev = nullptr;
    {
      std::lock_guard<std::mutex> lk(s->ctx->mu);
      HIPCHK(hipSetDevice(s->ctx->device));
      HIPCHK(s->ctx->out.ensure(bytes));
      dev = s->ctx->out.p;
    }
    rc = xrp_surrogate_render_device(s, p, views, n_views, ds_sur, tol, dev, st);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(s->ctx->mu);
    HIPCHK(hipMemcpy(out_host, dev, bytes, hipMemcpyDeviceToHost));
    return XRP_OK;
  });
}

void xrp_surrogate_destroy(xrp_surrogate* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->ctx->mu);
    (void)hipSetDevice(s->ctx->device);
    (void)hipStreamSynchronize(s->ctx->stream);
  }
  delete s;
}

int xrp_diag_mfma_bf16(xrp_context* ctx, const float* a, const float* b, float* d) {
  return guarded([&]() -> int {
    if (!ctx || !a || !b || !d) return fail(XRP_ERR_NULL, "NULL argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIPCHK(hipSetDevice(ctx->device));
    DevBuf buf;
    HIPCHK(buf.ensure(3 * 256 * sizeof(float)));
    float* p = (float*)buf.p;
    hipError_t he = hipMemcpy(p, a, 256 * sizeof(float), hipMemcpyHostToDevice);
    if (he == hipSuccess) he = hipMemcpy(p + 256, b, 256 * sizeof(float), hipMemcpyHostToDevice);
    int e = he == hipSuccess ? xrp::launch_mfma_probe(p, p + 256, p + 512, ctx->stream) : 0;
    if (he == hipSuccess && e == 0) he = hipStreamSynchronize(ctx->stream);
    if (he == hipSuccess && e == 0)
      he = hipMemcpy(d, p + 512, 256 * sizeof(float), hipMemcpyDeviceToHost);
    buf.release();
    if (e) return fail(XRP_ERR_DEVICE, hipGetErrorString((hipError_t)e));
    if (he != hipSuccess) return fail(XRP_ERR_DEVICE, hipGetErrorString(he));
    return XRP_OK;
  });
}

int xrp_diag_sincos_error(xrp_context* ctx, double* max_sin, double* max_cos) {
  return guarded([&]() -> int {
    if (!ctx || !max_sin || !max_cos) return fail(XRP_ERR_NULL, "NULL argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIPCHK(hipSetDevice(ctx->device));
    DevBuf out;
    HIPCHK(out.ensure(2 * sizeof(unsigned)));
    hipError_t he = hipMemsetAsync(out.p, 0, 2 * sizeof(unsigned), ctx->stream);
    if (he == hipSuccess) {
      int e = xrp::launch_sincos_error((unsigned*)out.p, ctx->stream);
      if (e) he = (hipError_t)e;
    }
    unsigned h[2] = {0, 0};
    if (he == hipSuccess)
      he = hipMemcpyAsync(h, out.p, sizeof h, hipMemcpyDeviceToHost, ctx->stream);
    if (he == hipSuccess) he = hipStreamSynchronize(ctx->stream);
    out.release();
    if (he != hipSuccess) return fail(XRP_ERR_DEVICE, hipGetErrorString(he));
    float fs, fc;
    std::memcpy(&fs, &h[0], sizeof fs);
    std::memcpy(&fc, &h[1], sizeof fc);
    *max_sin = fs;
    *max_cos = fc;
    return XRP_OK;
  });
}

int xrp_diag_deform_error(xrp_context* ctx, double* max_tanh, double* max_sigmoid,
                          double* max_expneg) {
  return guarded([&]() -> int {
    if (!ctx || !max_tanh || !max_sigmoid || !max_expneg)
      return fail(XRP_ERR_NULL, "NULL argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIPCHK(hipSetDevice(ctx->device));
    DevBuf out;
    HIPCHK(out.ensure(3 * sizeof(unsigned long long)));
    hipError_t he = hipMemsetAsync(out.p, 0, 3 * sizeof(unsigned long long), ctx->stream);
    if (he == hipSuccess) {
      int e = xrp::launch_deform_error((unsigned long long*)out.p, ctx->stream);
      if (e) he = (hipError_t)e;
    }
    unsigned long long h[3] = {0, 0, 0};
    if (he == hipSuccess)
      he = hipMemcpyAsync(h, out.p, sizeof h, hipMemcpyDeviceToHost, ctx->stream);
    if (he == hipSuccess) he = hipStreamSynchronize(ctx->stream);
    out.release();
    if (he != hipSuccess) return fail(XRP_ERR_DEVICE, hipGetErrorString(he));
    std::memcpy(max_tanh, &h[0], sizeof(double));
    std::memcpy(max_sigmoid, &h[1], sizeof(double));
    std::memcpy(max_expneg, &h[2], sizeof(double));
    return XRP_OK;
  });
}

int xrp_diag_l2_bandwidth(xrp_context* ctx, double* gbps) {
  return guarded([&]() -> int {
    if (!ctx || !gbps) return fail(XRP_ERR_NULL, "NULL argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIPCHK(hipSetDevice(ctx->device));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, ctx->device));
    const int workgroups = 8 * std::max(1, prop.multiProcessorCount);
    const int iters = 4096;
    const size_t bytes = xrp::l2_probe_buffer_bytes();
    DevBuf buf, out;
    HIPCHK(buf.ensure(bytes));
    HIPCHK(out.ensure((size_t)workgroups * 256 * sizeof(unsigned)));
    hipError_t he = hipMemsetAsync(buf.p, 1, bytes, ctx->stream);
    float best = 0.0f;
    for (int rep = 0; rep < 4 && he == hipSuccess; ++rep) {
      he = hipEventRecord(ctx->ev0, ctx->stream);
      int e = he == hipSuccess ? xrp::launch_l2_probe(buf.p, workgroups, iters,
                                                      (unsigned*)out.p, ctx->stream)
                               : 0;
      if (e) he = (hipError_t)e;
      if (he == hipSuccess) he = hipEventRecord(ctx->ev1, ctx->stream);
      if (he == hipSuccess) he = hipEventSynchronize(ctx->ev1);
      float ms = 0.0f;
      if (he == hipSuccess) he = hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
      if (he == hipSuccess && rep > 0 && ms > 0.0f && (best == 0.0f || ms < best)) best = ms;
    }
    buf.release();
    out.release();
    if (he != hipSuccess) return fail(XRP_ERR_DEVICE, hipGetErrorString(he));
    const double read = (double)workgroups * 256.0 * iters * 16.0;
    *gbps = best > 0.0f ? read / (best * 1e-3) / 1e9 : 0.0;
    return XRP_OK;
  });
}

}  // extern "C"
