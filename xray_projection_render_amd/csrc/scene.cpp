// scene.cpp - object/deformation loading with the reference's decoder
// rules, flattening to the device node table, clip boxes and the FP32
// guard-band model.
#include "scene.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <limits>
#include <sstream>

namespace xrp {

using json::Value;

Box3 Box3::none() {
  Box3 b;
  for (int k = 0; k < 3; ++k) {
    b.lo[k] = INFINITY;
    b.hi[k] = -INFINITY;
  }
  return b;
}
Box3 Box3::all() {
  Box3 b;
  for (int k = 0; k < 3; ++k) {
    b.lo[k] = -INFINITY;
    b.hi[k] = INFINITY;
  }
  return b;
}
bool Box3::finite() const {
  for (int k = 0; k < 3; ++k)
    if (!std::isfinite(lo[k]) || !std::isfinite(hi[k])) return false;
  return true;
}

static Box3 unite(const Box3& a, const Box3& b) {
  if (a.empty()) return b;
  if (b.empty()) return a;
  Box3 r;
  for (int k = 0; k < 3; ++k) {
    r.lo[k] = std::min(a.lo[k], b.lo[k]);
    r.hi[k] = std::max(a.hi[k], b.hi[k]);
  }
  return r;
}
static Box3 intersect(const Box3& a, const Box3& b) {
  Box3 r;
  for (int k = 0; k < 3; ++k) {
    r.lo[k] = std::max(a.lo[k], b.lo[k]);
    r.hi[k] = std::min(a.hi[k], b.hi[k]);
  }
  return r.empty() ? Box3::none() : r;
}

// ---------------------------------------------------------------------------
// Go decoder view of numbers: yaml.v3 keeps integer literals as int, Go's
// encoding/json decodes every number into float64.
// ---------------------------------------------------------------------------
namespace {
struct Num {
  bool json;
  bool is_float(const Value* v) const {
    return v && (v->type == Value::Float || (json && v->type == Value::Int));
  }
  bool is_int(const Value* v) const {
    return v && !json && v->type == Value::Int;
  }
  // data[k].(float64)
  double float64(const Value* v, const char* msg) const {
    if (!is_float(v)) throw SceneError(msg);
    return v->number();
  }
  // ToFloat64 (objects.go:261-270), toFloat64 (deformations.go:345-354)
  double to_float64(const Value* v, const char* msg) const {
    if (!is_float(v) && !is_int(v)) throw SceneError(msg);
    return v->number();
  }
  // ToVec (objects.go:272-282): ints and floats copied, others left 0;
  // more than three entries index out of range (a panic in Go).
  void to_vec(const Value* v, double out[3], const char* msg) const {
    if (!v || v->type != Value::Arr) throw SceneError(msg);
    if (v->a->size() > 3)
      throw SceneError("index out of range: vector has more than 3 entries");
    out[0] = out[1] = out[2] = 0.0;
    for (size_t k = 0; k < v->a->size(); ++k) {
      const Value* e = &(*v->a)[k];
      if (is_float(e) || is_int(e)) out[k] = e->number();
    }
  }
};

enum Allowed { TOP, CHILD };

const char* kChildTypes[] = {"sphere", "cube", "box", "cylinder",
                             "parallelepiped", "gyroid",
                             "tessellated_obj_coll", "voxel_grid"};
const char* kTopTypes[] = {"sphere", "cube", "box", "cylinder",
                           "parallelepiped", "gyroid", "object_collection",
                           "tessellated_obj_coll", "voxel_grid"};

bool type_allowed(const std::string& t, Allowed al) {
  if (al == TOP) {
    for (auto* n : kTopTypes)
      if (t == n) return true;
  } else {
    for (auto* n : kChildTypes)
      if (t == n) return true;
  }
  return false;
}

// mgl64 Mat3.Det/Inv restated (cofactor expansion, column-major), with
// FloatEqual(det, 0) reducing to |det| < 1e-20.
void mat3_inv(const double m[9], double out[9]) {
  double det = m[0] * m[4] * m[8] + m[3] * m[7] * m[2] + m[6] * m[1] * m[5] -
               m[6] * m[4] * m[2] - m[3] * m[1] * m[8] - m[0] * m[7] * m[5];
  if (det == 0.0 || std::fabs(det) < 1e-20) {
    std::memset(out, 0, 9 * sizeof(double));
    return;
  }
  double r[9] = {
      m[4] * m[8] - m[5] * m[7], m[2] * m[7] - m[1] * m[8],
      m[1] * m[5] - m[2] * m[4], m[5] * m[6] - m[3] * m[8],
      m[0] * m[8] - m[2] * m[6], m[2] * m[3] - m[0] * m[5],
      m[3] * m[7] - m[4] * m[6], m[1] * m[6] - m[0] * m[7],
      m[0] * m[4] - m[1] * m[3],
  };
  double inv = 1 / det;
  for (int k = 0; k < 9; ++k) out[k] = r[k] * inv;
}

std::vector<float> read_raw_voxels(const std::string& path, const int res[3],
                                   const std::string& dtype) {
  // objects.go:892-958
  int bpe;
  if (dtype == "uint8")
    bpe = 1;
  else if (dtype == "uint16")
    bpe = 2;
  else if (dtype == "uint32" || dtype == "float32")
    bpe = 4;
  else if (dtype == "float64")
    bpe = 8;
  else
    throw SceneError("unsupported data type: " + dtype);
  std::ifstream f(path, std::ios::binary);
  if (!f) throw SceneError("error reading file: open " + path);
  std::string data((std::istreambuf_iterator<char>(f)),
                   std::istreambuf_iterator<char>());
  long long expected = (long long)res[0] * res[1] * res[2] * bpe;
  if ((long long)data.size() != expected) {
    std::ostringstream os;
    os << "file size (" << data.size() << ") does not match expected size ("
       << expected << ") for type " << dtype;
    throw SceneError(os.str());
  }
  size_t n = (size_t)res[0] * res[1] * res[2];
  std::vector<float> out(n);
  const unsigned char* b = (const unsigned char*)data.data();
  for (size_t k = 0; k < n; ++k) {
    double v;
    if (bpe == 1) {
      v = (double)b[k] / 255.0;
    } else if (bpe == 2) {
      uint16_t x = (uint16_t)(b[2 * k] | (b[2 * k + 1] << 8));
      v = (double)x / 65535.0;
    } else if (dtype == "uint32") {
      uint32_t x = (uint32_t)b[4 * k] | ((uint32_t)b[4 * k + 1] << 8) |
                   ((uint32_t)b[4 * k + 2] << 16) |
                   ((uint32_t)b[4 * k + 3] << 24);
      v = (double)x / 4294967295.0;
    } else if (dtype == "float32") {
      float x;
      std::memcpy(&x, b + 4 * k, 4);
      v = (double)x;
    } else {
      std::memcpy(&v, b + 8 * k, 8);
    }
    // The device grid is float32.  Densities that are not exactly
    // representable in float32 are rounded once here.
    out[k] = (float)v;
  }
  return out;
}

HNode parse_object(const Value& d, const Num& N, Allowed al,
                   const std::string& base_dir);

HNode parse_collection(const Value& d, const Num& N,
                       const std::string& base_dir) {
  // objects.go:381-420
  HNode n;
  n.kind = K_COLL;
  const Value* g = d.get("greedy_dens_eval");
  if (g && g->type == Value::Bool) n.greedy = g->b;
  const Value* objs = d.get("objects");
  if (!objs || objs->type != Value::Arr)
    throw SceneError("objects is not a list");
  for (auto& od : *objs->a) n.kids.push_back(parse_object(od, N, CHILD, base_dir));
  return n;
}

void parse_bounds(const Value& d, const Num& N, double out[6]) {
  const char* keys[6] = {"xmin", "xmax", "ymin", "ymax", "zmin", "zmax"};
  for (int k = 0; k < 6; ++k) {
    std::string msg = std::string(keys[k]) + " is not a float64";
    out[k] = N.to_float64(d.get(keys[k]), msg.c_str());
  }
}

HNode parse_object(const Value& d, const Num& N, Allowed al,
                   const std::string& base_dir) {
  // objects.go:658-689 (top level) and 391-410 (collection children)
  if (d.type != Value::Obj)
    throw SceneError("interface conversion: object is not a map");
  const Value* tv = d.get("type");
  std::string t = (tv && tv->type == Value::String) ? tv->s : std::string();
  if (!tv || tv->type != Value::String || !type_allowed(t, al)) {
    if (al == TOP)
      throw SceneError("unknown object type `" +
                       (tv && tv->type == Value::String ? t : std::string("<nil>")) +
                       "`");
    throw SceneError("unknown object type");
  }
  HNode n;
  if (t == "sphere") {  // objects.go:42-61
    n.kind = K_SPHERE;
    const Value* c = d.get("center");
    if (!c || c->type != Value::Arr) throw SceneError("center is not a Vec3");
    if (c->a->size() > 3)
      throw SceneError("index out of range: center has more than 3 entries");
    for (size_t k = 0; k < c->a->size(); ++k) {
      std::string msg = "center[" + std::to_string(k) + "] is not a float64";
      n.c[k] = N.to_float64(&(*c->a)[k], msg.c_str());
    }
    n.a = N.float64(d.get("radius"), "radius is not a float64");
    n.rho = N.float64(d.get("rho"), "rho is not a float64");
  } else if (t == "cube") {  // objects.go:100-117
    n.kind = K_BOX;
    n.is_cube = true;
    const Value* c = d.get("center");
    if (!c || c->type != Value::Arr) throw SceneError("center is not a Vec3");
    if (c->a->size() > 3)
      throw SceneError("index out of range: center has more than 3 entries");
    for (size_t k = 0; k < c->a->size(); ++k) {
      const Value* e = &(*c->a)[k];
      if (!N.is_float(e))
        throw SceneError("interface conversion: cube center entry is not float64");
      n.c[k] = e->number();
    }
    double side = N.float64(d.get("side"), "side is not a float64");
    n.v[0] = n.v[1] = n.v[2] = side;
    n.a = side;
    n.rho = N.float64(d.get("rho"), "rho is not a float64");
  } else if (t == "box") {  // objects.go:148-169
    n.kind = K_BOX;
    N.to_vec(d.get("center"), n.c, "center is not a Vec3");
    N.to_vec(d.get("sides"), n.v, "sides is not a Vec3");
    n.rho = N.to_float64(d.get("rho"), "rho is not a float64");
  } else if (t == "parallelepiped") {  // objects.go:209-245
    n.kind = K_PPED;
    double v0[3], v1[3], v2[3];
    N.to_vec(d.get("origin"), n.c, "origin is not a Vec3");
    N.to_vec(d.get("v0"), v0, "v0 is not a Vec3");
    N.to_vec(d.get("v1"), v1, "v1 is not a Vec3");
    N.to_vec(d.get("v2"), v2, "v2 is not a Vec3");
    n.rho = N.to_float64(d.get("rho"), "rho is not a float64");
    double m[9] = {v0[0], v0[1], v0[2], v1[0], v1[1], v1[2], v2[0], v2[1], v2[2]};
    std::memcpy(n.pv, m, sizeof m);
    mat3_inv(m, n.minv);
  } else if (t == "cylinder") {  // objects.go:306-331
    n.kind = K_CYL;
    N.to_vec(d.get("p0"), n.c, "p0 is not a Vec3");
    N.to_vec(d.get("p1"), n.v, "p0 is not a Vec3");
    n.a = N.float64(d.get("radius"), "radius is not a float64");
    const Value* r = d.get("rho");
    n.rho = r ? N.to_float64(r, "rho is not a float64") : 1.0;
  } else if (t == "gyroid") {  // objects.go:983-1012
    n.kind = K_GYROID;
    const Value* c = d.get("center");
    if (!c || c->type != Value::Arr) throw SceneError("center is not a Vec3");
    if (c->a->size() > 3)
      throw SceneError("index out of range: center has more than 3 entries");
    for (size_t k = 0; k < c->a->size(); ++k) {
      std::string msg = "center[" + std::to_string(k) + "] is not a float64";
      n.c[k] = N.to_float64(&(*c->a)[k], msg.c_str());
    }
    n.a = N.float64(d.get("scale"), "scale is not a float64");
    n.b = N.float64(d.get("thickness"), "thickness is not a float64");
    n.rho = N.float64(d.get("rho"), "rho is not a float64");
  } else if (t == "object_collection") {
    n = parse_collection(d, N, base_dir);
  } else if (t == "tessellated_obj_coll") {  // objects.go:536-566, 479-510
    n.kind = K_TESS;
    const Value* ucd = d.get("uc");
    if (!ucd || ucd->type != Value::Obj) throw SceneError("uc is not a map");
    const Value* objs = ucd->get("objects");
    if (!objs || objs->type != Value::Obj)
      throw SceneError("objects is not a map");
    HNode coll = parse_collection(*objs, N, base_dir);
    coll.greedy = true;  // objects.go:487
    parse_bounds(*ucd, N, n.ucb);
    parse_bounds(d, N, n.bounds);
    n.kids.push_back(std::move(coll));
  } else if (t == "voxel_grid") {  // objects.go:715-787
    n.kind = K_VOXEL;
    const Value* path = d.get("path");
    if (path && path->type == Value::String) {
      std::string p = path->s;
      size_t dot = p.rfind('.');
      std::string ext = dot == std::string::npos ? p : p.substr(dot + 1);
      for (auto& ch : ext) ch = (char)std::tolower(ch);
      if (ext != "raw") throw SceneError("only raw files are supported");
      const Value* res = d.get("resolution");
      if (!res || res->type != Value::Arr)
        throw SceneError(
            "resolution must be provided for raw files as a list of 3 integers");
      if (res->a->size() != 3)
        throw SceneError("resolution must be a list of 3 integers");
      int r3[3];
      for (int k = 0; k < 3; ++k) {
        const Value* e = &(*res->a)[k];
        if (!(N.is_int(e) || N.is_float(e)))
          throw SceneError("resolution[" + std::to_string(k) +
                           "] has unsupported type (need int-like)");
        r3[k] = (int)e->number();
      }
      if (r3[0] <= 0 || r3[1] <= 0 || r3[2] <= 0)
        throw SceneError("resolution entries must be positive");
      const Value* dt = d.get("dtype");
      std::string dtype =
          (dt && dt->type == Value::String) ? dt->s : std::string("uint8");
      std::string full = (!p.empty() && p[0] == '/') || base_dir.empty()
                             ? p
                             : base_dir + "/" + p;
      n.vox = std::make_shared<std::vector<float>>(
          read_raw_voxels(full, r3, dtype));
      n.nx = r3[0];
      n.ny = r3[1];
      n.nz = r3[2];
      n.path = p;
    } else {
      for (const char* k : {"nx", "ny", "nz"})
        if (!N.is_int(d.get(k)))
          throw SceneError(std::string(k) + " is not an int");
      // objects.go:773-775: an inline grid also needs a string path,
      // which the branch above already took, so this form never loads.
      throw SceneError("path is not a string");
    }
  }
  return n;
}

std::vector<double> float_list(const Value* v, const Num& N, const char* msg,
                               size_t need) {
  if (!v || v->type != Value::Arr) throw SceneError(msg);
  std::vector<double> out;
  for (auto& e : *v->a) {
    if (!N.is_float(&e))
      throw SceneError("interface conversion: deformation entry is not float64");
    out.push_back(e.number());
  }
  if (out.size() < need)
    throw SceneError("index out of range: need " + std::to_string(need) +
                     " entries");
  return out;
}

void parse_deformation(const Value& d, const Num& N, std::vector<HDeform>& out,
                       int depth) {
  // deformations.go:308-343 and the FromMap of each type
  if (depth > 16) throw SceneError("composed deformation nested too deeply");
  if (d.type != Value::Obj) throw SceneError("deformation type is nil");
  const Value* tv = d.get("type");
  if (!tv || tv->is_null()) throw SceneError("deformation type is nil");
  std::string t = tv->type == Value::String ? tv->s : std::string("<non-string>");
  HDeform f;
  if (t == "gaussian") {
    f.kind = D_GAUSS;
    auto a = float_list(d.get("amplitudes"), N, "amplitudes must be a list", 3);
    auto s = float_list(d.get("sigmas"), N, "sigmas must be a list", 3);
    auto c = float_list(d.get("centers"), N, "centers must be a list", 3);
    for (int k = 0; k < 3; ++k) {
      f.p[k] = a[k];
      f.q[k] = s[k];
      f.r[k] = c[k];
    }
  } else if (t == "affine") {
    f.kind = D_AFFINE;
    const Value* m = d.get("matrix");
    if (!m || m->type != Value::Arr) throw SceneError("matrix must be a list");
    if (m->a->size() != 3) throw SceneError("matrix must have 3 rows");
    for (int i = 0; i < 3; ++i) {
      const Value& row = (*m->a)[i];
      if (row.type != Value::Arr) throw SceneError("matrix row must be a list");
      if (row.a->size() != 3) throw SceneError("matrix row must have 3 elements");
      for (int j = 0; j < 3; ++j) {
        const Value* e = &(*row.a)[j];
        if (!N.is_float(e))
          throw SceneError("interface conversion: matrix entry is not float64");
        f.p[i * 3 + j] = e->number();
      }
    }
  } else if (t == "linear") {
    f.kind = D_LINEAR;
    auto s = float_list(d.get("strains"), N, "strains must be a list", 6);
    for (int k = 0; k < 6; ++k) f.p[k] = s[k];
  } else if (t == "rigid") {
    f.kind = D_RIGID;
    auto s = float_list(d.get("displacements"), N,
                        "displacements must be a list", 3);
    for (int k = 0; k < 3; ++k) f.p[k] = s[k];
  } else if (t == "sigmoid") {
    f.kind = D_SIGMOID;
    f.p[0] = N.to_float64(d.get("amplitude"), "amplitude must be a float");
    f.p[1] = N.to_float64(d.get("center"), "center must be a float");
    f.p[2] = N.to_float64(d.get("lengthscale"), "lengthscale must be a float");
    const Value* dir = d.get("direction");
    if (!dir || dir->type != Value::String)
      throw SceneError("direction must be a string");
    if (dir->s == "x")
      f.axis = 0;
    else if (dir->s == "y")
      f.axis = 1;
    else if (dir->s == "z")
      f.axis = 2;
    else
      throw SceneError("Invalid direction");
    if (tv->type != Value::String) throw SceneError("type must be a string");
  } else if (t == "tanh_shear") {
    // Extension without a deformations.go counterpart:
    // p[direction] += amplitude * tanh(rate * p[source]).
    f.kind = D_TANH;
    f.p[0] = N.to_float64(d.get("amplitude"), "amplitude must be a float");
    f.p[1] = N.to_float64(d.get("rate"), "rate must be a float");
    const Value* dir = d.get("direction");
    const Value* src = d.get("source");
    if (!dir || dir->type != Value::String || !src || src->type != Value::String)
      throw SceneError("direction and source must be strings");
    auto axis_of = [](const std::string& a) {
      return a == "x" ? 0 : (a == "y" ? 1 : (a == "z" ? 2 : -1));
    };
    f.axis = axis_of(dir->s);
    f.src = axis_of(src->s);
    if (f.axis < 0 || f.src < 0 || f.axis == f.src) throw SceneError("Invalid direction");
  } else if (t == "composed") {
    const Value* lst = d.get("deformations");
    if (!lst || lst->type != Value::Arr)
      throw SceneError("deformations must be a list");
    for (auto& sub : *lst->a) parse_deformation(sub, N, out, depth + 1);
    return;
  } else {
    throw SceneError("unknown deformation type " + t);
  }
  out.push_back(f);
}

// ---------------------------------------------------------------------------
// Boxes
// ---------------------------------------------------------------------------
bool may_be_positive(const HNode& n) {
  switch (n.kind) {
    case K_COLL:
      for (auto& k : n.kids)
        if (may_be_positive(k)) return true;
      return false;
    case K_TESS: return may_be_positive(n.kids[0]);
    case K_VOXEL:
      for (float v : *n.vox)
        if (!(v <= 0.0f)) return true;
      return false;
    default: return !(n.rho <= 0.0);
  }
}

Box3 leaf_box(const HNode& n) {
  Box3 b;
  switch (n.kind) {
    case K_SPHERE:
      for (int k = 0; k < 3; ++k) {
        b.lo[k] = n.c[k] - std::fabs(n.a);
        b.hi[k] = n.c[k] + std::fabs(n.a);
      }
      return b;
    case K_BOX:
      for (int k = 0; k < 3; ++k) {
        b.lo[k] = n.c[k] - 0.5 * std::fabs(n.v[k]);
        b.hi[k] = n.c[k] + 0.5 * std::fabs(n.v[k]);
      }
      return b;
    case K_PPED: {
      b = Box3::none();
      for (int m = 0; m < 8; ++m) {
        double p[3];
        for (int k = 0; k < 3; ++k)
          p[k] = n.c[k] + ((m & 1) ? n.pv[0 + k] : 0.0) +
                 ((m & 2) ? n.pv[3 + k] : 0.0) + ((m & 4) ? n.pv[6 + k] : 0.0);
        for (int k = 0; k < 3; ++k) {
          b.lo[k] = std::min(b.lo[k], p[k]);
          b.hi[k] = std::max(b.hi[k], p[k]);
        }
      }
      return b;
    }
    case K_CYL:
      for (int k = 0; k < 3; ++k) {
        b.lo[k] = std::min(n.c[k], n.v[k]) - std::fabs(n.a);
        b.hi[k] = std::max(n.c[k], n.v[k]) + std::fabs(n.a);
      }
      return b;
    case K_GYROID: return Box3::all();
    case K_VOXEL:
      for (int k = 0; k < 3; ++k) {
        b.lo[k] = -1.0;
        b.hi[k] = 1.0;
      }
      return b;
  }
  return Box3::all();
}

Box3 node_box(const HNode& n, bool as_child) {
  if (n.kind == K_COLL) {
    Box3 b = Box3::none();
    for (auto& k : n.kids)
      if (may_be_positive(k)) b = unite(b, node_box(k, true));
    return b;
  }
  if (n.kind == K_TESS) {
    if (!may_be_positive(n.kids[0])) return Box3::none();
    Box3 b;
    for (int k = 0; k < 3; ++k) {
      b.lo[k] = n.bounds[2 * k];
      b.hi[k] = n.bounds[2 * k + 1];
    }
    return b.empty() ? Box3::none() : b;
  }
  (void)as_child;
  return leaf_box(n);
}

bool inv3_rowmajor(const double m[9], double out[9]) {
  double det = m[0] * (m[4] * m[8] - m[5] * m[7]) -
               m[1] * (m[3] * m[8] - m[5] * m[6]) +
               m[2] * (m[3] * m[7] - m[4] * m[6]);
  if (!(std::fabs(det) > 1e-300)) return false;
  out[0] = (m[4] * m[8] - m[5] * m[7]) / det;
  out[1] = (m[2] * m[7] - m[1] * m[8]) / det;
  out[2] = (m[1] * m[5] - m[2] * m[4]) / det;
  out[3] = (m[5] * m[6] - m[3] * m[8]) / det;
  out[4] = (m[0] * m[8] - m[2] * m[6]) / det;
  out[5] = (m[2] * m[3] - m[0] * m[5]) / det;
  out[6] = (m[3] * m[7] - m[4] * m[6]) / det;
  out[7] = (m[1] * m[6] - m[0] * m[7]) / det;
  out[8] = (m[0] * m[4] - m[1] * m[3]) / det;
  return true;
}

void def_matrix(const HDeform& f, double M[9]) {
  if (f.kind == D_AFFINE) {
    std::memcpy(M, f.p, 9 * sizeof(double));
  } else {  // LINEAR: I + E
    const double* s = f.p;
    double m[9] = {1 + s[0], s[5], s[4], s[5], 1 + s[1], s[3], s[4], s[3], 1 + s[2]};
    std::memcpy(M, m, sizeof m);
  }
}

// Box of points p whose deformed image D(p) can lie in b.
Box3 preimage(const HDeform& f, const Box3& b) {
  if (b.empty()) return b;
  Box3 r = b;
  switch (f.kind) {
    case D_RIGID:
      for (int k = 0; k < 3; ++k) {
        r.lo[k] = b.lo[k] - f.p[k];
        r.hi[k] = b.hi[k] - f.p[k];
      }
      return r;
    case D_SIGMOID: {
      double A = f.p[0];
      int a = f.axis;
      r.lo[a] = b.lo[a] - std::max(0.0, A);
      r.hi[a] = b.hi[a] - std::min(0.0, A);
      return r;
    }
    case D_TANH: {  // displacement in [-|A|, |A|] along the axis
      double A = std::fabs(f.p[0]);
      int a = f.axis;
      r.lo[a] = b.lo[a] - A;
      r.hi[a] = b.hi[a] + A;
      return r;
    }
    case D_GAUSS:
      for (int k = 0; k < 3; ++k) {
        r.lo[k] = b.lo[k] - std::max(0.0, f.p[k]);
        r.hi[k] = b.hi[k] - std::min(0.0, f.p[k]);
      }
      return r;
    case D_AFFINE:
    case D_LINEAR: {
      if (!b.finite()) return Box3::all();
      double M[9], Mi[9];
      def_matrix(f, M);
      if (!inv3_rowmajor(M, Mi)) return Box3::all();
      r = Box3::none();
      for (int m = 0; m < 8; ++m) {
        double c[3] = {(m & 1) ? b.hi[0] : b.lo[0], (m & 2) ? b.hi[1] : b.lo[1],
                       (m & 4) ? b.hi[2] : b.lo[2]};
        for (int i = 0; i < 3; ++i) {
          double v = Mi[3 * i] * c[0] + Mi[3 * i + 1] * c[1] + Mi[3 * i + 2] * c[2];
          r.lo[i] = std::min(r.lo[i], v);
          r.hi[i] = std::max(r.hi[i], v);
        }
      }
      for (int i = 0; i < 3; ++i) {  // covers the rounding of Mi
        double pad = 1e-9 * (1.0 + std::max(std::fabs(r.lo[i]), std::fabs(r.hi[i])));
        r.lo[i] -= pad;
        r.hi[i] += pad;
      }
      return r;
    }
  }
  return Box3::all();
}

double vlen3(const double* v) {
  return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
}

// objects.go MinFeatureSize methods (74-76, 123-125, 181-183, 257-259,
// 352-354, 440-446, 584-586, 857-860, 1034-1036).
double min_feature(const HNode& n) {
  switch (n.kind) {
    case K_SPHERE:
    case K_CYL: return n.a;
    case K_BOX: return 0.1 * std::fmin(n.v[0], std::fmin(n.v[1], n.v[2]));
    case K_PPED:
      return 0.2 * std::fmin(vlen3(n.pv), std::fmin(vlen3(n.pv + 3), vlen3(n.pv + 6)));
    case K_GYROID: return n.a * n.b * 0.1;
    case K_COLL: {
      double out = INFINITY;
      for (auto& k : n.kids) out = std::fmin(out, min_feature(k));
      return out;
    }
    case K_TESS: return min_feature(n.kids[0]);
    case K_VOXEL: return 2.0 / (double)std::max(n.nx, std::max(n.ny, n.nz));
  }
  return INFINITY;
}

// Object-space bounding sphere of a bounded primitive, stored in FP32 with
// the radius inflated past the rounding of the stored centre and radius, so
// the sphere encloses the primitive's closure (per-wave culling of the root
// collection's children tests it against each ray's line).
void bounding_sphere(const HNode& n, DevNode& dn) {
  double c[3] = {0, 0, 0}, r = 0.0;
  switch (n.kind) {
    case K_SPHERE:
      for (int k = 0; k < 3; ++k) c[k] = n.c[k];
      r = std::fabs(n.a);
      break;
    case K_BOX:
      for (int k = 0; k < 3; ++k) {
        c[k] = n.c[k];
        r += 0.25 * n.v[k] * n.v[k];
      }
      r = std::sqrt(r);
      break;
    case K_PPED:
      for (int k = 0; k < 3; ++k)
        c[k] = n.c[k] + 0.5 * (n.pv[k] + n.pv[3 + k] + n.pv[6 + k]);
      for (int m = 0; m < 8; ++m) {
        double e = 0.0;
        for (int k = 0; k < 3; ++k) {
          double q = n.c[k] + ((m & 1) ? n.pv[k] : 0.0) + ((m & 2) ? n.pv[3 + k] : 0.0) +
                     ((m & 4) ? n.pv[6 + k] : 0.0) - c[k];
          e += q * q;
        }
        r = std::max(r, std::sqrt(e));
      }
      break;
    case K_CYL: {
      double h = 0.0;
      for (int k = 0; k < 3; ++k) {
        c[k] = 0.5 * (n.c[k] + n.v[k]);
        h += 0.25 * (n.v[k] - n.c[k]) * (n.v[k] - n.c[k]);
      }
      r = std::sqrt(h + n.a * n.a);
      break;
    }
    default:
      return;
  }
  double m = std::fabs(r);
  for (int k = 0; k < 3; ++k) {
    dn.bsph[k] = (float)c[k];
    m = std::max(m, std::fabs(c[k]));
  }
  // FP32 rounding of centre and radius is <= 2^-24 of m each; the relative
  // and absolute terms also cover the host's FP64 rounding
  const double rr = r * (1.0 + 1e-6) + 4e-7 * m + 1e-12;
  dn.bsph[3] = (float)rr;
  if ((double)dn.bsph[3] < rr) dn.bsph[3] = std::nextafter(dn.bsph[3], INFINITY);
  if (!std::isfinite(dn.bsph[3])) dn.bsph[3] = INFINITY;
}

void flatten(const HNode& n, int depth, std::vector<DevNode>& out, Scene& s) {
  int idx = (int)out.size();
  out.emplace_back();
  DevNode dn;
  std::memset(&dn, 0, sizeof dn);
  dn.kind = n.kind;
  dn.depth = depth;
  s.max_depth = std::max(s.max_depth, depth);
  switch (n.kind) {
    case K_SPHERE:
      for (int k = 0; k < 3; ++k) {
        dn.d[k] = n.c[k];
        dn.f[k] = (float)n.c[k];
      }
      dn.d[3] = n.a;
      dn.d[4] = n.rho;
      dn.f[3] = (float)(n.a * n.a);
      dn.f[4] = (float)n.rho;
      break;
    case K_BOX:
      for (int k = 0; k < 3; ++k) {
        dn.d[k] = n.c[k];
        dn.d[3 + k] = 0.5 * n.v[k];
        dn.f[k] = (float)n.c[k];
        dn.f[3 + k] = (float)(0.5 * n.v[k]);
      }
      dn.d[6] = n.rho;
      dn.f[6] = (float)n.rho;
      break;
    case K_PPED:
      for (int k = 0; k < 3; ++k) {
        dn.d[k] = n.c[k];
        dn.f[k] = (float)n.c[k];
      }
      for (int k = 0; k < 9; ++k) {
        dn.d[3 + k] = n.minv[k];
        dn.f[3 + k] = (float)n.minv[k];
      }
      dn.d[12] = n.rho;
      dn.f[12] = (float)n.rho;
      break;
    case K_CYL: {
      double v0 = n.v[0] - n.c[0], v1 = n.v[1] - n.c[1], v2 = n.v[2] - n.c[2];
      double vv = v0 * v0 + v1 * v1 + v2 * v2;
      for (int k = 0; k < 3; ++k) {
        dn.d[k] = n.c[k];
        dn.f[k] = (float)n.c[k];
      }
      dn.d[3] = v0;
      dn.d[4] = v1;
      dn.d[5] = v2;
      dn.d[6] = vv;
      dn.d[7] = n.a;
      dn.d[8] = n.rho;
      dn.f[3] = (float)v0;
      dn.f[4] = (float)v1;
      dn.f[5] = (float)v2;
      dn.f[6] = (float)vv;
      dn.f[7] = (float)(1.0 / vv);
      dn.f[8] = (float)(n.a * n.a);
      dn.f[9] = (float)n.rho;
      break;
    }
    case K_GYROID:
      for (int k = 0; k < 3; ++k) {
        dn.d[k] = n.c[k];
        dn.f[k] = (float)n.c[k];
      }
      dn.d[3] = n.a;
      dn.d[4] = n.b;
      dn.d[5] = n.rho;
      dn.f[3] = (float)(1.0 / n.a);
      dn.f[4] = (float)n.b;
      dn.f[5] = (float)n.rho;
      dn.f[6] = (float)(1.0 / (n.a * 6.283185307179586476925286766559));
      break;
    case K_VOXEL:
      dn.nx = n.nx;
      dn.ny = n.ny;
      dn.nz = n.nz;
      dn.f[0] = (float)((n.nx - 1) * 0.5);
      dn.f[1] = (float)((n.ny - 1) * 0.5);
      dn.f[2] = (float)((n.nz - 1) * 0.5);
      s.has_voxels = true;
      s.voxel_bytes += n.vox->size() * sizeof(float);
      break;
    case K_TESS:
      if (depth + 1 > kMaxDepth)
        throw SceneError("tessellations nested more than " +
                         std::to_string(kMaxDepth) +
                         " deep are not supported by the device evaluator");
      for (int k = 0; k < 6; ++k) {
        dn.d[k] = n.bounds[k];
        dn.f[k] = (float)n.bounds[k];
      }
      for (int k = 0; k < 3; ++k) {
        double lo = n.ucb[2 * k], hi = n.ucb[2 * k + 1];
        double dx = hi - lo;
        dn.d[6 + k] = lo;
        dn.d[9 + k] = hi;
        dn.f[6 + k] = (float)lo;
        dn.f[9 + k] = (float)dx;
        dn.f[12 + k] = (float)(1.0 / dx);
      }
      break;
    case K_COLL:
      dn.greedy = n.greedy ? 1 : 0;
      break;
  }
  bounding_sphere(n, dn);
  if (n.kind == K_TESS) {
    dn.first = (int)out.size();
    flatten(n.kids[0], depth + 1, out, s);
  } else if (n.kind == K_COLL) {
    dn.first = (int)out.size();
    for (auto& k : n.kids) flatten(k, depth, out, s);
  }
  dn.next = (int)out.size();
  out[idx] = dn;
}

void collect_voxels(const HNode& n, std::vector<const HNode*>& out) {
  if (n.kind == K_VOXEL) out.push_back(&n);
  for (auto& k : n.kids) collect_voxels(k, out);
}

}  // namespace

void finalize_scene(Scene& s) {
  s.nodes.clear();
  s.max_depth = 0;
  s.has_voxels = false;
  s.voxel_bytes = 0;
  flatten(s.root, 0, s.nodes, s);
  s.ddefs.clear();
  if (s.defs.size() > (size_t)kMaxDeform)
    throw SceneError("more than " + std::to_string(kMaxDeform) +
                     " chained deformations");
  for (auto& f : s.defs) {
    DevDeform dd;
    std::memset(&dd, 0, sizeof dd);
    dd.kind = f.kind;
    dd.axis = f.axis;
    switch (f.kind) {
      case D_GAUSS:
        for (int k = 0; k < 3; ++k) {
          dd.d[k] = f.p[k];
          dd.d[3 + k] = f.q[k];
          dd.d[6 + k] = f.r[k];
          dd.f[k] = (float)f.p[k];
          dd.f[3 + k] = (float)(1.0 / (2 * f.q[k] * f.q[k]));
          dd.f[6 + k] = (float)f.r[k];
        }
        break;
      case D_AFFINE:
        for (int k = 0; k < 9; ++k) {
          dd.d[k] = f.p[k];
          dd.f[k] = (float)f.p[k];
        }
        break;
      case D_LINEAR:
        for (int k = 0; k < 6; ++k) {
          dd.d[k] = f.p[k];
          dd.f[k] = (float)f.p[k];
        }
        break;
      case D_RIGID:
        for (int k = 0; k < 3; ++k) {
          dd.d[k] = f.p[k];
          dd.f[k] = (float)f.p[k];
        }
        break;
      case D_SIGMOID:
        dd.d[0] = f.p[0];
        dd.d[1] = f.p[1];
        dd.d[2] = f.p[2];
        dd.f[0] = (float)f.p[0];
        dd.f[1] = (float)f.p[1];
        dd.f[2] = (float)(1.0 / f.p[2]);
        break;
      case D_TANH:
        dd.src = f.src;
        dd.d[0] = f.p[0];
        dd.d[1] = f.p[1];
        dd.f[0] = (float)f.p[0];
        dd.f[1] = (float)f.p[1];
        break;
    }
    s.ddefs.push_back(dd);
  }
  s.obj_box = node_box(s.root, false);
  Box3 w = s.obj_box;
  for (size_t k = s.defs.size(); k-- > 0;) w = preimage(s.defs[k], w);
  if (!w.empty()) {
    for (int k = 0; k < 3; ++k) {  // covers FP64 rounding of positions
      double pad = 1e-9 * (1.0 + std::max(std::fabs(w.lo[k]), std::fabs(w.hi[k])));
      if (std::isfinite(pad)) {
        w.lo[k] -= pad;
        w.hi[k] += pad;
      }
    }
  }
  s.world_box = w;
  s.min_feature = min_feature(s.root);
}

namespace {
Value vnum(double x) {
  Value v;
  v.type = Value::Float;
  v.f = x;
  return v;
}
Value vint(long long x) {
  Value v;
  v.type = Value::Int;
  v.i = x;
  v.f = (double)x;
  return v;
}
Value vstr(const std::string& s) {
  Value v;
  v.type = Value::String;
  v.s = s;
  return v;
}
Value vvec(const double* p) {
  Value v;
  v.type = Value::Arr;
  v.a = std::make_shared<json::Array>();
  for (int k = 0; k < 3; ++k) v.a->push_back(vnum(p[k]));
  return v;
}
Value vmap(std::vector<std::pair<std::string, Value>> kv) {
  std::sort(kv.begin(), kv.end(),
            [](const auto& a, const auto& b) { return a.first < b.first; });
  Value v;
  v.type = Value::Obj;
  v.o = std::make_shared<json::Object>(std::move(kv));
  return v;
}
Value bounds_entries(const double* b, std::vector<std::pair<std::string, Value>>& kv) {
  const char* k[6] = {"xmin", "xmax", "ymin", "ymax", "zmin", "zmax"};
  for (int i = 0; i < 6; ++i) kv.emplace_back(k[i], vnum(b[i]));
  return Value();
}
}  // namespace

// ToMap of each object type (objects.go:33-40, 91-98, 139-146, 198-207,
// 296-304, 370-379, 466-477, 523-534, 704-713, 973-981).
Value to_map(const HNode& n) {
  switch (n.kind) {
    case K_SPHERE:
      return vmap({{"type", vstr("sphere")}, {"center", vvec(n.c)},
                   {"radius", vnum(n.a)}, {"rho", vnum(n.rho)}});
    case K_BOX:
      if (n.is_cube)
        return vmap({{"type", vstr("cube")}, {"center", vvec(n.c)}, {"side", vnum(n.a)},
                     {"rho", vnum(n.rho)}});
      return vmap({{"type", vstr("box")}, {"center", vvec(n.c)}, {"sides", vvec(n.v)},
                   {"rho", vnum(n.rho)}});
    case K_PPED:
      return vmap({{"type", vstr("parallelepiped")}, {"origin", vvec(n.c)},
                   {"v0", vvec(n.pv)}, {"v1", vvec(n.pv + 3)}, {"v2", vvec(n.pv + 6)},
                   {"rho", vnum(n.rho)}});
    case K_CYL:
      return vmap({{"type", vstr("cylinder")}, {"p0", vvec(n.c)}, {"p1", vvec(n.v)},
                   {"radius", vnum(n.a)}, {"rho", vnum(n.rho)}});
    case K_GYROID:
      return vmap({{"type", vstr("gyroid")}, {"center", vvec(n.c)}, {"scale", vnum(n.a)},
                   {"thickness", vnum(n.b)}, {"rho", vnum(n.rho)}});
    case K_COLL: {
      Value arr;
      arr.type = Value::Arr;
      arr.a = std::make_shared<json::Array>();
      for (auto& k : n.kids) arr.a->push_back(to_map(k));
      return vmap({{"type", vstr("object_collection")}, {"objects", arr}});
    }
    case K_TESS: {
      std::vector<std::pair<std::string, Value>> uc{{"type", vstr("unit_cell")},
                                                    {"objects", to_map(n.kids[0])}};
      bounds_entries(n.ucb, uc);
      std::vector<std::pair<std::string, Value>> t{{"type", vstr("tessellated_obj_coll")},
                                                   {"uc", vmap(uc)}};
      bounds_entries(n.bounds, t);
      return vmap(t);
    }
    case K_VOXEL:
      return vmap({{"type", vstr("voxel_grid")}, {"nx", vint(n.nx)}, {"ny", vint(n.ny)},
                   {"nz", vint(n.nz)}, {"dtype", vstr("float64")}, {"path", vstr(n.path)}});
  }
  return Value();
}

void append_deformation(const Value& d, NumberMode numbers,
                        std::vector<HDeform>& out) {
  parse_deformation(d, Num{numbers == NUMBERS_JSON}, out, 0);
}

std::unique_ptr<Scene> load_scene(const Value& object, const Value* deformation,
                                  NumberMode numbers,
                                  const std::string& base_dir) {
  Num N{numbers == NUMBERS_JSON};
  auto s = std::make_unique<Scene>();
  s->root = parse_object(object, N, TOP, base_dir);
  if (deformation && !deformation->is_null())
    parse_deformation(*deformation, N, s->defs, 0);
  finalize_scene(*s);
  return s;
}

std::unique_ptr<Scene> voxel_scene(const float* rho, int nx, int ny, int nz,
                                   const Value* deformation,
                                   NumberMode numbers) {
  auto s = std::make_unique<Scene>();
  s->root.kind = K_VOXEL;
  s->root.nx = nx;
  s->root.ny = ny;
  s->root.nz = nz;
  s->root.vox = std::make_shared<std::vector<float>>(
      rho, rho + (size_t)nx * ny * nz);
  if (deformation && !deformation->is_null())
    parse_deformation(*deformation, Num{numbers == NUMBERS_JSON}, s->defs, 0);
  finalize_scene(*s);
  return s;
}

// ---------------------------------------------------------------------------
// Guard bands.  u = 2^-24 is the FP32 unit roundoff; delta bounds the
// per-component error of an FP32 position against the FP64 position the
// reference computes; B bounds |coordinate| of evaluated points.  Each band
// is kappa times a first-order error bound of the FP32 test quantity.
// ---------------------------------------------------------------------------
// Upper bound on the spectral norm of a 3x3 matrix: rho(S)^(1/2), S = M^T M,
// is at most (||S^n||_inf)^(1/2n) for every n (an induced norm bounds the
// spectral radius); n = 2^6 is within 3^(1/128) of the norm.
static double spectral_norm_bound(const double* M) {
  double S[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double a = 0;
      for (int k = 0; k < 3; ++k) a += M[3 * k + i] * M[3 * k + j];
      S[3 * i + j] = a;
    }
  double log_scale = 0.0;  // S^n = exp(log_scale) * S (renormalised)
  int n = 1;
  for (int it = 0; it < 6; ++it) {
    double T[9];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double a = 0;
        for (int k = 0; k < 3; ++k) a += S[3 * i + k] * S[3 * k + j];
        T[3 * i + j] = a;
      }
    double mx = 0;
    for (double t : T) mx = std::max(mx, std::fabs(t));
    if (!(mx > 0) || !std::isfinite(mx)) return 0.0;
    for (int k = 0; k < 9; ++k) S[k] = T[k] / mx;
    log_scale = 2 * log_scale + std::log(mx);
    n *= 2;
  }
  double row = 0;
  for (int i = 0; i < 3; ++i)
    row = std::max(row, std::fabs(S[3 * i]) + std::fabs(S[3 * i + 1]) + std::fabs(S[3 * i + 2]));
  // ||S^n||_inf = exp(log_scale) * row; the bound is its 1/(2n) power
  return std::exp((log_scale + std::log(row)) / (2.0 * n)) * (1 + 1e-9);
}

// $XRP_SEAMLESS=0 treats every tessellation as seamed (for comparison; the
// images are the same either way).
static bool seamless_enabled() {
  static const bool on = [] {
    const char* e = std::getenv("XRP_SEAMLESS");
    return !(e && std::strcmp(e, "0") == 0);
  }();
  return on;
}

double compute_guards(Scene& s, double world_bound, double step, double* lipschitz,
                      double s_err) {
  const double u = 1.0 / 16777216.0;
  const double kappa = 4.0;
  const double s3c = std::sqrt(3.0);
  double B = world_bound;
  // FP64 terms: ray set-up, and s_err, how far the march's anchors
  // s = smin + k ds can be from the reference's accumulated s_k.
  double delta = u * (2.0 * B + 4.0 * step) + 1e-13 * (1.0 + B) + s_err;
  // Euclidean Lipschitz bound of the chain: how far object-space positions
  // of two samples can be apart per unit of ray length between them.
  double lip_chain = 1.0;
  // Bound on the chain's second derivative along a ray (unit speed):
  // through a map with Lipschitz bound l and |D''[v, v]| <= h |v|^2 the
  // bound becomes l K + h L^2 (chain rule), L the Lipschitz bound so far.
  double curv_chain = 0.0;
  for (auto& f : s.defs) {
    double h = 0.0, l_prev = lip_chain;
    switch (f.kind) {
      case D_RIGID: {
        double t = std::max({std::fabs(f.p[0]), std::fabs(f.p[1]), std::fabs(f.p[2])});
        B = B + t;
        delta = delta + 2 * u * B;
        break;
      }
      case D_AFFINE:
      case D_LINEAR: {
        double M[9];
        def_matrix(f, M);
        double L = 0, F = 0;
        for (int i = 0; i < 3; ++i) {
          L = std::max(L, std::fabs(M[3 * i]) + std::fabs(M[3 * i + 1]) + std::fabs(M[3 * i + 2]));
          for (int j = 0; j < 3; ++j) F += M[3 * i + j] * M[3 * i + j];
        }
        delta = L * delta + 6 * u * L * B + 2 * u * (L + 1) * B;
        B = L * B;
        const double sn = spectral_norm_bound(M);
        lip_chain *= sn > 0 ? std::min(sn, std::sqrt(F)) : std::sqrt(F);
        break;
      }
      case D_SIGMOID: {
        // Lipschitz A / (4 ls); the argument (q - c) / ls rounded (at most
        // u arg, through a slope of at most 1/4, kept whole here), the
        // hardware logistic function's kHwSigmoidErr, the product with A
        // and the add.
        double A = std::fabs(f.p[0]), c = std::fabs(f.p[1]), ls = std::fabs(f.p[2]);
        double arg = (B + c) / ls;
        delta = delta * (1 + A / (4 * ls)) + A * (u * arg + kHwSigmoidErr + 4 * u) +
                4 * u * (B + A);
        B = B + A;
        lip_chain *= 1 + A / (4 * ls);  // max slope of A / (1 + e^(-x/ls))
        h = A / (ls * ls) * 0.0963;     // max |sigma''| = 1 / (6 sqrt 3)
        break;
      }
      case D_TANH: {
        // Lipschitz |A| k; the hardware tanh's kHwTanhErr; FP32 rounding of
        // the amplitude, the rate product and the add.
        double A = std::fabs(f.p[0]), k = std::fabs(f.p[1]);
        delta = delta * (1 + A * k) + A * k * 2 * u * B + A * (kHwTanhErr + 4 * u) +
                2 * u * (B + A);
        B = B + A;
        // Euclidean: the Jacobian is I + a e_axis e_source^T with |a| <= A k
        // (axis != source), a shear whose largest singular value is
        // (a + sqrt(a^2 + 4)) / 2, increasing in |a|
        lip_chain *= (1 + 1e-12) * (A * k + std::sqrt(A * k * A * k + 4.0)) / 2.0;
        h = A * k * k * 0.7699;  // max |tanh''| = 4 / (3 sqrt 3)
        break;
      }
      case D_GAUSS: {
        double lip = 0, amax = 0;
        for (int k = 0; k < 3; ++k) {
          double A = std::fabs(f.p[k]);
          double sg = std::fabs(f.q[k]);
          lip = std::max(lip, A * 1.06 / sg);
          amax = std::max(amax, A);
        }
        // the hardware e^-x's kHwExpNegErr, the exponent r^2 / (2 s^2)
        // rounded (relative 4u, through x e^-x <= 1/e: 1.5u), the product
        // with A and the add
        delta = delta * (1 + lip) + amax * (kHwExpNegErr + 4 * u) + 4 * u * (B + amax);
        B = B + amax;
        // |grad A e^(-r^2/2s^2)| <= |A| e^(-1/2) / s per component, three
        // components: sqrt(3) * 0.6066 < 1.06
        lip_chain *= 1 + lip;
        // |d^2/ds^2 e^(-r^2/2s^2)| <= |v|^2 / s^2 (|t^2 - 1| e^(-t^2/2) <= 1)
        for (int k = 0; k < 3; ++k)
          h = std::max(h, s3c * std::fabs(f.p[k]) / (f.q[k] * f.q[k]));
        break;
      }
    }
    const double l_this = lip_chain / l_prev;
    curv_chain = l_this * curv_chain + h * l_prev * l_prev;
  }
  const double delta_obj = delta;

  // Seamless tessellations: every child of the unit cell's collection is a
  // gyroid whose period 2 pi scale divides the cell along each axis (within
  // 1e-9 of a whole number of periods), so the wrap shifts a point by whole
  // periods and the density function is the same on both sides of a seam.
  // Returns how far from whole periods the cell side is, as an object-space
  // shift (the wrap into either neighbouring cell moves the gyroid's
  // argument by at most that much more than whole periods: added to the
  // children's position error), or -1 when the tessellation has seams.
  auto seamless = [&](const DevNode& t) -> double {
    if (t.first < 0 || t.first >= (int)s.nodes.size()) return -1.0;
    const DevNode& c = s.nodes[t.first];
    if (c.kind != K_COLL || c.first >= c.next) return -1.0;
    double shift = 0.0;
    for (int k = c.first; k < c.next; k = s.nodes[k].next) {
      const DevNode& g = s.nodes[k];
      if (g.kind != K_GYROID) return -1.0;
      const double period = 2.0 * M_PI * std::fabs(g.d[3]);
      for (int a = 0; a < 3; ++a) {
        const double side = t.d[9 + a] - t.d[6 + a];
        const double m = side / period;
        if (!(m >= 0.5) || std::fabs(m - std::round(m)) > 1e-9 * m) return -1.0;
        shift = std::max(shift, std::fabs(side - std::round(m) * period));
      }
    }
    return std::sqrt(3.0) * shift * (1.0 + 1e-12);
  };

  // Walk the node table; per-depth (delta, B) pairs.
  double dl[kMaxDepth + 2], Bl[kMaxDepth + 2];
  dl[0] = delta_obj;
  Bl[0] = B;
  const double s3 = std::sqrt(3.0);
  for (auto& n : s.nodes) {
    double d = dl[n.depth], b = Bl[n.depth];
    for (float& g : n.g) g = 0.0f;
    switch (n.kind) {
      case K_SPHERE: {
        double r = std::fabs(n.d[3]);
        double cm = std::max({std::fabs(n.d[0]), std::fabs(n.d[1]), std::fabs(n.d[2])});
        double e = d + u * (cm + 1.1 * r);
        n.g[0] = (float)(kappa * (2 * s3 * 1.1 * r * e + 3 * e * e + 9 * u * r * r));
        // distance slack: position error plus FP32 rounding of |p - c| - r
        n.g[4] = (float)r;
        n.g[5] = (float)(2 * d + 16 * u * (b + cm + 2 * r));
        break;
      }
      case K_BOX: {
        double hm = 0, cm = 0;
        for (int k = 0; k < 3; ++k) {
          double h = std::fabs(n.d[3 + k]);
          n.g[k] = (float)(kappa * (d + u * (std::fabs(n.d[k]) + 3 * h)));
          hm = std::max(hm, h);
          cm = std::max(cm, std::fabs(n.d[k]));
        }
        n.g[5] = (float)(2 * d + 16 * u * (b + cm + hm));
        break;
      }
      case K_PPED: {
        double om = std::max({std::fabs(n.d[0]), std::fabs(n.d[1]), std::fabs(n.d[2])});
        double W = b + om;
        for (int i = 0; i < 3; ++i) {
          double rs = 0, ro = 0;
          for (int j = 0; j < 3; ++j) {
            double mij = std::fabs(n.d[3 + i + 3 * j]);  // column-major
            rs += mij;
            ro += mij * std::fabs(n.d[j]);
          }
          n.g[i] = (float)(kappa * (rs * d + u * ro + 6 * u * W * rs));
        }
        // distance: (parameter-space distance - band) / ||M^-1||_F
        double F = 0;
        for (int k = 0; k < 9; ++k) F += n.d[3 + k] * n.d[3 + k];
        n.g[4] = (float)(1.0 / (std::sqrt(F) * (1 + 1e-6)));
        n.g[5] = std::max({n.g[0], n.g[1], n.g[2]}) * 2.0f;
        break;
      }
      case K_CYL: {
        double v1n = std::fabs(n.d[3]) + std::fabs(n.d[4]) + std::fabs(n.d[5]);
        double L = std::sqrt(n.d[6]), vv = n.d[6], r = std::fabs(n.d[7]);
        double pm = std::max({std::fabs(n.d[0]), std::fabs(n.d[1]), std::fabs(n.d[2])});
        double W = b + pm;
        double gt = v1n * (d + u * pm) + 6 * u * W * v1n + u * vv;
        double e = d + u * pm + 3 * u * (L + r) + L * gt / vv + 3 * u * L;
        n.g[1] = (float)(kappa * gt);
        n.g[0] = (float)(kappa * (2 * s3 * 1.1 * r * e + 3 * e * e + 9 * u * r * r));
        // distance: radial | |e| - r | and axial t / |v| past the caps
        n.g[4] = (float)r;
        n.g[5] = (float)(2 * (e + gt / L) + 16 * u * (W + L + r));
        n.g[6] = (float)(1.0 / L);
        break;
      }
      case K_GYROID: {
        double sc = std::fabs(n.d[3]);
        double cm = std::max({std::fabs(n.d[0]), std::fabs(n.d[1]), std::fabs(n.d[2])});
        double Q = (b + cm) / sc;
        double eq = (d + u * cm + 2 * u * (b + cm)) / sc + 2 * u * Q;
        // six sine/cosine factors, each off by the argument error eq and
        // the hardware unit's kHwSinCosErr, plus the products and sums
        n.g[0] = (float)(kappa * (6 * eq + 6 * kHwSinCosErr + 16 * u));
        // |grad g| of g = sin x cos y + sin y cos z + sin z cos x is at most
        // 1.77 (in q = p / scale units): |grad g|^2 peaks at 3 on a 721^3
        // grid of the 2 pi period, and between grid points it grows by at
        // most |grad |grad g|^2| * h sqrt(3)/2 <= 2 |H|_F |grad g| * 0.00756
        // <= 2 sqrt(12) sqrt(6) * 0.00756 = 0.128; sqrt(3.128) < 1.77.
        n.g[4] = (float)(sc / (1.77 * (1 + 1e-5)));
        // Second-order reach along the ray (eval_block.hpp gyroid_reach):
        // g(s) = G(q(s) / scale), |q'| <= L, |q''| <= K (the chain's bounds),
        // so |g''| <= |H G| (L / scale)^2 + |grad G| K / scale.  The
        // Hessian's spectral norm is at most 2.0 on a 241^3 grid of the
        // period and, its entries' gradients being at most 2 per component,
        // grows by at most sqrt(27 * 4) * h sqrt(3) / 2 < 0.235 between grid
        // points: |H G| <= 2.25 (tests/test_reach_bound.py repeats the
        // grid).  dg/ds = grad G . v / scale is off by at most e1: each
        // gradient component is two products of sine and cosine values, each
        // value within eq + the unit's error, and v is within a few ulps of L.
        {
          const double inv = 1.0 / sc;
          const double eg = 7.1 * (eq + kHwSinCosErr) + 6 * u;
          const double M = (1 + 1e-5) * (2.25 * (lip_chain * inv) * (lip_chain * inv) +
                                         1.77 * curv_chain * inv);
          n.g[3] = (float)inv;
          n.g[5] = (float)lip_chain;
          n.g[6] = std::isfinite(M) && M < 1e30 ? (float)M : 0.0f;
          n.g[1] = n.g[6] > 0.0f ? (float)(1.0 / (double)n.g[6]) : 0.0f;
          n.g[7] = (float)(2 * lip_chain * inv * (eg + 128 * u));
        }
        break;
      }
      case K_VOXEL:
        n.g[0] = (float)(kappa * (d + 2 * u));
        break;
      case K_TESS: {
        double bm = 0, um = 0, dxmin = INFINITY, dxmax = 0;
        for (int k = 0; k < 6; ++k) bm = std::max(bm, std::fabs(n.d[k]));
        for (int k = 0; k < 3; ++k) {
          um = std::max({um, std::fabs(n.d[6 + k]), std::fabs(n.d[9 + k])});
          double dx = n.d[9 + k] - n.d[6 + k];
          dxmin = std::min(dxmin, std::fabs(dx));
          dxmax = std::max(dxmax, std::fabs(dx));
        }
        double bb = std::min(b, bm + 1.0);  // points past the bounds return early
        n.g[1] = (float)(kappa * (d + u * (bb + bm)));
        const double shift = seamless_enabled() ? seamless(n) : -1.0;
        n.g[2] = shift >= 0.0 ? 1.0f : 0.0f;
        double F = (bb + um) / dxmin;
        n.g[0] = (float)(kappa * ((d + u * um + 2 * u * (bb + um)) / dxmin + 2 * u * F));
        dl[n.depth + 1] = d + 4 * u * (bb + um + dxmax) + std::max(shift, 0.0);
        Bl[n.depth + 1] = um;
        break;
      }
      default: break;
    }
  }
  if (lipschitz) *lipschitz = lip_chain;
  return delta_obj;
}

}  // namespace xrp
