// host_driver.cpp - RenderProjections / FreeString: the reference's JSON
// API (api.go:82-180) driving its render() flow (main.go:310-636) on the
// GPU: scene and deformation files, the ds rule, camera angles, PNG frames,
// transforms.json, object.json and the optional volume.raw.
#include <sys/stat.h>
#include <zlib.h>

#include <algorithm>
#include <cerrno>
#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <functional>
#include <mutex>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/xrp.h"
#include "camera.hpp"
#include "json.hpp"
#include "scene.hpp"
#include "yaml_lite.hpp"

namespace xrp {
const HNode& scene_root(const xrp_scene* s);         // api.cpp
int set_error(int code, const std::string& message);  // api.cpp
}

namespace {

using xrp::json::Value;

class Fatal : public std::runtime_error {
 public:
  using std::runtime_error::runtime_error;
};

// ---------------------------------------------------------------------------
// Go encoding/json output
// ---------------------------------------------------------------------------
// strconv.AppendFloat(f, fmt, -1, 64) with encoding/json's choice of 'e'
// below 1e-6 or from 1e21 on, and its e-09 -> e-9 clean-up.
std::string go_float(double f) {
  if (!std::isfinite(f)) throw Fatal("json: unsupported value: " + std::to_string(f));
  if (f == 0) return std::signbit(f) ? "-0" : "0";
  double a = std::fabs(f);
  bool e = a < 1e-6 || a >= 1e21;
  char buf[64];
  auto r = std::to_chars(buf, buf + sizeof buf, f,
                         e ? std::chars_format::scientific : std::chars_format::fixed);
  std::string s(buf, r.ptr);
  if (e) {
    size_t n = s.size();
    if (n >= 4 && s[n - 4] == 'e' && s[n - 3] == '-' && s[n - 2] == '0') s.erase(n - 2, 1);
  }
  return s;
}

std::string go_string(const std::string& s) {
  std::string o = "\"";
  for (unsigned char c : s) {
    switch (c) {
      case '"': o += "\\\""; break;
      case '\\': o += "\\\\"; break;
      case '\n': o += "\\n"; break;
      case '\r': o += "\\r"; break;
      case '\t': o += "\\t"; break;
      case '<': o += "\\u003c"; break;
      case '>': o += "\\u003e"; break;
      case '&': o += "\\u0026"; break;
      default:
        if (c < 0x20) {
          char b[8];
          std::snprintf(b, sizeof b, "\\u%04x", c);
          o += b;
        } else {
          o += (char)c;
        }
    }
  }
  return o + "\"";
}

// json.MarshalIndent(v, "", "  ") (indent > 0) or json.Marshal (indent < 0).
void encode(const Value& v, std::string& o, int indent, int depth) {
  auto nl = [&](int d) {
    if (indent < 0) return;
    o += '\n';
    o.append((size_t)(d * indent), ' ');
  };
  switch (v.type) {
    case Value::Null: o += "null"; break;
    case Value::Bool: o += v.b ? "true" : "false"; break;
    case Value::Int: o += std::to_string(v.i); break;
    case Value::Float: o += go_float(v.f); break;
    case Value::String: o += go_string(v.s); break;
    case Value::Arr:
      if (v.a->empty()) {
        o += "[]";
        break;
      }
      o += '[';
      for (size_t k = 0; k < v.a->size(); ++k) {
        if (k) o += ',';
        nl(depth + 1);
        encode((*v.a)[k], o, indent, depth + 1);
      }
      nl(depth);
      o += ']';
      break;
    case Value::Obj:
      if (v.o->empty()) {
        o += "{}";
        break;
      }
      o += '{';
      for (size_t k = 0; k < v.o->size(); ++k) {
        if (k) o += ',';
        nl(depth + 1);
        o += go_string((*v.o)[k].first);
        o += indent < 0 ? ":" : ": ";
        encode((*v.o)[k].second, o, indent, depth + 1);
      }
      nl(depth);
      o += '}';
      break;
  }
}

std::string marshal(const Value& v, bool indent) {
  std::string o;
  encode(v, o, indent ? 2 : -1, 0);
  return o;
}

Value vnum(double f) {
  Value v;
  v.type = Value::Float;
  v.f = f;
  return v;
}
Value vint(long long i) {
  Value v;
  v.type = Value::Int;
  v.i = i;
  return v;
}
Value vstr(const std::string& s) {
  Value v;
  v.type = Value::String;
  v.s = s;
  return v;
}
Value vbool(bool b) {
  Value v;
  v.type = Value::Bool;
  v.b = b;
  return v;
}
Value varr() {
  Value v;
  v.type = Value::Arr;
  v.a = std::make_shared<xrp::json::Array>();
  return v;
}
Value vobj() {
  Value v;
  v.type = Value::Obj;
  v.o = std::make_shared<xrp::json::Object>();
  return v;
}

// ---------------------------------------------------------------------------
// Go path/filepath (Unix)
// ---------------------------------------------------------------------------
std::string clean(const std::string& p) {
  if (p.empty()) return ".";
  bool rooted = p[0] == '/';
  std::vector<std::string> parts;
  std::stringstream ss(p);
  std::string seg;
  while (std::getline(ss, seg, '/')) {
    if (seg.empty() || seg == ".") continue;
    if (seg == "..") {
      if (!parts.empty() && parts.back() != "..")
        parts.pop_back();
      else if (!rooted)
        parts.push_back("..");
      continue;
    }
    parts.push_back(seg);
  }
  std::string out = rooted ? "/" : "";
  for (size_t k = 0; k < parts.size(); ++k) out += (k ? "/" : "") + parts[k];
  return out.empty() ? "." : out;
}

std::string join(const std::string& a, const std::string& b) {
  if (a.empty()) return b.empty() ? "" : clean(b);
  if (b.empty()) return clean(a);
  return clean(a + "/" + b);
}

void split(const std::string& p, std::string& dir, std::string& file) {
  size_t k = p.rfind('/');
  dir = k == std::string::npos ? "" : p.substr(0, k + 1);
  file = k == std::string::npos ? p : p.substr(k + 1);
}

std::string base(std::string p) {
  if (p.empty()) return ".";
  while (p.size() > 1 && p.back() == '/') p.pop_back();
  size_t k = p.rfind('/');
  if (k != std::string::npos) p = p.substr(k + 1);
  return p.empty() ? "/" : p;
}

std::string dir_of(const std::string& p) {
  std::string d, f;
  split(p, d, f);
  return clean(d.empty() ? "." : d);
}

void mkdir_all(const std::string& p) {
  std::string cur;
  std::stringstream ss(p);
  std::string seg;
  if (!p.empty() && p[0] == '/') cur = "/";
  while (std::getline(ss, seg, '/')) {
    if (seg.empty()) continue;
    cur += seg + "/";
    if (::mkdir(cur.c_str(), 0755) != 0 && errno != EEXIST)
      throw Fatal("cannot create directory " + cur);
  }
}

// fmt.Sprintf(pattern, i) for patterns holding one integer verb
// (%d with flags and width) and %% escapes.
std::string sprintf_int(const std::string& pat, int i) {
  std::string out;
  int verbs = 0;
  for (size_t k = 0; k < pat.size(); ++k) {
    if (pat[k] != '%') {
      out += pat[k];
      continue;
    }
    size_t e = k + 1;
    if (e < pat.size() && pat[e] == '%') {
      out += '%';
      k = e;
      continue;
    }
    while (e < pat.size() && std::strchr("+- 0#", pat[e])) ++e;
    while (e < pat.size() && std::isdigit((unsigned char)pat[e])) ++e;
    if (e >= pat.size() || pat[e] != 'd' || verbs++)
      throw Fatal("fname_pattern must hold one integer verb such as %03d");
    std::string spec = pat.substr(k, e - k + 1);
    char buf[64];
    std::snprintf(buf, sizeof buf, spec.c_str(), i);
    out += buf;
    k = e;
  }
  return out;
}

// ---------------------------------------------------------------------------
// PNG (image/png's choice: RGB when every alpha is opaque, else RGBA)
// ---------------------------------------------------------------------------
void put32(std::string& s, uint32_t v) {
  for (int k = 3; k >= 0; --k) s += (char)((v >> (8 * k)) & 0xff);
}

void chunk(std::string& png, const char* type, const std::string& data) {
  put32(png, (uint32_t)data.size());
  std::string td = std::string(type, 4) + data;
  png += td;
  put32(png, (uint32_t)crc32(0, (const Bytef*)td.data(), (uInt)td.size()));
}

// pix: [i][j] (gray, alpha) pairs; PNG pixel (x=i, y=res-1-j), main.go:498.
// image.RGBA holds alpha-premultiplied samples and PNG stores straight alpha,
// so a fully transparent pixel is written as (0, 0, 0, 0).
void write_png(const std::string& path, const unsigned char* pix, int res) {
  bool opaque = true;
  for (size_t k = 0; k < (size_t)res * res && opaque; ++k) opaque = pix[2 * k + 1] == 255;
  const int ch = opaque ? 3 : 4;
  std::vector<unsigned char> raw((size_t)res * (1 + (size_t)res * ch));
  size_t o = 0;
  for (int y = 0; y < res; ++y) {
    raw[o++] = 0;  // filter: none
    int j = res - 1 - y;
    for (int x = 0; x < res; ++x) {
      const unsigned char* px = pix + 2 * ((size_t)x * res + j);
      const unsigned char g = px[1] == 0 ? 0 : px[0];
      raw[o++] = g;
      raw[o++] = g;
      raw[o++] = g;
      if (ch == 4) raw[o++] = px[1];
    }
  }
  uLongf zlen = compressBound((uLong)raw.size());
  std::string z(zlen, '\0');
  if (compress2((Bytef*)z.data(), &zlen, raw.data(), (uLong)raw.size(), Z_DEFAULT_COMPRESSION) !=
      Z_OK)
    throw Fatal("png: compression failed");
  z.resize(zlen);
  std::string png("\x89PNG\r\n\x1a\n", 8), ihdr;
  put32(ihdr, (uint32_t)res);
  put32(ihdr, (uint32_t)res);
  ihdr += (char)8;
  ihdr += (char)(opaque ? 2 : 6);
  ihdr += std::string("\0\0\0", 3);
  chunk(png, "IHDR", ihdr);
  chunk(png, "IDAT", z);
  chunk(png, "IEND", "");
  std::ofstream f(path, std::ios::binary);
  if (!f) throw Fatal("error creating output file " + path);
  f.write(png.data(), (std::streamsize)png.size());
}

// ---------------------------------------------------------------------------
// Parameters (api.go:41-64) with Go's decoding: keys match field tags
// case-insensitively, absent fields keep their zero value, mistyped ones
// fail the whole decode.
// ---------------------------------------------------------------------------
struct Params {
  std::string input, output_dir, fname_pattern, transforms_file, deformation_file,
      integration, log_level;
  long long resolution = 0, num_images = 0, jobs_modulo = 0, job_num = 0;
  bool out_of_plane = false, transparency = false, export_volume = false;
  double ds = 0, R = 0, fov = 0, time_label = 0, polar_angle = 0, density_multiplier = 0,
         flat_field = 0;
  std::vector<std::pair<double, double>> angles;
};

const Value* field(const Value& o, const char* name) {
  for (auto& kv : *o.o) {
    if (kv.first.size() != std::strlen(name)) continue;
    bool eq = true;
    for (size_t k = 0; k < kv.first.size() && eq; ++k)
      eq = std::tolower((unsigned char)kv.first[k]) == std::tolower((unsigned char)name[k]);
    if (eq) return &kv.second;
  }
  return nullptr;
}

// encoding/json's name for the kind of a mistyped JSON value
std::string kind(const Value& v) {
  switch (v.type) {
    case Value::Bool: return "bool";
    case Value::Int:
    case Value::Float: return "number";
    case Value::String: return "string";
    case Value::Arr: return "array";
    case Value::Obj: return "object";
    default: return "null";
  }
}

[[noreturn]] void type_error(const std::string& what, const char* name, const char* type) {
  throw Fatal("Failed to parse parameters: json: cannot unmarshal " + what +
              " into Go struct field RenderParams." + name + " of type " + type);
}

void get(const Value& o, const char* name, std::string& out) {
  const Value* v = field(o, name);
  if (!v || v->is_null()) return;
  if (v->type != Value::String) type_error(kind(*v), name, "string");
  out = v->s;
}
void get(const Value& o, const char* name, long long& out) {
  const Value* v = field(o, name);
  if (!v || v->is_null()) return;
  if (!v->is_number()) type_error(kind(*v), name, "int");
  double d = v->number();
  if (d != std::floor(d)) type_error("number " + go_float(d), name, "int");
  out = (long long)d;
}
void get(const Value& o, const char* name, double& out) {
  const Value* v = field(o, name);
  if (!v || v->is_null()) return;
  if (!v->is_number()) type_error(kind(*v), name, "float64");
  out = v->number();
}
void get(const Value& o, const char* name, bool& out) {
  const Value* v = field(o, name);
  if (!v || v->is_null()) return;
  if (v->type != Value::Bool) type_error(kind(*v), name, "bool");
  out = v->b;
}

Params parse_params(const char* text) {
  Value o;
  try {
    o = xrp::json::parse(text ? text : "");
  } catch (const std::exception& e) {
    throw Fatal(std::string("Failed to parse parameters: ") + e.what());
  }
  if (o.type != Value::Obj)
    throw Fatal("Failed to parse parameters: json: cannot unmarshal value into Go value of "
                "type main.RenderParams");
  Params p;
  get(o, "input", p.input);
  get(o, "output_dir", p.output_dir);
  get(o, "fname_pattern", p.fname_pattern);
  get(o, "resolution", p.resolution);
  get(o, "num_images", p.num_images);
  get(o, "out_of_plane", p.out_of_plane);
  get(o, "ds", p.ds);
  get(o, "R", p.R);
  get(o, "fov", p.fov);
  get(o, "jobs_modulo", p.jobs_modulo);
  get(o, "job_num", p.job_num);
  get(o, "transforms_file", p.transforms_file);
  get(o, "deformation_file", p.deformation_file);
  get(o, "time_label", p.time_label);
  get(o, "transparency", p.transparency);
  get(o, "export_volume", p.export_volume);
  get(o, "polar_angle", p.polar_angle);
  get(o, "density_multiplier", p.density_multiplier);
  get(o, "flat_field", p.flat_field);
  get(o, "integration", p.integration);
  get(o, "log_level", p.log_level);
  if (const Value* ca = field(o, "camera_angles")) {
    if (ca->type == Value::Arr) {
      for (auto& a : *ca->a) {
        if (a.type != Value::Obj) type_error(kind(a), "camera_angles", "main.CameraAngle");
        double az = 0, po = 0;
        if (const Value* v = field(a, "azimuthal")) {
          if (!v->is_number()) type_error(kind(*v), "camera_angles.azimuthal", "float64");
          az = v->number();
        }
        if (const Value* v = field(a, "polar")) {
          if (!v->is_number()) type_error(kind(*v), "camera_angles.polar", "float64");
          po = v->number();
        }
        p.angles.emplace_back(az, po);
      }
    } else if (!ca->is_null()) {
      type_error(kind(*ca), "camera_angles", "[]main.CameraAngle");
    }
  }
  return p;
}

// A failed read is log.Fatal in the reference (main.go:57-59, 93-95), which
// its API turns into the panic text "fatal: <msg>" (api.go:34-37).
std::string read_file(const std::string& path, const char* what) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw Fatal(std::string("fatal: ") + what);
  return std::string((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
}

// main.go:63-76 / 97-110: the last four characters pick the decoder.
// yaml.v3 keeps integer literals as int; encoding/json makes every number
// a float64; any other extension decodes nothing (an empty map).
Value load_doc(const std::string& path, const std::string& text) {
  std::string ext = path.size() >= 4 ? path.substr(path.size() - 4) : path;
  if (ext == "yaml") return xrp::yaml::parse(text);
  if (ext != "json") return vobj();
  Value v = xrp::json::parse(text);
  std::function<void(Value&)> to_float = [&](Value& x) {
    if (x.type == Value::Int) {
      x.type = Value::Float;
      x.f = (double)x.i;
    } else if (x.type == Value::Arr) {
      for (auto& e : *x.a) to_float(e);
    } else if (x.type == Value::Obj) {
      for (auto& kv : *x.o) to_float(kv.second);
    }
  };
  to_float(v);
  return v;
}

// JSON text that keeps the int/float distinction (floats always carry a
// '.' or an exponent), read back with XRP_NUMBERS_YAML.
void typed(const Value& v, std::string& o) {
  switch (v.type) {
    case Value::Float: {
      if (!std::isfinite(v.f)) throw Fatal("non-finite number in scene description");
      char buf[64];
      auto r = std::to_chars(buf, buf + sizeof buf, v.f);
      std::string t(buf, r.ptr);
      if (t.find_first_of(".eE") == std::string::npos) t += ".0";
      o += t;
      break;
    }
    case Value::Arr:
      o += '[';
      for (size_t k = 0; k < v.a->size(); ++k) {
        if (k) o += ',';
        typed((*v.a)[k], o);
      }
      o += ']';
      break;
    case Value::Obj:
      o += '{';
      for (size_t k = 0; k < v.o->size(); ++k) {
        if (k) o += ',';
        o += go_string((*v.o)[k].first) + ":";
        typed((*v.o)[k].second, o);
      }
      o += '}';
      break;
    default: encode(v, o, -1, 0);
  }
}

std::string typed_json(const Value& v) {
  std::string o;
  typed(v, o);
  return o;
}

xrp_context* shared_context(std::string& err) {
  static std::mutex mu;
  static xrp_context* ctx = nullptr;
  std::lock_guard<std::mutex> g(mu);
  if (!ctx && xrp_context_create(0, &ctx) != XRP_OK) {
    err = xrp_last_error();
    ctx = nullptr;
  }
  return ctx;
}

void check_rc(int rc) {
  if (rc != XRP_OK) throw Fatal(xrp_last_error());
}

struct SceneGuard {
  xrp_scene* s = nullptr;
  ~SceneGuard() { xrp_scene_destroy(s); }
};

void render(const Params& p) {
  // load_object (main.go:90-120), load_deformation (main.go:50-85)
  std::string obj_json, def_json;
  try {
    obj_json = typed_json(load_doc(p.input, read_file(p.input, "error reading object file")));
    if (!p.deformation_file.empty())
      def_json = typed_json(load_doc(
          p.deformation_file, read_file(p.deformation_file, "error reading deformation file")));
  } catch (const Fatal&) {
    throw;
  } catch (const std::exception& e) {
    throw Fatal(std::string("Error unmarshalling scene file: ") + e.what());
  }
  SceneGuard sc;
  // raw voxel paths resolve against the working directory, as os.ReadFile.
  // A bad deformation is fatal in the reference (main.go:338-341); a bad
  // object there renders a nil object and panics later, here it stops now.
  int rc = xrp_scene_create(obj_json.c_str(), def_json.empty() ? nullptr : def_json.c_str(),
                            XRP_NUMBERS_YAML, nullptr, &sc.s);
  if (rc == XRP_ERR_DEFORMATION)
    throw Fatal(std::string("fatal: Error loading deformation: ") + xrp_last_error());
  if (rc != XRP_OK) throw Fatal(std::string("Error loading object: ") + xrp_last_error());
  mkdir_all(p.output_dir.empty() ? "." : p.output_dir);

  double ds = p.ds;
  if (ds < 0) check_rc(xrp_scene_auto_ds(sc.s, &ds));  // main.go:350-353

  // camera angles (main.go:356-371, 242-257)
  std::vector<std::pair<double, double>> angles = p.angles;
  if (angles.empty()) {
    if (p.jobs_modulo < 1 && p.num_images > p.job_num)
      throw Fatal("jobs_modulo must be at least 1");
    std::mt19937_64 rng{std::random_device{}()};
    std::uniform_real_distribution<double> uni(0.0, 1.0);
    for (long long i = p.job_num; i < p.num_images; i += p.jobs_modulo) {
      double dth = 360.0 / (double)p.num_images;
      double th = (double)i * dth + 90.0;
      double phi = p.out_of_plane ? std::acos(uni(rng) * 2 - 1) * 180.0 / M_PI : p.polar_angle;
      angles.emplace_back(th, phi);
    }
  }
  const int res = (int)p.resolution;
  const double res_f = (double)res;

  // transforms.json (main.go:380-388, 447-459, 520-534)
  Value tp = vobj();
  tp.o->emplace_back("flat_field", vnum(std::exp(-p.flat_field)));
  tp.o->emplace_back("camera_angle_x", vnum(p.fov * M_PI / 180.0));
  double fl = 0;
  if (!angles.empty()) fl = 1 / std::tan(p.fov / 2 * M_PI / 180) * res_f / 2.0;
  tp.o->emplace_back("fl_x", vnum(fl));
  tp.o->emplace_back("fl_y", vnum(fl));
  tp.o->emplace_back("w", vint(res));
  tp.o->emplace_back("h", vint(res));
  tp.o->emplace_back("cx", vnum(res_f / 2.0));
  tp.o->emplace_back("cy", vnum(res_f / 2.0));
  Value frames = varr();

  if (!angles.empty()) {
    if (res <= 0) throw Fatal("resolution must be positive");
    std::string err;
    xrp_context* ctx = shared_context(err);
    if (!ctx) throw Fatal("GPU renderer unavailable: " + err);
    xrp_render_params rp;
    xrp_default_params(&rp);
    rp.resolution = res;
    rp.ds = ds;
    rp.R = p.R;
    rp.fov_deg = p.fov;
    rp.flat_field = p.flat_field;
    rp.density_multiplier = p.density_multiplier;
    rp.integrator = p.integration == "simple" ? XRP_INTEGRATE_SIMPLE  // api.go:128-132
                                              : XRP_INTEGRATE_HIERARCHICAL;
    rp.output_format = XRP_OUTPUT_GRAY8_ALPHA8;
    rp.transparency = p.transparency ? 1 : 0;
    const size_t img = (size_t)res * res * 2;
    const size_t batch = std::max<size_t>(1, ((size_t)256 << 20) / img);
    std::vector<unsigned char> pix;
    for (size_t v0 = 0; v0 < angles.size(); v0 += batch) {
      size_t nv = std::min(batch, angles.size() - v0);
      std::vector<xrp_view_angle> va(nv);
      for (size_t k = 0; k < nv; ++k) va[k] = {angles[v0 + k].first, angles[v0 + k].second};
      pix.resize(nv * img);
      check_rc(xrp_render(ctx, sc.s, &rp, va.data(), (int)nv, pix.data(), nullptr));
      for (size_t k = 0; k < nv; ++k) {
        size_t i_img = v0 + k;
        std::string filename = join(p.output_dir, sprintf_int(p.fname_pattern, (int)i_img));
        write_png(filename, pix.data() + k * img, res);
        double eye[3], cam[16];
        xrp::camera_from_angles(angles[i_img].first, angles[i_img].second, p.R, eye, cam);
        Value tm = varr();
        for (int r = 0; r < 4; ++r) {
          Value row = varr();
          for (int c = 0; c < 4; ++c) row.a->push_back(vnum(cam[r * 4 + c]));
          tm.a->push_back(row);
        }
        std::string dname, fname;
        split(filename, dname, fname);
        Value fr = vobj();
        fr.o->emplace_back("file_path", vstr(join(base(dname), fname)));
        fr.o->emplace_back("time", vnum(p.time_label));
        fr.o->emplace_back("transform_matrix", tm);
        frames.a->push_back(fr);
      }
    }
  }
  tp.o->emplace_back("frames", frames);
  {
    std::ofstream f(p.transforms_file, std::ios::binary);
    if (!f) throw Fatal("Error writing JSON to file");
    f << marshal(tp, true);
  }
  {
    // object.json next to the output directory (main.go:536-546)
    std::ofstream f(join(dir_of(p.output_dir), "object.json"), std::ios::binary);
    if (!f) throw Fatal("Error writing object.json to file");
    f << marshal(xrp::to_map(xrp::scene_root(sc.s)), true);
  }
  if (p.export_volume) {  // main.go:549-635
    if (res <= 0) throw Fatal("resolution must be positive");
    std::string err;
    xrp_context* ctx = shared_context(err);
    if (!ctx) throw Fatal("GPU renderer unavailable: " + err);
    std::vector<uint8_t> vol((size_t)res * res * res);
    check_rc(xrp_export_volume(ctx, sc.s, res, p.density_multiplier, vol.data()));
    std::ofstream f(join(dir_of(p.output_dir), "volume.raw"), std::ios::binary);
    if (!f) throw Fatal("Error writing volume.raw to file");
    f.write((const char*)vol.data(), (std::streamsize)vol.size());
  }
}

char* cstring(const std::string& s) {
  char* out = (char*)std::malloc(s.size() + 1);
  if (out) std::memcpy(out, s.c_str(), s.size() + 1);
  return out;
}

char* result(bool ok, const std::string& err, long long num_images, const std::string& dir) {
  Value r = vobj();  // RenderResult (api.go:67-72), error omitted when empty
  r.o->emplace_back("success", vbool(ok));
  if (!err.empty()) r.o->emplace_back("error", vstr(err));
  r.o->emplace_back("num_images", vint(num_images));
  r.o->emplace_back("output_dir", vstr(dir));
  return cstring(marshal(r, false));
}

}  // namespace

extern "C" {

int xrp_scene_load_files(const char* object_path, const char* deformation_path,
                         xrp_scene** out) {
  if (!out || !object_path) return xrp::set_error(XRP_ERR_NULL, "NULL argument");
  *out = nullptr;
  std::string obj_json, def_json;
  const char* reading = object_path;
  try {
    std::string text = read_file(object_path, "error reading object file");
    reading = nullptr;
    obj_json = typed_json(load_doc(object_path, text));
    if (deformation_path && deformation_path[0]) {
      reading = deformation_path;
      text = read_file(deformation_path, "error reading deformation file");
      reading = nullptr;
      def_json = typed_json(load_doc(deformation_path, text));
    }
  } catch (const std::exception& e) {
    if (reading) return xrp::set_error(XRP_ERR_IO, std::string("cannot read ") + reading);
    return xrp::set_error(XRP_ERR_SCENE, std::string("Error unmarshalling: ") + e.what());
  }
  return xrp_scene_create(obj_json.c_str(), def_json.empty() ? nullptr : def_json.c_str(),
                          XRP_NUMBERS_YAML, nullptr, out);
}

char* RenderProjections(const char* json_params) {
  Params p;
  try {
    p = parse_params(json_params);
  } catch (const std::exception& e) {
    return result(false, e.what(), 0, "");
  }
  // api.go:97-103
  if (p.ds == 0)
    return result(false, "ds is 0; use a negative value for automatic step-size selection", 0,
                  "");
  if (p.density_multiplier == 0)
    return result(false,
                  "density_multiplier is 0; all densities will be zero and the render will "
                  "produce a blank image",
                  0, "");
  try {
    render(p);
  } catch (const std::exception& e) {
    return result(false, std::string("render failed: ") + e.what(), 0, "");
  }
  return result(true, "", p.num_images, p.output_dir);
}

void FreeString(char* s) { std::free(s); }

}  // extern "C"
