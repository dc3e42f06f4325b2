// scene.hpp - scene loading, flattening and guard-band sizing (host).
#pragma once
#include <memory>
#include <string>
#include <vector>

#include "dev_types.hpp"
#include "json.hpp"

namespace xrp {

enum NumberMode { NUMBERS_JSON = 0, NUMBERS_YAML = 1 };

class SceneError : public std::runtime_error {
 public:
  using std::runtime_error::runtime_error;
};

// Host object tree, one record per reference object.
struct HNode {
  int kind = 0;        // NodeKind
  bool is_cube = false;
  double c[3] = {0, 0, 0};   // centre / origin / p0
  double v[3] = {0, 0, 0};   // sides / p1
  double a = 0, b = 0, rho = 0;
  double pv[9] = {0};        // parallelepiped columns v0 v1 v2
  double minv[9] = {0};      // parallelepiped M^-1 (column-major)
  double bounds[6] = {0};    // tessellation bounds
  double ucb[6] = {0};       // unit cell bounds
  bool greedy = false;
  std::vector<HNode> kids;   // collection children (tess: the UC collection)
  std::shared_ptr<std::vector<float>> vox;  // [z][x][y]
  int nx = 0, ny = 0, nz = 0;
  std::string path;          // voxel grid raw file as written in the scene
};

struct HDeform {
  int kind = 0;  // DefKind
  int axis = 0;
  int src = 0;  // D_TANH source axis
  double p[9] = {0};
  double q[3] = {0};
  double r[3] = {0};
};

struct Box3 {
  double lo[3], hi[3];
  bool empty() const {
    return !(lo[0] <= hi[0] && lo[1] <= hi[1] && lo[2] <= hi[2]);
  }
  bool finite() const;
  static Box3 none();
  static Box3 all();
};

// A loaded scene: the object tree, its deformation chain (composed
// deformations flattened in application order) and derived data.
struct Scene {
  HNode root;
  std::vector<HDeform> defs;
  std::vector<DevNode> nodes;    // flattened preorder (device layout)
  std::vector<DevDeform> ddefs;
  Box3 obj_box;                  // where the object can be non-zero
  Box3 world_box;                // preimage under the deformation chain
  double min_feature = 0;        // MinFeatureSize of the root
  bool has_voxels = false;
  int max_depth = 0;
  size_t voxel_bytes = 0;
};

// Parses an object description with the reference's decoder rules.
std::unique_ptr<Scene> load_scene(const json::Value& object,
                                  const json::Value* deformation,
                                  NumberMode numbers,
                                  const std::string& base_dir);
// Voxel grid scene from caller memory ([z][x][y], float32 densities).
std::unique_ptr<Scene> voxel_scene(const float* rho, int nx, int ny, int nz,
                                   const json::Value* deformation,
                                   NumberMode numbers);

// Appends one deformation description (composed ones flattened).
void append_deformation(const json::Value& d, NumberMode numbers,
                        std::vector<HDeform>& out);

// Builds nodes/ddefs/boxes for a scene whose root/defs are filled.
void finalize_scene(Scene& s);

// The object as the reference's ToMap methods describe it (objects.go), for
// object.json; maps are emitted with sorted keys like Go's encoder.
json::Value to_map(const HNode& n);

// FP32 position error model: returns the per-component bound on
// |p32 - p64| of object-space positions and writes guard bands into
// s.nodes.  world_bound is max |coordinate| over sampled points, step the
// longest FP32 stride from an FP64 anchor (chunk * ds), s_err the error of
// the anchors' FP64 ray parameter.  Also writes the distance-bound slacks
// (g[4..6]) of each node and, into *lipschitz, a Euclidean Lipschitz bound
// of the deformation chain.
double compute_guards(Scene& s, double world_bound, double step,
                      double* lipschitz = nullptr, double s_err = 0.0);

}  // namespace xrp
