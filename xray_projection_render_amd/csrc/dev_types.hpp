// dev_types.hpp - flattened scene layout shared by the host flattener and
// the gfx950 kernels.
//
// A scene is a preorder array of DevNode (one 256-byte record per object,
// collection, tessellation or voxel grid) plus a short deformation chain.
// Every wave evaluates the same node at the same time, so node fields are
// wave-uniform and live in scalar registers (s_load from the constant
// cache); no per-lane copy is made.
#pragma once
#include <stdint.h>

namespace xrp {

enum NodeKind : int32_t {
  K_SPHERE = 1,
  K_BOX = 2,     // cube and box (objects.go:115 builds a Box for a cube)
  K_PPED = 3,
  K_CYL = 4,
  K_GYROID = 5,
  K_COLL = 6,    // object_collection / unit-cell collection
  K_TESS = 7,    // tessellated_obj_coll (+ its unit cell bounds)
  K_VOXEL = 8,
};

enum DefKind : int32_t {
  D_GAUSS = 1, D_AFFINE = 2, D_LINEAR = 3, D_RIGID = 4, D_SIGMOID = 5,
  D_TANH = 6,  // extension: p[axis] += amplitude * tanh(rate * p[source])
};

// Largest tessellation nesting the device evaluator unrolls.  Depth 0 is
// the root; a tessellation's unit-cell collection sits one level deeper.
constexpr int kMaxDepth = 3;
constexpr int kMaxDeform = 8;

// Absolute error bound of the hardware sine and cosine (v_sin_f32 and
// v_cos_f32, argument in revolutions) over every FP32 argument in
// [-0.5, 0.5]: the largest error against FP64 sin/cos(2 pi t) over all of
// them is 1.254e-7 on MI355X (xrp_diag_sincos_error re-measures it).  The
// FP32 gyroid evaluators reduce their argument to that range exactly and
// the guard bands carry this bound.
constexpr double kHwSinCosErr = 1.3e-7;

// The FP32 deformations evaluate tanh, the logistic function 1 / (1 + e^-t)
// and e^-x (x >= 0) through the hardware base-2 exponential and reciprocal
// (eval_device.hpp tanh_hw, sigmoid_hw, expneg_hw).  Their largest absolute
// errors against FP64 over every FP32 argument (tanh on [-20, 20],
// beyond which both round to +-1; the logistic function on [-100, 100];
// e^-x on [0, 110]; outside, the results are the exact limits) are at most
// these bounds on MI355X (xrp_diag_deform_error re-measures them); the
// deformations' position error bounds carry them.
constexpr double kHwTanhErr = 3.0e-7;
constexpr double kHwSigmoidErr = 2.0e-7;
constexpr double kHwExpNegErr = 2.0e-7;

// 256-byte node.  f[] holds FP32 copies for the guarded march, d[] the
// FP64 values the exact evaluator uses (same values the reference uses).
// g[] holds the FP32 guard bands, recomputed on the host for every render
// from that render's position error bound (see guards.cpp).
struct alignas(16) DevNode {
  int32_t kind;
  int32_t next;      // index one past this node's subtree
  int32_t first;     // COLL: first child; TESS: index of its collection
  int32_t greedy;    // COLL: greedy evaluation (objects.go:426)
  int32_t depth;     // nesting level (see kMaxDepth)
  int32_t nx, ny, nz;  // VOXEL dims
  float f[16];
  float g[8];
  double d[14];
  union {
    const float* vox;  // VOXEL: device pointer, [z][x][y] layout
    float bsph[4];     // SPHERE/BOX/PPED/CYL: object-space bounding sphere
                       // (centre, radius), enclosing the primitive's closure
  };
};
static_assert(sizeof(DevNode) == 256, "DevNode must stay 256 bytes");

// Parameter slots.
//  SPHERE f: cx cy cz r2 rho        d: cx cy cz r rho
//  BOX    f: cx cy cz hx hy hz rho  d: cx cy cz hx hy hz rho    (h = 0.5*side)
//  PPED   f: ox oy oz m00..m22 rho  d: ox oy oz m[9] rho  (m = M^-1 column-major)
//  CYL    f: p0(3) v(3) vv inv_vv r2 rho   d: p0(3) v(3) vv r rho
//  GYROID f: cx cy cz inv_scale thick rho inv_scale/(2 pi)
//         d: cx cy cz scale thick rho
//  TESS   f: bounds(6) ucmin(3) dx dy dz inv_dx inv_dy inv_dz
//         d: bounds(6) ucmin(3) ucmax(3)   (dx = ucmax - ucmin, objects.go:574)
//  VOXEL  f: (nx-1)/2 (ny-1)/2 (nz-1)/2
//  guards: SPHERE/CYL q-band in [0]; CYL t-band in [1]; BOX/PPED per axis
//          [0..2]; GYROID [0]; TESS wrap (fraction units) [0], bounds [1];
//          VOXEL bounds [0]; COLL sum-near-zero band [0] (hierarchical).
//  g[] also holds, per node kind, the safe-radius terms (eval_block.hpp):
//          GYROID [3] 1/scale, [4] radius factor, [5] the deformation chain's
//          Lipschitz bound, [6] the curvature bound M of the second-order
//          reach (0: none), [1] 1/M, [7] the slope's error bound; TESS [2] 1 when the
//          content is seamless (a whole number of gyroid periods per cell).

struct alignas(16) DevDeform {
  int32_t kind;
  int32_t axis;   // SIGMOID, TANH: displaced axis
  int32_t src;    // TANH: source axis
  int32_t pad;
  float f[12];
  double d[12];
};
//  GAUSS  f/d: amp(3) sigma(3) center(3)     (f: sigma slots hold 1/(2 s^2))
//  AFFINE f/d: m[9] row-major
//  LINEAR f/d: strains[6]
//  RIGID  f/d: disp[3]
//  SIGMOID f/d: amplitude center lengthscale   (f[2] = 1/lengthscale)
//  TANH   f/d: amplitude rate

}  // namespace xrp
