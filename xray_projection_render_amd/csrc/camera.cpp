// camera.cpp - FP64 camera matrices and sample grids (host).
//
// computeCameraFromAngles (main.go:226-239) builds
// inverse(LookAt(eye, origin, +z)) with github.com/go-gl/mathgl v1.1.0
// (go.mod:6), whose source is not vendored.  Its published algorithms are
// restated here: Vec3.Normalize (v * (1/|v|)), LookAtV (basis s, u, -f
// rows times a translation), Mat4.Mul4, Mat4.Det/Inv by cofactor
// expansion scaled by 1/det.  Storage is column-major as in mathgl.
#include "camera.hpp"

#include <cmath>
#include <cstring>

namespace xrp {

namespace {
void normalize(const double v[3], double out[3]) {
  double l = 1.0 / std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  out[0] = v[0] * l;
  out[1] = v[1] * l;
  out[2] = v[2] * l;
}

void cross(const double a[3], const double b[3], double out[3]) {
  out[0] = a[1] * b[2] - a[2] * b[1];
  out[1] = a[2] * b[0] - a[0] * b[2];
  out[2] = a[0] * b[1] - a[1] * b[0];
}

void mul4(const double a[16], const double b[16], double out[16]) {
  for (int c = 0; c < 4; ++c)
    for (int r = 0; r < 4; ++r)
      out[c * 4 + r] = a[0 * 4 + r] * b[c * 4 + 0] + a[1 * 4 + r] * b[c * 4 + 1] +
                       a[2 * 4 + r] * b[c * 4 + 2] + a[3 * 4 + r] * b[c * 4 + 3];
}

void inv4(const double m[16], double out[16]) {
  double det =
      m[0] * m[5] * m[10] * m[15] - m[0] * m[5] * m[11] * m[14] -
      m[0] * m[6] * m[9] * m[15] + m[0] * m[6] * m[11] * m[13] +
      m[0] * m[7] * m[9] * m[14] - m[0] * m[7] * m[10] * m[13] -
      m[1] * m[4] * m[10] * m[15] + m[1] * m[4] * m[11] * m[14] +
      m[1] * m[6] * m[8] * m[15] - m[1] * m[6] * m[11] * m[12] -
      m[1] * m[7] * m[8] * m[14] + m[1] * m[7] * m[10] * m[12] +
      m[2] * m[4] * m[9] * m[15] - m[2] * m[4] * m[11] * m[13] -
      m[2] * m[5] * m[8] * m[15] + m[2] * m[5] * m[11] * m[12] +
      m[2] * m[7] * m[8] * m[13] - m[2] * m[7] * m[9] * m[12] -
      m[3] * m[4] * m[9] * m[14] + m[3] * m[4] * m[10] * m[13] +
      m[3] * m[5] * m[8] * m[14] - m[3] * m[5] * m[10] * m[12] -
      m[3] * m[6] * m[8] * m[13] + m[3] * m[6] * m[9] * m[12];
  if (det == 0.0 || std::fabs(det) < 1e-20) {
    std::memset(out, 0, 16 * sizeof(double));
    return;
  }
  // Adjugate entries, column-major.
  double r[16] = {
      -m[7] * m[10] * m[13] + m[6] * m[11] * m[13] + m[7] * m[9] * m[14] -
          m[5] * m[11] * m[14] - m[6] * m[9] * m[15] + m[5] * m[10] * m[15],
      m[3] * m[10] * m[13] - m[2] * m[11] * m[13] - m[3] * m[9] * m[14] +
          m[1] * m[11] * m[14] + m[2] * m[9] * m[15] - m[1] * m[10] * m[15],
      -m[3] * m[6] * m[13] + m[2] * m[7] * m[13] + m[3] * m[5] * m[14] -
          m[1] * m[7] * m[14] - m[2] * m[5] * m[15] + m[1] * m[6] * m[15],
      m[3] * m[6] * m[9] - m[2] * m[7] * m[9] - m[3] * m[5] * m[10] +
          m[1] * m[7] * m[10] + m[2] * m[5] * m[11] - m[1] * m[6] * m[11],
      m[7] * m[10] * m[12] - m[6] * m[11] * m[12] - m[7] * m[8] * m[14] +
          m[4] * m[11] * m[14] + m[6] * m[8] * m[15] - m[4] * m[10] * m[15],
      -m[3] * m[10] * m[12] + m[2] * m[11] * m[12] + m[3] * m[8] * m[14] -
          m[0] * m[11] * m[14] - m[2] * m[8] * m[15] + m[0] * m[10] * m[15],
      m[3] * m[6] * m[12] - m[2] * m[7] * m[12] - m[3] * m[4] * m[14] +
          m[0] * m[7] * m[14] + m[2] * m[4] * m[15] - m[0] * m[6] * m[15],
      -m[3] * m[6] * m[8] + m[2] * m[7] * m[8] + m[3] * m[4] * m[10] -
          m[0] * m[7] * m[10] - m[2] * m[4] * m[11] + m[0] * m[6] * m[11],
      -m[7] * m[9] * m[12] + m[5] * m[11] * m[12] + m[7] * m[8] * m[13] -
          m[4] * m[11] * m[13] - m[5] * m[8] * m[15] + m[4] * m[9] * m[15],
      m[3] * m[9] * m[12] - m[1] * m[11] * m[12] - m[3] * m[8] * m[13] +
          m[0] * m[11] * m[13] + m[1] * m[8] * m[15] - m[0] * m[9] * m[15],
      -m[3] * m[5] * m[12] + m[1] * m[7] * m[12] + m[3] * m[4] * m[13] -
          m[0] * m[7] * m[13] - m[1] * m[4] * m[15] + m[0] * m[5] * m[15],
      m[3] * m[5] * m[8] - m[1] * m[7] * m[8] - m[3] * m[4] * m[9] +
          m[0] * m[7] * m[9] + m[1] * m[4] * m[11] - m[0] * m[5] * m[11],
      m[6] * m[9] * m[12] - m[5] * m[10] * m[12] - m[6] * m[8] * m[13] +
          m[4] * m[10] * m[13] + m[5] * m[8] * m[14] - m[4] * m[9] * m[14],
      -m[2] * m[9] * m[12] + m[1] * m[10] * m[12] + m[2] * m[8] * m[13] -
          m[0] * m[10] * m[13] - m[1] * m[8] * m[14] + m[0] * m[9] * m[14],
      m[2] * m[5] * m[12] - m[1] * m[6] * m[12] - m[2] * m[4] * m[13] +
          m[0] * m[6] * m[13] + m[1] * m[4] * m[14] - m[0] * m[5] * m[14],
      -m[2] * m[5] * m[8] + m[1] * m[6] * m[8] + m[2] * m[4] * m[9] -
          m[0] * m[6] * m[9] - m[1] * m[4] * m[10] + m[0] * m[5] * m[10],
  };
  double s = 1 / det;
  for (int k = 0; k < 16; ++k) out[k] = r[k] * s;
}
}  // namespace

void camera_from_angles(double az_deg, double polar_deg, double R,
                        double eye[3], double cam[16]) {
  double th = az_deg * M_PI / 180;    // mgl64.DegToRad
  double phi = polar_deg * M_PI / 180;
  eye[0] = R * std::cos(th) * std::sin(phi);
  eye[1] = R * std::sin(th) * std::sin(phi);
  eye[2] = std::cos(phi) * R;
  const double up[3] = {0, 0, 1};
  double fwd[3] = {0 - eye[0], 0 - eye[1], 0 - eye[2]};
  double f[3], upn[3], s0[3], s[3], u[3];
  normalize(fwd, f);
  normalize(up, upn);
  cross(f, upn, s0);
  normalize(s0, s);
  cross(s, f, u);
  const double M[16] = {s[0], u[0], -f[0], 0, s[1], u[1], -f[1], 0,
                        s[2], u[2], -f[2], 0, 0,    0,    0,     1};
  const double T[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0,
                        -eye[0], -eye[1], -eye[2], 1};
  double view[16], inv[16];
  mul4(M, T, view);
  inv4(view, inv);
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) cam[r * 4 + c] = inv[c * 4 + r];
}

double focal_length(double fov_deg) {
  return 1 / std::tan(fov_deg / 2 * M_PI / 180);  // main.go:457
}

std::vector<double> simple_samples(double smin, double smax, double ds) {
  std::vector<double> s;
  for (double x = smin; x < smax; x += ds) s.push_back(x);  // main.go:147
  return s;
}

std::vector<double> coarse_samples(double smin, double smax, double DS) {
  std::vector<double> s;
  for (double right = smin + DS; right <= smax; right += DS)  // main.go:171-196
    s.push_back(right);
  return s;
}

}  // namespace xrp
