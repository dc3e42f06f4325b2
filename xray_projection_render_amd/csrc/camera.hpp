// camera.hpp - FP64 cameras and per-ray sample grids (host).
#pragma once
#include <vector>

namespace xrp {

// main.go:46
constexpr double kCubeHalfDiagonal = 1.74;

// eye and row-major cam2world (cam[r*4+c] = camera.At(r, c)).
void camera_from_angles(double az_deg, double polar_deg, double R,
                        double eye[3], double cam[16]);
double focal_length(double fov_deg);
// s_k of the simple integrator: smin, smin+ds, ... while s < smax.
std::vector<double> simple_samples(double smin, double smax, double ds);
// Coarse right endpoints of the hierarchical integrator.
std::vector<double> coarse_samples(double smin, double smax, double DS);

}  // namespace xrp
