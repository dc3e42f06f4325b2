/*
 * xrp_oracle.h - FP64 CPU restatement of the reference renderer's
 * per-pixel line-integral path.  TEST INFRASTRUCTURE ONLY: the oracle is
 * what the HIP library is compared with, and the timed CPU baseline; no product
 * code links, loads or calls it.
 *
 * Every function restates a reference function (file:line given beside
 * it) with the same order of IEEE double operations, compiled with
 * -ffp-contract=off so no multiply-add is fused.
 */
#ifndef XRP_ORACLE_H
#define XRP_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- object tree (objects/objects.go) ---------------------------------- */
typedef struct orc_node orc_node;

orc_node *orc_sphere(const double c[3], double radius, double rho);
orc_node *orc_cube(const double c[3], double side, double rho);
orc_node *orc_box(const double c[3], const double sides[3], double rho);
orc_node *orc_parallelepiped(const double origin[3], const double v0[3],
                             const double v1[3], const double v2[3], double rho);
orc_node *orc_cylinder(const double p0[3], const double p1[3], double radius,
                       double rho);
orc_node *orc_gyroid(const double c[3], double scale, double thickness,
                     double rho);
orc_node *orc_collection(int greedy);
int orc_collection_add(orc_node *coll, orc_node *child);
/* bounds = {xmin,xmax,ymin,ymax,zmin,zmax}; takes ownership of coll and
 * forces its greedy flag (objects.go:487). */
orc_node *orc_unit_cell(orc_node *coll, const double bounds[6]);
orc_node *orc_tessellated(orc_node *unit_cell, const double bounds[6]);
/* rho is copied; layout [z][x][y] (objects.go:837-845). */
orc_node *orc_voxel_grid(const double *rho, int nx, int ny, int nz);
void orc_node_free(orc_node *n);
double orc_density(const orc_node *n, double x, double y, double z);
double orc_min_feature_size(const orc_node *n);

/* ---- deformations (deformations/deformations.go) ----------------------- */
typedef struct orc_deform orc_deform;

orc_deform *orc_def_gaussian(const double amp[3], const double sigma[3],
                             const double center[3]);
orc_deform *orc_def_affine(const double m[9]); /* row-major M[i][j] */
orc_deform *orc_def_linear(const double strains[6]);
orc_deform *orc_def_rigid(const double disp[3]);
orc_deform *orc_def_sigmoid(double amplitude, double center, double lengthscale,
                            int axis);
/* p[axis] += amplitude * tanh(rate * p[source]); no reference counterpart. */
orc_deform *orc_def_tanh(double amplitude, double rate, int axis, int source);
orc_deform *orc_def_composed(void);
int orc_def_composed_add(orc_deform *c, orc_deform *child);
void orc_deform_free(orc_deform *d);
void orc_deform_apply(const orc_deform *d, double p[3]);

/* ---- render core (main.go) --------------------------------------------- */
typedef struct {
  const orc_node *obj;     /* lat[0] */
  const orc_deform *def;   /* df[0] or NULL */
  double density_multiplier;
  double flat_field;
} orc_scene;

enum { ORC_SIMPLE = 0, ORC_HIERARCHICAL = 1 };

double orc_scene_density(const orc_scene *s, double x, double y, double z);
double orc_integrate_simple(const orc_scene *s, const double origin[3],
                            const double direction[3], double ds, double smin,
                            double smax);
double orc_integrate_hierarchical(const orc_scene *s, const double origin[3],
                                  const double direction[3], double DS,
                                  double smin, double smax);
/* cam is row-major: cam[r*4+c] = camera.At(r, c). */
void orc_camera_from_angles(double az_deg, double polar_deg, double R,
                            double eye[3], double cam[16]);
/* Ray of pixel (i, j) for one view, as main.go:457-465 builds it. */
void orc_pixel_ray(const double eye[3], const double cam[16], int res,
                   double fov_deg, int i, int j, double origin[3],
                   double direction[3]);
/* Renders rows [row_begin, row_end) of every view; out is [view][i][j]
 * with stride res*res per view (rows outside the range untouched).
 * nthreads workers pull rows dynamically. */
void orc_render(const orc_scene *s, const double *az_deg,
                const double *polar_deg, int nviews, int res, double ds,
                double R, double fov_deg, int integrator, int nthreads,
                int row_begin, int row_end, double *out);
/* Volume export density grid, main.go:549-615: out[k*res*res+i*res+j]. */
void orc_export_volume(const orc_scene *s, int res, int nthreads, double *out);

#ifdef __cplusplus
}
#endif
#endif
