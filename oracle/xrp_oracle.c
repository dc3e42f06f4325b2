/*
 * xrp_oracle.c - FP64 CPU oracle for the per-pixel X-ray line integral.
 *
 * TEST INFRASTRUCTURE: this file is what the tests compare the HIP library
 * with, and the timed CPU baseline.  Only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg load it; the HIP library never does.
 *
 * It restates the reference renderer (igrega348/xray_projection_render)
 * function by function, with the same order of double operations.  Built
 * with -ffp-contract=off (see oracle/Makefile) so a*b+c is never fused.
 *
 * Third-party arithmetic: the reference computes cameras with
 * github.com/go-gl/mathgl v1.1.0 (go.mod:6), which is not present here.
 * Its published algorithms are restated below (LookAtV, Mat4.Mul4,
 * Mat4.Det/Inv by cofactor expansion, TransformCoordinate, Mat3.Inv,
 * Vec3.Normalize).  Their last-ulp rounding versus mathgl is parity
 * unpinned; the camera is pinned against the closed-form gluLookAt basis
 * in tests/test_oracle_closed_forms.py.  Go's math.Sin/Cos/Exp are replaced by
 * the C library's (both within 1 ulp of the true value).
 */
#include "xrp_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

/* main.go:46 */
#define CUBE_HALF_DIAGONAL 1.74

/* Planted mistakes for tests/test_oracle_hierarchical.py, which compiles
 * this file with -DORC_MUTANT=n into a temporary library and checks that
 * the integrator pins reject each one.  0 (the default, and the only value
 * oracle/Makefile uses) is the faithful restatement.
 *   1  hierarchical: the right endpoint's rho*ds term dropped (main.go:191)
 *   2  hierarchical: first fine sample at left, before left += ds (main.go:182)
 *   3  hierarchical: fine walk up to and including right (main.go:183)
 *   4  hierarchical: rho*DS added on a transition as well (main.go:192-193)
 *   5  simple: samples while s <= smax (main.go:147)
 *   6  simple: flat field left out (main.go:146) */
#ifndef ORC_MUTANT
#define ORC_MUTANT 0
#endif

enum {
  N_SPHERE = 1, N_CUBE, N_BOX, N_PPED, N_CYL, N_GYROID,
  N_COLL, N_UC, N_TESS, N_VOXEL
};

struct orc_node {
  int type;
  double c[3];        /* sphere/cube/box/gyroid centre, pped origin, cyl p0 */
  double v[3];        /* box sides, cyl p1 */
  double a, b, rho;   /* radius | side | scale,thickness */
  double minv[9];     /* parallelepiped M^-1, column-major (mgl64.Mat3) */
  double pv[9];       /* parallelepiped v0,v1,v2 */
  double bounds[6];   /* unit cell / tessellation bounds */
  int greedy;
  orc_node **kids;
  int nkids, capkids;
  orc_node *uc;       /* tessellation -> unit cell; unit cell -> collection */
  double *vox;
  int nx, ny, nz;
};

static orc_node *node_new(int type) {
  orc_node *n = (orc_node *)calloc(1, sizeof(orc_node));
  if (n) n->type = type;
  return n;
}

orc_node *orc_sphere(const double c[3], double radius, double rho) {
  orc_node *n = node_new(N_SPHERE);
  memcpy(n->c, c, sizeof n->c);
  n->a = radius;
  n->rho = rho;
  return n;
}

/* objects.go:100-117: a cube builds a Box with equal sides at load. */
orc_node *orc_cube(const double c[3], double side, double rho) {
  orc_node *n = node_new(N_CUBE);
  memcpy(n->c, c, sizeof n->c);
  n->v[0] = n->v[1] = n->v[2] = side;
  n->a = side;
  n->rho = rho;
  return n;
}

orc_node *orc_box(const double c[3], const double sides[3], double rho) {
  orc_node *n = node_new(N_BOX);
  memcpy(n->c, c, sizeof n->c);
  memcpy(n->v, sides, sizeof n->v);
  n->rho = rho;
  return n;
}

/* mgl64 Mat3.Det / Mat3.Inv (cofactor expansion, column-major storage);
 * FloatEqual(det, 0) with a zero operand reduces to |det| < 1e-20. */
static void mat3_inv(const double m[9], double out[9]) {
  double det = m[0] * m[4] * m[8] + m[3] * m[7] * m[2] + m[6] * m[1] * m[5] -
               m[6] * m[4] * m[2] - m[3] * m[1] * m[8] - m[0] * m[7] * m[5];
  if (det == 0.0 || fabs(det) < 1e-20) {
    memset(out, 0, 9 * sizeof(double));
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

/* objects.go:209-245 (load-time inverse at 243). */
orc_node *orc_parallelepiped(const double origin[3], const double v0[3],
                             const double v1[3], const double v2[3],
                             double rho) {
  orc_node *n = node_new(N_PPED);
  memcpy(n->c, origin, sizeof n->c);
  double m[9] = {v0[0], v0[1], v0[2], v1[0], v1[1], v1[2], v2[0], v2[1], v2[2]};
  memcpy(n->pv, m, sizeof m);
  mat3_inv(m, n->minv);
  n->rho = rho;
  return n;
}

orc_node *orc_cylinder(const double p0[3], const double p1[3], double radius,
                       double rho) {
  orc_node *n = node_new(N_CYL);
  memcpy(n->c, p0, sizeof n->c);
  memcpy(n->v, p1, sizeof n->v);
  n->a = radius;
  n->rho = rho;
  return n;
}

orc_node *orc_gyroid(const double c[3], double scale, double thickness,
                     double rho) {
  orc_node *n = node_new(N_GYROID);
  memcpy(n->c, c, sizeof n->c);
  n->a = scale;
  n->b = thickness;
  n->rho = rho;
  return n;
}

orc_node *orc_collection(int greedy) {
  orc_node *n = node_new(N_COLL);
  n->greedy = greedy;
  return n;
}

int orc_collection_add(orc_node *coll, orc_node *child) {
  if (!coll || coll->type != N_COLL || !child) return -1;
  if (coll->nkids == coll->capkids) {
    int cap = coll->capkids ? 2 * coll->capkids : 8;
    orc_node **k = (orc_node **)realloc(coll->kids, cap * sizeof(orc_node *));
    if (!k) return -1;
    coll->kids = k;
    coll->capkids = cap;
  }
  coll->kids[coll->nkids++] = child;
  return 0;
}

/* objects.go:479-510: greedy evaluation is forced in a unit cell (487). */
orc_node *orc_unit_cell(orc_node *coll, const double bounds[6]) {
  orc_node *n = node_new(N_UC);
  coll->greedy = 1;
  n->uc = coll;
  memcpy(n->bounds, bounds, sizeof n->bounds);
  return n;
}

orc_node *orc_tessellated(orc_node *unit_cell, const double bounds[6]) {
  orc_node *n = node_new(N_TESS);
  n->uc = unit_cell;
  memcpy(n->bounds, bounds, sizeof n->bounds);
  return n;
}

orc_node *orc_voxel_grid(const double *rho, int nx, int ny, int nz) {
  orc_node *n = node_new(N_VOXEL);
  size_t cnt = (size_t)nx * ny * nz;
  n->vox = (double *)malloc(cnt * sizeof(double));
  memcpy(n->vox, rho, cnt * sizeof(double));
  n->nx = nx;
  n->ny = ny;
  n->nz = nz;
  return n;
}

void orc_node_free(orc_node *n) {
  if (!n) return;
  for (int k = 0; k < n->nkids; ++k) orc_node_free(n->kids[k]);
  free(n->kids);
  orc_node_free(n->uc);
  free(n->vox);
  free(n);
}

/* objects.go:63-72 */
static double sphere_density(const orc_node *s, double x, double y, double z) {
  x = x - s->c[0];
  y = y - s->c[1];
  z = z - s->c[2];
  double r_2 = x * x + y * y + z * z;
  if (r_2 < s->a * s->a) return s->rho;
  return 0.0;
}

/* objects.go:171-179 (cube: 119-121 -> Box) */
static double box_density(const orc_node *b, double x, double y, double z) {
  x = fabs(x - b->c[0]);
  y = fabs(y - b->c[1]);
  z = fabs(z - b->c[2]);
  if (x < 0.5 * b->v[0] && y < 0.5 * b->v[1] && z < 0.5 * b->v[2])
    return b->rho;
  return 0.0;
}

/* objects.go:247-255; mgl64 Vec3.Sub then Mat3.Mul3x1 */
static double pped_density(const orc_node *p, double x, double y, double z) {
  double w0 = x - p->c[0], w1 = y - p->c[1], w2 = z - p->c[2];
  const double *m = p->minv;
  double u = m[0] * w0 + m[3] * w1 + m[6] * w2;
  double v = m[1] * w0 + m[4] * w1 + m[7] * w2;
  double t = m[2] * w0 + m[5] * w1 + m[8] * w2;
  if (u > 0.0 && u < 1.0 && v > 0.0 && v < 1.0 && t > 0.0 && t < 1.0)
    return p->rho;
  return 0.0;
}

/* objects.go:334-350; Dot = a0*b0 + a1*b1 + a2*b2, Len = sqrt(Dot(v,v)) */
static double cylinder_density(const orc_node *cyl, double x, double y,
                               double z) {
  double v0 = cyl->v[0] - cyl->c[0], v1 = cyl->v[1] - cyl->c[1],
         v2 = cyl->v[2] - cyl->c[2];
  double w0 = x - cyl->c[0], w1 = y - cyl->c[1], w2 = z - cyl->c[2];
  double c = (w0 * v0 + w1 * v1 + w2 * v2) / (v0 * v0 + v1 * v1 + v2 * v2);
  if (c < 0.0 || c > 1.0) return 0.0;
  double d0 = w0 - v0 * c, d1 = w1 - v1 * c, d2 = w2 - v2 * c;
  double d = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
  if (d < cyl->a) return cyl->rho;
  return 0.0;
}

/* objects.go:1014-1032 */
static double gyroid_density(const orc_node *g, double x, double y, double z) {
  x = (x - g->c[0]) / g->a;
  y = (y - g->c[1]) / g->a;
  z = (z - g->c[2]) / g->a;
  double gv = sin(x) * cos(y) + sin(y) * cos(z) + sin(z) * cos(x);
  if (fabs(gv) < g->b) return g->rho;
  return 0.0;
}

/* objects.go:789-855 */
static double voxel_density(const orc_node *v, double x, double y, double z) {
  if (x < -1 || x > 1 || y < -1 || y > 1 || z < -1 || z > 1) return 0.0;
  x = (x + 1) / 2;
  y = (y + 1) / 2;
  z = (z + 1) / 2;
  x = x * (double)(v->nx - 1);
  y = y * (double)(v->ny - 1);
  z = z * (double)(v->nz - 1);
  long x0 = (long)floor(x), y0 = (long)floor(y), z0 = (long)floor(z);
  long x1 = x0 + 1, y1 = y0 + 1, z1 = z0 + 1;
  if (x0 < 0) x0 = 0;
  if (y0 < 0) y0 = 0;
  if (z0 < 0) z0 = 0;
  if (x1 >= v->nx) x1 = v->nx - 1;
  if (y1 >= v->ny) y1 = v->ny - 1;
  if (z1 >= v->nz) z1 = v->nz - 1;
  double wx = x - (double)x0, wy = y - (double)y0, wz = z - (double)z0;
  long NX = v->nx, NY = v->ny;
  const double *r = v->vox;
  double v000 = r[z0 * NX * NY + x0 * NY + y0];
  double v001 = r[z1 * NX * NY + x0 * NY + y0];
  double v010 = r[z0 * NX * NY + x0 * NY + y1];
  double v011 = r[z1 * NX * NY + x0 * NY + y1];
  double v100 = r[z0 * NX * NY + x1 * NY + y0];
  double v101 = r[z1 * NX * NY + x1 * NY + y0];
  double v110 = r[z0 * NX * NY + x1 * NY + y1];
  double v111 = r[z1 * NX * NY + x1 * NY + y1];
  double v00 = v000 * (1 - wz) + v001 * wz;
  double v01 = v010 * (1 - wz) + v011 * wz;
  double v10 = v100 * (1 - wz) + v101 * wz;
  double v11 = v110 * (1 - wz) + v111 * wz;
  double v0 = v00 * (1 - wy) + v01 * wy;
  double v1 = v10 * (1 - wy) + v11 * wy;
  return v0 * (1 - wx) + v1 * wx;
}

static double collection_density(const orc_node *oc, double x, double y,
                                 double z);

/* objects.go:458-464 */
static double uc_density(const orc_node *uc, double x, double y, double z) {
  const double *b = uc->bounds;
  if (x < b[0] || x > b[1] || y < b[2] || y > b[3] || z < b[4] || z > b[5])
    return 0.0;
  return collection_density(uc->uc, x, y, z);
}

/* objects.go:568-582 */
static double tess_density(const orc_node *l, double x, double y, double z) {
  const double *b = l->bounds;
  if (x < b[0] || x > b[1] || y < b[2] || y > b[3] || z < b[4] || z > b[5])
    return 0.0;
  const double *u = l->uc->bounds;
  double dx = u[1] - u[0];
  x = x - dx * floor((x - u[0]) / dx);
  double dy = u[3] - u[2];
  y = y - dy * floor((y - u[2]) / dy);
  double dz = u[5] - u[4];
  z = z - dz * floor((z - u[4]) / dz);
  return uc_density(l->uc, x, y, z);
}

double orc_density(const orc_node *n, double x, double y, double z) {
  switch (n->type) {
    case N_SPHERE: return sphere_density(n, x, y, z);
    case N_CUBE:
    case N_BOX: return box_density(n, x, y, z);
    case N_PPED: return pped_density(n, x, y, z);
    case N_CYL: return cylinder_density(n, x, y, z);
    case N_GYROID: return gyroid_density(n, x, y, z);
    case N_COLL: return collection_density(n, x, y, z);
    case N_UC: return uc_density(n, x, y, z);
    case N_TESS: return tess_density(n, x, y, z);
    case N_VOXEL: return voxel_density(n, x, y, z);
  }
  return 0.0;
}

/* objects.go:422-438: sum then clamp to [0,1]; greedy returns the first
 * positive child unclamped. */
static double collection_density(const orc_node *oc, double x, double y,
                                 double z) {
  double density = 0.0;
  for (int k = 0; k < oc->nkids; ++k) {
    double rho = orc_density(oc->kids[k], x, y, z);
    if (oc->greedy && rho > 0.0) return rho;
    density += rho;
  }
  if (density < 0.0)
    density = 0.0;
  else if (density > 1.0)
    density = 1.0;
  return density;
}

static double vlen(const double v[3]) {
  return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
}

/* objects.go:74-76, 123-125, 181-183, 257-259, 352-354, 440-446, 584-586,
 * 857-860, 1034-1036 */
double orc_min_feature_size(const orc_node *n) {
  switch (n->type) {
    case N_SPHERE:
    case N_CYL: return n->a;
    case N_CUBE:
    case N_BOX: return 0.1 * fmin(n->v[0], fmin(n->v[1], n->v[2]));
    case N_PPED:
      return 0.2 * fmin(vlen(n->pv), fmin(vlen(n->pv + 3), vlen(n->pv + 6)));
    case N_GYROID: return n->a * n->b * 0.1;
    case N_COLL: {
      double out = INFINITY;
      for (int k = 0; k < n->nkids; ++k)
        out = fmin(out, orc_min_feature_size(n->kids[k]));
      return out;
    }
    case N_UC: return orc_min_feature_size(n->uc);
    case N_TESS: return orc_min_feature_size(n->uc->uc);
    case N_VOXEL: {
      int m = n->nx > n->ny ? n->nx : n->ny;
      m = m > n->nz ? m : n->nz;
      return 2.0 / (double)m;
    }
  }
  return INFINITY;
}

/* ---- deformations -------------------------------------------------------- */
enum { D_GAUSS = 1, D_AFFINE, D_LINEAR, D_RIGID, D_SIGMOID, D_COMPOSED, D_TANH };

struct orc_deform {
  int type;
  double p[9];
  double q[3], r[3];
  int axis;
  orc_deform **kids;
  int nkids, capkids;
};

static orc_deform *def_new(int type) {
  orc_deform *d = (orc_deform *)calloc(1, sizeof(orc_deform));
  if (d) d->type = type;
  return d;
}

orc_deform *orc_def_gaussian(const double amp[3], const double sigma[3],
                             const double center[3]) {
  orc_deform *d = def_new(D_GAUSS);
  memcpy(d->p, amp, 3 * sizeof(double));
  memcpy(d->q, sigma, 3 * sizeof(double));
  memcpy(d->r, center, 3 * sizeof(double));
  return d;
}
orc_deform *orc_def_affine(const double m[9]) {
  orc_deform *d = def_new(D_AFFINE);
  memcpy(d->p, m, 9 * sizeof(double));
  return d;
}
orc_deform *orc_def_linear(const double strains[6]) {
  orc_deform *d = def_new(D_LINEAR);
  memcpy(d->p, strains, 6 * sizeof(double));
  return d;
}
orc_deform *orc_def_rigid(const double disp[3]) {
  orc_deform *d = def_new(D_RIGID);
  memcpy(d->p, disp, 3 * sizeof(double));
  return d;
}
orc_deform *orc_def_sigmoid(double amplitude, double center, double lengthscale,
                            int axis) {
  orc_deform *d = def_new(D_SIGMOID);
  d->p[0] = amplitude;
  d->p[1] = center;
  d->p[2] = lengthscale;
  d->axis = axis;
  return d;
}
/* Extension without a counterpart in deformations.go (parity unpinned by
 * the reference): p[axis] += amplitude * tanh(rate * p[source]). */
orc_deform *orc_def_tanh(double amplitude, double rate, int axis, int source) {
  orc_deform *d = def_new(D_TANH);
  d->p[0] = amplitude;
  d->p[1] = rate;
  d->axis = axis;
  d->q[0] = (double)source;
  return d;
}
orc_deform *orc_def_composed(void) { return def_new(D_COMPOSED); }
int orc_def_composed_add(orc_deform *c, orc_deform *child) {
  if (!c || c->type != D_COMPOSED || !child) return -1;
  if (c->nkids == c->capkids) {
    int cap = c->capkids ? 2 * c->capkids : 4;
    orc_deform **k = (orc_deform **)realloc(c->kids, cap * sizeof(*k));
    if (!k) return -1;
    c->kids = k;
    c->capkids = cap;
  }
  c->kids[c->nkids++] = child;
  return 0;
}
void orc_deform_free(orc_deform *d) {
  if (!d) return;
  for (int k = 0; k < d->nkids; ++k) orc_deform_free(d->kids[k]);
  free(d->kids);
  free(d);
}

void orc_deform_apply(const orc_deform *d, double p[3]) {
  double x = p[0], y = p[1], z = p[2];
  switch (d->type) {
    case D_GAUSS: { /* deformations.go:29-38 */
      double x0 = x - d->r[0], y0 = y - d->r[1], z0 = z - d->r[2];
      double r2 = x0 * x0 + y0 * y0 + z0 * z0;
      double dx = d->p[0] * exp(-r2 / (2 * d->q[0] * d->q[0]));
      double dy = d->p[1] * exp(-r2 / (2 * d->q[1] * d->q[1]));
      double dz = d->p[2] * exp(-r2 / (2 * d->q[2] * d->q[2]));
      p[0] = x + dx;
      p[1] = y + dy;
      p[2] = z + dz;
      return;
    }
    case D_AFFINE: { /* deformations.go:87-92 */
      const double *m = d->p;
      p[0] = m[0] * x + m[1] * y + m[2] * z;
      p[1] = m[3] * x + m[4] * y + m[5] * z;
      p[2] = m[6] * x + m[7] * y + m[8] * z;
      return;
    }
    case D_LINEAR: { /* deformations.go:136-141 */
      const double *s = d->p;
      p[0] = x + s[0] * x + s[5] * y + s[4] * z;
      p[1] = y + s[5] * x + s[1] * y + s[3] * z;
      p[2] = z + s[4] * x + s[3] * y + s[2] * z;
      return;
    }
    case D_RIGID: /* deformations.go:173-175 */
      p[0] = x + d->p[0];
      p[1] = y + d->p[1];
      p[2] = z + d->p[2];
      return;
    case D_SIGMOID: { /* deformations.go:210-222 */
      double q = p[d->axis];
      p[d->axis] = q + d->p[0] / (1 + exp(-(q - d->p[1]) / d->p[2]));
      return;
    }
    case D_TANH: {
      double q = p[d->axis];
      p[d->axis] = q + d->p[0] * tanh(d->p[1] * p[(int)d->q[0]]);
      return;
    }
    case D_COMPOSED: /* deformations.go:269-274 */
      for (int k = 0; k < d->nkids; ++k) orc_deform_apply(d->kids[k], p);
      return;
  }
}

/* ---- render core ---------------------------------------------------------- */
/* main.go:137-140 (deform: 123-133) */
double orc_scene_density(const orc_scene *s, double x, double y, double z) {
  if (s->def) {
    double p[3] = {x, y, z};
    orc_deform_apply(s->def, p);
    x = p[0];
    y = p[1];
    z = p[2];
  }
  return orc_density(s->obj, x, y, z) * s->density_multiplier;
}

/* mgl64 Vec3.Normalize: l = 1/Len; v*l */
static void normalize(const double v[3], double out[3]) {
  double l = 1.0 / vlen(v);
  out[0] = v[0] * l;
  out[1] = v[1] * l;
  out[2] = v[2] * l;
}

/* main.go:144-154 */
double orc_integrate_simple(const orc_scene *sc, const double origin[3],
                            const double direction[3], double ds, double smin,
                            double smax) {
  double d[3];
  normalize(direction, d);
#if ORC_MUTANT == 6
  double T = 0.0;
#else
  double T = sc->flat_field;
#endif
#if ORC_MUTANT == 5
  for (double s = smin; s <= smax; s += ds) {
#else
  for (double s = smin; s < smax; s += ds) {
#endif
    double x = origin[0] + d[0] * s;
    double y = origin[1] + d[1] * s;
    double z = origin[2] + d[2] * s;
    T += orc_scene_density(sc, x, y, z) * ds;
  }
  return exp(-T);
}

/* main.go:159-199.  The two clipping probes at smin/smax (162-169) only
 * emit warnings and do not change T, so they are not evaluated here. */
double orc_integrate_hierarchical(const orc_scene *sc, const double origin[3],
                                  const double direction[3], double DS,
                                  double smin, double smax) {
  double d[3];
  normalize(direction, d);
  double right = smin + DS;
  double left = smin;
  double ds = DS / 10.0;
  double prev_rho = 0.0;
  double T = sc->flat_field;
  while (right <= smax) {
    double x = origin[0] + d[0] * right;
    double y = origin[1] + d[1] * right;
    double z = origin[2] + d[2] * right;
    double rho = orc_scene_density(sc, x, y, z);
    if ((rho == 0) != (prev_rho == 0)) {
#if ORC_MUTANT != 2
      left += ds;
#endif
#if ORC_MUTANT == 3
      while (left <= right) {
#else
      while (left < right) {
#endif
        double xl = origin[0] + d[0] * left;
        double yl = origin[1] + d[1] * left;
        double zl = origin[2] + d[2] * left;
        T += orc_scene_density(sc, xl, yl, zl) * ds;
        left += ds;
      }
#if ORC_MUTANT != 1
      T += rho * ds;
#endif
#if ORC_MUTANT == 4
      T += rho * DS;
#endif
    } else {
      T += rho * DS;
    }
    prev_rho = rho;
    left = right;
    right += DS;
  }
  return exp(-T);
}

/* mgl64 Mat4 is column-major: element (r, c) at m[c*4+r]. */
static void mat4_mul4(const double a[16], const double b[16], double out[16]) {
  for (int c = 0; c < 4; ++c)
    for (int r = 0; r < 4; ++r)
      out[c * 4 + r] = a[0 * 4 + r] * b[c * 4 + 0] + a[1 * 4 + r] * b[c * 4 + 1] +
                       a[2 * 4 + r] * b[c * 4 + 2] + a[3 * 4 + r] * b[c * 4 + 3];
}

/* Cofactor-expansion determinant and inverse of a 4x4 (the adjugate
 * scaled by 1/det), as mgl64 Mat4.Det / Mat4.Inv compute them. */
static void mat4_inv(const double m[16], double out[16]) {
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
  if (det == 0.0 || fabs(det) < 1e-20) {
    memset(out, 0, 16 * sizeof(double));
    return;
  }
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
  double inv = 1 / det;
  for (int k = 0; k < 16; ++k) out[k] = r[k] * inv;
}

static void cross(const double a[3], const double b[3], double out[3]) {
  out[0] = a[1] * b[2] - a[2] * b[1];
  out[1] = a[2] * b[0] - a[0] * b[2];
  out[2] = a[0] * b[1] - a[1] * b[0];
}

/* main.go:226-239 with mgl64.LookAtV(eye, 0, +z) then Mat4.Inv. */
void orc_camera_from_angles(double az_deg, double polar_deg, double R,
                            double eye[3], double cam[16]) {
  double th = az_deg * M_PI / 180;
  double phi = polar_deg * M_PI / 180;
  eye[0] = R * cos(th) * sin(phi);
  eye[1] = R * sin(th) * sin(phi);
  eye[2] = cos(phi) * R;
  double up[3] = {0, 0, 1}, upn[3];
  double fwd[3] = {0 - eye[0], 0 - eye[1], 0 - eye[2]};
  double f[3], s0[3], s[3], u[3];
  normalize(fwd, f);
  normalize(up, upn);
  cross(f, upn, s0);
  normalize(s0, s);
  cross(s, f, u);
  double M[16] = {s[0], u[0], -f[0], 0, s[1], u[1], -f[1], 0,
                  s[2], u[2], -f[2], 0, 0,    0,    0,     1};
  double Tm[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0,
                   -eye[0], -eye[1], -eye[2], 1};
  double view[16], inv[16];
  mat4_mul4(M, Tm, view);
  mat4_inv(view, inv);
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) cam[r * 4 + c] = inv[c * 4 + r];
}

/* main.go:457-465: camera-space point (i/(res/2)-1, j/(res/2)-1, -f),
 * mgl64.TransformCoordinate (Mul4x1, then scale by 1/w), minus eye. */
void orc_pixel_ray(const double eye[3], const double cam[16], int res,
                   double fov_deg, int i, int j, double origin[3],
                   double direction[3]) {
  double res_f = (double)res;
  double f = 1 / tan(fov_deg / 2 * M_PI / 180);
  double v[4] = {(double)i / (res_f / 2) - 1, (double)j / (res_f / 2) - 1, -f,
                 1};
  double t[4];
  for (int r = 0; r < 4; ++r)
    t[r] = cam[r * 4 + 0] * v[0] + cam[r * 4 + 1] * v[1] +
           cam[r * 4 + 2] * v[2] + cam[r * 4 + 3] * v[3];
  double iw = 1 / t[3];
  for (int k = 0; k < 3; ++k) {
    origin[k] = eye[k];
    direction[k] = t[k] * iw - eye[k];
  }
}

typedef struct {
  const orc_scene *s;
  const double *az, *polar;
  int nviews, res, integrator, row_begin, row_end;
  double ds, R, fov;
  double *out;
  double *eyes, *cams;
  atomic_long next;
} render_job;

static void *render_worker(void *arg) {
  render_job *jb = (render_job *)arg;
  long rows = jb->row_end - jb->row_begin;
  long total = rows * jb->nviews;
  double smin = jb->R - CUBE_HALF_DIAGONAL, smax = jb->R + CUBE_HALF_DIAGONAL;
  for (;;) {
    long item = atomic_fetch_add(&jb->next, 1);
    if (item >= total) break;
    int view = (int)(item / rows);
    int i = jb->row_begin + (int)(item % rows);
    const double *eye = jb->eyes + 3 * view, *cam = jb->cams + 16 * view;
    double *row = jb->out + (size_t)view * jb->res * jb->res + (size_t)i * jb->res;
    for (int j = 0; j < jb->res; ++j) {
      double o[3], d[3];
      orc_pixel_ray(eye, cam, jb->res, jb->fov, i, j, o, d);
      row[j] = jb->integrator == ORC_HIERARCHICAL
                   ? orc_integrate_hierarchical(jb->s, o, d, jb->ds, smin, smax)
                   : orc_integrate_simple(jb->s, o, d, jb->ds, smin, smax);
    }
  }
  return NULL;
}

void orc_render(const orc_scene *s, const double *az_deg,
                const double *polar_deg, int nviews, int res, double ds,
                double R, double fov_deg, int integrator, int nthreads,
                int row_begin, int row_end, double *out) {
  render_job jb;
  jb.s = s;
  jb.az = az_deg;
  jb.polar = polar_deg;
  jb.nviews = nviews;
  jb.res = res;
  jb.integrator = integrator;
  jb.row_begin = row_begin;
  jb.row_end = row_end;
  jb.ds = ds;
  jb.R = R;
  jb.fov = fov_deg;
  jb.out = out;
  jb.eyes = (double *)malloc(sizeof(double) * 3 * (nviews > 0 ? nviews : 1));
  jb.cams = (double *)malloc(sizeof(double) * 16 * (nviews > 0 ? nviews : 1));
  for (int v = 0; v < nviews; ++v)
    orc_camera_from_angles(az_deg[v], polar_deg[v], R, jb.eyes + 3 * v,
                           jb.cams + 16 * v);
  atomic_init(&jb.next, 0);
  if (nthreads < 1) nthreads = 1;
  pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * nthreads);
  for (int t = 0; t < nthreads; ++t)
    pthread_create(&th[t], NULL, render_worker, &jb);
  for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
  free(th);
  free(jb.eyes);
  free(jb.cams);
}

typedef struct {
  const orc_scene *s;
  int res;
  double *out;
  atomic_long next;
} volume_job;

static void *volume_worker(void *arg) {
  volume_job *jb = (volume_job *)arg;
  long res = jb->res;
  for (;;) {
    long k = atomic_fetch_add(&jb->next, 1);
    if (k >= res) break;
    for (long i = 0; i < res; ++i)
      for (long j = 0; j < res; ++j) {
        double x = (double)i / (double)res * 2.0 - 1.0;
        double y = (double)j / (double)res * 2.0 - 1.0;
        double z = (double)k / (double)res * 2.0 - 1.0;
        jb->out[k * res * res + i * res + j] = orc_scene_density(jb->s, x, y, z);
      }
  }
  return NULL;
}

/* main.go:586-615 (and computeVoxel 208-214) */
void orc_export_volume(const orc_scene *s, int res, int nthreads, double *out) {
  volume_job jb;
  jb.s = s;
  jb.res = res;
  jb.out = out;
  atomic_init(&jb.next, 0);
  if (nthreads < 1) nthreads = 1;
  pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * nthreads);
  for (int t = 0; t < nthreads; ++t)
    pthread_create(&th[t], NULL, volume_worker, &jb);
  for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
  free(th);
}
