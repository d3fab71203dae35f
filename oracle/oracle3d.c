/* oracle3d.c — CPU oracle for the 3D nine-field elastic curvilinear solver (FP64).
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load this library; the product (libswl.so) never does.
 *
 * Plain-C restatement of the reference's 3D CPU path, in the same floating-point
 * operation order, so an FP64 run reproduces the reference bit for bit:
 *   validation      src/three_d/CurvilinearSolver.cpp:258-299, Grid3D::validateGeometry :117-145,
 *                   SurfaceGrid3D::validate src/three_d/SurfaceGrid3D.cpp:6-15,
 *                   validateMaterial3D include/Seis3D/MaterialModels3D.hpp:31-40,
 *                   receiverNodes3D include/Seis3D/ReceiverGeometry3D.hpp:8-42
 *   geometry        Grid3D::buildGeometry CurvilinearSolver.cpp:146-190 (FD4 coordinate
 *                   derivatives :47-54, cross-product inverse metrics, surface normals),
 *                   bilinear DEM include/Seis3D/SurfaceGrid3D.hpp:23-31
 *   setup           CpuSimulation3D::initialize CurvilinearSolver.cpp:301-442 (metric norm,
 *                   CFL plane distances, edges, layers, sponge include/Seis3D/AbsorbingBoundary3D.hpp,
 *                   dt plan include/Seis3D/TimeStepPlan3D.hpp, Gaussian source weights,
 *                   isotropic-moment pattern include/SeisCommon/SourceMechanism.hpp:43-56)
 *   stepping        CpuSimulation3D::advance :445-513 with evaluateRhs :70-89, combine :91-101,
 *                   projectFreeSurface :65-68 and Point::rhs / surfaceDerivative /
 *                   surfaceMomentum / energyDensity (src/three_d/PointKernels3D.hpp:30-120)
 *   diagnostics     Impl::updateDiagnostics :214-250
 *
 * Pinning: every function here is compared bit for bit with the reference build in
 * oracle/_ref (tests/test_oracle_3d.py), on analytic terrain with a sponge and a force,
 * on DEM terrain with layers and an isotropic moment, and on a flat grid; the flat grid
 * is also checked against the closed forms (identity metrics, J = 1, h/sqrt(3) CFL length).
 * Host-memory admission (fitsMemory3D) is not restated: it depends on the machine.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define G3 3
#define PI3 3.14159265358979323846

size_t orc3d_params_size(void) { return sizeof(orc_params3d); }

/* FiniteDifference.hpp:9-25 */
static int dir3(int step, int stage, int axis) {
    static const int flags[8][3] = {{0, 0, 0}, {1, 1, 0}, {1, 1, 1}, {0, 0, 1},
                                    {0, 1, 0}, {1, 0, 0}, {1, 0, 1}, {0, 1, 1}};
    return (flags[step % 8][axis] + stage) % 2;
}
static int stencil3(int dir, int depth, int *off, double *c) {
    const int n = depth == 1 ? 2 : depth == 2 ? 3 : 5;
    const double fwd[5] = {-.30874, -.6326, 1.233, -.3334, .04168};
    const double low[3] = {-7.0 / 6.0, 8.0 / 6.0, -1.0 / 6.0};
    for (int l = 0; l < n; ++l) {
        const int pos = n == 5 ? l - 1 : l;
        const double v = n == 5 ? fwd[l] : n == 3 ? low[l] : l == 0 ? -1.0 : 1.0;
        off[l] = dir ? -pos : pos;
        c[l] = dir ? -v : v;
    }
    return n;
}

typedef struct {
    orc_params3d p;
    double *dem; /* owned copy of the DEM */
    int nx, ny, nz;
    size_t count;
    double h, T; /* spacing, resolved terrain transition depth */
    int terrain;
    double *coord, *metric, *jac, *normal; /* 3N, 9N, N, 3N */
    double lambda, mu;                      /* homogeneous moduli */
    double *mlam, *mmu, *mrho;              /* per node when layered, else NULL */
    double *damping, *source, *grad[3];
    double dir[3], amplitude;
    double *state[9], *base[9], *k[4][9], *stage[9];
    double dt;
    int step, steps, threads;
    int nrec;
    int64_t *rec_index;
    double *rec_pos;
    double *rec_vx, *rec_vy, *rec_vz; /* [r * (steps+1) + n] */
    int rec_len;
    /* diagnostics (Diagnostics3D) */
    double energy, peak, source_peak, max_velocity, traction_ratio, display[5];
    int finite;
    double min_jacobian, max_metric_norm, min_length, max_edge, wave_dt, damping_dt, source_dt, max_damping;
    char fault[256];
} O3;

static size_t idx3(const O3 *s, int i, int j, int k) {
    return (size_t)i + (size_t)s->nx * ((size_t)j + (size_t)s->ny * (size_t)k);
}
static void seterr(char *e, int n, const char *m) {
    if (e && n > 0) {
        strncpy(e, m, (size_t)n - 1);
        e[n - 1] = 0;
    }
}
static double clampd(double v, double lo, double hi) { return v < lo ? lo : v > hi ? hi : v; }
/* std::hypot(x, y, z) as the C++17 library computes it: scaled by the largest magnitude */
static double hyp3(double x, double y, double z) {
    x = fabs(x), y = fabs(y), z = fabs(z);
    const double a = x < y ? (y < z ? z : y) : (x < z ? z : x);
    if (a == 0) return 0;
    return a * sqrt((x / a) * (x / a) + (y / a) * (y / a) + (z / a) * (z / a));
}

/* SurfaceGrid3D.hpp:23-31: bilinear with edge-cell linear extension */
static double dem_elevation(const O3 *s, double x, double y) {
    const orc_params3d *p = &s->p;
    const double u = x / p->dem_spacing, v = y / p->dem_spacing;
    const int i = (int)floor(clampd(u, 0.0, (double)(p->dem_nx - 2)));
    const int j = (int)floor(clampd(v, 0.0, (double)(p->dem_ny - 2)));
    const double a = u - i, b = v - j;
#define AT(ix, iy) s->dem[(size_t)(iy) * p->dem_nx + (ix)]
    return (1 - b) * ((1 - a) * AT(i, j) + a * AT(i + 1, j)) + b * ((1 - a) * AT(i, j + 1) + a * AT(i + 1, j + 1));
#undef AT
}
static double dem_displacement(const O3 *s, double x, double y) { return s->p.dem_datum - dem_elevation(s, x, y); }

static int has_dem(const orc_params3d *p) { return p->dem_nx > 0 && p->dem_elevations != NULL; }

/* ReceiverGeometry3D.hpp:8-42 (limits) and :26-42 (node list); returns count or -1 */
static int receiver_capacity(int n, int start, int stride) { return stride > 0 ? 1 + (n - 1 - start) / stride : 0; }
static int receiver_nodes(const orc_params3d *p, int *nodes /* 3 per node, may be NULL */, char *e, int el) {
    const int maxr = 1024;
    int lx = 0, ly = 0, total = 0;
    const int kind = p->receiver_kind;
    int ok = !(p->nx < 8 || p->ny < 8 || p->nz < 8 || p->nx > 4096 || p->ny > 4096 || p->nz > 4096 || kind < 0 ||
               kind > 3 || p->receiver_z < 0 || p->receiver_z >= p->nz);
    int x = 0, y = 0;
    if (ok) {
        if (kind == 0) {
            lx = p->nx - 4, ly = 1, total = p->nx - 4 < maxr ? p->nx - 4 : maxr;
        } else {
            x = p->receiver_start_x == -1 ? p->nx / 2 : p->receiver_start_x;
            y = p->receiver_start_y == -1 ? p->ny / 2 : p->receiver_start_y;
            if (x < 0 || x >= p->nx || y < 0 || y >= p->ny) ok = 0;
            else {
                lx = kind == 2 ? 1 : receiver_capacity(p->nx, x, p->receiver_stride_x);
                ly = kind == 1 ? 1 : receiver_capacity(p->ny, y, p->receiver_stride_y);
                const long long t = (long long)lx * ly;
                total = (int)(t < maxr ? t : maxr);
            }
        }
    }
    if (!ok || !total) {
        seterr(e, el, "Invalid receiver grid, mode, start, active stride or depth index.");
        return -1;
    }
    if (kind == 0) {
        if (p->nx - 4 > maxr) {
            seterr(e, el, "Auto X line exceeds 1024 traces; select a custom line with a larger receiver stride.");
            return -1;
        }
        int n = 0;
        for (int i = 2; i < p->nx - 2; ++i, ++n)
            if (nodes) nodes[3 * n] = i, nodes[3 * n + 1] = p->ny / 2, nodes[3 * n + 2] = p->receiver_z;
        return n;
    }
    const int cx = kind == 2 ? 1 : p->receiver_count_x, cy = kind == 1 ? 1 : p->receiver_count_y;
    if (cx < 1 || cy < 1 || cx > lx || cy > ly || (long long)cx * cy > maxr) {
        seterr(e, el,
               "Receiver counts exceed the start/stride capacity or the 1024-trace total limit. Review capacity or fit counts.");
        return -1;
    }
    int n = 0;
    for (int j = 0; j < cy; ++j)
        for (int i = 0; i < cx; ++i, ++n)
            if (nodes) {
                nodes[3 * n] = x + i * (kind == 2 ? 0 : p->receiver_stride_x);
                nodes[3 * n + 1] = y + j * (kind == 1 ? 0 : p->receiver_stride_y);
                nodes[3 * n + 2] = p->receiver_z;
            }
    return n;
}

/* CurvilinearSolver.cpp:117-145 (the grid copy's transition depth is resolved in place) */
static int validate_geometry(O3 *s, char *e, int el) {
    const orc_params3d *p = &s->p;
    const int nx = s->nx, ny = s->ny, nz = s->nz;
    const double dx = p->spacing;
    if (nx < 8 + 2 * G3 || ny < 8 + 2 * G3 || nz < 8 + 2 * G3 || nx > 4096 + 2 * G3 || ny > 4096 + 2 * G3 ||
        nz > 4096 + 2 * G3 || !isfinite(dx) || dx <= 0.0 || !isfinite(p->topography_amplitude) ||
        !isfinite(p->topography_wavelength) || p->topography_wavelength <= 0.0 || !isfinite(s->T) || s->T < 0.0) {
        seterr(e, el, "Grid dimensions, ghost width, spacing or topography parameters are invalid.");
        return 0;
    }
    if ((size_t)nx * ny * nz > (size_t)306 * 306 * 306) {
        seterr(e, el, "Grid exceeds the 28,652,616 padded-node total budget; redistribute X/Y/Z nodes.");
        return 0;
    }
    const double bottom = (nz - 2 * G3 - 1) * dx;
    if (has_dem(p)) {
        /* SurfaceGrid3D::validate (SurfaceGrid3D.cpp:6-15) */
        const double width = (p->dem_nx - 1) * p->dem_spacing, height = (p->dem_ny - 1) * p->dem_spacing;
        if (p->dem_nx < 2 || p->dem_ny < 2 || p->dem_nx > 4096 || p->dem_ny > 4096 ||
            (size_t)p->dem_nx * p->dem_ny > 4000000 || !isfinite(p->dem_spacing) || p->dem_spacing <= 0 ||
            !isfinite(p->dem_origin_x) || !isfinite(p->dem_origin_y) || !isfinite(p->dem_datum) || !isfinite(width) ||
            !isfinite(height)) {
            seterr(e, el,
                   "DEM array requires 2..4096 rows/columns, at most 4 million samples, positive metre spacing and finite origin/datum.");
            return 0;
        }
        for (size_t q = 0; q < (size_t)p->dem_nx * p->dem_ny; ++q)
            if (!isfinite(s->dem[q]) || fabs(s->dem[q]) > 1e7) {
                seterr(e, el, "DEM contains missing, nonfinite or unreasonable elevations.");
                return 0;
            }
        if ((nx - 2 * G3 - 1) * dx > width + 1e-8 * p->dem_spacing || (ny - 2 * G3 - 1) * dx > height + 1e-8 * p->dem_spacing) {
            seterr(e, el, "Imported DEM does not cover the staged X/Y model span; reduce dimensions/spacing or export a larger region.");
            return 0;
        }
    }
    double relief = fabs(p->topography_amplitude);
    if (has_dem(p)) {
        relief = 0;
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const double d = fabs(dem_displacement(s, (i - G3) * dx, (j - G3) * dx));
                relief = relief > d ? relief : d;
            }
    }
    if (s->T == 0.0) {
        double m = 0.35 * bottom;
        if (6.0 * dx > m) m = 6.0 * dx;
        if (4.0 * relief + 2.0 * dx > m) m = 4.0 * relief + 2.0 * dx;
        s->T = bottom < m ? bottom : m;
    }
    if (s->T > bottom || (s->terrain && (s->T < 4.0 * dx || s->T <= 2.6 * relief))) {
        seterr(e, el, "Terrain transition must fit the model depth and exceed 2.6 times relief amplitude (at least four cells).");
        return 0;
    }
    return 1;
}

static void cross3(const double *a, const double *b, double *o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}
static double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* CurvilinearSolver.cpp:47-54 */
static double coord_derivative(const O3 *s, int c, int i, int j, int k, int axis) {
    const int di = axis == 0, dj = axis == 1, dk = axis == 2;
    const double h = s->h;
#define AT(x, y, z) s->coord[3 * idx3(s, (x), (y), (z)) + c]
    return (AT(i - 2 * di, j - 2 * dj, k - 2 * dk) - 8.0 * AT(i - di, j - dj, k - dk) + 8.0 * AT(i + di, j + dj, k + dk) -
            AT(i + 2 * di, j + 2 * dj, k + 2 * dk)) /
           (12.0 * h);
#undef AT
}

/* Grid3D::buildGeometry, CurvilinearSolver.cpp:146-190 */
static int build_geometry(O3 *s, char *e, int el) {
    if (!validate_geometry(s, e, el)) return 0;
    const orc_params3d *p = &s->p;
    const double dx = s->h;
    for (int k = 0; k < s->nz; ++k)
        for (int j = 0; j < s->ny; ++j)
            for (int i = 0; i < s->nx; ++i) {
                const double x = (i - G3) * dx, y = (j - G3) * dx, depth = (k - G3) * dx;
                double *c = s->coord + 3 * idx3(s, i, j, k);
                if (!s->terrain) {
                    c[0] = x, c[1] = y, c[2] = depth;
                    continue;
                }
                const double px = 2.0 * PI3 * x / p->topography_wavelength;
                const double py = 2.0 * PI3 * y / p->topography_wavelength;
                const double surface =
                    has_dem(p) ? dem_displacement(s, x, y) : p->topography_amplitude * cos(px) * cos(py);
                const double t = clampd(depth / s->T, 0.0, 1.0);
                const double fade =
                    depth >= s->T ? 0.0 : 1.0 - pow(t, 5) * (126.0 + t * (-420.0 + t * (540.0 + t * (-315.0 + 70.0 * t))));
                c[0] = x, c[1] = y, c[2] = depth + surface * fade;
            }
    for (int k = G3; k < s->nz - G3; ++k)
        for (int j = G3; j < s->ny - G3; ++j)
            for (int i = G3; i < s->nx - G3; ++i) {
                const size_t q = idx3(s, i, j, k);
                double *m = s->metric + 9 * q;
                if (!s->terrain) {
                    for (int a = 0; a < 9; ++a) m[a] = (a % 4 == 0) ? 1 : 0;
                    s->jac[q] = 1;
                    if (k == G3) s->normal[3 * q] = 0, s->normal[3 * q + 1] = 0, s->normal[3 * q + 2] = 1;
                    continue;
                }
                double rx[3], ry[3], rz[3];
                for (int c = 0; c < 3; ++c) {
                    rx[c] = coord_derivative(s, c, i, j, k, 0);
                    ry[c] = coord_derivative(s, c, i, j, k, 1);
                    rz[c] = coord_derivative(s, c, i, j, k, 2);
                }
                double ec[3], zc[3], xc[3];
                cross3(ry, rz, ec);
                cross3(rz, rx, zc);
                cross3(rx, ry, xc);
                const double jac = dot3(rx, ec);
                if (!(jac > 1.0e-10) || !isfinite(jac)) {
                    seterr(e, el, "Curvilinear mapping has a non-positive or non-finite Jacobian.");
                    return 0;
                }
                s->jac[q] = jac;
                for (int c = 0; c < 3; ++c) {
                    m[c] = ec[c] / jac;
                    m[3 + c] = zc[c] / jac;
                    m[6 + c] = xc[c] / jac;
                }
                if (k == G3) {
                    const double length = sqrt(dot3(xc, xc));
                    if (!(length > 0.0) || !isfinite(length)) {
                        seterr(e, el, "Could not compute topographic surface normal.");
                        return 0;
                    }
                    for (int c = 0; c < 3; ++c) s->normal[3 * q + c] = xc[c] / length;
                }
            }
    return 1;
}

/* MaterialModels3D.hpp:19-31 */
static double material_depth(const O3 *s, size_t q) {
    const double *xyz = s->coord + 3 * q;
    if (s->p.layer_depth_mode == 0) return ((int)(q / ((size_t)s->nx * s->ny)) - G3) * s->h;
    if (s->p.layer_depth_mode == 2) return xyz[2];
    const size_t plane = (size_t)s->nx * s->ny, surf = (size_t)G3 * plane + q % plane;
    return xyz[2] - s->coord[3 * surf + 2];
}
static int layer_at(const orc_params3d *p, double depth) {
    double bottom = 0;
    for (int i = 0; i + 1 < p->n_layers; ++i) {
        bottom += p->layer_thickness[i];
        if (depth < bottom) return i;
    }
    return p->n_layers - 1;
}

/* AbsorbingBoundary3D.hpp:5-21 */
static double absorbing_rate(const orc_params3d *p, int i, int j, int k, double vp) {
    if (!p->absorbing) return 0;
    const int nodes[3] = {p->nx, p->ny, p->nz};
    int w[3];
    for (int a = 0; a < 3; ++a) w[a] = p->absorbing_cells[a] < nodes[a] / 3 ? p->absorbing_cells[a] : nodes[a] / 3;
    const int dist[3] = {i < p->nx - 1 - i ? i : p->nx - 1 - i, j < p->ny - 1 - j ? j : p->ny - 1 - j, p->nz - 1 - k};
    double rate = 0;
    for (int a = 0; a < 3; ++a)
        if (w[a] > 0) {
            const double f = clampd(1.0 - (double)dist[a] / w[a], 0.0, 1.0);
            rate += p->absorbing_strength * vp / (w[a] * p->spacing) * f * f;
        }
    return rate;
}

/* CurvilinearSolver.cpp:258-299 */
static int validate_config(const orc_params3d *p, char *e, int el) {
    if (!isfinite(p->cfl) || p->cfl < .02 || p->cfl > .8) {
        seterr(e, el, "CFL coefficient must be finite and in 0.02..0.8.");
        return 0;
    }
    if (p->max_steps < 0 || p->max_steps > 20000) {
        seterr(e, el, "Maximum steps must be 0 (duration) or 1..20,000.");
        return 0;
    }
    if (!isfinite(p->absorbing_strength) || p->absorbing_strength < 0.1 || p->absorbing_strength > 10.0) {
        seterr(e, el, "Sponge strength must be in 0.1..10.");
        return 0;
    }
    for (int a = 0; a < 3; ++a)
        if (p->absorbing_cells[a] < 1 || p->absorbing_cells[a] > 64) {
            seterr(e, el, "Sponge widths must be 1..64 cells.");
            return 0;
        }
    const double sc[9] = {p->spacing, p->vp, p->vs, p->density, p->dominant_frequency, p->duration,
                          p->topography_amplitude, p->topography_wavelength, p->terrain_transition_depth};
    for (int a = 0; a < 9; ++a)
        if (!isfinite(sc[a])) {
            seterr(e, el, "Simulation parameters must be finite.");
            return 0;
        }
    if (p->nx < 8 || p->ny < 8 || p->nz < 8 || p->nx > 4096 || p->ny > 4096 || p->nz > 4096 || p->spacing <= 0.0 ||
        p->vp <= 0.0 || p->vs <= 0.0 || p->density <= 0.0 || p->vp * p->vp <= 4.0 * p->vs * p->vs / 3.0 ||
        p->dominant_frequency <= 0.0 || p->dominant_frequency > 1.0e6 || p->duration <= 0.0 || p->duration > 10.0 ||
        p->topography_wavelength <= 0.0 || p->terrain_transition_depth < 0.0 || p->threads < 0 || p->threads > 64 ||
        (!has_dem(p) && p->topography_amplitude != 0.0 && p->topography_wavelength < 4.0 * p->spacing)) {
        seterr(e, el, "Require 8..4096 nodes per axis within the total node/memory budget, positive finite parameters and resolved topography.");
        return 0;
    }
    if ((size_t)(p->nx + 2 * G3) * (p->ny + 2 * G3) * (p->nz + 2 * G3) > (size_t)306 * 306 * 306) {
        seterr(e, el, "Grid exceeds the 28,652,616 padded-node total budget; redistribute X/Y/Z nodes.");
        return 0;
    }
    if (p->source_kind != 0 && p->source_kind != 1) {
        seterr(e, el, "Unknown source mechanism.");
        return 0;
    }
    if (p->source_kind == 0) {
        const double norm = hyp3(p->source_direction[0], p->source_direction[1], p->source_direction[2]);
        if (!isfinite(norm) || norm <= 0) {
            seterr(e, el, "Force direction must be finite and nonzero.");
            return 0;
        }
    }
    const double amp = p->source_kind == 1 ? p->source_moment : p->source_force;
    if (!isfinite(p->source_force) || !isfinite(p->source_moment) || amp <= 0) {
        seterr(e, el, "Source amplitudes must be finite and the active force (N) or moment (N*m) must be positive.");
        return 0;
    }
    const int sx = p->source_x < 0 ? p->nx / 2 : p->source_x, sy = p->source_y < 0 ? p->ny / 2 : p->source_y,
              sz = p->source_z;
    if (p->source_kind == 1 && (p->nx < 11 || p->ny < 11 || p->nz < 11 || sx < 5 || sx > p->nx - 6 || sy < 5 ||
                                sy > p->ny - 6 || sz < 5 || sz > p->nz - 6)) {
        seterr(e, el, "Isotropic moment needs 11+ nodes per axis and source center five nodes inside each face.");
        return 0;
    }
    if (p->source_x < -1 || p->source_y < -1 || sx < 2 || sx > p->nx - 3 || sy < 2 || sy > p->ny - 3 || sz < 2 ||
        sz > p->nz - 3) {
        seterr(e, el, "Source must stay two nodes inside every physical face; X/Y may be -1 for center.");
        return 0;
    }
    if (p->layered) {
        if (p->layer_depth_mode < 0 || p->layer_depth_mode > 2) {
            seterr(e, el, "Invalid layer depth coordinate mode.");
            return 0;
        }
        if (p->n_layers < 2 || p->n_layers > 8) {
            seterr(e, el, "Layered material requires 2..8 layers.");
            return 0;
        }
        for (int i = 0; i < p->n_layers; ++i) {
            const double t = p->layer_thickness[i], vp = p->layer_vp[i], vs = p->layer_vs[i], d = p->layer_density[i];
            if (!isfinite(t) || !isfinite(vp) || !isfinite(vs) || !isfinite(d) || vp <= 0 || vs <= 0 || d <= 0 ||
                vp * vp <= 4 * vs * vs / 3 || !isfinite(d * vp * vp) || d * vs * vs <= 0 ||
                (i + 1 == p->n_layers ? t != 0 : t <= 0)) {
                seterr(e, el, "Invalid layer: positive finite Vp/Vs/density and bulk modulus required; last thickness must be zero (auto).");
                return 0;
            }
        }
    }
    return receiver_nodes(p, NULL, e, el) < 0 ? 0 : 1;
}

static void free3(O3 *s) {
    if (!s) return;
    free(s->dem);
    free(s->coord), free(s->metric), free(s->jac), free(s->normal);
    free(s->mlam), free(s->mmu), free(s->mrho), free(s->damping), free(s->source);
    for (int a = 0; a < 3; ++a) free(s->grad[a]);
    for (int f = 0; f < 9; ++f) {
        free(s->state[f]), free(s->base[f]), free(s->stage[f]);
        for (int r = 0; r < 4; ++r) free(s->k[r][f]);
    }
    free(s->rec_index), free(s->rec_pos), free(s->rec_vx), free(s->rec_vy), free(s->rec_vz);
    free(s);
}

static double *zalloc(size_t n) { return (double *)calloc(n ? n : 1, sizeof(double)); }

/* ---- Point kernels (PointKernels3D.hpp) ---- */
static double lam_at(const O3 *s, size_t q) { return s->mlam ? s->mlam[q] : s->lambda; }
static double mu_at(const O3 *s, size_t q) { return s->mmu ? s->mmu[q] : s->mu; }
static double rho_at(const O3 *s, size_t q) { return s->mrho ? s->mrho[q] : s->p.density; }

static double derivative3(const O3 *s, const double *v, size_t q, int axis, int dir, int depth) {
    const ptrdiff_t st = axis == 0 ? 1 : axis == 1 ? s->nx : (ptrdiff_t)s->nx * s->ny;
    int off[5];
    double c[5];
    const int n = stencil3(dir, depth, off, c);
    double r = 0;
    for (int l = 0; l < n; ++l) r += c[l] * v[(ptrdiff_t)q + off[l] * st];
    return r / s->h;
}
static void stress9(double *const *w, size_t q, double *t) {
    t[0] = w[3][q], t[4] = w[4][q], t[8] = w[5][q];
    t[1] = t[3] = w[6][q], t[2] = t[6] = w[7][q], t[5] = t[7] = w[8][q];
}
/* project :36-51 */
static void project3(O3 *s, double *const *w) {
    for (int j = G3; j < s->ny - G3; ++j)
        for (int i = G3; i < s->nx - G3; ++i) {
            const size_t q = idx3(s, i, j, G3);
            const double *n = s->normal + 3 * q;
            double sg[9], out[9], t[3] = {0, 0, 0};
            stress9(w, q, sg);
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) t[a] += sg[3 * a + b] * n[b];
            const double nn = t[0] * n[0] + t[1] * n[1] + t[2] * n[2];
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) out[3 * a + b] = sg[3 * a + b] - n[a] * t[b] - t[a] * n[b] + nn * n[a] * n[b];
            w[3][q] = out[0], w[4][q] = out[4], w[5][q] = out[8];
            w[6][q] = 0.5 * (out[1] + out[3]), w[7][q] = 0.5 * (out[2] + out[6]), w[8][q] = 0.5 * (out[5] + out[7]);
        }
}
/* surfaceDerivative :52-66 */
static void surface_derivative(const O3 *s, double *const *w, int i, int j, double *res, int dx_, int dy_) {
    const size_t q = idx3(s, i, j, G3);
    const double *m = s->metric + 9 * q;
    const double lambda = lam_at(s, q), mu = mu_at(s, q);
    const double c[3] = {m[6], m[7], m[8]};
    const double c2 = c[0] * c[0] + c[1] * c[1] + c[2] * c[2];
    const double length = sqrt(c2), n[3] = {c[0] / length, c[1] / length, c[2] / length};
    double g[9] = {0}, tr[3] = {0, 0, 0};
    for (int v = 0; v < 3; ++v) {
        const double dx = derivative3(s, w[v], q, 0, dx_, -1), dy = derivative3(s, w[v], q, 1, dy_, -1);
        for (int a = 0; a < 3; ++a) g[3 * v + a] = m[a] * dx + m[3 + a] * dy;
    }
    const double trace = g[0] + g[4] + g[8];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) tr[a] += ((a == b ? lambda * trace : 0) + mu * (g[3 * a + b] + g[3 * b + a])) * c[b];
    const double normal = tr[0] * n[0] + tr[1] * n[1] + tr[2] * n[2];
    for (int a = 0; a < 3; ++a) res[a] = -(tr[a] - n[a] * normal) / (mu * c2) - n[a] * normal / ((lambda + 2 * mu) * c2);
}
/* tractionFlux :67-72 */
static double traction_flux(const O3 *s, double *const *w, size_t q, int axis, int comp) {
    static const int fl[3][3] = {{3, 6, 7}, {6, 4, 8}, {7, 8, 5}};
    const double *m = s->metric + 9 * q + 3 * axis;
    double v = 0;
    for (int a = 0; a < 3; ++a) v += m[a] * w[fl[comp][a]][q];
    return s->jac[q] * v;
}
/* surfaceMomentum :73-91 */
static double surface_momentum(const O3 *s, double *const *w, size_t q, int comp, const int *dirs, int depth) {
    double sum = 0;
    for (int axis = 0; axis < 3; ++axis) {
        int off[5];
        double c[5];
        const int n = stencil3(dirs[axis], -1, off, c);
        const ptrdiff_t st = axis == 0 ? 1 : axis == 1 ? s->nx : (ptrdiff_t)s->nx * s->ny;
        for (int l = 0; l < n; ++l) {
            int o = off[l];
            double sign = 1;
            if (axis == 2) {
                const int z = depth + o;
                if (z == 0) continue;
                if (z < 0) o = -z - depth, sign = -1;
            }
            sum += sign * c[l] * traction_flux(s, w, (size_t)((ptrdiff_t)q + o * st), axis, comp);
        }
    }
    return sum / (s->h * s->jac[q] * rho_at(s, q));
}
/* Point::rhs :92-112 */
static void rhs_point(const O3 *s, double *const *w, double *const *out, size_t q, double source, int step, int stage) {
    const double lambda = lam_at(s, q), mu = mu_at(s, q), density = rho_at(s, q);
    double grad[9][3];
    const double *m = s->metric + 9 * q;
    const int depth = (int)(q / ((size_t)s->nx * s->ny)) - G3;
    const int dirs[3] = {dir3(step, stage, 0), dir3(step, stage, 1), dir3(step, stage, 2)};
    double dzs[3] = {0, 0, 0};
    if (depth == 0) surface_derivative(s, w, (int)(q % s->nx), (int)((q / s->nx) % s->ny), dzs, dirs[0], dirs[1]);
    for (int f = 0; f < 9; ++f) {
        double d[3];
        for (int a = 0; a < 3; ++a)
            d[a] = f < 3 && a == 2 && depth == 0 ? dzs[f]
                                                 : derivative3(s, w[f], q, a, dirs[a], f < 3 && a == 2 ? depth : -1);
        for (int a = 0; a < 3; ++a) grad[f][a] = m[a] * d[0] + m[3 + a] * d[1] + m[6 + a] * d[2];
    }
    out[0][q] = (grad[3][0] + grad[6][1] + grad[7][2]) / density;
    out[1][q] = (grad[6][0] + grad[4][1] + grad[8][2]) / density;
    out[2][q] = (grad[7][0] + grad[8][1] + grad[5][2]) / density;
    if (depth <= (dirs[2] ? 3 : 1))
        for (int a = 0; a < 3; ++a) out[a][q] = surface_momentum(s, w, q, a, dirs, depth);
    for (int a = 0; a < 3; ++a)
        out[a][q] += source * (s->grad[a] ? s->grad[a][q] : s->source[q] * s->dir[a]) * (s->p.density / density);
    const double div = grad[0][0] + grad[1][1] + grad[2][2];
    out[3][q] = lambda * div + 2 * mu * grad[0][0];
    out[4][q] = lambda * div + 2 * mu * grad[1][1];
    out[5][q] = lambda * div + 2 * mu * grad[2][2];
    out[6][q] = mu * (grad[0][1] + grad[1][0]);
    out[7][q] = mu * (grad[0][2] + grad[2][0]);
    out[8][q] = mu * (grad[1][2] + grad[2][1]);
    for (int f = 0; f < 9; ++f) out[f][q] -= s->damping[q] * w[f][q];
}
/* energyDensity :113-120 */
static double energy_density(const O3 *s, double *const *w, size_t q) {
    const double lambda = lam_at(s, q), mu = mu_at(s, q), density = rho_at(s, q);
    const double mean = (w[3][q] + w[4][q] + w[5][q]) / 3;
    double v2 = 0, dev = 0;
    for (int c = 0; c < 3; ++c) {
        v2 += w[c][q] * w[c][q];
        const double t = w[c + 3][q] - mean;
        dev += t * t;
    }
    for (int c = 6; c < 9; ++c) dev += 2 * w[c][q] * w[c][q];
    const double bulk = lambda + 2 * mu / 3;
    return (0.5 * density * v2 + dev / (4 * mu) + mean * mean / (2 * bulk)) * s->jac[q] * s->h * s->h * s->h;
}

/* evaluateRhs :70-89 */
static void eval_rhs(O3 *s, double *const *w, double *const *out, double time, int step, int stage) {
    for (int f = 0; f < 9; ++f) memset(out[f], 0, s->count * sizeof(double));
    const orc_params3d *p = &s->p;
    const double scale = s->amplitude / (p->density * s->h * s->h * s->h);
    const double tau = PI3 * p->dominant_frequency * (time - 1.5 / p->dominant_frequency);
    const double source = scale * (1.0 - 2.0 * tau * tau) * exp(-tau * tau);
    const int nzp = s->nz;
#pragma omp parallel for schedule(static) num_threads(s->threads) if (s->threads > 1)
    for (int k = G3; k < nzp - G3; ++k)
        for (int j = G3; j < s->ny - G3; ++j)
            for (int i = G3; i < s->nx - G3; ++i) rhs_point(s, w, out, idx3(s, i, j, k), source, step, stage);
}
/* combine :91-101: base + wa a (+ wb b + wc c + wd d, or + 0.0 each when absent) */
static void combine3(O3 *s, double *const *out, double *const *base, double *const *a, double wa, double *const *b,
                     double wb, double *const *c, double wc, double *const *d, double wd) {
#pragma omp parallel for schedule(static) num_threads(s->threads) if (s->threads > 1)
    for (int f = 0; f < 9; ++f)
        for (size_t i = 0; i < s->count; ++i)
            out[f][i] = base[f][i] + wa * a[f][i] + (b ? wb * b[f][i] : 0.0) + (c ? wc * c[f][i] : 0.0) +
                        (d ? wd * d[f][i] : 0.0);
}

/* Impl::updateDiagnostics :214-250 */
static void update_diagnostics(O3 *s) {
    s->max_velocity = 0, s->traction_ratio = 0, s->finite = 1, s->energy = 0;
    for (int c = 0; c < 5; ++c) s->display[c] = 0;
    double *const *w = s->state;
    double *const *k4 = s->k[3];
    double rate_norm = 0, rate_traction = 0;
    for (int k = G3; k < s->nz - G3; ++k)
        for (int j = G3; j < s->ny - G3; ++j)
            for (int i = G3; i < s->nx - G3; ++i) {
                const size_t q = idx3(s, i, j, k);
                for (int f = 0; f < 9; ++f) s->finite = s->finite && isfinite(w[f][q]);
                const double v2 = w[0][q] * w[0][q] + w[1][q] * w[1][q] + w[2][q] * w[2][q];
                const double sv = sqrt(v2);
                s->max_velocity = s->max_velocity > sv ? s->max_velocity : sv;
                const double disp[5] = {fabs(w[2][q]), fabs(w[0][q]), fabs(w[1][q]), sqrt(v2),
                                        fabs((w[3][q] + w[4][q] + w[5][q]) / 3.0)};
                for (int c = 0; c < 5; ++c) s->display[c] = s->display[c] > disp[c] ? s->display[c] : disp[c];
                s->energy += energy_density(s, w, q);
                s->finite = s->finite && isfinite(v2);
                if (k == G3) {
                    const double *n = s->normal + 3 * q;
                    const double t[3] = {k4[3][q] * n[0] + k4[6][q] * n[1] + k4[7][q] * n[2],
                                         k4[6][q] * n[0] + k4[4][q] * n[1] + k4[8][q] * n[2],
                                         k4[7][q] * n[0] + k4[8][q] * n[1] + k4[5][q] * n[2]};
                    const double tt = sqrt(dot3(t, t));
                    rate_traction = rate_traction > tt ? rate_traction : tt;
                    double norm2 = 0.0;
                    for (int f = 3; f < 9; ++f) norm2 += (f >= 6 ? 2.0 : 1.0) * k4[f][q] * k4[f][q];
                    const double nn = sqrt(norm2);
                    rate_norm = rate_norm > nn ? rate_norm : nn;
                }
            }
    s->traction_ratio = rate_traction / (rate_norm > 1.0e-30 ? rate_norm : 1.0e-30);
    s->finite = s->finite && isfinite(s->energy) && isfinite(s->traction_ratio);
    s->peak = s->peak > s->energy ? s->peak : s->energy;
}

void *orc3d_create(const orc_params3d *pin, char *e, int el) {
    seterr(e, el, "");
    if (!validate_config(pin, e, el)) return NULL;
    O3 *s = (O3 *)calloc(1, sizeof(O3));
    s->p = *pin;
    const orc_params3d *p = &s->p;
    if (has_dem(pin)) {
        const size_t n = (size_t)pin->dem_nx * pin->dem_ny;
        s->dem = (double *)malloc(n * sizeof(double));
        memcpy(s->dem, pin->dem_elevations, n * sizeof(double));
    }
    s->p.dem_elevations = s->dem; /* the session's own copy */
    s->nx = p->nx + 2 * G3, s->ny = p->ny + 2 * G3, s->nz = p->nz + 2 * G3;
    s->count = (size_t)s->nx * s->ny * s->nz;
    s->h = p->spacing;
    s->T = p->terrain_transition_depth;
    s->terrain = has_dem(pin) || p->topography_amplitude != 0;
#ifdef _OPENMP
    s->threads = p->threads == 0 ? (omp_get_num_procs() < 8 ? omp_get_num_procs() : 8) : p->threads;
    if (s->threads < 1) s->threads = 1;
#else
    s->threads = 1;
#endif
    const size_t N = s->count;
    s->coord = zalloc(3 * N), s->metric = zalloc(9 * N), s->jac = zalloc(N), s->normal = zalloc(3 * N);
    if (!build_geometry(s, e, el)) goto fail;
    {
        s->min_jacobian = DBL_MAX;
        double maxnorm = 0.0;
        for (int k = G3; k < s->nz - G3; ++k)
            for (int j = G3; j < s->ny - G3; ++j)
                for (int i = G3; i < s->nx - G3; ++i) {
                    const size_t q = idx3(s, i, j, k);
                    s->min_jacobian = s->min_jacobian < s->jac[q] ? s->min_jacobian : s->jac[q];
                    double n2 = 0.0;
                    for (int a = 0; a < 9; ++a) n2 += s->metric[9 * q + a] * s->metric[9 * q + a];
                    const double nr = sqrt(n2);
                    maxnorm = maxnorm > nr ? maxnorm : nr;
                }
        s->max_metric_norm = maxnorm;
        /* CFL plane distances :338-351 */
        double minlen = !s->terrain ? p->spacing / sqrt(3.0) : DBL_MAX;
        if (s->terrain)
            for (int k = G3; k < s->nz - G3; ++k)
                for (int j = G3; j < s->ny - G3; ++j)
                    for (int i = G3; i < s->nx - G3; ++i) {
                        const double *ctr = s->coord + 3 * idx3(s, i, j, k);
                        for (int sk = -1; sk <= 1; sk += 2)
                            for (int sj = -1; sj <= 1; sj += 2)
                                for (int si = -1; si <= 1; si += 2) {
                                    const double *a = s->coord + 3 * idx3(s, i + si, j, k);
                                    const double *b = s->coord + 3 * idx3(s, i, j + sj, k);
                                    const double *c = s->coord + 3 * idx3(s, i, j, k + sk);
                                    double ab[3], ac[3], ap[3], nrm[3];
                                    for (int x = 0; x < 3; ++x) ab[x] = b[x] - a[x], ac[x] = c[x] - a[x], ap[x] = ctr[x] - a[x];
                                    cross3(ab, ac, nrm);
                                    const double length = sqrt(dot3(nrm, nrm));
                                    const double dist = fabs(dot3(ap, nrm)) / length;
                                    if (!isfinite(dist) || dist <= 0) {
                                        seterr(e, el, "Degenerate adjacent cell plane in CFL geometry.");
                                        goto fail;
                                    }
                                    minlen = minlen < dist ? minlen : dist;
                                }
                    }
        s->min_length = minlen;
        double maxedge = !s->terrain ? p->spacing : 0;
        if (s->terrain)
            for (int k = G3; k < s->nz - G3; ++k)
                for (int j = G3; j < s->ny - G3; ++j)
                    for (int i = G3; i < s->nx - G3; ++i) {
                        const double *a = s->coord + 3 * idx3(s, i, j, k);
                        for (int ax = 0; ax < 3; ++ax) {
                            const int ni = i + (ax == 0), nj = j + (ax == 1), nk = k + (ax == 2);
                            if (ni >= s->nx - G3 || nj >= s->ny - G3 || nk >= s->nz - G3) continue;
                            const double *b = s->coord + 3 * idx3(s, ni, nj, nk);
                            double d[3];
                            for (int c = 0; c < 3; ++c) d[c] = b[c] - a[c];
                            const double ed = sqrt(dot3(d, d));
                            maxedge = maxedge > ed ? maxedge : ed;
                        }
                    }
        s->max_edge = maxedge;
        s->lambda = p->density * (p->vp * p->vp - 2.0 * p->vs * p->vs);
        s->mu = p->density * p->vs * p->vs;
        if (!isfinite(s->lambda) || !isfinite(s->mu) || s->mu <= 0.0 || !isfinite(maxnorm)) {
            seterr(e, el, "Derived elastic coefficients or grid metrics are out of range.");
            goto fail;
        }
        double maxvp = p->vp;
        if (p->layered) {
            maxvp = 0;
            for (int l = 0; l < p->n_layers; ++l) maxvp = maxvp > p->layer_vp[l] ? maxvp : p->layer_vp[l];
            s->mlam = zalloc(N), s->mmu = zalloc(N), s->mrho = zalloc(N);
            for (size_t q = 0; q < N; ++q) {
                const int id = layer_at(p, material_depth(s, q));
                const double d = p->layer_density[id], vp = p->layer_vp[id], vs = p->layer_vs[id];
                s->mlam[q] = d * (vp * vp - 2 * vs * vs);
                s->mmu[q] = d * vs * vs;
                s->mrho[q] = d;
            }
        }
        s->damping = zalloc(N);
        double maxdamp = 0;
        for (int k = G3; k < s->nz - G3; ++k)
            for (int j = G3; j < s->ny - G3; ++j)
                for (int i = G3; i < s->nx - G3; ++i) {
                    const double r = absorbing_rate(p, i - G3, j - G3, k - G3, maxvp);
                    s->damping[idx3(s, i, j, k)] = r;
                    maxdamp = maxdamp > r ? maxdamp : r;
                }
        s->max_damping = maxdamp;
        /* TimeStepPlan3D.hpp:8-20 */
        const double in[8] = {p->spacing, p->cfl, p->duration, p->dominant_frequency, maxvp, maxnorm, minlen, maxdamp};
        for (int a = 0; a < 8; ++a)
            if (!isfinite(in[a])) {
                seterr(e, el, "Time-step inputs must be finite.");
                goto fail;
            }
        if (maxvp <= 0 || maxnorm <= 0 || minlen <= 0 || maxdamp < 0) {
            seterr(e, el, "Invalid time-step parameters.");
            goto fail;
        }
        const double geo = p->spacing / maxnorm < minlen ? p->spacing / maxnorm : minlen;
        s->wave_dt = p->cfl * geo / maxvp;
        s->damping_dt = maxdamp > 0 ? .5 / maxdamp : DBL_MAX;
        s->source_dt = 1 / (50 * p->dominant_frequency);
        double safe = s->wave_dt < s->damping_dt ? s->wave_dt : s->damping_dt;
        if (s->source_dt < safe) safe = s->source_dt;
        double dsteps = ceil(p->duration / safe);
        if (dsteps < 1.0) dsteps = 1.0;
        if (!isfinite(dsteps) || (p->max_steps == 0 && dsteps > 20000)) {
            seterr(e, el, "Time-step limit: at most 20,000 steps. Set a maximum step cap.");
            goto fail;
        }
        s->dt = p->duration / dsteps;
        s->steps = (int)(p->max_steps > 0 ? (dsteps < (double)p->max_steps ? dsteps : (double)p->max_steps) : dsteps);
        if (!(s->dt > 0) || !isfinite(s->dt)) {
            seterr(e, el, "Invalid time-step parameters.");
            goto fail;
        }
        for (int f = 0; f < 9; ++f) {
            s->state[f] = zalloc(N), s->base[f] = zalloc(N), s->stage[f] = zalloc(N);
            for (int r = 0; r < 4; ++r) s->k[r][f] = zalloc(N);
        }
        /* Gaussian source weights :395-403 and moment pattern :404-412 */
        const int sx = p->source_x < 0 ? p->nx / 2 : p->source_x, sy = p->source_y < 0 ? p->ny / 2 : p->source_y,
                  sz = p->source_z;
        s->source = zalloc(N);
        double sum = 0.0;
        for (int k = -2; k <= 2; ++k)
            for (int j = -2; j <= 2; ++j)
                for (int i = -2; i <= 2; ++i) sum += exp(-0.5 * (i * i + j * j + k * k));
        for (int k = -2; k <= 2; ++k)
            for (int j = -2; j <= 2; ++j)
                for (int i = -2; i <= 2; ++i) {
                    const size_t q = idx3(s, G3 + sx + i, G3 + sy + j, G3 + sz + k);
                    s->source[q] = exp(-0.5 * (i * i + j * j + k * k)) / (sum * s->jac[q]);
                }
        if (p->source_kind == 1) {
            for (int a = 0; a < 3; ++a) s->grad[a] = zalloc(N);
            const ptrdiff_t st[3] = {1, s->nx, (ptrdiff_t)s->nx * s->ny};
            for (int k = sz - 4; k <= sz + 4; ++k)
                for (int j = sy - 4; j <= sy + 4; ++j)
                    for (int i = sx - 4; i <= sx + 4; ++i) {
                        const size_t q = idx3(s, i + G3, j + G3, k + G3);
                        double r[3] = {0, 0, 0};
                        for (int c = 0; c < 3; ++c)
                            for (int a = 0; a < 3; ++a) {
                                const ptrdiff_t pq = (ptrdiff_t)q, d = st[a];
#define FLUX(x) (s->jac[(x)] * s->metric[9 * (size_t)(x) + 3 * a + c] * s->source[(x)])
                                r[c] -= (FLUX(pq - 2 * d) - 8 * FLUX(pq - d) + 8 * FLUX(pq + d) - FLUX(pq + 2 * d)) /
                                        (12 * p->spacing * s->jac[q]);
#undef FLUX
                            }
                        for (int a = 0; a < 3; ++a) s->grad[a][q] = r[a];
                    }
            s->dir[0] = s->dir[1] = s->dir[2] = 0;
        } else {
            const double *d = p->source_direction;
            const double norm = hyp3(d[0], d[1], d[2]);
            for (int a = 0; a < 3; ++a) s->dir[a] = d[a] / norm;
        }
        s->amplitude = p->source_kind == 1 ? p->source_moment : p->source_force;
        /* receivers :413-424 */
        const int nr = receiver_nodes(p, NULL, e, el);
        int *nodes = (int *)malloc(sizeof(int) * 3 * (size_t)(nr > 0 ? nr : 1));
        receiver_nodes(p, nodes, e, el);
        s->nrec = nr;
        s->rec_len = s->steps + 1;
        s->rec_index = (int64_t *)calloc((size_t)(nr > 0 ? nr : 1), sizeof(int64_t));
        s->rec_pos = zalloc(3 * (size_t)nr);
        s->rec_vx = zalloc((size_t)nr * s->rec_len), s->rec_vy = zalloc((size_t)nr * s->rec_len),
        s->rec_vz = zalloc((size_t)nr * s->rec_len);
        for (int r = 0; r < nr; ++r) {
            const size_t q = idx3(s, nodes[3 * r] + G3, nodes[3 * r + 1] + G3, nodes[3 * r + 2] + G3);
            s->rec_index[r] = (int64_t)q;
            for (int c = 0; c < 3; ++c) s->rec_pos[3 * r + c] = s->coord[3 * q + c];
        }
        free(nodes);
    }
    update_diagnostics(s);
    seterr(e, el, "");
    return s;
fail:
    free3(s);
    return NULL;
}

void orc3d_destroy(void *h) { free3((O3 *)h); }

int orc3d_advance(void *h, int count, char *e, int el) {
    O3 *s = (O3 *)h;
    seterr(e, el, "");
    if (!s) {
        seterr(e, el, "Initialize a model before stepping.");
        return 0;
    }
    if (count < 0 || count > 20000) {
        seterr(e, el, "Invalid step batch.");
        return 0;
    }
    if (s->fault[0]) {
        seterr(e, el, s->fault);
        return 0;
    }
    if (!s->finite) {
        seterr(e, el, "Fields are non-finite. Rebuild the session.");
        return 0;
    }
    const orc_params3d *p = &s->p;
    const int end = s->step + (count < s->steps - s->step ? count : s->steps - s->step);
    for (; s->step < end; ++s->step) {
        const double e0 = s->energy, pk0 = s->peak, mv0 = s->max_velocity, tr0 = s->traction_ratio;
        double disp0[5];
        memcpy(disp0, s->display, sizeof disp0);
        const int fin0 = s->finite;
        for (int f = 0; f < 9; ++f) memcpy(s->base[f], s->state[f], s->count * sizeof(double));
        const double t = s->step * s->dt, dt = s->dt;
        eval_rhs(s, s->state, s->k[0], t, s->step, 0);
        combine3(s, s->stage, s->base, s->k[0], 0.5 * dt, NULL, 0, NULL, 0, NULL, 0);
        project3(s, s->stage);
        eval_rhs(s, s->stage, s->k[1], t + 0.5 * dt, s->step, 1);
        combine3(s, s->stage, s->base, s->k[1], 0.5 * dt, NULL, 0, NULL, 0, NULL, 0);
        project3(s, s->stage);
        eval_rhs(s, s->stage, s->k[2], t + 0.5 * dt, s->step, 2);
        combine3(s, s->stage, s->base, s->k[2], dt, NULL, 0, NULL, 0, NULL, 0);
        project3(s, s->stage);
        eval_rhs(s, s->stage, s->k[3], t + dt, s->step, 3);
        combine3(s, s->state, s->base, s->k[0], dt / 6.0, s->k[1], dt / 3.0, s->k[2], dt / 3.0, s->k[3], dt / 6.0);
        project3(s, s->state);
        update_diagnostics(s);
        if ((s->step + 1) * s->dt <= 3.0 / p->dominant_frequency)
            s->source_peak = s->source_peak < s->energy ? s->energy : s->source_peak;
        const int growth =
            (s->step + 1) * s->dt > 3.0 / p->dominant_frequency && s->source_peak > 0 && s->energy > 1e4 * s->source_peak;
        if (!s->finite || growth) {
            for (int f = 0; f < 9; ++f) memcpy(s->state[f], s->base[f], s->count * sizeof(double));
            s->energy = e0, s->peak = pk0, s->max_velocity = mv0, s->traction_ratio = tr0, s->finite = fin0;
            memcpy(s->display, disp0, sizeof disp0);
            snprintf(s->fault, sizeof s->fault, "%s",
                     growth ? "Energy grew above 10,000 times the source-window peak. Last accepted step retained; rebuild with smaller CFL or gentler terrain."
                            : "Non-finite RK4 step rejected. Last accepted fields retained; rebuild with smaller CFL.");
            seterr(e, el, s->fault);
            return 0;
        }
        for (int r = 0; r < s->nrec; ++r) {
            const size_t q = (size_t)s->rec_index[r], n = (size_t)s->step + 1;
            s->rec_vz[(size_t)r * s->rec_len + n] = s->state[2][q];
            s->rec_vx[(size_t)r * s->rec_len + n] = s->state[0][q];
            s->rec_vy[(size_t)r * s->rec_len + n] = s->state[1][q];
        }
    }
    if (!s->finite) {
        seterr(e, el, "Simulation produced non-finite fields; stop and reduce the time/grid distortion.");
        return 0;
    }
    return 1;
}

int orc3d_step(void *h) { return ((O3 *)h)->step; }
/* Test hook: replaces the source-window peak energy that the late-growth fault rule compares against
   (CurvilinearSolver.cpp:477-488), so a test can drive that branch without an unstable model. */
void orc3d_set_source_peak(void *h, double peak) { ((O3 *)h)->source_peak = peak; }
int orc3d_total(void *h) { return ((O3 *)h)->steps; }
double orc3d_dt(void *h) { return ((O3 *)h)->dt; }
void orc3d_layout(void *h, int64_t *o) {
    const O3 *s = (const O3 *)h;
    o[0] = s->nx, o[1] = s->ny, o[2] = s->nz, o[3] = s->nrec, o[4] = s->step + 1; /* samples recorded so far */
}
void orc3d_diagnostics(void *h, double *o) {
    const O3 *s = (const O3 *)h;
    const double v[18] = {s->energy,        s->peak,          s->source_peak,     s->max_velocity, s->traction_ratio,
                          (double)s->finite, s->min_jacobian, s->max_metric_norm, s->min_length,   s->max_edge,
                          s->wave_dt,       s->damping_dt,    s->source_dt,       s->display[0],   s->display[1],
                          s->display[2],    s->display[3],    s->display[4]};
    memcpy(o, v, sizeof v);
}
void orc3d_receivers(void *h, double *vx, double *vy, double *vz, double *position, int64_t *index) {
    const O3 *s = (const O3 *)h;
    const size_t n = (size_t)s->step + 1; /* samples so far, t = 0 included; packed per receiver */
    for (int r = 0; r < s->nrec; ++r) {
        if (vx) memcpy(vx + r * n, s->rec_vx + (size_t)r * s->rec_len, n * sizeof(double));
        if (vy) memcpy(vy + r * n, s->rec_vy + (size_t)r * s->rec_len, n * sizeof(double));
        if (vz) memcpy(vz + r * n, s->rec_vz + (size_t)r * s->rec_len, n * sizeof(double));
    }
    if (position) memcpy(position, s->rec_pos, 3 * (size_t)s->nrec * sizeof(double));
    if (index) memcpy(index, s->rec_index, (size_t)s->nrec * sizeof(int64_t));
}
void orc3d_field(void *h, int f, double *out) {
    const O3 *s = (const O3 *)h;
    memcpy(out, s->state[f], s->count * sizeof(double));
}
int orc3d_model(void *h, int which, double *out) {
    const O3 *s = (const O3 *)h;
    const size_t N = s->count;
    switch (which) {
    case 0: memcpy(out, s->metric, 9 * N * sizeof(double)); return 1;
    case 1: memcpy(out, s->jac, N * sizeof(double)); return 1;
    case 2: memcpy(out, s->normal, 3 * N * sizeof(double)); return 1;
    case 3: memcpy(out, s->coord, 3 * N * sizeof(double)); return 1;
    case 4: for (size_t q = 0; q < N; ++q) out[q] = lam_at(s, q); return 1;
    case 5: for (size_t q = 0; q < N; ++q) out[q] = mu_at(s, q); return 1;
    case 6: for (size_t q = 0; q < N; ++q) out[q] = rho_at(s, q); return 1;
    case 7: memcpy(out, s->damping, N * sizeof(double)); return 1;
    case 8: memcpy(out, s->source, N * sizeof(double)); return 1;
    case 9:
    case 10:
    case 11:
        if (s->grad[which - 9]) memcpy(out, s->grad[which - 9], N * sizeof(double));
        else memset(out, 0, N * sizeof(double));
        return 1;
    default: return 0;
    }
}
