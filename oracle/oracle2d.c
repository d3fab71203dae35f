/*
 * oracle2d.c — CPU oracle for the 2D P-SV elastic / Qp-Qs GSLS curvilinear solver.
 *
 * TEST INFRASTRUCTURE ONLY: a plain-C restatement of the reference's FP64 CPU path
 * (src/two_d/CurvilinearSolver2D.cpp initialize :151-540, advance :541-604,
 * sample :120-147; src/two_d/PointKernels2D.hpp; src/two_d/CfsPml2D.hpp;
 * src/common/Gsls.cpp; include/SeisCommon/SourceMechanism.hpp;
 * src/two_d/SurfaceProfile.cpp).  It is the parity checker for the HIP engine and
 * the CPU baseline; nothing in the product links or calls it.
 *
 * Parity pins (tests/test_oracle_*.py): the reference build in oracle/_ref (same
 * inputs, every public output), docs/data/precision-2d-2026-10-28/{storage.csv,
 * dry_loess_traces.csv}, the 43370-step known answer, scipy's PCHIP, flat-metric
 * identities, zero-net-force of the moment pattern and energy/traction invariants.
 */
#include "oracle.h"
#include <complex.h>
#include <float.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PI 3.14159265358979323846
#define MAXF 18

typedef struct orc2d orc2d;
struct orc2d {
    orc_params2d p, logical;
    int prec, nx, nz, offset_x, wx, wz, fields, mech, explosive, threads;
    size_t count;
    int64_t cpml;
    double h, flatten, max_metric, min_jac, max_edge, amplitude;
    /* FP64 construction arrays (kept for inspection) */
    double *coord, *nrm, *a, *b, *jac, *vp, *vs, *rho, *lam, *mu, *src, *damp, *grad[2], *dP[3], *dS[3], *cx, *cz;
    double rates[3], dir[3];
    double dt;
    int total, step;
    double energy, peak, max_velocity, traction;
    int finite;
    double wave_dt, damping_dt, source_dt, cpml_dt, relaxation_dt, ppw, max_q_error, weight_sum, rms_radius;
    int nrec;
    int32_t *rec_x, *rec_z;
    double *rec_pos;
    int64_t *probe;
    double *rec_vx, *rec_vz, *prev;
    size_t rec_len, rec_cap;
    char fault[256];
    void *ctx;   /* SFX(ctx_t) of the active precision */
    void *mesh;  /* SFX(mesh_t) */
    void *state; /* SFX(state_t) */
};

static int orc_direction(int step, int stage, int axis) {
    /* FiniteDifference.hpp:9-13 */
    static const int flags[8][3] = {{0, 0, 0}, {1, 1, 0}, {1, 1, 1}, {0, 0, 1},
                                    {0, 1, 0}, {1, 0, 0}, {1, 0, 1}, {0, 1, 1}};
    return (flags[step % 8][axis] + stage) % 2;
}

/* PointKernels2D.hpp:50-56: compact CFS-PML slot, 4 memories per active direction;
   X strips first ([k][2w]), then the bottom Z strip ([k][nx]); corners own both. */
static int64_t orc_pml_slot(int nx, int nz, int wx, int wz, int64_t cpml, int64_t q, int axis) {
    if (!cpml) return -1;
    const int i = (int)(q % nx) - 3, k = (int)(q / nx) - 3, nxc = nx - 6, nzc = nz - 6;
    if (axis == 0) {
        const int j = i < wx ? i : i >= nxc - wx ? wx + i - (nxc - wx) : -1;
        return j < 0 ? -1 : 4 * ((int64_t)k * 2 * wx + j);
    }
    return k < nzc - wz ? -1 : 4 * ((int64_t)(2 * wx) * nzc + (int64_t)(k - (nzc - wz)) * nxc + i);
}
int64_t orc2d_pml_slot(int nx_pad, int nz_pad, int width_x, int width_z, int64_t q, int axis) {
    const int64_t cpml = 4 * (2 * (int64_t)width_x * (nz_pad - 6) + (int64_t)width_z * (nx_pad - 6));
    return orc_pml_slot(nx_pad, nz_pad, width_x, width_z, cpml, q, axis);
}

/* CurvilinearSolver2D.cpp:31-34 */
static double orc_pulse(double t, double f) {
    const double a = PI * f * (t - 1.5 / f);
    return (1 - 2 * a * a) * exp(-a * a);
}

/* ---- FP64 instantiation ---- */
#define R double
#define SFX(x) x##_f64
#define ORC_SQRT sqrt
#define ORC_RMAX 1.7976931348623157e308
#include "oracle2d_step.inc"
#undef R
#undef SFX
#undef ORC_SQRT
#undef ORC_RMAX
/* ---- FP32 instantiation ---- */
#define R float
#define SFX(x) x##_f32
#define ORC_SQRT sqrtf
#define ORC_RMAX 3.4028234663852886e38
#include "oracle2d_step.inc"
#undef R
#undef SFX
#undef ORC_SQRT
#undef ORC_RMAX

static void set_error(char *error, int len, const char *fmt, ...) {
    if (!error || len <= 0) return;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(error, (size_t)len, fmt, ap);
    va_end(ap);
}
static double dmin(double a, double b) { return a < b ? a : b; }
static double dmax(double a, double b) { return a > b ? a : b; }
static double dclamp(double x, double lo, double hi) { return x < lo ? lo : x > hi ? hi : x; }
static int imin(int a, int b) { return a < b ? a : b; }
static int imax(int a, int b) { return a > b ? a : b; }
static int iclamp(int x, int lo, int hi) { return x < lo ? lo : x > hi ? hi : x; }

/* P0811R3 lerp (the C++20 std::lerp contract used by the reference's sample()). */
static double orc_lerp(double a, double b, double t) {
    if ((a <= 0 && b >= 0) || (a >= 0 && b <= 0)) return t * b + (1 - t) * a;
    if (t == 1) return b;
    const double x = a + t * (b - a);
    return (t > 1) == (b > a) ? (b < x ? x : b) : (b > x ? x : b);
}
/* Scaled three-argument hypot (std::hypot(x,y,z)), used to normalise the force direction. */
static double orc_hypot3(double x, double y, double z) {
    x = fabs(x);
    y = fabs(y);
    z = fabs(z);
    const double a = x < y ? (y < z ? z : y) : (x < z ? z : x);
    if (a == 0) return 0;
    return a * sqrt((x / a) * (x / a) + (y / a) * (y / a) + (z / a) * (z / a));
}

/* ---------------- PCHIP terrain (SurfaceProfile.cpp:8-65) ---------------- */
typedef struct {
    int n;
    const double *xz;
    double *slope;
} pchip_t;
static double pchip_end(double h0, double h1, double d0, double d1) {
    const double m = ((2 * h0 + h1) * d0 - h0 * d1) / (h0 + h1);
    if (m * d0 <= 0) return 0.0;
    if (d0 * d1 < 0 && fabs(m) > 3 * fabs(d0)) return 3 * d0;
    return m;
}
static int pchip_init(pchip_t *c, int n, const double *xz, char *err, int len) {
    c->n = n;
    c->xz = xz;
    c->slope = NULL;
    if (n < 2 || n > 1000000) return set_error(err, len, "Surface profile requires 2..1000000 samples."), 0;
    double *h = malloc(sizeof(double) * (size_t)(n - 1)), *d = malloc(sizeof(double) * (size_t)(n - 1));
    c->slope = malloc(sizeof(double) * (size_t)n);
    for (int i = 0; i < n; ++i) {
        if (!isfinite(xz[2 * i]) || !isfinite(xz[2 * i + 1]) || (i && xz[2 * i] <= xz[2 * i - 2])) {
            free(h), free(d);
            return set_error(err, len, "Surface samples must be finite with strictly increasing x."), 0;
        }
        if (i) {
            h[i - 1] = xz[2 * i] - xz[2 * i - 2];
            d[i - 1] = (xz[2 * i + 1] - xz[2 * i - 1]) / h[i - 1];
        }
    }
    c->slope[0] = d[0];
    c->slope[n - 1] = d[n - 2];
    if (n > 2) {
        c->slope[0] = pchip_end(h[0], h[1], d[0], d[1]);
        const int m = n - 1;
        c->slope[n - 1] = pchip_end(h[m - 1], h[m - 2], d[m - 1], d[m - 2]);
        for (int i = 1; i + 1 < n; ++i) {
            if (d[i - 1] * d[i] <= 0) c->slope[i] = 0;
            else {
                const double a = 2 * h[i] + h[i - 1], b = h[i] + 2 * h[i - 1];
                c->slope[i] = (a + b) / (a / d[i - 1] + b / d[i]);
            }
        }
    }
    free(h);
    free(d);
    return 1;
}
static void pchip_sample(const pchip_t *c, double x, double *value, double *slope) {
    const double *p = c->xz;
    const int n = c->n;
    if (x <= p[0]) {
        *value = p[1] + (x - p[0]) * c->slope[0];
        *slope = c->slope[0];
        return;
    }
    if (x >= p[2 * (n - 1)]) {
        *value = p[2 * (n - 1) + 1] + (x - p[2 * (n - 1)]) * c->slope[n - 1];
        *slope = c->slope[n - 1];
        return;
    }
    int lo = 0, hi = n; /* first index with x < p[idx] */
    while (lo < hi) {
        const int mid = (lo + hi) / 2;
        if (x < p[2 * mid]) hi = mid;
        else lo = mid + 1;
    }
    const int i = lo - 1;
    const double h = p[2 * (i + 1)] - p[2 * i], t = (x - p[2 * i]) / h;
    const double a = p[2 * i + 1], b = p[2 * (i + 1) + 1], m = c->slope[i], nn = c->slope[i + 1];
    *value = (2 * t * t * t - 3 * t * t + 1) * a + (t * t * t - 2 * t * t + t) * h * m + (-2 * t * t * t + 3 * t * t) * b +
             (t * t * t - t * t) * h * nn;
    *slope = (6 * t * t - 6 * t) * (a - b) / h + (3 * t * t - 4 * t + 1) * m + (3 * t * t - 2 * t) * nn;
}

/* ---------------- GSLS fit (Gsls.cpp:11-78) ---------------- */
typedef struct {
    int mech;
    double rates[3], wp[3], ws[3]; /* relaxation rates and normalised strengths for Qp / Qs */
} gsls_fit_t;
typedef struct {
    double pR, sR, pU, sU, dP[3], dS[3], max_q_error;
} gsls_mat_t;

static int gsls_strengths(double q, const gsls_fit_t *m, double fmin, double fmax, double out[3]) {
    enum { S = 81 };
    double design[S][3];
    memset(design, 0, sizeof design);
    const double target = 1 / q;
    for (int i = 0; i < S; ++i) {
        const double w = 2 * PI * fmin * pow(fmax / fmin, (double)i / (S - 1));
        for (int l = 0; l < m->mech; ++l) {
            const double x = w / m->rates[l];
            design[i][l] = (x - target * x * x) / (1 + x * x);
        }
    }
    double best[3] = {0, 0, 0}, best_cost = INFINITY;
    for (int mask = 1; mask < (1 << m->mech); ++mask) {
        int ids[3] = {0, 0, 0}, n = 0;
        for (int l = 0; l < m->mech; ++l)
            if (mask & (1 << l)) ids[n++] = l;
        double a[3][4];
        memset(a, 0, sizeof a);
        for (int i = 0; i < n; ++i)
            for (int row = 0; row < S; ++row) {
                a[i][n] += design[row][ids[i]] * target;
                for (int j = 0; j < n; ++j) a[i][j] += design[row][ids[i]] * design[row][ids[j]];
            }
        int valid = 1;
        for (int k = 0; k < n; ++k) {
            int pivot = k;
            for (int i = k + 1; i < n; ++i)
                if (fabs(a[i][k]) > fabs(a[pivot][k])) pivot = i;
            if (fabs(a[pivot][k]) < 1e-20) {
                valid = 0;
                break;
            }
            for (int j = 0; j <= n; ++j) {
                const double t = a[k][j];
                a[k][j] = a[pivot][j];
                a[pivot][j] = t;
            }
            const double d = a[k][k];
            for (int j = k; j <= n; ++j) a[k][j] /= d;
            for (int i = 0; i < n; ++i)
                if (i != k) {
                    const double f = a[i][k];
                    for (int j = k; j <= n; ++j) a[i][j] -= f * a[k][j];
                }
        }
        double cand[3] = {0, 0, 0};
        for (int i = 0; i < n && valid; ++i) {
            valid = isfinite(a[i][n]) && a[i][n] >= 0;
            cand[ids[i]] = a[i][n];
        }
        if (!valid) continue;
        double cost = 0;
        for (int row = 0; row < S; ++row) {
            double r = -target;
            for (int l = 0; l < m->mech; ++l) r += design[row][l] * cand[l];
            cost += r * r;
        }
        if (cost < best_cost) {
            best_cost = cost;
            memcpy(best, cand, sizeof best);
        }
    }
    if (!isfinite(best_cost)) return 0;
    memcpy(out, best, sizeof best);
    return 1;
}
static double gsls_scale(const gsls_fit_t *m, const double *w8, double fref, double rho, double v) {
    const double complex w = CMPLX(0, 2 * PI * fref);
    double complex shape = 1.0;
    for (int l = 0; l < m->mech; ++l) shape += w8[l] * w / (w + m->rates[l]);
    const double complex one = 1.0;
    const double inv = creal(one / csqrt(shape));
    return rho * v * v * inv * inv;
}
static double gsls_quality(const gsls_mat_t *g, const gsls_fit_t *m, int shear, double f) {
    const double complex s = CMPLX(0, 2 * PI * f);
    double complex c = shear ? g->sR : g->pR;
    for (int l = 0; l < m->mech; ++l) c += (shear ? g->dS[l] : g->dP[l]) * s / (s + m->rates[l]);
    return creal(c) / fabs(cimag(c));
}
static int gsls_fit(const orc_params2d *p, double qp, double qs, gsls_fit_t *m, char *err, int len) {
    m->mech = p->mechanisms;
    for (int l = 0; l < m->mech; ++l)
        m->rates[l] = 2 * PI *
                      (m->mech == 1 ? sqrt(p->q_min_frequency * p->q_max_frequency)
                                    : p->q_min_frequency * pow(p->q_max_frequency / p->q_min_frequency, (double)l / (m->mech - 1)));
    if (!gsls_strengths(qp, m, p->q_min_frequency, p->q_max_frequency, m->wp) ||
        !gsls_strengths(qs, m, p->q_min_frequency, p->q_max_frequency, m->ws))
        return set_error(err, len, "No passive GSLS fit for this Q/frequency band."), 0;
    return 1;
}
static int gsls_material(const orc_params2d *p, const gsls_fit_t *m, double vp, double vs, double rho, double qp, double qs,
                         int q_error, gsls_mat_t *g, char *err, int len) {
    g->pR = gsls_scale(m, m->wp, p->q_reference_frequency, rho, vp);
    g->sR = gsls_scale(m, m->ws, p->q_reference_frequency, rho, vs);
    g->pU = g->pR;
    g->sU = g->sR;
    if (!(g->pR > 4 * g->sR / 3)) return set_error(err, len, "Qp/Qs produce nonpositive relaxed bulk modulus."), 0;
    for (int l = 0; l < m->mech; ++l) {
        g->dP[l] = g->pR * m->wp[l];
        g->dS[l] = g->sR * m->ws[l];
        if (!(g->dP[l] + 1e-12 * g->pR >= 4 * g->dS[l] / 3))
            return set_error(err, len, "Qp/Qs violate passive bulk relaxation (Qp cannot be arbitrarily large relative to Qs)."), 0;
        g->pU += g->dP[l];
        g->sU += g->dS[l];
    }
    if (!isfinite(g->pU) || !isfinite(g->sU)) return set_error(err, len, "GSLS moduli overflow."), 0;
    g->max_q_error = 0;
    if (q_error)
        for (int i = 0; i <= 160; ++i) {
            const double f = p->q_min_frequency * pow(p->q_max_frequency / p->q_min_frequency, i / 160.);
            const double e1 = fabs(gsls_quality(g, m, 0, f) / qp - 1), e2 = fabs(gsls_quality(g, m, 1, f) / qs - 1);
            g->max_q_error = dmax(g->max_q_error, dmax(e1, e2));
        }
    return 1;
}

/* ---------------- setup ---------------- */
static int valid_material(double vp, double vs, double rho) {
    return isfinite(vp) && isfinite(vs) && isfinite(rho) && vp > 0 && vs > 0 && rho > 0 && vp * vp > 4 * vs * vs / 3 &&
           isfinite(rho * vp * vp);
}

static void orc2d_free(orc2d *s) {
    if (!s) return;
    double **arrays[] = {&s->coord, &s->nrm, &s->a, &s->b, &s->jac, &s->vp, &s->vs, &s->rho, &s->lam, &s->mu,
                         &s->src, &s->damp, &s->grad[0], &s->grad[1], &s->dP[0], &s->dP[1], &s->dP[2], &s->dS[0],
                         &s->dS[1], &s->dS[2], &s->cx, &s->cz, &s->rec_pos, &s->rec_vx, &s->rec_vz, &s->prev};
    for (size_t i = 0; i < sizeof arrays / sizeof arrays[0]; ++i) free(*arrays[i]);
    free(s->rec_x);
    free(s->rec_z);
    free(s->probe);
    if (s->state) {
        if (s->prec) {
            state_t_f32 *st = s->state;
            for (int k = 0; k < 7; ++k) {
                free(st->f[k][0]);
                free(st->p[k]);
            }
            mesh_t_f32 *m = s->mesh;
            void *ptrs[] = {m->a, m->b, m->jac, m->nrm, m->lam, m->mu, m->rho, m->src, m->damp, m->grad[0], m->grad[1],
                            m->dP[0], m->dP[1], m->dP[2], m->dS[0], m->dS[1], m->dS[2], m->cx, m->cz};
            for (size_t i = 0; i < sizeof ptrs / sizeof ptrs[0]; ++i) free(ptrs[i]);
        } else {
            state_t_f64 *st = s->state;
            for (int k = 0; k < 7; ++k) {
                free(st->f[k][0]);
                free(st->p[k]);
            }
        }
    }
    free(s->state);
    free(s->mesh);
    free(s->ctx);
    free(s);
}

#define REQUIRE(cond, msg)                                                                                             \
    do {                                                                                                               \
        if (!(cond)) {                                                                                                 \
            set_error(error, error_len, "%s", msg);                                                                    \
            orc2d_free(s);                                                                                             \
            return NULL;                                                                                               \
        }                                                                                                              \
    } while (0)

static float *to_f32(const double *x, size_t n, int *ok) {
    if (!x) return NULL;
    float *r = malloc(sizeof(float) * (n ? n : 1));
    for (size_t i = 0; i < n; ++i) {
        r[i] = (float)x[i];
        if (!isfinite(r[i])) *ok = 0;
    }
    return r;
}

void *orc2d_create(const orc_params2d *in, char *error, int error_len) {
    orc2d *s = calloc(1, sizeof(orc2d));
    if (error && error_len) error[0] = 0;
    orc_params2d p = *in;
    const orc_params2d logical = *in;
    s->logical = logical;
    /* CurvilinearSolver2D.cpp:161-227 (validation subset that shapes the numerics) */
    REQUIRE(p.precision == 0 || p.precision == 1, "Unknown 2D calculation precision.");
    REQUIRE(p.nx >= 8 && p.nz >= 8 && p.nx <= 16384 && p.nz <= 16384, "2D dimensions must be in 8..16384 (subject to RAM/VRAM).");
    REQUIRE(p.spacing > 0 && p.frequency > 0 && p.frequency <= 1e6 && p.duration > 0 && p.duration <= 10 && p.cfl >= .02 &&
                p.cfl <= .8 && p.wavelength > 0 && p.flatten_depth >= 0 &&
                (p.source_kind == 1 ? p.moment_per_length : p.force_per_length) > 0,
            "Invalid 2D spacing, frequency, duration, CFL or source amplitude.");
    REQUIRE(valid_material(p.vp, p.vs, p.density), "Require positive material values and positive elastic bulk modulus.");
    REQUIRE(p.sample_interval_ms == 0 || (p.sample_interval_ms >= .01 && p.sample_interval_ms <= 1000),
            "Record interval must be 0 (each step) or 0.01..1000 ms.");
    REQUIRE(p.max_steps >= 0 && p.max_steps <= 200000, "Step cap must be 0..200000 (0 = full duration); CPU workers must be 0..64.");
    const int explosive = p.source_kind == 1;
    if (!explosive) {
        const double nrm = orc_hypot3(p.source_direction[0], p.source_direction[1], p.source_direction[2]);
        REQUIRE(isfinite(nrm) && nrm > 0 && p.source_direction[1] == 0,
                "2D force direction must be finite, nonzero and inside the X-Z plane.");
    }
    REQUIRE(isfinite(p.source_sigma_m) && isfinite(p.source_radius_m) &&
                ((p.source_sigma_m == 0 && p.source_radius_m == 0) ||
                 (!explosive && p.source_sigma_m > 0 && p.source_radius_m >= p.source_sigma_m)),
            "Physical source kernel requires a directional force and 0 < sigma <= radius, or both zero for legacy.");
    int sx = p.source_x < 0 ? p.nx / 2 : p.source_x;
    if (explosive)
        REQUIRE(p.nx >= 11 && p.nz >= 11 && sx >= 5 && sx <= p.nx - 6 && p.source_z >= 5 && p.source_z <= p.nz - 6,
                "Isotropic moment needs 11+ nodes per axis and source center five nodes inside each face.");
    REQUIRE(p.source_x >= -1 && sx >= 2 && sx <= p.nx - 3 && p.source_z >= 0 && p.source_z <= p.nz - 3,
            "2D force source may be on the free surface; keep two cells from the sides and bottom.");
    REQUIRE(p.receiver_start >= 0 && p.receiver_count > 0 && p.receiver_count <= 1024 && p.receiver_stride > 0 &&
                p.receiver_depth >= 0 && p.receiver_depth < p.nz &&
                (long long)p.receiver_start + ((long long)p.receiver_count - 1) * p.receiver_stride < p.nx,
            "Receiver line is outside the 2D model or exceeds 1024 positions.");
    REQUIRE(p.sponge_x >= 1 && p.sponge_x <= 512 && p.sponge_z >= 1 && p.sponge_z <= 512 && p.sponge_strength >= .1 &&
                p.sponge_strength <= 10,
            "Sponge widths must be 1..512 cells (effective width <= axis nodes / 3) and strength 0.1..10.");
    REQUIRE(p.absorber == 0 || p.absorber == 1, "Unknown 2D absorber.");
    REQUIRE(isfinite(p.pml_reflection) && p.pml_reflection >= 1e-8 && p.pml_reflection <= .1 && p.pml_kappa_max >= 1 &&
                p.pml_kappa_max <= 20 && p.pml_frequency >= 0 && p.pml_frequency <= 1e6,
            "CFS-PML requires R in 1e-8..0.1, kappa in 1..20, frequency in 0..1e6 Hz.");
    if (p.layered) {
        REQUIRE(p.n_layers >= 1 && p.n_layers <= ORC_MAX_LAYERS, "Require 1..16 material layers.");
        for (int l = 0; l < p.n_layers; ++l)
            REQUIRE(valid_material(p.layer_vp[l], p.layer_vs[l], p.layer_density[l]) && isfinite(p.layer_thickness[l]) &&
                        p.layer_thickness[l] >= 0 && (l + 1 == p.n_layers || p.layer_thickness[l] > 0),
                    "Invalid layer thickness or material.");
    }
    gsls_fit_t fit[ORC_MAX_LAYERS];
    gsls_mat_t rheo[ORC_MAX_LAYERS];
    const int nmat = p.layered ? p.n_layers : 1;
    if (p.attenuation) {
        REQUIRE(p.mechanisms >= 1 && p.mechanisms <= 3, "GSLS mechanism count must be 1..3.");
        REQUIRE(p.q_min_frequency >= .01 && p.q_max_frequency > p.q_min_frequency && p.q_max_frequency <= 1e6 &&
                    p.q_reference_frequency >= p.q_min_frequency && p.q_reference_frequency <= p.q_max_frequency,
                "GSLS requires 0.01 <= fmin < fmax <= 1e6 Hz and fref within the band.");
        REQUIRE(p.qp >= 5 && p.qp <= 1e9 && p.qs >= 5 && p.qs <= 1e9,
                "Qp and Qs must be finite and in 5..1e9; disable attenuation for the elastic limit.");
        for (int l = 0; l < nmat; ++l) {
            const double qp = p.layered ? p.layer_qp[l] : p.qp, qs = p.layered ? p.layer_qs[l] : p.qs;
            const double vp = p.layered ? p.layer_vp[l] : p.vp, vs = p.layered ? p.layer_vs[l] : p.vs,
                         rho = p.layered ? p.layer_density[l] : p.density;
            REQUIRE(qp >= 5 && qp <= 1e9 && qs >= 5 && qs <= 1e9, "Layer Qp/Qs must be in 5..1e9.");
            char e2[256];
            if (!gsls_fit(&p, qp, qs, &fit[l], e2, sizeof e2) || !gsls_material(&p, &fit[l], vp, vs, rho, qp, qs, 1, &rheo[l], e2, sizeof e2))
                REQUIRE(0, e2);
        }
    }
    /* Detail::domainLayout (CfsPml2D.hpp:10-14) */
    s->wx = iclamp(p.sponge_x, 1, imax(1, p.nx / 3));
    s->wz = iclamp(p.sponge_z, 1, imax(1, p.nz / 3));
    int dnx = p.nx, dnz = p.nz;
    s->offset_x = 0;
    if (p.absorbing && p.external_absorption) {
        s->offset_x = s->wx;
        dnx += 2 * s->wx;
        dnz += s->wz;
    }
    REQUIRE(dnx <= 16384 && dnz <= 16384,
            "2D computational dimensions including external absorbing strips must not exceed 16384 per axis.");
    p.nx = dnx;
    p.nz = dnz;
    p.sponge_x = s->wx;
    p.sponge_z = s->wz;
    p.external_absorption = 0;
    sx += s->offset_x;
    p.source_x = sx;
    p.receiver_start += s->offset_x;
    s->p = p;
    s->prec = p.precision;
#ifdef _OPENMP
    s->threads = p.threads ? p.threads : imin(8, omp_get_num_procs());
#else
    s->threads = 1;
#endif
    const int gnx = p.nx + 6, gnz = p.nz + 6;
    const size_t count = (size_t)gnx * gnz;
    s->nx = gnx;
    s->nz = gnz;
    s->count = count;
    s->h = p.spacing;
    const int off = s->offset_x;
    const double bottom = (logical.nz - 1) * p.spacing;
    double *sz = malloc(sizeof(double) * gnx), *sslope = malloc(sizeof(double) * gnx);
    double relief = fabs(p.relief);
    if (p.imported_surface) {
        pchip_t prof;
        char e2[256];
        if (!pchip_init(&prof, p.n_surface, p.surface_xz, e2, sizeof e2)) {
            free(prof.slope), free(sz), free(sslope);
            REQUIRE(0, e2);
        }
        if (!(p.surface_xz[0] <= 0 && p.surface_xz[2 * (p.n_surface - 1)] >= (logical.nx - 1) * p.spacing)) {
            free(prof.slope), free(sz), free(sslope);
            REQUIRE(0, "Imported terrain must cover x=0 through the full physical model span.");
        }
        relief = 0;
        for (int i = 0; i < gnx; ++i) {
            double v, d;
            pchip_sample(&prof, (i - 3 - off) * p.spacing, &v, &d);
            sz[i] = p.surface_datum - v;
            sslope[i] = -d;
            if (i >= off && i < off + logical.nx + 6) relief = dmax(relief, fabs(sz[i]));
        }
        free(prof.slope);
    } else {
        for (int i = 0; i < gnx; ++i) {
            const double phase = 2 * PI * (i - 3 - off) * p.spacing / p.wavelength;
            sz[i] = p.relief * cos(phase);
            sslope[i] = -p.relief * 2 * PI / p.wavelength * sin(phase);
        }
    }
    s->flatten = p.flatten_depth > 0 ? p.flatten_depth
                                     : dmin(bottom, dmax(dmax(.35 * bottom, 6 * p.spacing), 4 * relief + 2 * p.spacing));
    if (!(s->flatten <= bottom && (relief == 0 || ((p.imported_surface || p.wavelength >= 4 * p.spacing) &&
                                                   s->flatten >= 4 * p.spacing && s->flatten > 2.6 * relief)))) {
        free(sz), free(sslope);
        REQUIRE(0, "Terrain must be resolved and flatten depth must exceed 2.6 times relief within the model.");
    }
    if (off) { /* C1 continuation into the external strips, CurvilinearSolver2D.cpp:317-330 */
        for (int side = 0; side < 2; ++side) {
            const int edge = 3 + off + (side ? logical.nx - 1 : 0);
            const double z = sz[edge], slope = sslope[edge];
            const double length =
                dmin(s->wx * p.spacing / 3., .5 * (s->flatten / 2.6 - fabs(z)) / dmax(fabs(slope), 1e-30));
            for (int i = side ? edge + 1 : 0; i < (side ? gnx : edge); ++i) {
                const double distance = abs(i - edge) * p.spacing, decay = exp(-distance / length);
                sz[i] = z + (side ? 1. : -1.) * slope * length * (1 - decay);
                sslope[i] = slope * decay;
            }
        }
    }
    double **alloc[] = {&s->a, &s->b, &s->jac, &s->vp, &s->vs, &s->rho, &s->lam, &s->mu, &s->src, &s->damp};
    for (size_t i = 0; i < sizeof alloc / sizeof alloc[0]; ++i) *alloc[i] = calloc(count, sizeof(double));
    s->coord = calloc(2 * count, sizeof(double));
    s->nrm = calloc(2 * count, sizeof(double));
    s->mech = p.attenuation ? p.mechanisms : 0;
    for (int l = 0; l < s->mech; ++l) {
        s->dP[l] = calloc(count, sizeof(double));
        s->dS[l] = calloc(count, sizeof(double));
    }
    s->min_jac = DBL_MAX;
    s->max_metric = 0;
    double max_vp = 0, min_vs = DBL_MAX, max_sigma = 0;
    int ok = 1;
    char merr[256] = "";
    for (int k = 0; k < gnz && ok; ++k)
        for (int i = 0; i < gnx; ++i) {
            const size_t q = (size_t)i + (size_t)gnx * k;
            const double x = (i - 3 - off) * p.spacing, eta = (k - 3) * p.spacing, t = dclamp(eta / s->flatten, 0.0, 1.0);
            const double fade = 1 - pow(t, 5) * (126 + t * (-420 + t * (540 + t * (-315 + 70 * t))));
            const double fd = eta > 0 && eta < s->flatten ? -630 * pow(t, 4) * pow(1 - t, 4) / s->flatten : 0;
            const double zx = sslope[i] * fade, j = 1 + sz[i] * fd;
            if (!(j > 1e-9 && isfinite(j))) {
                ok = 0;
                snprintf(merr, sizeof merr, "Nonpositive curvilinear Jacobian.");
                break;
            }
            s->coord[2 * q] = x;
            s->coord[2 * q + 1] = eta + sz[i] * fade;
            s->a[q] = -zx / j;
            s->b[q] = 1 / j;
            s->jac[q] = j;
            const double n = hypot(zx, 1.0);
            s->nrm[2 * q] = -zx / n;
            s->nrm[2 * q + 1] = 1 / n;
            double vp = p.vp, vs = p.vs, rho = p.density;
            int mi = 0;
            if (p.layered) {
                double boundary = 0;
                const double md = off ? dclamp(eta, 0., bottom) : dmax(0., eta);
                for (int l = 0; l < p.n_layers; ++l) {
                    boundary += p.layer_thickness[l];
                    if (l + 1 == p.n_layers || md < boundary) {
                        vp = p.layer_vp[l];
                        vs = p.layer_vs[l];
                        rho = p.layer_density[l];
                        mi = l;
                        break;
                    }
                }
            }
            if (p.node_vp) {
                const int li = iclamp(i - 3 - off, 0, logical.nx - 1), lk = iclamp(k - 3, 0, logical.nz - 1);
                const size_t lq = (size_t)li + (size_t)logical.nx * lk;
                vp = p.node_vp[lq];
                vs = p.node_vs[lq];
                rho = p.node_density[lq];
            }
            s->vp[q] = vp;
            s->vs[q] = vs;
            s->rho[q] = rho;
            s->mu[q] = rho * vs * vs;
            s->lam[q] = rho * vp * vp - 2 * s->mu[q];
            double speed_max = vp, speed_min = vs;
            if (p.attenuation) {
                gsls_mat_t gm = rheo[mi];
                if (p.node_vp) {
                    const double qp = p.layered ? p.layer_qp[mi] : p.qp, qs = p.layered ? p.layer_qs[mi] : p.qs;
                    if (!gsls_material(&p, &fit[mi], vp, vs, rho, qp, qs, 0, &gm, merr, sizeof merr)) {
                        ok = 0;
                        break;
                    }
                }
                s->mu[q] = gm.sU;
                s->lam[q] = gm.pU - 2 * gm.sU;
                for (int l = 0; l < s->mech; ++l) {
                    s->dP[l][q] = gm.dP[l];
                    s->dS[l][q] = gm.dS[l];
                }
                speed_max = sqrt(gm.pU / rho);
                speed_min = sqrt(gm.sR / rho);
            }
            if (i >= 3 && i < gnx - 3 && k >= 3 && k < gnz - 3) {
                s->min_jac = dmin(s->min_jac, j);
                s->max_metric = dmax(s->max_metric, sqrt(1 + s->a[q] * s->a[q] + s->b[q] * s->b[q]));
                max_vp = dmax(max_vp, speed_max);
                min_vs = dmin(min_vs, speed_min);
            }
        }
    free(sz);
    free(sslope);
    REQUIRE(ok, merr);
    s->max_edge = 0;
    const int sponge = p.absorbing && p.absorber == 0;
    for (int k = 0; k < p.nz; ++k)
        for (int i = 0; i < p.nx; ++i) {
            const size_t q = (size_t)(i + 3) + (size_t)gnx * (k + 3);
            if (i + 1 < p.nx)
                s->max_edge = dmax(s->max_edge, hypot(s->coord[2 * (q + 1)] - s->coord[2 * q],
                                                      s->coord[2 * (q + 1) + 1] - s->coord[2 * q + 1]));
            if (k + 1 < p.nz) s->max_edge = dmax(s->max_edge, s->coord[2 * (q + gnx) + 1] - s->coord[2 * q + 1]);
            if (sponge) {
                const int wx = imin(p.sponge_x, p.nx / 3), wz = imin(p.sponge_z, p.nz / 3);
                const double x = dmax(0.0, 1.0 - imin(i, p.nx - 1 - i) / (double)wx),
                             z = dmax(0.0, 1.0 - (p.nz - 1 - k) / (double)wz);
                s->damp[q] = p.sponge_strength * max_vp / p.spacing * (x * x / wx + z * z / wz);
                max_sigma = dmax(max_sigma, s->damp[q]);
            }
        }
    double cpml_rate = 0;
    if (p.absorbing && p.absorber == 1) { /* CfsPml2D.hpp:31-44 */
        const int wx = imin(p.sponge_x, p.nx / 3), wz = imin(p.sponge_z, p.nz / 3);
        const double freq = p.pml_frequency > 0 ? p.pml_frequency : p.frequency;
        s->cx = malloc(sizeof(double) * 3 * p.nx);
        s->cz = malloc(sizeof(double) * 3 * p.nz);
        for (int pass = 0; pass < 2; ++pass) {
            const int nodes = pass ? p.nz : p.nx, width = pass ? wz : wx, both = !pass;
            double *r = pass ? s->cz : s->cx;
            const double d0 = -3 * max_vp * log(p.pml_reflection) / (2 * width * p.spacing);
            for (int i = 0; i < nodes; ++i) {
                const int distance = both ? imin(i, nodes - 1 - i) : nodes - 1 - i;
                const double t = dclamp(1. - (double)distance / width, 0., 1.);
                r[3 * i] = d0 * t * t;
                r[3 * i + 1] = 1 + (p.pml_kappa_max - 1) * t * t;
                r[3 * i + 2] = t > 0 ? 3.14159265358979323846 * freq * (1 - t) : 0;
                cpml_rate = dmax(cpml_rate, r[3 * i] / r[3 * i + 1] + r[3 * i + 2]);
            }
        }
        s->cpml = 4 * (2 * (int64_t)s->wx * p.nz + (int64_t)s->wz * p.nx);
    }
    /* source weights, CurvilinearSolver2D.cpp:419-442 */
    double weight = 0;
    if (p.source_sigma_m > 0) {
        const size_t cq = (size_t)(sx + 3) + (size_t)gnx * (p.source_z + 3);
        const double cx0 = s->coord[2 * cq], cz0 = s->coord[2 * cq + 1];
        double bd = INFINITY;
        for (int k = 0; k < p.nz; ++k)
            for (int i = 0; i < p.nx; ++i) {
                const size_t q = (size_t)(i + 3) + (size_t)gnx * (k + 3);
                const double d = hypot(s->coord[2 * q] - cx0, s->coord[2 * q + 1] - cz0);
                if (k == 0 || k == p.nz - 1 || i == 0 || i == p.nx - 1) bd = dmin(bd, d);
                if (d <= p.source_radius_m) {
                    s->src[q] = exp(-.5 * d * d / (p.source_sigma_m * p.source_sigma_m));
                    weight += s->src[q] * s->jac[q];
                }
            }
        REQUIRE(bd > p.source_radius_m, "Physical source support intersects a physical computational boundary.");
        REQUIRE(weight > 0 && isfinite(weight), "Invalid physical source normalization.");
    } else {
        for (int dk = -2; dk <= 2; ++dk)
            for (int di = -2; di <= 2; ++di) {
                if (p.source_z + dk < 0) continue;
                const size_t q = (size_t)(sx + 3 + di) + (size_t)gnx * (p.source_z + 3 + dk);
                s->src[q] = exp(-.5 * (di * di + dk * dk));
                weight += s->src[q] * s->jac[q];
            }
    }
    for (size_t q = 0; q < count; ++q) s->src[q] = s->src[q] / weight;
    s->explosive = explosive;
    if (explosive) { /* SourceMechanism.hpp:43-56 with the 2D mapped flux */
        s->grad[0] = calloc(count, sizeof(double));
        s->grad[1] = calloc(count, sizeof(double));
        for (int k = p.source_z - 4; k <= p.source_z + 4; ++k)
            for (int i = sx - 4; i <= sx + 4; ++i) {
                const ptrdiff_t q = (ptrdiff_t)(i + 3) + (ptrdiff_t)gnx * (k + 3);
                double res[2] = {0, 0};
                for (int c = 0; c < 2; ++c)
                    for (int ax = 0; ax < 2; ++ax) {
                        const ptrdiff_t st = ax == 0 ? 1 : gnx;
                        double fl[4];
                        const ptrdiff_t at[4] = {q - 2 * st, q - st, q + st, q + 2 * st};
                        for (int u = 0; u < 4; ++u) {
                            const ptrdiff_t r = at[u];
                            const double metric = ax == 0 ? (c == 0 ? 1. : 0.) : (c == 0 ? s->a[r] : s->b[r]);
                            fl[u] = s->jac[r] * metric * s->src[r];
                        }
                        res[c] -= (fl[0] - 8 * fl[1] + 8 * fl[2] - fl[3]) / (12 * p.spacing * s->jac[q]);
                    }
                s->grad[0][q] = res[0];
                s->grad[1][q] = res[1];
            }
    }
    if (p.source_sigma_m > 0 || p.source_z < 2) {
        const size_t cq = (size_t)(sx + 3) + (size_t)gnx * (p.source_z + 3);
        double second = 0;
        for (size_t q = 0; q < count; ++q)
            if (s->src[q] > 0) {
                const double w = s->src[q] * s->jac[q], dx = s->coord[2 * q] - s->coord[2 * cq], dz = s->coord[2 * q + 1] - s->coord[2 * cq + 1];
                s->weight_sum += w;
                second += w * (dx * dx + dz * dz);
            }
        s->rms_radius = sqrt(second / s->weight_sum);
    }
    /* dt plan, CurvilinearSolver2D.cpp:460-481 */
    s->ppw = min_vs / (2.5 * p.frequency * s->max_edge);
    s->wave_dt = p.cfl * p.spacing / (max_vp * s->max_metric);
    s->damping_dt = max_sigma > 0 ? .5 / max_sigma : INFINITY;
    s->source_dt = 1 / (50 * p.frequency);
    s->cpml_dt = cpml_rate > 0 ? .5 / cpml_rate : INFINITY;
    s->relaxation_dt = p.attenuation ? .25 / fit[0].rates[p.mechanisms - 1] : INFINITY;
    s->fields = p.attenuation ? 6 + 4 * p.mechanisms : 5;
    for (int l = 0; l < nmat && p.attenuation; ++l) s->max_q_error = dmax(s->max_q_error, rheo[l].max_q_error);
    const double safe = dmin(dmin(dmin(s->wave_dt, s->damping_dt), dmin(s->source_dt, s->relaxation_dt)), s->cpml_dt);
    const double full = ceil(p.duration / safe);
    REQUIRE(isfinite(full) && full <= 10000000, "Requested duration needs too many steps; increase spacing or lower frequency.");
    s->dt = p.duration / full;
    s->total = p.max_steps ? imin((int)full, p.max_steps) : (int)full;
    REQUIRE(s->total <= 200000,
            "2D session exceeds 200000 accepted steps; set an explicit cap or shorter duration. A cap ends before the requested duration.");
    /* receivers, CurvilinearSolver2D.cpp:482-497 */
    s->nrec = p.receiver_count;
    s->rec_x = malloc(sizeof(int32_t) * s->nrec);
    s->rec_z = malloc(sizeof(int32_t) * s->nrec);
    s->rec_pos = malloc(sizeof(double) * 2 * s->nrec);
    s->probe = malloc(sizeof(int64_t) * s->nrec);
    s->prev = calloc(2 * (size_t)s->nrec, sizeof(double));
    const double interval = p.sample_interval_ms > 0 ? p.sample_interval_ms * .001 : s->dt;
    s->rec_cap = (size_t)floor(s->total * s->dt / interval) + 2;
    s->rec_vx = calloc(s->rec_cap * s->nrec, sizeof(double));
    s->rec_vz = calloc(s->rec_cap * s->nrec, sizeof(double));
    s->rec_len = 1;
    for (int r = 0; r < s->nrec; ++r) {
        s->rec_x[r] = logical.receiver_start + r * p.receiver_stride;
        s->rec_z[r] = p.receiver_depth;
        const size_t mq = (size_t)(s->rec_x[r] + 3 + off) + (size_t)gnx * (s->rec_z[r] + 3);
        s->rec_pos[2 * r] = s->coord[2 * mq];
        s->rec_pos[2 * r + 1] = s->coord[2 * mq + 1];
        s->probe[r] = (int64_t)mq;
    }
    /* unit source direction (SourceMechanism.hpp:32-38) */
    if (explosive) s->dir[0] = s->dir[1] = s->dir[2] = 0;
    else {
        const double n = orc_hypot3(p.source_direction[0], p.source_direction[1], p.source_direction[2]);
        for (int c = 0; c < 3; ++c) s->dir[c] = p.source_direction[c] / n;
    }
    for (int l = 0; l < s->mech; ++l) s->rates[l] = fit[0].rates[l];
    s->amplitude = explosive ? p.moment_per_length : p.force_per_length;
    /* precision conversion and quantised validation, :498-503 */
    if (s->prec) {
        int fin = 1;
        mesh_t_f32 *m = calloc(1, sizeof(mesh_t_f32));
        s->mesh = m;
        float *coord32 = to_f32(s->coord, 2 * count, &fin);
        m->nrm = to_f32(s->nrm, 2 * count, &fin);
        m->a = to_f32(s->a, count, &fin);
        m->b = to_f32(s->b, count, &fin);
        m->jac = to_f32(s->jac, count, &fin);
        float *vp32 = to_f32(s->vp, count, &fin), *vs32 = to_f32(s->vs, count, &fin);
        m->rho = to_f32(s->rho, count, &fin);
        m->lam = to_f32(s->lam, count, &fin);
        m->mu = to_f32(s->mu, count, &fin);
        m->src = to_f32(s->src, count, &fin);
        m->damp = to_f32(s->damp, count, &fin);
        m->cx = to_f32(s->cx, s->cx ? 3 * (size_t)p.nx : 0, &fin);
        m->cz = to_f32(s->cz, s->cz ? 3 * (size_t)p.nz : 0, &fin);
        for (int c = 0; c < 2; ++c) m->grad[c] = to_f32(s->grad[c], count, &fin);
        for (int l = 0; l < s->mech; ++l) {
            m->dP[l] = to_f32(s->dP[l], count, &fin);
            m->dS[l] = to_f32(s->dS[l], count, &fin);
        }
        for (size_t q = 0; q < count && fin; ++q)
            if (!((double)m->mu[q] > 0 && (double)m->lam[q] + 2 * (double)m->mu[q] / 3 > 0 && (double)m->rho[q] > 0 &&
                  (double)vp32[q] > 0 && (double)vs32[q] > 0 && (double)m->jac[q] > 0 && (double)m->b[q] > 0))
                fin = 2;
        free(coord32), free(vp32), free(vs32);
        REQUIRE(fin != 0, "2D FP32 model/state coefficient overflow; select FP64 or revise model.");
        REQUIRE(fin == 1, "Invalid quantized material or geometry; select FP64 or revise model.");
        REQUIRE(isfinite((float)p.spacing) && (float)p.spacing > 0 && isfinite((float)s->dt) && (float)s->dt > 0,
                "FP32 spacing/time step is not representable; select FP64.");
        m->h = (float)p.spacing;
        for (int l = 0; l < 3; ++l) m->rates[l] = (float)s->rates[l];
        for (int c = 0; c < 3; ++c) m->dir[c] = (float)s->dir[c];
        ctx_t_f32 *g = calloc(1, sizeof(ctx_t_f32));
        *g = (ctx_t_f32){m, gnx, gnz, s->fields, s->mech, s->wx, s->wz, count, s->cpml};
        s->ctx = g;
        state_t_f32 *st = calloc(1, sizeof(state_t_f32));
        for (int k = 0; k < 7; ++k) {
            float *block = calloc((size_t)s->fields * count, sizeof(float));
            for (int f = 0; f < s->fields; ++f) st->f[k][f] = block + (size_t)f * count;
            st->p[k] = s->cpml ? calloc((size_t)s->cpml, sizeof(float)) : NULL;
        }
        s->state = st;
    } else {
        for (size_t q = 0; q < count; ++q)
            REQUIRE(s->mu[q] > 0 && s->lam[q] + 2 * s->mu[q] / 3 > 0 && s->rho[q] > 0 && s->vp[q] > 0 && s->vs[q] > 0 &&
                        s->jac[q] > 0 && s->b[q] > 0,
                    "Invalid quantized material or geometry; select FP64 or revise model.");
        mesh_t_f64 *m = calloc(1, sizeof(mesh_t_f64));
        *m = (mesh_t_f64){s->a, s->b, s->jac, s->nrm, s->lam, s->mu, s->rho, s->src, s->damp, {s->grad[0], s->grad[1]},
                          {s->dP[0], s->dP[1], s->dP[2]}, {s->dS[0], s->dS[1], s->dS[2]}, s->cx, s->cz, p.spacing,
                          {s->rates[0], s->rates[1], s->rates[2]}, {s->dir[0], s->dir[1], s->dir[2]}};
        s->mesh = m;
        ctx_t_f64 *g = calloc(1, sizeof(ctx_t_f64));
        *g = (ctx_t_f64){m, gnx, gnz, s->fields, s->mech, s->wx, s->wz, count, s->cpml};
        s->ctx = g;
        state_t_f64 *st = calloc(1, sizeof(state_t_f64));
        for (int k = 0; k < 7; ++k) {
            double *block = calloc((size_t)s->fields * count, sizeof(double));
            for (int f = 0; f < s->fields; ++f) st->f[k][f] = block + (size_t)f * count;
            st->p[k] = s->cpml ? calloc((size_t)s->cpml, sizeof(double)) : NULL;
        }
        s->state = st;
    }
    s->finite = 1;
    return s;
}

void orc2d_destroy(void *h) { orc2d_free((orc2d *)h); }

/* CurvilinearSolver2D.cpp:120-147 */
static void orc2d_sample(orc2d *s) {
    double *vx = malloc(sizeof(double) * s->nrec), *vz = malloc(sizeof(double) * s->nrec);
    for (int r = 0; r < s->nrec; ++r) {
        if (s->prec) {
            state_t_f32 *st = s->state;
            vx[r] = st->f[0][0][s->probe[r]];
            vz[r] = st->f[0][1][s->probe[r]];
        } else {
            state_t_f64 *st = s->state;
            vx[r] = st->f[0][0][s->probe[r]];
            vz[r] = st->f[0][1][s->probe[r]];
        }
    }
    const double interval = s->p.sample_interval_ms > 0 ? s->p.sample_interval_ms * .001 : s->dt;
    const double end = (s->step + 1) * s->dt;
    for (size_t n = s->rec_len; n * interval <= end + 1e-9 * s->dt; ++n) {
        const double a = dclamp((n * interval - s->step * s->dt) / s->dt, 0.0, 1.0);
        if (n >= s->rec_cap) { /* grow */
            const size_t cap = s->rec_cap * 2;
            double *nx = calloc(cap * s->nrec, sizeof(double)), *nz = calloc(cap * s->nrec, sizeof(double));
            for (int r = 0; r < s->nrec; ++r) {
                memcpy(nx + (size_t)r * cap, s->rec_vx + (size_t)r * s->rec_cap, sizeof(double) * s->rec_len);
                memcpy(nz + (size_t)r * cap, s->rec_vz + (size_t)r * s->rec_cap, sizeof(double) * s->rec_len);
            }
            free(s->rec_vx), free(s->rec_vz);
            s->rec_vx = nx;
            s->rec_vz = nz;
            s->rec_cap = cap;
        }
        for (int r = 0; r < s->nrec; ++r) {
            s->rec_vx[(size_t)r * s->rec_cap + n] = orc_lerp(s->prev[2 * r], vx[r], a);
            s->rec_vz[(size_t)r * s->rec_cap + n] = orc_lerp(s->prev[2 * r + 1], vz[r], a);
        }
        s->rec_len = n + 1;
    }
    for (int r = 0; r < s->nrec; ++r) {
        s->prev[2 * r] = vx[r];
        s->prev[2 * r + 1] = vz[r];
    }
    free(vx);
    free(vz);
}

int orc2d_advance(void *hh, int count, char *error, int error_len) {
    orc2d *s = hh;
    if (error && error_len) error[0] = 0;
    if (s->fault[0]) return set_error(error, error_len, "%s", s->fault), 0;
    while (count-- > 0 && s->step < s->total) {
        const double before_peak = s->peak, be = s->energy, bv = s->max_velocity, bt = s->traction;
        const int bf = s->finite;
        if (s->prec) {
            rk_step_f32(s, s->state);
            inspect_f32(s, s->state);
        } else {
            rk_step_f64(s, s->state);
            inspect_f64(s, s->state);
        }
        if (!s->finite || ((s->step + 1) * s->dt > 3 / s->p.frequency && before_peak > 0 && s->energy > before_peak * 10000)) {
            if (s->prec) rollback_f32(s, s->state);
            else rollback_f64(s, s->state);
            s->energy = be, s->max_velocity = bv, s->traction = bt, s->finite = bf, s->peak = before_peak;
            snprintf(s->fault, sizeof s->fault,
                     "2D step rejected: nonfinite fields or excessive late energy growth. Rebuild the session.");
            return set_error(error, error_len, "%s", s->fault), 0;
        }
        orc2d_sample(s);
        ++s->step;
        s->peak = dmax(before_peak, s->energy);
    }
    return 1;
}

int orc2d_reset(void *hh) {
    orc2d *s = hh;
    for (int k = 0; k < 7; ++k) {
        if (s->prec) {
            state_t_f32 *st = s->state;
            memset(st->f[k][0], 0, sizeof(float) * s->count * s->fields);
            if (s->cpml) memset(st->p[k], 0, sizeof(float) * (size_t)s->cpml);
        } else {
            state_t_f64 *st = s->state;
            memset(st->f[k][0], 0, sizeof(double) * s->count * s->fields);
            if (s->cpml) memset(st->p[k], 0, sizeof(double) * (size_t)s->cpml);
        }
    }
    s->step = 0;
    s->fault[0] = 0;
    s->energy = s->peak = s->max_velocity = s->traction = 0;
    s->finite = 1;
    s->rec_len = 1;
    memset(s->rec_vx, 0, sizeof(double) * s->rec_cap * s->nrec);
    memset(s->rec_vz, 0, sizeof(double) * s->rec_cap * s->nrec);
    memset(s->prev, 0, sizeof(double) * 2 * s->nrec);
    return 1;
}

int orc2d_step(void *h) { return ((orc2d *)h)->step; }
/* Test hook: replaces the accepted peak energy that the late-growth fault rule compares against
   (CurvilinearSolver2D.cpp:576-578), so a test can drive that branch without an unstable model. */
void orc2d_set_peak(void *h, double peak) { ((orc2d *)h)->peak = peak; }
int orc2d_total(void *h) { return ((orc2d *)h)->total; }
double orc2d_dt(void *h) { return ((orc2d *)h)->dt; }
double orc2d_record_dt(void *h) {
    orc2d *s = h;
    return s->p.sample_interval_ms > 0 ? s->p.sample_interval_ms * .001 : s->dt;
}
void orc2d_layout(void *h, int64_t *o) {
    orc2d *s = h;
    o[0] = s->nx, o[1] = s->nz, o[2] = s->offset_x, o[3] = s->wx, o[4] = s->wz, o[5] = s->fields, o[6] = s->cpml;
    o[7] = s->nrec, o[8] = (int64_t)s->rec_len, o[9] = s->mech;
}
void orc2d_diagnostics(void *h, double *o) {
    orc2d *s = h;
    const double v[] = {s->energy,     s->peak,          s->max_velocity, s->traction,  s->finite,    s->wave_dt,
                        s->damping_dt, s->source_dt,     s->cpml_dt,      s->relaxation_dt, s->ppw,   s->max_q_error,
                        s->max_metric, s->min_jac,       s->max_edge,     s->flatten,   s->weight_sum, s->rms_radius};
    memcpy(o, v, sizeof v);
}
void orc2d_receivers(void *h, double *vx, double *vz, double *pos, int32_t *xz) {
    orc2d *s = h;
    for (int r = 0; r < s->nrec; ++r) {
        if (vx) memcpy(vx + (size_t)r * s->rec_len, s->rec_vx + (size_t)r * s->rec_cap, sizeof(double) * s->rec_len);
        if (vz) memcpy(vz + (size_t)r * s->rec_len, s->rec_vz + (size_t)r * s->rec_cap, sizeof(double) * s->rec_len);
        if (pos) pos[2 * r] = s->rec_pos[2 * r], pos[2 * r + 1] = s->rec_pos[2 * r + 1];
        if (xz) xz[2 * r] = s->rec_x[r], xz[2 * r + 1] = s->rec_z[r];
    }
}
void orc2d_probe_indices(void *h, int64_t *out) {
    orc2d *s = h;
    memcpy(out, s->probe, sizeof(int64_t) * s->nrec);
}
void orc2d_field(void *h, int f, double *out) {
    orc2d *s = h;
    if (f < 0 || f >= s->fields) return;
    if (s->prec) {
        state_t_f32 *st = s->state;
        for (size_t q = 0; q < s->count; ++q) out[q] = st->f[0][f][q];
    } else {
        state_t_f64 *st = s->state;
        memcpy(out, st->f[0][f], sizeof(double) * s->count);
    }
}
void orc2d_pml(void *h, double *out) {
    orc2d *s = h;
    if (s->prec) {
        state_t_f32 *st = s->state;
        for (int64_t q = 0; q < s->cpml; ++q) out[q] = st->p[0][q];
    } else if (s->cpml) {
        state_t_f64 *st = s->state;
        memcpy(out, st->p[0], sizeof(double) * (size_t)s->cpml);
    }
}
int orc2d_model(void *h, int which, double *out) {
    orc2d *s = h;
    const double *src = NULL;
    size_t n = s->count;
    switch (which) {
    case 0: src = s->a; break;
    case 1: src = s->b; break;
    case 2: src = s->jac; break;
    case 3: src = s->lam; break;
    case 4: src = s->mu; break;
    case 5: src = s->rho; break;
    case 6: src = s->src; break;
    case 7: src = s->damp; break;
    case 8: src = s->grad[0]; break;
    case 9: src = s->grad[1]; break;
    case 10: src = s->nrm, n = 2 * s->count; break;
    case 11: src = s->coord, n = 2 * s->count; break;
    case 12: src = s->vp; break;
    case 13: src = s->vs; break;
    case 14: src = s->cx, n = 3 * (size_t)s->p.nx; break;
    case 15: src = s->cz, n = 3 * (size_t)s->p.nz; break;
    default:
        if (which >= 16 && which < 19) src = s->dP[which - 16];
        else if (which >= 19 && which < 22) src = s->dS[which - 19];
    }
    if (!src) return 0;
    memcpy(out, src, sizeof(double) * n);
    return 1;
}
size_t orc2d_params_size(void) { return sizeof(orc_params2d); }
