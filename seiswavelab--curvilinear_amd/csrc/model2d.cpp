// model2d.cpp — FP64 2D session construction (see model2d.hpp).
#include "model2d.hpp"
#include <algorithm>
#include <cmath>
#include <complex>
#include <limits>
#include <stdexcept>

namespace swl {
namespace {
constexpr double kPi = 3.14159265358979323846;
void need(bool ok, const char *message) {
    if (!ok) throw std::runtime_error(message);
}
bool validMaterial(double vp, double vs, double rho) {
    return std::isfinite(vp) && std::isfinite(vs) && std::isfinite(rho) && vp > 0 && vs > 0 && rho > 0 &&
           vp * vp > 4 * vs * vs / 3 && std::isfinite(rho * vp * vp);
}
// FiniteDifference-independent helpers --------------------------------------------------------
std::array<double, 3> nnlsStrengths(double q, const GslsFit &f, double fmin, double fmax) {
    // exhaustive active-set NNLS over 81 log-spaced samples of 1/Q (Gsls.cpp:11-27)
    constexpr int samples = 81;
    double design[samples][3]{};
    const double target = 1 / q;
    for (int i = 0; i < samples; ++i) {
        const double w = 2 * kPi * fmin * std::pow(fmax / fmin, double(i) / (samples - 1));
        for (int l = 0; l < f.mechanisms; ++l) {
            const double x = w / f.rates[l];
            design[i][l] = (x - target * x * x) / (1 + x * x);
        }
    }
    std::array<double, 3> best{};
    double bestCost = std::numeric_limits<double>::infinity();
    for (int mask = 1; mask < (1 << f.mechanisms); ++mask) {
        int ids[3]{}, n = 0;
        for (int l = 0; l < f.mechanisms; ++l)
            if (mask & (1 << l)) ids[n++] = l;
        double g[3][4]{};
        for (int i = 0; i < n; ++i)
            for (int row = 0; row < samples; ++row) {
                g[i][n] += design[row][ids[i]] * target;
                for (int j = 0; j < n; ++j) g[i][j] += design[row][ids[i]] * design[row][ids[j]];
            }
        bool valid = true;
        for (int k = 0; k < n && valid; ++k) {
            int pivot = k;
            for (int i = k + 1; i < n; ++i)
                if (std::abs(g[i][k]) > std::abs(g[pivot][k])) pivot = i;
            if (std::abs(g[pivot][k]) < 1e-20) {
                valid = false;
                break;
            }
            for (int j = 0; j <= n; ++j) std::swap(g[k][j], g[pivot][j]);
            const double d = g[k][k];
            for (int j = k; j <= n; ++j) g[k][j] /= d;
            for (int i = 0; i < n; ++i)
                if (i != k) {
                    const double s = g[i][k];
                    for (int j = k; j <= n; ++j) g[i][j] -= s * g[k][j];
                }
        }
        std::array<double, 3> cand{};
        for (int i = 0; i < n && valid; ++i) {
            valid = std::isfinite(g[i][n]) && g[i][n] >= 0;
            cand[ids[i]] = g[i][n];
        }
        if (!valid) continue;
        double cost = 0;
        for (int row = 0; row < samples; ++row) {
            double r = -target;
            for (int l = 0; l < f.mechanisms; ++l) r += design[row][l] * cand[l];
            cost += r * r;
        }
        if (cost < bestCost) bestCost = cost, best = cand;
    }
    need(std::isfinite(bestCost), "No passive GSLS fit for this Q/frequency band.");
    return best;
}
double velocityScale(const GslsFit &f, const std::array<double, 3> &weights, double fref) {
    const std::complex<double> w{0, 2 * kPi * fref};
    std::complex<double> shape = 1.;
    for (int l = 0; l < f.mechanisms; ++l) shape += weights[l] * w / (w + f.rates[l]);
    return std::real(1. / std::sqrt(shape));
}
double relaxedModulus(double inv, double rho, double v) { return rho * v * v * inv * inv; }
double qualityFactor(const GslsFit &f, const GslsMaterial &m, bool shear, double freq) {
    const std::complex<double> s{0, 2 * kPi * freq};
    std::complex<double> c = shear ? m.sRelaxed : m.pRelaxed;
    for (int l = 0; l < f.mechanisms; ++l) c += (shear ? m.deltaS[l] : m.deltaP[l]) * s / (s + f.rates[l]);
    return c.real() / std::abs(c.imag());
}
double hypot3(double x, double y, double z) {
    x = std::abs(x), y = std::abs(y), z = std::abs(z);
    const double a = x < y ? (y < z ? z : y) : (x < z ? z : x);
    if (a == 0) return 0;
    return a * std::sqrt((x / a) * (x / a) + (y / a) * (y / a) + (z / a) * (z / a));
}
} // namespace

double lerpExact(double a, double b, double t) {
    if ((a <= 0 && b >= 0) || (a >= 0 && b <= 0)) return t * b + (1 - t) * a;
    if (t == 1) return b;
    const double x = a + t * (b - a);
    return (t > 1) == (b > a) ? (b < x ? x : b) : (b > x ? x : b);
}

GslsFit gslsFit(int mechanisms, double fmin, double fmax, double qp, double qs, double fref) {
    GslsFit f;
    f.mechanisms = mechanisms;
    for (int l = 0; l < mechanisms; ++l)
        f.rates[l] = 2 * kPi *
                     (mechanisms == 1 ? std::sqrt(fmin * fmax) : fmin * std::pow(fmax / fmin, double(l) / (mechanisms - 1)));
    f.strengthP = nnlsStrengths(qp, f, fmin, fmax);
    f.strengthS = nnlsStrengths(qs, f, fmin, fmax);
    f.scaleFref = fref, f.invP = velocityScale(f, f.strengthP, fref), f.invS = velocityScale(f, f.strengthS, fref);
    return f;
}

GslsMaterial gslsMaterial(const GslsFit &f, double fref, double fmin, double fmax, double vp, double vs, double rho,
                          double qp, double qs, bool qError) {
    GslsMaterial m;
    const bool cached = f.scaleFref == fref;
    m.pRelaxed = relaxedModulus(cached ? f.invP : velocityScale(f, f.strengthP, fref), rho, vp);
    m.sRelaxed = relaxedModulus(cached ? f.invS : velocityScale(f, f.strengthS, fref), rho, vs);
    m.pUnrelaxed = m.pRelaxed, m.sUnrelaxed = m.sRelaxed;
    need(m.pRelaxed > 4 * m.sRelaxed / 3, "Qp/Qs produce nonpositive relaxed bulk modulus.");
    for (int l = 0; l < f.mechanisms; ++l) {
        m.deltaP[l] = m.pRelaxed * f.strengthP[l];
        m.deltaS[l] = m.sRelaxed * f.strengthS[l];
        need(m.deltaP[l] + 1e-12 * m.pRelaxed >= 4 * m.deltaS[l] / 3,
             "Qp/Qs violate passive bulk relaxation (Qp cannot be arbitrarily large relative to Qs).");
        m.pUnrelaxed += m.deltaP[l];
        m.sUnrelaxed += m.deltaS[l];
    }
    need(std::isfinite(m.pUnrelaxed) && std::isfinite(m.sUnrelaxed), "GSLS moduli overflow.");
    if (qError)
        for (int i = 0; i <= 160; ++i) {
            const double fr = fmin * std::pow(fmax / fmin, i / 160.);
            m.maxRelativeQError = std::max({m.maxRelativeQError, std::abs(qualityFactor(f, m, false, fr) / qp - 1),
                                            std::abs(qualityFactor(f, m, true, fr) / qs - 1)});
        }
    return m;
}

Pchip::Pchip(const std::vector<std::array<double, 2>> &points) : p_(points) {
    need(points.size() >= 2 && points.size() <= 1000000, "Surface profile requires 2..1000000 samples.");
    const size_t n = points.size();
    std::vector<double> h(n - 1), d(n - 1);
    for (size_t i = 0; i < n; ++i) {
        need(std::isfinite(points[i][0]) && std::isfinite(points[i][1]) && (!i || points[i][0] > points[i - 1][0]),
             "Surface samples must be finite with strictly increasing x.");
        if (i) {
            h[i - 1] = points[i][0] - points[i - 1][0];
            d[i - 1] = (points[i][1] - points[i - 1][1]) / h[i - 1];
            need(std::isfinite(h[i - 1]) && std::isfinite(d[i - 1]), "Surface sample differences overflow.");
        }
    }
    m_.assign(n, 0);
    m_.front() = d.front();
    m_.back() = d.back();
    if (n == 2) return;
    const auto end = [](double h0, double h1, double d0, double d1) {
        const double m = ((2 * h0 + h1) * d0 - h0 * d1) / (h0 + h1);
        if (m * d0 <= 0) return 0.0;
        if (d0 * d1 < 0 && std::abs(m) > 3 * std::abs(d0)) return 3 * d0;
        return m;
    };
    m_.front() = end(h[0], h[1], d[0], d[1]);
    m_.back() = end(h[n - 2], h[n - 3], d[n - 2], d[n - 3]);
    for (size_t i = 1; i + 1 < n; ++i) {
        if (d[i - 1] * d[i] <= 0) m_[i] = 0;
        else {
            const double a = 2 * h[i] + h[i - 1], b = h[i] + 2 * h[i - 1];
            m_[i] = (a + b) / (a / d[i - 1] + b / d[i]);
        }
    }
    for (double s : m_) need(std::isfinite(s), "Surface interpolation coefficients overflow.");
}

std::array<double, 2> Pchip::sample(double x) const {
    if (x <= p_.front()[0]) return {p_.front()[1] + (x - p_.front()[0]) * m_.front(), m_.front()};
    if (x >= p_.back()[0]) return {p_.back()[1] + (x - p_.back()[0]) * m_.back(), m_.back()};
    const auto upper = std::upper_bound(p_.begin(), p_.end(), x, [](double v, const auto &pt) { return v < pt[0]; });
    const size_t i = static_cast<size_t>(upper - p_.begin() - 1);
    const double h = p_[i + 1][0] - p_[i][0], t = (x - p_[i][0]) / h;
    const double a = p_[i][1], b = p_[i + 1][1], m = m_[i], n = m_[i + 1];
    return {(2 * t * t * t - 3 * t * t + 1) * a + (t * t * t - 2 * t * t + t) * h * m + (-2 * t * t * t + 3 * t * t) * b +
                (t * t * t - t * t) * h * n,
            (6 * t * t - 6 * t) * (a - b) / h + (3 * t * t - 4 * t + 1) * m + (3 * t * t - 2 * t) * n};
}

// SeisWave::Rheology::validateSettings (src/common/Gsls.cpp:28-33)
void validateGsls(int mechanisms, double fmin, double fmax, double fref, double qp, double qs) {
    need(mechanisms >= 1 && mechanisms <= 3, "GSLS mechanism count must be 1..3.");
    need(std::isfinite(fmin) && std::isfinite(fmax) && std::isfinite(fref) && fmin > 0 && fmax > fmin && fmax <= 1e6 &&
             fmin >= .01 && fref >= fmin && fref <= fmax,
         "GSLS requires 0.01 <= fmin < fmax <= 1e6 Hz and fref within the band.");
    for (double q : {qp, qs})
        need(std::isfinite(q) && q >= 5 && q <= 1e9, "Qp and Qs must be finite and in 5..1e9; disable attenuation for the elastic limit.");
}

namespace {
// domainLayout (CfsPml2D.hpp:10-14): expanded nx, nz, offsetX, widthX, widthZ
std::array<int, 5> domainLayout2D(const swl_params2d &p) {
    const int wx = std::clamp(p.sponge_x, 1, std::max(1, p.nx / 3)), wz = std::clamp(p.sponge_z, 1, std::max(1, p.nz / 3));
    if (p.absorbing && p.external_absorption) return {p.nx + 2 * wx, p.nz + wz, wx, wx, wz};
    return {p.nx, p.nz, 0, wx, wz};
}
// cpmlScalars (CfsPml2D.hpp:25-29)
double cpmlScalars2D(const swl_params2d &p) {
    if (!(p.absorbing && p.absorber == 1)) return 0;
    const auto d = domainLayout2D(p);
    return 4.0 * (2.0 * d[3] * d[1] + static_cast<double>(d[4]) * d[0]);
}
// setUniformSurfaceReceivers (SurfaceSurvey2D.hpp:8-25), surface depth 0
bool uniformSurfaceReceivers(swl_params2d &p, int limit, int sideMargin) {
    if (p.nx < 2 || p.nx > 16384 || p.nz < 2 || limit < 1) return false;
    const int margin = std::clamp(sideMargin, 0, (p.nx - 2) / 2), span = p.nx - 1 - 2 * margin;
    const int most = std::min(limit, span + 1), wanted = std::clamp(p.receiver_count, 1, most);
    int count = 1, stride = 1, start = margin + span / 2;
    if (wanted > 1) {
        const int minimumStride = std::max(1, (span + most - 2) / (most - 1));
        stride = std::max(minimumStride, std::max(1, static_cast<int>(std::lround(static_cast<double>(span) / (wanted - 1)))));
        count = span / stride + 1;
        start = margin + (span - (count - 1) * stride) / 2;
    }
    p.receiver_start = start, p.receiver_count = count, p.receiver_stride = stride, p.receiver_depth = 0;
    return true;
}
} // namespace

bool synchronizeSurfaceReceivers(swl_params2d &p, int limit, int margin) {
    if (!p.auto_surface_receivers) return true;
    if (p.nx < 2 || p.nx > 16384 || p.nz < 2 || p.nz > 16384 || p.receiver_target_count < 0 || p.receiver_target_count > limit)
        return false;
    const int target = p.receiver_target_count > 0 ? p.receiver_target_count : p.receiver_count;
    if (target < 1 || target > limit) return false;
    const int old = p.receiver_count;
    p.receiver_count = target;
    if (!uniformSurfaceReceivers(p, limit, margin)) {
        p.receiver_count = old;
        return false;
    }
    p.receiver_target_count = target;
    return true;
}

MemoryBudget2D memoryBudget2D(const swl_params2d &p, int backend) {
    const auto d = domainLayout2D(p);
    const double count = static_cast<double>(d[0] + 6) * (d[1] + 6);
    const bool device = backend != 0; // the RK state lives on the accelerator (CUDA in the reference, HIP here)
    const double scalar = p.precision == 1 ? 4.0 : 8.0;
    const double gradients = p.source_kind == 1 ? 2 : 0;
    const double samples = p.sample_interval_ms > 0 ? std::floor(p.duration / (p.sample_interval_ms * .001)) + 1
                                                    : (p.max_steps ? p.max_steps + 1 : 200000 + 1);
    const double records = p.receiver_count * samples * 2 * sizeof(double);
    const double qExtra = p.attenuation ? 4.0 + 2.0 * p.mechanisms : 0;
    const double stateExtra = p.attenuation ? 7.0 * (1 + 4 * p.mechanisms) : 0;
    const double pmlState = 7. * cpmlScalars2D(p) * scalar;
    const double pmlProfiles = p.absorbing && p.absorber == 1 ? 3. * (d[0] + d[1]) * scalar : 0;
    const double hostPerPoint = std::max(((device ? 18. : 60.) + gradients + qExtra + (device ? 0 : stateExtra)) * scalar,
                                         (14. + gradients + qExtra) * sizeof(double) + 2 * scalar);
    MemoryBudget2D b;
    b.host = pmlProfiles + (device ? 0 : pmlState) + count * hostPerPoint + records;
    if (backend == 1)
        b.device = pmlState + pmlProfiles + count * (48. + gradients + (p.attenuation ? 2.0 * p.mechanisms : 0) + stateExtra) * scalar;
    else if (backend == 2)
        b.device = hipDeviceBytes2D(p);
    return b;
}

double hipDeviceBytes2D(const swl_params2d &p, int rows) {
    // the allocations of Engine2D (engine2d_host.inl) for this session's padded grid
    const auto d = domainLayout2D(p);
    const double nx = d[0] + 6, nz = d[1] + 6, nodes = static_cast<double>(d[0]) * d[1];
    const double word = p.precision == 1 ? 4 : 8;
    const int mech = p.attenuation ? p.mechanisms : 0, fields = mech ? 6 + 4 * mech : 5;
    const double ld = std::ceil((nx + 61) / 64) * 64, count = std::ceil((61 + ld * nz) / 64) * 64;
    const bool sponge = p.absorbing && p.absorber == 0, explosive = p.source_kind == 1;
    const double cpml = cpmlScalars2D(p), nrec = std::max(p.receiver_count, 1);
    const double tiles = std::ceil(d[0] / 64.0) * std::ceil(d[1] / 4.0);
    if (p.max_steps > 0) rows = std::min(rows, p.max_steps); // a run never reserves more rows than it has steps
    double bytes = 5.0 * fields * count * word;                                                 // 4 rotating states + accumulator
    bytes += count * word * (7 + 2 + (sponge ? 1 : 0) + (explosive ? 2 : 0) + 2 * mech);     // mesh arrays (normals: 2)
    bytes += 8.0 * (6 * nx + 2 * nx + 2 * nz);                                                // FP64 surface row, terrain profiles
    if (cpml) bytes += 5.0 * cpml * word + 3.0 * (d[0] + d[1]) * word * (word == 4 ? 2 : 1);  // memories, profiles
    bytes += (3.0 + 2 * mech) * nz * word;                                                     // per-row material tables
    bytes += 8.0 * nodes;                                                                      // border-node list (bound)
    bytes += 32.0 * (tiles + std::ceil(nodes / 256) + 1) + 32.0 * 128;                         // inspection partials
    bytes += 4.0 * nodes;                                                                      // display frame
    bytes += nrec * (16 + 40) + 16.0 + rows * (4 + 2 * nrec) * 8.0;                            // probes, gather, run rows
    return bytes;
}

Model2D buildModel2D(const swl_params2d &requested, int device) {
    Model2D s;
    swl_params2d p = requested;
    // validation, CurvilinearSolver2D.cpp:161-227
    need(p.precision == 0 || p.precision == 1, "Unknown 2D calculation precision.");
    need(p.nx >= 8 && p.nz >= 8 && p.nx <= 16384 && p.nz <= 16384, "2D dimensions must be in 8..16384 (subject to RAM/VRAM).");
    need(p.receiver_target_count >= 0 && p.receiver_target_count <= 1024, "Desired surface receiver count must be 0..1024.");
    if (p.auto_surface_receivers) { // CurvilinearSolver2D.cpp:164-167: the layout becomes part of the session's parameters
        const int margin = p.absorbing && !p.external_absorption ? std::clamp(p.sponge_x, 1, std::max(1, p.nx / 3)) : 0;
        need(synchronizeSurfaceReceivers(p, 1024, margin), "Invalid automatic surface receiver layout.");
    }
    const swl_params2d logical = p;
    for (double v : {p.spacing, p.relief, p.wavelength, p.flatten_depth, p.frequency, p.duration, p.cfl,
                     p.force_per_length, p.moment_per_length, p.sponge_strength, p.surface_datum, p.sample_interval_ms})
        need(std::isfinite(v), "2D parameters must be finite.");
    const bool explosive = p.source_kind == 1;
    need(p.source_kind == 0 || p.source_kind == 1, "Unknown source mechanism.");
    need(p.spacing > 0 && p.frequency > 0 && p.frequency <= 1e6 && p.duration > 0 && p.duration <= 10 && p.cfl >= .02 &&
             p.cfl <= .8 && p.wavelength > 0 && p.flatten_depth >= 0 &&
             (explosive ? p.moment_per_length : p.force_per_length) > 0,
         "Invalid 2D spacing, frequency, duration, CFL or source amplitude.");
    need(validMaterial(p.vp, p.vs, p.density), "Require positive material values and positive elastic bulk modulus.");
    need(p.sample_interval_ms == 0 || (p.sample_interval_ms >= .01 && p.sample_interval_ms <= 1000),
         "Record interval must be 0 (each step) or 0.01..1000 ms.");
    need(p.max_steps >= 0 && p.max_steps <= 200000 && p.threads >= 0 && p.threads <= 64,
         "Step cap must be 0..200000 (0 = full duration); CPU workers must be 0..64.");
    if (!explosive) {
        const double norm = hypot3(p.source_direction[0], p.source_direction[1], p.source_direction[2]);
        need(std::isfinite(norm) && norm > 0 && p.source_direction[1] == 0,
             "2D force direction must be finite, nonzero and inside the X-Z plane.");
    }
    need(std::isfinite(p.source_sigma_m) && std::isfinite(p.source_radius_m) &&
             ((p.source_sigma_m == 0 && p.source_radius_m == 0) ||
              (!explosive && p.source_sigma_m > 0 && p.source_radius_m >= p.source_sigma_m)),
         "Physical source kernel requires a directional force and 0 < sigma <= radius, or both zero for legacy.");
    int sx = p.source_x < 0 ? p.nx / 2 : p.source_x;
    if (explosive)
        need(p.nx >= 11 && p.nz >= 11 && sx >= 5 && sx <= p.nx - 6 && p.source_z >= 5 && p.source_z <= p.nz - 6,
             "Isotropic moment needs 11+ nodes per axis and source center five nodes inside each face.");
    need(p.source_x >= -1 && sx >= 2 && sx <= p.nx - 3 && p.source_z >= 0 && p.source_z <= p.nz - 3,
         "2D force source may be on the free surface; keep two cells from the sides and bottom.");
    need(p.receiver_start >= 0 && p.receiver_count > 0 && p.receiver_count <= 1024 && p.receiver_stride > 0 &&
             p.receiver_depth >= 0 && p.receiver_depth < p.nz &&
             static_cast<long long>(p.receiver_start) + (static_cast<long long>(p.receiver_count) - 1) * p.receiver_stride <
                 p.nx,
         "Receiver line is outside the 2D model or exceeds 1024 positions.");
    need(p.sponge_x >= 1 && p.sponge_x <= 512 && p.sponge_z >= 1 && p.sponge_z <= 512 && p.sponge_strength >= .1 &&
             p.sponge_strength <= 10,
         "Sponge widths must be 1..512 cells (effective width <= axis nodes / 3) and strength 0.1..10.");
    need(p.absorber == 0 || p.absorber == 1, "Unknown 2D absorber.");
    need(std::isfinite(p.pml_reflection) && p.pml_reflection >= 1e-8 && p.pml_reflection <= .1 &&
             std::isfinite(p.pml_kappa_max) && p.pml_kappa_max >= 1 && p.pml_kappa_max <= 20 &&
             std::isfinite(p.pml_frequency) && p.pml_frequency >= 0 && p.pml_frequency <= 1e6,
         "CFS-PML requires R in 1e-8..0.1, kappa in 1..20, frequency in 0..1e6 Hz.");
    if (p.layered) {
        need(p.n_layers >= 1 && p.n_layers <= SWL_MAX_LAYERS_2D, "Require 1..16 material layers.");
        for (int l = 0; l < p.n_layers; ++l)
            need(validMaterial(p.layer_vp[l], p.layer_vs[l], p.layer_density[l]) && std::isfinite(p.layer_thickness[l]) &&
                     p.layer_thickness[l] >= 0 && (l + 1 == p.n_layers || p.layer_thickness[l] > 0),
                 "Invalid layer thickness or material.");
    }
    const int materials = p.layered ? p.n_layers : 1;
    std::vector<GslsFit> fits;
    std::vector<GslsMaterial> rheology;
    if (p.attenuation) {
        validateGsls(p.mechanisms, p.q_min_frequency, p.q_max_frequency, p.q_reference_frequency, p.qp, p.qs);
        for (int l = 0; l < materials; ++l) {
            const double qp = p.layered ? p.layer_qp[l] : p.qp, qs = p.layered ? p.layer_qs[l] : p.qs;
            const double vp = p.layered ? p.layer_vp[l] : p.vp, vs = p.layered ? p.layer_vs[l] : p.vs,
                         rho = p.layered ? p.layer_density[l] : p.density;
            need(std::isfinite(vp) && std::isfinite(vs) && std::isfinite(rho) && rho > 0 && vs > 0 && vp * vp > 4 * vs * vs / 3,
                 "Invalid GSLS reference velocities/density.");
            for (double q : {qp, qs}) need(std::isfinite(q) && q >= 5 && q <= 1e9, "Layer Qp/Qs must be in 5..1e9.");
            fits.push_back(gslsFit(p.mechanisms, p.q_min_frequency, p.q_max_frequency, qp, qs, p.q_reference_frequency));
            rheology.push_back(gslsMaterial(fits.back(), p.q_reference_frequency, p.q_min_frequency, p.q_max_frequency, vp,
                                            vs, rho, qp, qs, true));
        }
    }
    if (p.node_vp || p.node_vs || p.node_density)
        need(p.node_vp && p.node_vs && p.node_density, "Heterogeneous models need Vp, Vs and density per node.");
    // domain layout (CfsPml2D.hpp:10-14)
    s.widthX = std::clamp(p.sponge_x, 1, std::max(1, p.nx / 3));
    s.widthZ = std::clamp(p.sponge_z, 1, std::max(1, p.nz / 3));
    int dnx = p.nx, dnz = p.nz;
    if (p.absorbing && p.external_absorption) {
        s.offsetX = s.widthX;
        dnx += 2 * s.widthX;
        dnz += s.widthZ;
    }
    need(dnx <= 16384 && dnz <= 16384,
         "2D computational dimensions including external absorbing strips must not exceed 16384 per axis.");
    if (device >= 0) preflightMemory2D(logical, device); // the reference checks here, before anything is built
    p.nx = dnx, p.nz = dnz, p.sponge_x = s.widthX, p.sponge_z = s.widthZ, p.external_absorption = 0;
    sx += s.offsetX;
    p.source_x = sx;
    p.receiver_start += s.offsetX;
    s.p = p;
    s.logical = logical;
    const int gnx = p.nx + 6, gnz = p.nz + 6, off = s.offsetX;
    const size_t count = static_cast<size_t>(gnx) * gnz;
    s.nx = gnx, s.nz = gnz, s.count = count, s.h = p.spacing;
    const double bottom = (logical.nz - 1) * p.spacing;
    std::vector<std::array<double, 2>> surface(gnx);
    double relief = std::abs(p.relief);
    if (p.imported_surface) {
        need(p.n_surface >= 2 && p.surface_xz, "Surface profile requires 2..1000000 samples.");
        std::vector<std::array<double, 2>> pts(p.n_surface);
        for (int i = 0; i < p.n_surface; ++i) pts[i] = {p.surface_xz[2 * i], p.surface_xz[2 * i + 1]};
        const Pchip profile(pts);
        need(pts.front()[0] <= 0 && pts.back()[0] >= (logical.nx - 1) * p.spacing,
             "Imported terrain must cover x=0 through the full physical model span.");
        relief = 0;
        for (int i = 0; i < gnx; ++i) {
            const auto v = profile.sample((i - 3 - off) * p.spacing);
            surface[i] = {p.surface_datum - v[0], -v[1]};
            if (i >= off && i < off + logical.nx + 6) relief = std::max(relief, std::abs(surface[i][0]));
        }
    } else {
        for (int i = 0; i < gnx; ++i) {
            const double phase = 2 * kPi * (i - 3 - off) * p.spacing / p.wavelength;
            surface[i] = {p.relief * std::cos(phase), -p.relief * 2 * kPi / p.wavelength * std::sin(phase)};
        }
    }
    s.flatten = p.flatten_depth > 0 ? p.flatten_depth
                                    : std::min(bottom, std::max({.35 * bottom, 6 * p.spacing, 4 * relief + 2 * p.spacing}));
    need(s.flatten <= bottom && (relief == 0 || ((p.imported_surface || p.wavelength >= 4 * p.spacing) &&
                                                 s.flatten >= 4 * p.spacing && s.flatten > 2.6 * relief)),
         "Terrain must be resolved and flatten depth must exceed 2.6 times relief within the model.");
    if (off) { // C1 continuation of edge height/slope into the external strips (:317-330)
        for (int side = 0; side < 2; ++side) {
            const int edge = 3 + off + (side ? logical.nx - 1 : 0);
            const double z = surface[edge][0], slope = surface[edge][1];
            const double length = std::min(s.widthX * p.spacing / 3.,
                                           .5 * (s.flatten / 2.6 - std::abs(z)) / std::max(std::abs(slope), 1e-30));
            for (int i = side ? edge + 1 : 0; i < (side ? gnx : edge); ++i) {
                const double distance = std::abs(i - edge) * p.spacing, decay = std::exp(-distance / length);
                surface[i] = {z + (side ? 1. : -1.) * slope * length * (1 - decay), slope * decay};
            }
        }
    }
    for (auto *v : {&s.a, &s.b, &s.jac, &s.vp, &s.vs, &s.rho, &s.lambda, &s.mu, &s.source, &s.damping}) v->assign(count, 0);
    s.coordinates.assign(2 * count, 0);
    s.normals.assign(2 * count, 0);
    s.mechanisms = p.attenuation ? p.mechanisms : 0;
    for (int l = 0; l < s.mechanisms; ++l) s.deltaP[l].assign(count, 0), s.deltaS[l].assign(count, 0);
    if (p.attenuation) s.material.assign(count, 0);
    s.minJacobian = std::numeric_limits<double>::max();
    double maxVp = 0, minVs = std::numeric_limits<double>::max(), maxSigma = 0;
    // metric / material pass (:339-392)
    std::vector<double> &fade = s.fade, &fadeDer = s.fadeDerivative;
    fade.assign(gnz, 0), fadeDer.assign(gnz, 0);
    for (int k = 0; k < gnz; ++k) {
        const double eta = (k - 3) * p.spacing, t = std::clamp(eta / s.flatten, 0.0, 1.0);
        fade[k] = 1 - std::pow(t, 5) * (126 + t * (-420 + t * (540 + t * (-315 + 70 * t))));
        fadeDer[k] = eta > 0 && eta < s.flatten ? -630 * std::pow(t, 4) * std::pow(1 - t, 4) / s.flatten : 0;
    }
    s.surfaceDepth.resize(gnx), s.surfaceSlope.resize(gnx);
    for (int i = 0; i < gnx; ++i) s.surfaceDepth[i] = surface[i][0], s.surfaceSlope[i] = surface[i][1];
    // device sessions: coordinates, J, a, b and their CFL figures (min J, max metric) in FP64 on the GPU; the
    // first node with a bad Jacobian is reported here when the pass below reaches it
    const bool onDevice = device >= 0;
    const long long badNode = onDevice ? buildGeometry2DOnDevice(s, device) : -1;
    for (int k = 0; k < gnz; ++k)
        for (int i = 0; i < gnx; ++i) {
            const size_t q = static_cast<size_t>(i) + static_cast<size_t>(gnx) * k;
            const double x = (i - 3 - off) * p.spacing, eta = (k - 3) * p.spacing;
            const double zx = surface[i][1] * fade[k], j = 1 + surface[i][0] * fadeDer[k];
            if (onDevice) {
                need(static_cast<long long>(q) != badNode, "Nonpositive curvilinear Jacobian.");
            } else {
                need(j > 1e-9 && std::isfinite(j), "Nonpositive curvilinear Jacobian.");
                s.coordinates[2 * q] = x;
                s.coordinates[2 * q + 1] = eta + surface[i][0] * fade[k];
                s.a[q] = -zx / j;
                s.b[q] = 1 / j;
                s.jac[q] = j;
            }
            const double n = std::hypot(zx, 1.0);
            s.normals[2 * q] = -zx / n;
            s.normals[2 * q + 1] = 1 / n;
            double vp = p.vp, vs = p.vs, rho = p.density;
            int material = 0;
            if (p.layered) {
                double boundary = 0;
                const double depth = off ? std::clamp(eta, 0., bottom) : std::max(0., eta);
                for (int l = 0; l < p.n_layers; ++l) {
                    boundary += p.layer_thickness[l];
                    if (l + 1 == p.n_layers || depth < boundary) {
                        vp = p.layer_vp[l], vs = p.layer_vs[l], rho = p.layer_density[l], material = l;
                        break;
                    }
                }
            }
            if (p.node_vp) {
                const int li = std::clamp(i - 3 - off, 0, logical.nx - 1), lk = std::clamp(k - 3, 0, logical.nz - 1);
                const size_t lq = static_cast<size_t>(li) + static_cast<size_t>(logical.nx) * lk;
                vp = p.node_vp[lq], vs = p.node_vs[lq], rho = p.node_density[lq];
            }
            s.vp[q] = vp, s.vs[q] = vs, s.rho[q] = rho;
            s.mu[q] = rho * vs * vs;
            s.lambda[q] = rho * vp * vp - 2 * s.mu[q];
            double speedMax = vp, speedMin = vs;
            if (p.attenuation) {
                GslsMaterial m = rheology[material];
                if (p.node_vp) {
                    const double qp = p.layered ? p.layer_qp[material] : p.qp, qs = p.layered ? p.layer_qs[material] : p.qs;
                    m = gslsMaterial(fits[material], p.q_reference_frequency, p.q_min_frequency, p.q_max_frequency, vp, vs,
                                     rho, qp, qs, false);
                }
                s.mu[q] = m.sUnrelaxed;
                s.lambda[q] = m.pUnrelaxed - 2 * m.sUnrelaxed;
                s.material[q] = static_cast<uint8_t>(material);
                for (int l = 0; l < s.mechanisms; ++l) s.deltaP[l][q] = m.deltaP[l], s.deltaS[l][q] = m.deltaS[l];
                speedMax = std::sqrt(m.pUnrelaxed / rho);
                speedMin = std::sqrt(m.sRelaxed / rho);
            }
            if (i >= 3 && i < gnx - 3 && k >= 3 && k < gnz - 3) {
                if (!onDevice) {
                    s.minJacobian = std::min(s.minJacobian, j);
                    s.maxMetric = std::max(s.maxMetric, std::sqrt(1 + s.a[q] * s.a[q] + s.b[q] * s.b[q]));
                }
                maxVp = std::max(maxVp, speedMax);
                minVs = std::min(minVs, speedMin);
            }
        }
    const bool sponge = p.absorbing && p.absorber == 0;
    for (int k = 0; k < p.nz; ++k)
        for (int i = 0; i < p.nx; ++i) {
            const size_t q = static_cast<size_t>(i + 3) + static_cast<size_t>(gnx) * (k + 3);
            if (i + 1 < p.nx)
                s.maxEdge = std::max(s.maxEdge, std::hypot(s.coordinates[2 * (q + 1)] - s.coordinates[2 * q],
                                                           s.coordinates[2 * (q + 1) + 1] - s.coordinates[2 * q + 1]));
            if (k + 1 < p.nz) s.maxEdge = std::max(s.maxEdge, s.coordinates[2 * (q + gnx) + 1] - s.coordinates[2 * q + 1]);
            if (sponge) { // (:403-409)
                const int wx = std::min(p.sponge_x, p.nx / 3), wz = std::min(p.sponge_z, p.nz / 3);
                const double x = std::max(0.0, 1.0 - std::min(i, p.nx - 1 - i) / static_cast<double>(wx)),
                             z = std::max(0.0, 1.0 - (p.nz - 1 - k) / static_cast<double>(wz));
                s.damping[q] = p.sponge_strength * maxVp / p.spacing * (x * x / wx + z * z / wz);
                maxSigma = std::max(maxSigma, s.damping[q]);
            }
        }
    double cpmlRate = 0;
    if (p.absorbing && p.absorber == 1) { // CfsPml2D.hpp:31-44
        const int wx = std::min(p.sponge_x, p.nx / 3), wz = std::min(p.sponge_z, p.nz / 3);
        const double freq = p.pml_frequency > 0 ? p.pml_frequency : p.frequency;
        const auto profile = [&](int nodes, int width, bool both) {
            std::vector<double> r(3 * static_cast<size_t>(nodes));
            const double d0 = -3 * maxVp * std::log(p.pml_reflection) / (2 * width * p.spacing);
            for (int i = 0; i < nodes; ++i) {
                const int distance = both ? std::min(i, nodes - 1 - i) : nodes - 1 - i;
                const double t = std::clamp(1. - double(distance) / width, 0., 1.);
                r[3 * i] = d0 * t * t;
                r[3 * i + 1] = 1 + (p.pml_kappa_max - 1) * t * t;
                r[3 * i + 2] = t > 0 ? 3.14159265358979323846 * freq * (1 - t) : 0;
                cpmlRate = std::max(cpmlRate, r[3 * i] / r[3 * i + 1] + r[3 * i + 2]);
            }
            return r;
        };
        s.cpmlX = profile(p.nx, wx, true);
        s.cpmlZ = profile(p.nz, wz, false);
        s.cpmlCount = 4 * (2 * static_cast<int64_t>(s.widthX) * p.nz + static_cast<int64_t>(s.widthZ) * p.nx);
    }
    // source weights (:419-442)
    double weight = 0;
    const size_t centre = static_cast<size_t>(sx + 3) + static_cast<size_t>(gnx) * (p.source_z + 3);
    if (p.source_sigma_m > 0) {
        const double cx = s.coordinates[2 * centre], cz = s.coordinates[2 * centre + 1];
        double boundaryDistance = std::numeric_limits<double>::infinity();
        for (int k = 0; k < p.nz; ++k)
            for (int i = 0; i < p.nx; ++i) {
                const size_t q = static_cast<size_t>(i + 3) + static_cast<size_t>(gnx) * (k + 3);
                const double distance = std::hypot(s.coordinates[2 * q] - cx, s.coordinates[2 * q + 1] - cz);
                if (k == 0 || k == p.nz - 1 || i == 0 || i == p.nx - 1)
                    boundaryDistance = std::min(boundaryDistance, distance);
                if (distance <= p.source_radius_m) {
                    s.source[q] = std::exp(-.5 * distance * distance / (p.source_sigma_m * p.source_sigma_m));
                    weight += s.source[q] * s.jac[q];
                }
            }
        need(boundaryDistance > p.source_radius_m, "Physical source support intersects a physical computational boundary.");
        need(weight > 0 && std::isfinite(weight), "Invalid physical source normalization.");
    } else {
        for (int dk = -2; dk <= 2; ++dk)
            for (int di = -2; di <= 2; ++di) {
                if (p.source_z + dk < 0) continue;
                const size_t q = static_cast<size_t>(sx + 3 + di) + static_cast<size_t>(gnx) * (p.source_z + 3 + dk);
                s.source[q] = std::exp(-.5 * (di * di + dk * dk));
                weight += s.source[q] * s.jac[q];
            }
    }
    for (auto &v : s.source) v = v / weight;
    s.explosive = explosive;
    if (explosive) { // conservative FD4 moment pattern (SourceMechanism.hpp:43-56)
        s.gradX.assign(count, 0), s.gradZ.assign(count, 0);
        const auto flux = [&](ptrdiff_t q, int axis, int comp) {
            const double metric = axis == 0 ? (comp == 0 ? 1. : 0.) : (comp == 0 ? s.a[q] : s.b[q]);
            return s.jac[q] * metric * s.source[q];
        };
        for (int k = p.source_z - 4; k <= p.source_z + 4; ++k)
            for (int i = sx - 4; i <= sx + 4; ++i) {
                const ptrdiff_t q = static_cast<ptrdiff_t>(i + 3) + static_cast<ptrdiff_t>(gnx) * (k + 3);
                double force[2]{};
                for (int c = 0; c < 2; ++c)
                    for (int ax = 0; ax < 2; ++ax) {
                        const ptrdiff_t st = ax == 0 ? 1 : gnx;
                        force[c] -= (flux(q - 2 * st, ax, c) - 8 * flux(q - st, ax, c) + 8 * flux(q + st, ax, c) -
                                     flux(q + 2 * st, ax, c)) /
                                    (12 * p.spacing * s.jac[q]);
                    }
                s.gradX[q] = force[0], s.gradZ[q] = force[1];
            }
    }
    if (p.source_sigma_m > 0 || p.source_z < 2) {
        double second = 0;
        for (size_t q = 0; q < count; ++q)
            if (s.source[q] > 0) {
                const double w = s.source[q] * s.jac[q], dx = s.coordinates[2 * q] - s.coordinates[2 * centre],
                             dz = s.coordinates[2 * q + 1] - s.coordinates[2 * centre + 1];
                s.sourceWeightSum += w;
                second += w * (dx * dx + dz * dz);
            }
        s.sourceRmsRadius = std::sqrt(second / s.sourceWeightSum);
    }
    // dt plan (:460-481)
    s.ppw = minVs / (2.5 * p.frequency * s.maxEdge);
    s.waveDt = p.cfl * p.spacing / (maxVp * s.maxMetric);
    s.dampingDt = maxSigma > 0 ? .5 / maxSigma : std::numeric_limits<double>::infinity();
    s.sourceDt = 1 / (50 * p.frequency);
    s.cpmlDt = cpmlRate > 0 ? .5 / cpmlRate : std::numeric_limits<double>::infinity();
    s.relaxationDt = p.attenuation ? .25 / fits.front().rates[p.mechanisms - 1] : std::numeric_limits<double>::infinity();
    s.fields = p.attenuation ? 6 + 4 * p.mechanisms : 5;
    for (const auto &m : rheology) s.maxQError = std::max(s.maxQError, m.maxRelativeQError);
    const double safe = std::min({s.waveDt, s.dampingDt, s.sourceDt, s.relaxationDt, s.cpmlDt});
    const double full = std::ceil(p.duration / safe);
    need(std::isfinite(full) && full <= 10000000, "Requested duration needs too many steps; increase spacing or lower frequency.");
    s.dt = p.duration / full;
    s.total = p.max_steps ? std::min(static_cast<int>(full), p.max_steps) : static_cast<int>(full);
    need(s.total <= 200000,
         "2D session exceeds 200000 accepted steps; set an explicit cap or shorter duration. A cap ends before the requested duration.");
    // receivers (:482-497)
    const double interval = p.sample_interval_ms > 0 ? p.sample_interval_ms * .001 : s.dt;
    s.recordReserve = static_cast<size_t>(std::floor(s.total * s.dt / interval)) + 1;
    for (int r = 0; r < p.receiver_count; ++r) {
        const int x = logical.receiver_start + r * p.receiver_stride, z = p.receiver_depth;
        const size_t q = static_cast<size_t>(x + 3 + off) + static_cast<size_t>(gnx) * (z + 3);
        s.receiverX.push_back(x), s.receiverZ.push_back(z);
        s.receiverPosition.push_back(s.coordinates[2 * q]);
        s.receiverPosition.push_back(s.coordinates[2 * q + 1]);
        s.probeIndex.push_back(static_cast<int64_t>(q));
    }
    if (explosive) s.direction = {0, 0, 0};
    else {
        const double n = hypot3(p.source_direction[0], p.source_direction[1], p.source_direction[2]);
        for (int c = 0; c < 3; ++c) s.direction[c] = p.source_direction[c] / n;
    }
    if (p.attenuation)
        for (int l = 0; l < p.mechanisms; ++l) s.rates[l] = fits.front().rates[l];
    s.amplitude = explosive ? p.moment_per_length : p.force_per_length;
    // quantised validation in the run precision (:498-503)
    const bool fp32 = p.precision == 1;
    const auto quant = [&](double v) { return fp32 ? static_cast<double>(static_cast<float>(v)) : v; };
    if (fp32) {
        const auto finite32 = [](const std::vector<double> &v) {
            for (double x : v)
                if (!std::isfinite(static_cast<float>(x))) return false;
            return true;
        };
        bool ok = true;
        for (const auto *v : {&s.coordinates, &s.normals, &s.a, &s.b, &s.jac, &s.vp, &s.vs, &s.rho, &s.lambda, &s.mu,
                              &s.source, &s.damping, &s.cpmlX, &s.cpmlZ, &s.gradX, &s.gradZ})
            ok = ok && finite32(*v);
        for (int l = 0; l < s.mechanisms; ++l) ok = ok && finite32(s.deltaP[l]) && finite32(s.deltaS[l]);
        need(ok, "2D FP32 model/state coefficient overflow; select FP64 or revise model.");
    }
    for (size_t q = 0; q < count; ++q)
        need(quant(s.mu[q]) > 0 && quant(s.lambda[q]) + 2 * quant(s.mu[q]) / 3 > 0 && quant(s.rho[q]) > 0 && quant(s.vp[q]) > 0 &&
                 quant(s.vs[q]) > 0 && quant(s.jac[q]) > 0 && quant(s.b[q]) > 0,
             "Invalid quantized material or geometry; select FP64 or revise model.");
    if (fp32)
        need(std::isfinite(static_cast<float>(p.spacing)) && static_cast<float>(p.spacing) > 0 &&
                 std::isfinite(static_cast<float>(s.dt)) && static_cast<float>(s.dt) > 0,
             "FP32 spacing/time step is not representable; select FP64.");
    return s;
}

swl_mesh2d_desc meshDescriptor(const Model2D &m) {
    swl_mesh2d_desc d{};
    d.precision = m.p.precision;
    d.nx = m.nx, d.nz = m.nz, d.h = m.h;
    d.a = m.a.data(), d.b = m.b.data(), d.jac = m.jac.data(), d.normals = m.normals.data();
    d.lambda = m.lambda.data(), d.mu = m.mu.data(), d.rho = m.rho.data(), d.source = m.source.data();
    d.damping = m.damping.data();
    for (int c = 0; c < 3; ++c) d.source_direction[c] = m.direction[c];
    if (m.explosive) d.source_gradient[0] = m.gradX.data(), d.source_gradient[1] = m.gradZ.data();
    d.mechanisms = m.mechanisms, d.field_count = m.fields;
    if (m.cpmlCount) {
        d.cpml_x = m.cpmlX.data(), d.cpml_z = m.cpmlZ.data();
        d.cpml_width_x = std::min(m.p.sponge_x, m.p.nx / 3);
        d.cpml_width_z = std::min(m.p.sponge_z, m.p.nz / 3);
        d.cpml_count = m.cpmlCount;
    }
    for (int l = 0; l < 3; ++l) d.rates[l] = m.rates[l];
    for (int l = 0; l < m.mechanisms; ++l) d.delta_p[l] = m.deltaP[l].data(), d.delta_s[l] = m.deltaS[l].data();
    d.surface_depth = m.surfaceDepth.data(), d.surface_slope = m.surfaceSlope.data();
    d.fade = m.fade.data(), d.fade_derivative = m.fadeDerivative.data();
    return d;
}

} // namespace swl
