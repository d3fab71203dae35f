// model3d.cpp — FP64 host construction of a 3D session (validation, geometry, materials,
// sponge, dt plan, sources, receivers).  Follows src/three_d/CurvilinearSolver.cpp:47-54,
// :117-190, :258-442, include/Seis3D/{AbsorbingBoundary3D,TimeStepPlan3D,MaterialModels3D,
// ReceiverGeometry3D,SurfaceGrid3D}.hpp and include/SeisCommon/SourceMechanism.hpp:16-56 in the
// reference's operation order.  Compiled with g++ -ffp-contract=off (no fused multiply-adds).
#include "model3d.hpp"
#include <algorithm>
#include <cmath>
#include <limits>
#include <stdexcept>

namespace swl {
namespace {
constexpr int G = 3;
constexpr double kPi = 3.14159265358979323846;
constexpr int kMaxReceivers = 1024;

[[noreturn]] void fail(const char *m) { throw std::runtime_error(m); }

// std::hypot(x, y, z) of the C++17 library: scaled by the largest magnitude
double hypot3(double x, double y, double z) {
    x = std::fabs(x), y = std::fabs(y), z = std::fabs(z);
    const double a = x < y ? (y < z ? z : y) : (x < z ? z : x);
    if (a == 0) return 0;
    return a * std::sqrt((x / a) * (x / a) + (y / a) * (y / a) + (z / a) * (z / a));
}
using V3 = std::array<double, 3>;
V3 cross(const V3 &a, const V3 &b) { return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]}; }
double dot(const V3 &a, const V3 &b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

bool hasDem(const swl_params3d &p) { return p.dem_nx > 0 && p.dem_elevations != nullptr; }

// SurfaceGrid3D::elevation: bilinear, edge cells extended linearly
double demElevation(const swl_params3d &p, double x, double y) {
    const double u = x / p.dem_spacing, v = y / p.dem_spacing;
    const int i = static_cast<int>(std::floor(std::clamp(u, 0.0, static_cast<double>(p.dem_nx - 2))));
    const int j = static_cast<int>(std::floor(std::clamp(v, 0.0, static_cast<double>(p.dem_ny - 2))));
    const double a = u - i, b = v - j;
    const auto at = [&](int ix, int iy) { return p.dem_elevations[static_cast<size_t>(iy) * p.dem_nx + ix]; };
    return (1 - b) * ((1 - a) * at(i, j) + a * at(i + 1, j)) + b * ((1 - a) * at(i, j + 1) + a * at(i + 1, j + 1));
}
double demDisplacement(const swl_params3d &p, double x, double y) { return p.dem_datum - demElevation(p, x, y); }

// receiverLimits3D / receiverNodes3D: returns the node list or throws
std::vector<std::array<int, 3>> receiverNodes(const swl_params3d &p) {
    const int kind = p.receiver_kind;
    int lx = 0, ly = 0, total = 0, x = 0, y = 0;
    bool ok = !(p.nx < 8 || p.ny < 8 || p.nz < 8 || p.nx > 4096 || p.ny > 4096 || p.nz > 4096 || kind < 0 || kind > 3 ||
                p.receiver_z < 0 || p.receiver_z >= p.nz);
    const auto capacity = [](int n, int start, int stride) { return stride > 0 ? 1 + (n - 1 - start) / stride : 0; };
    if (ok) {
        if (kind == 0) {
            lx = p.nx - 4, ly = 1, total = std::min(kMaxReceivers, p.nx - 4);
        } else {
            x = p.receiver_start_x == -1 ? p.nx / 2 : p.receiver_start_x;
            y = p.receiver_start_y == -1 ? p.ny / 2 : p.receiver_start_y;
            if (x < 0 || x >= p.nx || y < 0 || y >= p.ny) ok = false;
            else {
                lx = kind == 2 ? 1 : capacity(p.nx, x, p.receiver_stride_x);
                ly = kind == 1 ? 1 : capacity(p.ny, y, p.receiver_stride_y);
                total = static_cast<int>(std::min<long long>(kMaxReceivers, static_cast<long long>(lx) * ly));
            }
        }
    }
    if (!ok || !total) fail("Invalid receiver grid, mode, start, active stride or depth index.");
    std::vector<std::array<int, 3>> nodes;
    if (kind == 0) {
        if (p.nx - 4 > kMaxReceivers) fail("Auto X line exceeds 1024 traces; select a custom line with a larger receiver stride.");
        for (int i = 2; i < p.nx - 2; ++i) nodes.push_back({i, p.ny / 2, p.receiver_z});
        return nodes;
    }
    const int cx = kind == 2 ? 1 : p.receiver_count_x, cy = kind == 1 ? 1 : p.receiver_count_y;
    if (cx < 1 || cy < 1 || cx > lx || cy > ly || static_cast<long long>(cx) * cy > kMaxReceivers)
        fail("Receiver counts exceed the start/stride capacity or the 1024-trace total limit. Review capacity or fit counts.");
    for (int j = 0; j < cy; ++j)
        for (int i = 0; i < cx; ++i)
            nodes.push_back({x + i * (kind == 2 ? 0 : p.receiver_stride_x), y + j * (kind == 1 ? 0 : p.receiver_stride_y),
                             p.receiver_z});
    return nodes;
}

// validateSimulationConfiguration3D, CurvilinearSolver.cpp:258-299
void validate(const swl_params3d &p) {
    if (!std::isfinite(p.cfl) || p.cfl < .02 || p.cfl > .8) fail("CFL coefficient must be finite and in 0.02..0.8.");
    if (p.max_steps < 0 || p.max_steps > 20000) fail("Maximum steps must be 0 (duration) or 1..20,000.");
    if (!std::isfinite(p.absorbing_strength) || p.absorbing_strength < 0.1 || p.absorbing_strength > 10.0)
        fail("Sponge strength must be in 0.1..10.");
    for (int a = 0; a < 3; ++a)
        if (p.absorbing_cells[a] < 1 || p.absorbing_cells[a] > 64) fail("Sponge widths must be 1..64 cells.");
    for (double v : {p.spacing, p.vp, p.vs, p.density, p.dominant_frequency, p.duration, p.topography_amplitude,
                     p.topography_wavelength, p.terrain_transition_depth})
        if (!std::isfinite(v)) fail("Simulation parameters must be finite.");
    if (p.nx < 8 || p.ny < 8 || p.nz < 8 || p.nx > 4096 || p.ny > 4096 || p.nz > 4096 || p.spacing <= 0.0 || p.vp <= 0.0 ||
        p.vs <= 0.0 || p.density <= 0.0 || p.vp * p.vp <= 4.0 * p.vs * p.vs / 3.0 || p.dominant_frequency <= 0.0 ||
        p.dominant_frequency > 1.0e6 || p.duration <= 0.0 || p.duration > 10.0 || p.topography_wavelength <= 0.0 ||
        p.terrain_transition_depth < 0.0 || p.threads < 0 || p.threads > 64 ||
        (!hasDem(p) && p.topography_amplitude != 0.0 && p.topography_wavelength < 4.0 * p.spacing))
        fail("Require 8..4096 nodes per axis within the total node/memory budget, positive finite parameters and resolved topography.");
    if (static_cast<size_t>(p.nx + 2 * G) * (p.ny + 2 * G) * (p.nz + 2 * G) > static_cast<size_t>(306) * 306 * 306)
        fail("Grid exceeds the 28,652,616 padded-node total budget; redistribute X/Y/Z nodes.");
    if (p.source_kind != 0 && p.source_kind != 1) fail("Unknown source mechanism.");
    if (p.source_kind == 0) {
        const double norm = hypot3(p.source_direction[0], p.source_direction[1], p.source_direction[2]);
        if (!std::isfinite(norm) || norm <= 0) fail("Force direction must be finite and nonzero.");
    }
    if (!std::isfinite(p.source_force) || !std::isfinite(p.source_moment) ||
        (p.source_kind == 1 ? p.source_moment : p.source_force) <= 0)
        fail("Source amplitudes must be finite and the active force (N) or moment (N*m) must be positive.");
    const int sx = p.source_x < 0 ? p.nx / 2 : p.source_x, sy = p.source_y < 0 ? p.ny / 2 : p.source_y, sz = p.source_z;
    if (p.source_kind == 1 && (p.nx < 11 || p.ny < 11 || p.nz < 11 || sx < 5 || sx > p.nx - 6 || sy < 5 ||
                               sy > p.ny - 6 || sz < 5 || sz > p.nz - 6))
        fail("Isotropic moment needs 11+ nodes per axis and source center five nodes inside each face.");
    if (p.source_x < -1 || p.source_y < -1 || sx < 2 || sx > p.nx - 3 || sy < 2 || sy > p.ny - 3 || sz < 2 || sz > p.nz - 3)
        fail("Source must stay two nodes inside every physical face; X/Y may be -1 for center.");
    if (p.layered) {
        if (p.layer_depth_mode < 0 || p.layer_depth_mode > 2) fail("Invalid layer depth coordinate mode.");
        if (p.n_layers < 2 || p.n_layers > 8) fail("Layered material requires 2..8 layers.");
        for (int i = 0; i < p.n_layers; ++i) {
            const double t = p.layer_thickness[i], vp = p.layer_vp[i], vs = p.layer_vs[i], d = p.layer_density[i];
            if (!std::isfinite(t) || !std::isfinite(vp) || !std::isfinite(vs) || !std::isfinite(d) || vp <= 0 || vs <= 0 ||
                d <= 0 || vp * vp <= 4 * vs * vs / 3 || !std::isfinite(d * vp * vp) || d * vs * vs <= 0 ||
                (i + 1 == p.n_layers ? t != 0 : t <= 0))
                fail("Invalid layer: positive finite Vp/Vs/density and bulk modulus required; last thickness must be zero (auto).");
        }
    }
    (void)receiverNodes(p);
}

// Grid3D::validateGeometry (CurvilinearSolver.cpp:117-145); resolves the automatic transition depth
void validateGeometry(Model3D &m) {
    const auto &p = m.p;
    const double dx = m.h;
    if (m.nx < 8 + 2 * G || m.ny < 8 + 2 * G || m.nz < 8 + 2 * G || m.nx > 4096 + 2 * G || m.ny > 4096 + 2 * G ||
        m.nz > 4096 + 2 * G || !std::isfinite(dx) || dx <= 0.0 || !std::isfinite(p.topography_amplitude) ||
        !std::isfinite(p.topography_wavelength) || p.topography_wavelength <= 0.0 || !std::isfinite(m.transition) ||
        m.transition < 0.0)
        fail("Grid dimensions, ghost width, spacing or topography parameters are invalid.");
    if (m.count > static_cast<size_t>(306) * 306 * 306)
        fail("Grid exceeds the 28,652,616 padded-node total budget; redistribute X/Y/Z nodes.");
    const double bottom = (m.nz - 2 * G - 1) * dx;
    if (hasDem(p)) {
        const double width = (p.dem_nx - 1) * p.dem_spacing, height = (p.dem_ny - 1) * p.dem_spacing;
        if (p.dem_nx < 2 || p.dem_ny < 2 || p.dem_nx > 4096 || p.dem_ny > 4096 ||
            static_cast<size_t>(p.dem_nx) * p.dem_ny > 4000000 || !std::isfinite(p.dem_spacing) || p.dem_spacing <= 0 ||
            !std::isfinite(p.dem_origin_x) || !std::isfinite(p.dem_origin_y) || !std::isfinite(p.dem_datum) ||
            !std::isfinite(width) || !std::isfinite(height))
            fail("DEM array requires 2..4096 rows/columns, at most 4 million samples, positive metre spacing and finite origin/datum.");
        for (double v : m.dem)
            if (!std::isfinite(v) || std::fabs(v) > 1e7) fail("DEM contains missing, nonfinite or unreasonable elevations.");
        if ((m.nx - 2 * G - 1) * dx > width + 1e-8 * p.dem_spacing || (m.ny - 2 * G - 1) * dx > height + 1e-8 * p.dem_spacing)
            fail("Imported DEM does not cover the staged X/Y model span; reduce dimensions/spacing or export a larger region.");
    }
    double relief = std::fabs(p.topography_amplitude);
    if (hasDem(p)) {
        relief = 0;
        for (int j = 0; j < m.ny; ++j)
            for (int i = 0; i < m.nx; ++i) relief = std::max(relief, std::fabs(demDisplacement(p, (i - G) * dx, (j - G) * dx)));
    }
    if (m.transition == 0.0) m.transition = std::min(bottom, std::max({0.35 * bottom, 6.0 * dx, 4.0 * relief + 2.0 * dx}));
    if (m.transition > bottom || (m.terrain && (m.transition < 4.0 * dx || m.transition <= 2.6 * relief)))
        fail("Terrain transition must fit the model depth and exceed 2.6 times relief amplitude (at least four cells).");
}

// Grid3D::buildGeometry (CurvilinearSolver.cpp:146-190).  The surface displacement depends on (i, j) and the
// fade on k only: both are evaluated once here (cos / pow on the host), the volume from them.
void geometryInputs(const Model3D &m, std::vector<double> &surface, std::vector<double> &fade) {
    const auto &p = m.p;
    const double dx = m.h;
    surface.assign(static_cast<size_t>(m.nx) * m.ny, 0.0), fade.assign(m.nz, 0.0);
    if (!m.terrain) return;
    for (int j = 0; j < m.ny; ++j)
        for (int i = 0; i < m.nx; ++i) {
            const double x = (i - G) * dx, y = (j - G) * dx;
            const double px = 2.0 * kPi * x / p.topography_wavelength, py = 2.0 * kPi * y / p.topography_wavelength;
            surface[static_cast<size_t>(i) + static_cast<size_t>(m.nx) * j] =
                hasDem(p) ? demDisplacement(p, x, y) : p.topography_amplitude * std::cos(px) * std::cos(py);
        }
    for (int k = 0; k < m.nz; ++k) {
        const double depth = (k - G) * dx;
        const double s = std::clamp(depth / m.transition, 0.0, 1.0);
        fade[k] = depth >= m.transition ? 0.0
                                        : 1.0 - std::pow(s, 5) * (126.0 + s * (-420.0 + s * (540.0 + s * (-315.0 + 70.0 * s))));
    }
}
void buildGeometry(Model3D &m, const std::vector<double> &surface, const std::vector<double> &fade) {
    const double dx = m.h;
    m.coordinates.assign(3 * m.count, 0.0), m.metrics.assign(9 * m.count, 0.0);
    m.jacobian.assign(m.count, 0.0), m.normals.assign(3 * m.count, 0.0);
    for (int k = 0; k < m.nz; ++k)
        for (int j = 0; j < m.ny; ++j)
            for (int i = 0; i < m.nx; ++i) {
                const double x = (i - G) * dx, y = (j - G) * dx, depth = (k - G) * dx;
                double *c = &m.coordinates[3 * m.index(i, j, k)];
                c[0] = x, c[1] = y;
                c[2] = m.terrain ? depth + surface[static_cast<size_t>(i) + static_cast<size_t>(m.nx) * j] * fade[k] : depth;
            }
    const auto coordDerivative = [&](int c, int i, int j, int k, int axis) {
        const int di = axis == 0, dj = axis == 1, dk = axis == 2;
        const auto at = [&](int x, int y, int z) { return m.coordinates[3 * m.index(x, y, z) + c]; };
        return (at(i - 2 * di, j - 2 * dj, k - 2 * dk) - 8.0 * at(i - di, j - dj, k - dk) + 8.0 * at(i + di, j + dj, k + dk) -
                at(i + 2 * di, j + 2 * dj, k + 2 * dk)) /
               (12.0 * dx);
    };
    for (int k = G; k < m.nz - G; ++k)
        for (int j = G; j < m.ny - G; ++j)
            for (int i = G; i < m.nx - G; ++i) {
                const size_t q = m.index(i, j, k);
                double *mt = &m.metrics[9 * q];
                if (!m.terrain) {
                    for (int a = 0; a < 9; ++a) mt[a] = a % 4 == 0 ? 1 : 0;
                    m.jacobian[q] = 1;
                    if (k == G) m.normals[3 * q + 2] = 1;
                    continue;
                }
                V3 rx{}, ry{}, rz{};
                for (int c = 0; c < 3; ++c) rx[c] = coordDerivative(c, i, j, k, 0), ry[c] = coordDerivative(c, i, j, k, 1),
                                              rz[c] = coordDerivative(c, i, j, k, 2);
                const V3 ec = cross(ry, rz), zc = cross(rz, rx), xc = cross(rx, ry);
                const double jac = dot(rx, ec);
                if (!(jac > 1.0e-10) || !std::isfinite(jac)) fail("Curvilinear mapping has a non-positive or non-finite Jacobian.");
                m.jacobian[q] = jac;
                for (int c = 0; c < 3; ++c) mt[c] = ec[c] / jac, mt[3 + c] = zc[c] / jac, mt[6 + c] = xc[c] / jac;
                if (k == G) {
                    const double length = std::sqrt(dot(xc, xc));
                    if (!(length > 0.0) || !std::isfinite(length)) fail("Could not compute topographic surface normal.");
                    for (int c = 0; c < 3; ++c) m.normals[3 * q + c] = xc[c] / length;
                }
            }
}

double materialDepth(const Model3D &m, size_t q) {
    const double *xyz = &m.coordinates[3 * q];
    if (m.p.layer_depth_mode == 0) return (static_cast<int>(q / (static_cast<size_t>(m.nx) * m.ny)) - G) * m.h;
    if (m.p.layer_depth_mode == 2) return xyz[2];
    const size_t plane = static_cast<size_t>(m.nx) * m.ny, surface = static_cast<size_t>(G) * plane + q % plane;
    return xyz[2] - m.coordinates[3 * surface + 2];
}
int layerAt(const swl_params3d &p, double depth) {
    double bottom = 0;
    for (int i = 0; i + 1 < p.n_layers; ++i) {
        bottom += p.layer_thickness[i];
        if (depth < bottom) return i;
    }
    return p.n_layers - 1;
}
// metric norm, Jacobian range, CFL plane distances, edges (CurvilinearSolver.cpp:328-362)
void cflGeometry(Model3D &m) {
    const auto &p = m.p;
    m.minJacobian = std::numeric_limits<double>::max();
    double maxNorm = 0.0;
    for (int k = G; k < m.nz - G; ++k)
        for (int j = G; j < m.ny - G; ++j)
            for (int i = G; i < m.nx - G; ++i) {
                const size_t q = m.index(i, j, k);
                m.minJacobian = std::min(m.minJacobian, m.jacobian[q]);
                double n2 = 0.0;
                for (int a = 0; a < 9; ++a) n2 += m.metrics[9 * q + a] * m.metrics[9 * q + a];
                maxNorm = std::max(maxNorm, std::sqrt(n2));
            }
    m.maxMetricNorm = maxNorm;
    double minLength = !m.terrain ? p.spacing / std::sqrt(3.0) : std::numeric_limits<double>::max();
    if (m.terrain)
        for (int k = G; k < m.nz - G; ++k)
            for (int j = G; j < m.ny - G; ++j)
                for (int i = G; i < m.nx - G; ++i) {
                    const double *ctr = &m.coordinates[3 * m.index(i, j, k)];
                    for (int sk : {-1, 1})
                        for (int sj : {-1, 1})
                            for (int si : {-1, 1}) {
                                const double *a = &m.coordinates[3 * m.index(i + si, j, k)];
                                const double *b = &m.coordinates[3 * m.index(i, j + sj, k)];
                                const double *c = &m.coordinates[3 * m.index(i, j, k + sk)];
                                V3 ab{}, ac{}, ap{};
                                for (int x = 0; x < 3; ++x) ab[x] = b[x] - a[x], ac[x] = c[x] - a[x], ap[x] = ctr[x] - a[x];
                                const V3 n = cross(ab, ac);
                                const double length = std::sqrt(dot(n, n));
                                const double distance = std::fabs(dot(ap, n)) / length;
                                if (!std::isfinite(distance) || distance <= 0) fail("Degenerate adjacent cell plane in CFL geometry.");
                                minLength = std::min(minLength, distance);
                            }
                }
    m.minLength = minLength;
    double maxEdge = !m.terrain ? p.spacing : 0;
    if (m.terrain)
        for (int k = G; k < m.nz - G; ++k)
            for (int j = G; j < m.ny - G; ++j)
                for (int i = G; i < m.nx - G; ++i) {
                    const double *a = &m.coordinates[3 * m.index(i, j, k)];
                    for (int axis = 0; axis < 3; ++axis) {
                        const int ni = i + (axis == 0), nj = j + (axis == 1), nk = k + (axis == 2);
                        if (ni >= m.nx - G || nj >= m.ny - G || nk >= m.nz - G) continue;
                        const double *b = &m.coordinates[3 * m.index(ni, nj, nk)];
                        V3 d{};
                        for (int c = 0; c < 3; ++c) d[c] = b[c] - a[c];
                        maxEdge = std::max(maxEdge, std::sqrt(dot(d, d)));
                    }
                }
    m.maxEdge = maxEdge;
}

} // namespace

// layer of node q as the material pass assigns it (MaterialModels3D.hpp:19-31)
int layerOfNode3D(const Model3D &m, size_t q) { return layerAt(m.p, materialDepth(m, q)); }

ModelMemory3D estimateModelMemory3D(const swl_params3d &p) {
    ModelMemory3D r;
    if (p.nx < 8 || p.ny < 8 || p.nz < 8 || p.nx > 4096 || p.ny > 4096 || p.nz > 4096) return r;
    r.points = static_cast<size_t>(p.nx + 2 * G) * (p.ny + 2 * G) * (p.nz + 2 * G);
    size_t receivers = 0;
    try {
        receivers = receiverNodes(p).size();
    } catch (const std::exception &) {
        return r;
    }
    const double mib = 1024 * 1024, moment = p.source_kind == 1 ? 3 : 0;
    const double receiverBytes = receivers * 3.0 * (p.max_steps > 0 ? p.max_steps + 1.0 : 20001.0) * sizeof(double);
    const double stats = (r.points + 127) / 128 * 72.0;
    r.cpuMiB = (r.points * ((p.layered ? 84.0 : 81.0) + moment) * sizeof(double) + receiverBytes) / mib;
    r.cudaHostMiB = (r.points * ((p.layered ? 21.0 : 18.0) + moment) * sizeof(double) + receiverBytes + stats) / mib;
    r.deviceMiB = (r.points * ((p.layered ? 81.0 : 78.0) + moment) * sizeof(double) + stats) / mib;
    const double demCopies = 3.0 * (hasDem(p) ? static_cast<double>(p.dem_nx) * p.dem_ny : 0.0) * sizeof(double) / mib;
    r.cpuMiB += demCopies, r.cudaHostMiB += demCopies;
    r.withinBudget = r.points <= static_cast<size_t>(306) * 306 * 306;
    return r;
}

double hipDeviceBytes3D(const swl_params3d &p, int precision, int rows) {
    // the allocations of Engine3D (engine3d_host.inl) for this session's padded grid
    const double nx = p.nx + 2 * G, ny = p.ny + 2 * G, nz = p.nz + 2 * G, N = nx * ny * nz;
    const double word = precision == 1 ? 4 : 8, tx = 128 / word, ty = 256 / tx;
    const double ld = std::ceil(nx / 64) * 64, count = std::ceil((61 + ld * ny * nz) / 64) * 64;
    double receivers = 0;
    try {
        receivers = static_cast<double>(receiverNodes(p).size());
    } catch (const std::exception &) {
    }
    const double nrec = std::max(receivers, 1.0);
    const double tiles = std::ceil(p.nx / tx) * std::ceil(p.ny / ty) * p.nz;
    if (p.max_steps > 0) rows = std::min(rows, p.max_steps); // a run never reserves more rows than it has steps
    double bytes = 4.0 * 9 * count * word;                        // 3 rotating states + accumulator
    bytes += count * word * (10 + 3 + 1 + 1 + (p.source_kind == 1 ? 3 : 0)); // metrics + J, materials, damping, source
    bytes += ld * ny * (9 + 1 + 1 + 3) * 8.0;                      // FP64 surface plane
    bytes += 2 * 88.0 * tiles + 128 * 88.0;                        // inspection partials
    bytes += 8.0 * N;                                              // field download buffer
    bytes += 4.0 * std::max({p.nx * p.nz, p.ny * p.nz, p.nx * p.ny}); // display slice
    bytes += nrec * 3 * 16 + 32.0 + 80.0 + rows * (10 + 3 * nrec) * 8.0; // probes, gather, run rows
    return bytes;
}

Model3D buildModel3D(const swl_params3d &requested, int device, int precision) {
    validate(requested);
    Model3D m;
    m.p = requested;
    if (hasDem(requested)) {
        m.dem.assign(requested.dem_elevations,
                     requested.dem_elevations + static_cast<size_t>(requested.dem_nx) * requested.dem_ny);
        m.p.dem_elevations = m.dem.data();
    } else {
        m.p.dem_elevations = nullptr, m.p.dem_nx = m.p.dem_ny = 0;
    }
    const auto &p = m.p;
    m.nx = p.nx + 2 * G, m.ny = p.ny + 2 * G, m.nz = p.nz + 2 * G;
    m.count = static_cast<size_t>(m.nx) * m.ny * m.nz;
    m.h = p.spacing;
    m.transition = p.terrain_transition_depth;
    m.terrain = hasDem(p) || p.topography_amplitude != 0;
    validateGeometry(m);
    if (device >= 0) preflightMemory3D(p, precision, device);
    std::vector<double> surface, fade;
    geometryInputs(m, surface, fade);
    if (device >= 0) {
        buildGeometry3DOnDevice(m, surface, fade, device); // the volume and the CFL geometry on the GPU (FP64)
    } else {
        buildGeometry(m, surface, fade);
        cflGeometry(m);
    }
    m.lambda = p.density * (p.vp * p.vp - 2.0 * p.vs * p.vs);
    m.mu = p.density * p.vs * p.vs;
    if (!std::isfinite(m.lambda) || !std::isfinite(m.mu) || m.mu <= 0.0 || !std::isfinite(m.maxMetricNorm))
        fail("Derived elastic coefficients or grid metrics are out of range.");
    // materials (:366-381)
    double maxVp = p.vp, minVs = p.vs, maxVs = p.vs, maxDensity = p.density;
    m.lambdaNode.assign(m.count, m.lambda), m.muNode.assign(m.count, m.mu), m.densityNode.assign(m.count, p.density);
    if (p.layered) {
        maxVp = maxVs = maxDensity = 0, minVs = std::numeric_limits<double>::max();
        for (int l = 0; l < p.n_layers; ++l) {
            maxVp = std::max(maxVp, p.layer_vp[l]), minVs = std::min(minVs, p.layer_vs[l]);
            maxVs = std::max(maxVs, p.layer_vs[l]), maxDensity = std::max(maxDensity, p.layer_density[l]);
        }
        for (size_t q = 0; q < m.count; ++q) {
            const int id = layerAt(p, materialDepth(m, q));
            const double d = p.layer_density[id], vp = p.layer_vp[id], vs = p.layer_vs[id];
            m.lambdaNode[q] = d * (vp * vp - 2 * vs * vs), m.muNode[q] = d * vs * vs, m.densityNode[q] = d;
        }
    }
    m.maxVp = maxVp, m.minVs = minVs, m.maxVs = maxVs, m.maxDensity = maxDensity;
    // sponge (AbsorbingBoundary3D.hpp:5-21): rate = sum over axes of strength*vp/(w h) * f^2
    int widths[3]{};
    const int nodes[3]{p.nx, p.ny, p.nz};
    for (int a = 0; a < 3; ++a) widths[a] = p.absorbing ? std::min(p.absorbing_cells[a], nodes[a] / 3) : 0;
    const auto term = [&](int a, int distance) {
        if (widths[a] <= 0) return 0.0;
        const double f = std::clamp(1.0 - static_cast<double>(distance) / widths[a], 0.0, 1.0);
        return p.absorbing_strength * maxVp / (widths[a] * p.spacing) * f * f;
    };
    m.damping.assign(m.count, 0.0);
    double maxDamping = 0;
    for (int k = G; k < m.nz - G; ++k)
        for (int j = G; j < m.ny - G; ++j)
            for (int i = G; i < m.nx - G; ++i) {
                const int li = i - G, lj = j - G, lk = k - G;
                double rate = 0;
                const int distance[3]{std::min(li, p.nx - 1 - li), std::min(lj, p.ny - 1 - lj), p.nz - 1 - lk};
                for (int a = 0; a < 3; ++a)
                    if (widths[a] > 0) rate += term(a, distance[a]);
                m.damping[m.index(i, j, k)] = rate;
                maxDamping = std::max(maxDamping, rate);
            }
    m.maxDamping = maxDamping;
    const double maxNorm = m.maxMetricNorm, minLength = m.minLength;
    // dt plan (TimeStepPlan3D.hpp:8-20)
    for (double v : {p.spacing, p.cfl, p.duration, p.dominant_frequency, maxVp, maxNorm, minLength, maxDamping})
        if (!std::isfinite(v)) fail("Time-step inputs must be finite.");
    if (maxVp <= 0 || maxNorm <= 0 || minLength <= 0 || maxDamping < 0) fail("Invalid time-step parameters.");
    m.waveDt = p.cfl * std::min(p.spacing / maxNorm, minLength) / maxVp;
    m.dampingDt = maxDamping > 0 ? .5 / maxDamping : std::numeric_limits<double>::max();
    m.sourceDt = 1 / (50 * p.dominant_frequency);
    const double safe = std::min({m.waveDt, m.dampingDt, m.sourceDt});
    const double steps = std::max(1.0, std::ceil(p.duration / safe));
    if (!std::isfinite(steps) || (p.max_steps == 0 && steps > 20000))
        fail("Time-step limit: at most 20,000 steps. Set a maximum step cap.");
    m.dt = p.duration / steps;
    m.total = static_cast<int>(p.max_steps > 0 ? std::min(steps, static_cast<double>(p.max_steps)) : steps);
    if (!(m.dt > 0) || !std::isfinite(m.dt)) fail("Invalid time-step parameters.");
    // Gaussian source weights (:395-403) and isotropic-moment pattern (:404-412)
    const int sx = p.source_x < 0 ? p.nx / 2 : p.source_x, sy = p.source_y < 0 ? p.ny / 2 : p.source_y, sz = p.source_z;
    m.sourceNode = {sx, sy, sz};
    m.source.assign(m.count, 0.0);
    double sum = 0.0;
    for (int k = -2; k <= 2; ++k)
        for (int j = -2; j <= 2; ++j)
            for (int i = -2; i <= 2; ++i) sum += std::exp(-0.5 * (i * i + j * j + k * k));
    for (int k = -2; k <= 2; ++k)
        for (int j = -2; j <= 2; ++j)
            for (int i = -2; i <= 2; ++i) {
                const size_t q = m.index(G + sx + i, G + sy + j, G + sz + k);
                m.source[q] = std::exp(-0.5 * (i * i + j * j + k * k)) / (sum * m.jacobian[q]);
            }
    if (p.source_kind == 1) {
        for (auto &g : m.gradient) g.assign(m.count, 0.0);
        const std::ptrdiff_t stride[3]{1, m.nx, static_cast<std::ptrdiff_t>(m.nx) * m.ny};
        for (int k = sz - 4; k <= sz + 4; ++k)
            for (int j = sy - 4; j <= sy + 4; ++j)
                for (int i = sx - 4; i <= sx + 4; ++i) {
                    const size_t q = m.index(i + G, j + G, k + G);
                    const auto flux = [&](std::ptrdiff_t x, int axis, int c) {
                        return m.jacobian[x] * m.metrics[9 * static_cast<size_t>(x) + 3 * axis + c] * m.source[x];
                    };
                    double r[3]{};
                    for (int c = 0; c < 3; ++c)
                        for (int a = 0; a < 3; ++a) {
                            const std::ptrdiff_t s = stride[a], x = static_cast<std::ptrdiff_t>(q);
                            r[c] -= (flux(x - 2 * s, a, c) - 8 * flux(x - s, a, c) + 8 * flux(x + s, a, c) - flux(x + 2 * s, a, c)) /
                                    (12 * p.spacing * m.jacobian[q]);
                        }
                    for (int a = 0; a < 3; ++a) m.gradient[a][q] = r[a];
                }
        m.direction = {0, 0, 0};
    } else {
        const double norm = hypot3(p.source_direction[0], p.source_direction[1], p.source_direction[2]);
        for (int a = 0; a < 3; ++a) m.direction[a] = p.source_direction[a] / norm;
    }
    m.amplitude = p.source_kind == 1 ? p.source_moment : p.source_force;
    // receivers (:413-424)
    for (const auto &node : receiverNodes(p)) {
        const size_t q = m.index(node[0] + G, node[1] + G, node[2] + G);
        m.receiverIndex.push_back(static_cast<int64_t>(q));
        for (int c = 0; c < 3; ++c) m.receiverPosition.push_back(m.coordinates[3 * q + c]);
    }
    // separable forms for the engine
    m.flatFrom = m.nz;
    for (int k = m.nz - G - 1; k >= G; --k) {
        bool identity = true;
        for (int j = G; j < m.ny - G && identity; ++j)
            for (int i = G; i < m.nx - G && identity; ++i) {
                const size_t q = m.index(i, j, k);
                for (int a = 0; a < 9; ++a) identity = identity && m.metrics[9 * q + a] == (a % 4 == 0 ? 1.0 : 0.0);
                identity = identity && m.jacobian[q] == 1.0;
            }
        if (!identity) break;
        m.flatFrom = k;
    }
    if (m.flatFrom <= G) m.flatFrom = G; // the surface plane always keeps its metric (closure)
    for (int a = 0; a < 3; ++a) {
        const int n = a == 0 ? m.nx : a == 1 ? m.ny : m.nz;
        m.dampingAxis[a].assign(n, 0.0);
        for (int x = G; x < n - G; ++x) {
            const int l = x - G, distance = a == 2 ? p.nz - 1 - l : std::min(l, nodes[a] - 1 - l);
            m.dampingAxis[a][x] = term(a, distance);
        }
    }
    m.materialByDepth = true;
    for (int k = 0; k < m.nz && m.materialByDepth; ++k) {
        const size_t q0 = m.index(0, 0, k);
        for (int j = 0; j < m.ny && m.materialByDepth; ++j)
            for (int i = 0; i < m.nx; ++i) {
                const size_t q = m.index(i, j, k);
                if (m.lambdaNode[q] != m.lambdaNode[q0] || m.muNode[q] != m.muNode[q0] || m.densityNode[q] != m.densityNode[q0]) {
                    m.materialByDepth = false;
                    break;
                }
            }
    }
    return m;
}

swl_mesh3d_desc meshDescriptor3D(const Model3D &m, int precision) {
    swl_mesh3d_desc d{};
    d.precision = precision;
    d.nx = m.nx, d.ny = m.ny, d.nz = m.nz;
    d.h = m.h;
    d.metrics = m.metrics.data(), d.jacobian = m.jacobian.data(), d.normals = m.normals.data();
    d.lambda = m.lambdaNode.data(), d.mu = m.muNode.data(), d.density = m.densityNode.data();
    d.density_global = m.p.density;
    d.damping = m.damping.data(), d.source = m.source.data();
    for (int a = 0; a < 3; ++a) {
        d.source_gradient[a] = m.gradient[a].empty() ? nullptr : m.gradient[a].data();
        d.source_direction[a] = m.direction[a];
        d.damping_axis[a] = m.dampingAxis[a].data();
    }
    d.flat_from = m.flatFrom;
    d.material_by_depth = m.materialByDepth;
    return d;
}

} // namespace swl
