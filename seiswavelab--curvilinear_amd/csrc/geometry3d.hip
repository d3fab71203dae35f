// geometry3d.hip — FP64 construction of the 3D curvilinear grid on the GPU.
//
// Grid3D::buildGeometry (src/three_d/CurvilinearSolver.cpp:146-190) and the CFL geometry of
// CpuSimulation3D::initialize (:328-362) as kernels: node coordinates, the FD4 covariant basis, the
// cross-product inverse metrics and Jacobian, the surface normals, and the reductions min J, max
// metric norm, min distance from a node to its 8 adjacent cell planes, max cell edge.  Every value is
// computed in FP64 in the host builder's operation order with no contraction (the library is built
// with -ffp-contract=off; sqrt and division are IEEE correctly rounded), so the arrays and the CFL
// figures are bit-identical to the host path.  The host supplies the per-column surface displacement
// and the per-plane fade (they involve cos / pow), the GPU does the volume work.
#include "common.hpp"
#include "model3d.hpp"

#include <cfloat>
#include <unistd.h>
#include <vector>

namespace swl {
namespace {

constexpr int G3 = 3, GB = 256;

struct Geo {
    int nx, ny, nz, terrain;
    double dx;
    double *coords, *metrics, *jac, *normals;
};
__device__ inline size_t gidx(const Geo &g, int i, int j, int k) {
    return (size_t)i + (size_t)g.nx * ((size_t)j + (size_t)g.ny * k);
}

__global__ void coords_kernel(Geo g, const double *surface, const double *fade) {
    const size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (q >= (size_t)g.nx * g.ny * g.nz) return;
    const int i = (int)(q % g.nx), j = (int)((q / g.nx) % g.ny), k = (int)(q / ((size_t)g.nx * g.ny));
    const double x = (i - G3) * g.dx, y = (j - G3) * g.dx, depth = (k - G3) * g.dx;
    double *c = g.coords + 3 * q;
    c[0] = x, c[1] = y, c[2] = g.terrain ? depth + surface[i + (size_t)g.nx * j] * fade[k] : depth;
}

// FD4 derivative of coordinate component c along `axis` (CurvilinearSolver.cpp:47-54)
__device__ inline double dcoord(const Geo &g, int c, int i, int j, int k, int axis) {
    const int di = axis == 0, dj = axis == 1, dk = axis == 2;
    auto at = [&](int x, int y, int z) { return g.coords[3 * gidx(g, x, y, z) + c]; };
    return (at(i - 2 * di, j - 2 * dj, k - 2 * dk) - 8.0 * at(i - di, j - dj, k - dk) + 8.0 * at(i + di, j + dj, k + dk) -
            at(i + 2 * di, j + 2 * dj, k + 2 * dk)) /
           (12.0 * g.dx);
}
__device__ inline void cross3(const double *a, const double *b, double *o) {
    o[0] = a[1] * b[2] - a[2] * b[1], o[1] = a[2] * b[0] - a[0] * b[2], o[2] = a[0] * b[1] - a[1] * b[0];
}
__device__ inline double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// interior nodes; `fail` keeps the first failing node in the host's loop order: 2q (Jacobian) or 2q+1 (normal)
__global__ void metrics_kernel(Geo g, unsigned long long *fail) {
    const int pi = g.nx - 2 * G3, pj = g.ny - 2 * G3, pk = g.nz - 2 * G3;
    const size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (t >= (size_t)pi * pj * pk) return;
    const int i = G3 + (int)(t % pi), j = G3 + (int)((t / pi) % pj), k = G3 + (int)(t / ((size_t)pi * pj));
    const size_t q = gidx(g, i, j, k);
    double *mt = g.metrics + 9 * q;
    if (!g.terrain) {
        for (int a = 0; a < 9; ++a) mt[a] = a % 4 == 0 ? 1 : 0;
        g.jac[q] = 1;
        if (k == G3) g.normals[3 * q + 2] = 1;
        return;
    }
    double rx[3], ry[3], rz[3];
    for (int c = 0; c < 3; ++c) rx[c] = dcoord(g, c, i, j, k, 0), ry[c] = dcoord(g, c, i, j, k, 1), rz[c] = dcoord(g, c, i, j, k, 2);
    double ec[3], zc[3], xc[3];
    cross3(ry, rz, ec), cross3(rz, rx, zc), cross3(rx, ry, xc);
    const double jac = dot3(rx, ec);
    if (!(jac > 1.0e-10) || !isfinite(jac)) {
        atomicMin(fail, 2ull * q);
        return;
    }
    g.jac[q] = jac;
    for (int c = 0; c < 3; ++c) mt[c] = ec[c] / jac, mt[3 + c] = zc[c] / jac, mt[6 + c] = xc[c] / jac;
    if (k == G3) {
        const double length = sqrt(dot3(xc, xc));
        if (!(length > 0.0) || !isfinite(length)) {
            atomicMin(fail, 2ull * q + 1);
            return;
        }
        for (int c = 0; c < 3; ++c) g.normals[3 * q + c] = xc[c] / length;
    }
}

// per block: min J, max metric norm, min cell-plane distance, max edge (and a degenerate-plane flag)
struct Cfl {
    double minJ, maxNorm, minLength, maxEdge;
    int degenerate, pad;
};
__global__ __launch_bounds__(GB) void cfl_kernel(Geo g, Cfl *partials) {
    __shared__ Cfl sh[GB];
    const int pi = g.nx - 2 * G3, pj = g.ny - 2 * G3, pk = g.nz - 2 * G3;
    const size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    Cfl r{DBL_MAX, 0.0, DBL_MAX, 0.0, 0, 0};
    if (t < (size_t)pi * pj * pk) {
        const int i = G3 + (int)(t % pi), j = G3 + (int)((t / pi) % pj), k = G3 + (int)(t / ((size_t)pi * pj));
        const size_t q = gidx(g, i, j, k);
        r.minJ = g.jac[q];
        double n2 = 0.0;
        for (int a = 0; a < 9; ++a) n2 += g.metrics[9 * q + a] * g.metrics[9 * q + a];
        r.maxNorm = sqrt(n2);
        if (g.terrain) {
            const double *ctr = g.coords + 3 * q;
            for (int sk = -1; sk <= 1; sk += 2)
                for (int sj = -1; sj <= 1; sj += 2)
                    for (int si = -1; si <= 1; si += 2) {
                        const double *a = g.coords + 3 * gidx(g, i + si, j, k);
                        const double *b = g.coords + 3 * gidx(g, i, j + sj, k);
                        const double *c = g.coords + 3 * gidx(g, i, j, k + sk);
                        double ab[3], ac[3], ap[3], n[3];
                        for (int x = 0; x < 3; ++x) ab[x] = b[x] - a[x], ac[x] = c[x] - a[x], ap[x] = ctr[x] - a[x];
                        cross3(ab, ac, n);
                        const double length = sqrt(dot3(n, n));
                        const double distance = fabs(dot3(ap, n)) / length;
                        if (!isfinite(distance) || distance <= 0) r.degenerate = 1;
                        else r.minLength = fmin(r.minLength, distance);
                    }
            const double *a = g.coords + 3 * q;
            for (int axis = 0; axis < 3; ++axis) {
                const int ni = i + (axis == 0), nj = j + (axis == 1), nk = k + (axis == 2);
                if (ni >= g.nx - G3 || nj >= g.ny - G3 || nk >= g.nz - G3) continue;
                const double *b = g.coords + 3 * gidx(g, ni, nj, nk);
                double d[3];
                for (int c = 0; c < 3; ++c) d[c] = b[c] - a[c];
                r.maxEdge = fmax(r.maxEdge, sqrt(dot3(d, d)));
            }
        }
    }
    sh[threadIdx.x] = r;
    __syncthreads();
    for (int s = GB / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            Cfl &x = sh[threadIdx.x];
            const Cfl &y = sh[threadIdx.x + s];
            x.minJ = fmin(x.minJ, y.minJ), x.maxNorm = fmax(x.maxNorm, y.maxNorm);
            x.minLength = fmin(x.minLength, y.minLength), x.maxEdge = fmax(x.maxEdge, y.maxEdge);
            x.degenerate |= y.degenerate;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = sh[0];
}

} // namespace

// Fills m.coordinates, m.metrics, m.jacobian, m.normals and m.minJacobian, m.maxMetricNorm, m.minLength,
// m.maxEdge on `device` (the values the host loops in model3d.cpp produce, bit for bit).
void buildGeometry3DOnDevice(Model3D &m, const std::vector<double> &surface, const std::vector<double> &fade, int device) {
    DeviceScope scope(device);
    const hipStream_t st = setup_stream(); // set-up work waits for this thread's own stream only
    const size_t N = m.count;
    SetupBuffer coords, metrics, jac, normals, dsurf, dfade, fail, parts;
    coords.allocate(3 * N * sizeof(double)), metrics.allocate(9 * N * sizeof(double));
    jac.allocate(N * sizeof(double)), normals.allocate(3 * N * sizeof(double));
    dsurf.upload(surface.data(), std::max<size_t>(surface.size(), 1) * sizeof(double));
    dfade.upload(fade.data(), std::max<size_t>(fade.size(), 1) * sizeof(double));
    const unsigned long long none = ~0ull;
    fail.upload(&none, sizeof none);
    Geo g{m.nx, m.ny, m.nz, m.terrain ? 1 : 0, m.h, coords.as<double>(), metrics.as<double>(), jac.as<double>(),
          normals.as<double>()};
    const size_t interior = (size_t)(m.nx - 2 * G3) * (m.ny - 2 * G3) * (m.nz - 2 * G3);
    const unsigned nb = (unsigned)((interior + GB - 1) / GB);
    coords_kernel<<<(unsigned)((N + GB - 1) / GB), GB, 0, st>>>(g, dsurf.as<double>(), dfade.as<double>());
    check(hipGetLastError(), "3D geometry launch");
    metrics_kernel<<<nb, GB, 0, st>>>(g, fail.as<unsigned long long>());
    check(hipGetLastError(), "3D geometry launch");
    unsigned long long first = none;
    check(hipMemcpyAsync(&first, fail.ptr, sizeof first, hipMemcpyDeviceToHost, st), "3D geometry");
    check(hipStreamSynchronize(st), "3D geometry");
    if (first != none)
        throw std::runtime_error(first & 1 ? "Could not compute topographic surface normal."
                                           : "Curvilinear mapping has a non-positive or non-finite Jacobian.");
    parts.allocate((size_t)nb * sizeof(Cfl));
    cfl_kernel<<<nb, GB, 0, st>>>(g, parts.as<Cfl>());
    check(hipGetLastError(), "3D CFL launch");
    std::vector<Cfl> h(nb);
    check(hipMemcpyAsync(h.data(), parts.ptr, h.size() * sizeof(Cfl), hipMemcpyDeviceToHost, st), "3D CFL");
    m.coordinates.resize(3 * N), m.metrics.resize(9 * N), m.jacobian.resize(N), m.normals.resize(3 * N);
    check(hipMemcpyAsync(m.coordinates.data(), coords.ptr, 3 * N * sizeof(double), hipMemcpyDeviceToHost, st), "3D geometry");
    check(hipMemcpyAsync(m.metrics.data(), metrics.ptr, 9 * N * sizeof(double), hipMemcpyDeviceToHost, st), "3D geometry");
    check(hipMemcpyAsync(m.jacobian.data(), jac.ptr, N * sizeof(double), hipMemcpyDeviceToHost, st), "3D geometry");
    check(hipMemcpyAsync(m.normals.data(), normals.ptr, 3 * N * sizeof(double), hipMemcpyDeviceToHost, st), "3D geometry");
    check(hipStreamSynchronize(st), "3D geometry");
    Cfl r{DBL_MAX, 0.0, DBL_MAX, 0.0, 0, 0};
    for (const auto &x : h) {
        r.minJ = std::min(r.minJ, x.minJ), r.maxNorm = std::max(r.maxNorm, x.maxNorm);
        r.minLength = std::min(r.minLength, x.minLength), r.maxEdge = std::max(r.maxEdge, x.maxEdge);
        r.degenerate |= x.degenerate;
    }
    if (m.terrain && r.degenerate) throw std::runtime_error("Degenerate adjacent cell plane in CFL geometry.");
    m.minJacobian = r.minJ, m.maxMetricNorm = r.maxNorm;
    m.minLength = m.terrain ? r.minLength : m.p.spacing / std::sqrt(3.0);
    m.maxEdge = m.terrain ? r.maxEdge : m.p.spacing;
}

// fitsMemory3D (include/Seis3D/ModelPresets3D.hpp:50-56) for a HIP session: the host estimate within 80 % of the
// free host memory and this engine's device bytes within 90 % of the device's free memory, else the reference's
// refusal (CurvilinearSolver.cpp:309).
void preflightMemory3D(const swl_params3d &p, int precision, int device) {
    const ModelMemory3D m = estimateModelMemory3D(p);
    const long pages = sysconf(_SC_AVPHYS_PAGES), page = sysconf(_SC_PAGESIZE);
    const double host = pages > 0 && page > 0 ? (double)pages * (double)page : 0;
    DeviceScope scope(device);
    size_t freeBytes = 0, total = 0;
    check(hipMemGetInfo(&freeBytes, &total), "memory query");
    const bool fits = m.withinBudget && (host == 0 || m.cudaHostMiB * 1024 * 1024 <= .8 * host) &&
                      hipDeviceBytes3D(p, precision) <= .9 * (double)freeBytes;
    if (!fits)
        throw std::runtime_error("Insufficient available memory for the requested grid. Current session retained; reduce "
                                 "dimensions or use Fit to available memory.");
}

} // namespace swl
