// geometry2d.hip — FP64 construction of the 2D curvilinear grid on the GPU.
//
// The metric part of Simulation::initialize (src/two_d/CurvilinearSolver2D.cpp:339-392) as one kernel over
// the padded grid: node coordinates (x, eta + H F), the closed-form Jacobian J = 1 + H F', the inverse
// metrics a = -H' F / J and b = 1 / J, and per-block partials of the CFL geometry (min J and
// max sqrt(1 + a^2 + b^2) over the computational nodes).  Every value is computed in FP64 in the host
// builder's operation order with no contraction (-ffp-contract=off; division and sqrt are IEEE correctly
// rounded), so arrays, min J, max metric and therefore dt and the step count are bit-identical to the host
// path.  The host supplies the per-column surface height / slope (cosine or PCHIP) and the per-row fade and
// its derivative (they involve cos and pow).  The surface normals and the max cell edge use hypot, whose
// last bit is library-specific, so they stay with the host.
#include "common.hpp"
#include "model2d.hpp"

#include <cfloat>
#include <iomanip>
#include <locale>
#include <sstream>
#include <unistd.h>
#include <vector>

namespace swl {
namespace {

constexpr int GB = 256;

struct Geo2 {
    int nx, nz, off;
    double h;
    const double *depth, *slope, *fade, *fadeDer; // H(x_i), H'(x_i) per padded column; F, F' per padded row
    double *coords, *a, *b, *jac;
};
struct Cfl2 {
    double minJ, maxMetric;
};

__global__ __launch_bounds__(GB) void geometry2d_kernel(Geo2 g, unsigned long long *fail, Cfl2 *partials) {
    __shared__ Cfl2 sh[GB];
    const size_t q = blockIdx.x * (size_t)GB + threadIdx.x;
    Cfl2 r{DBL_MAX, 0.0};
    if (q < (size_t)g.nx * g.nz) {
        const int i = (int)(q % g.nx), k = (int)(q / g.nx);
        const double x = (i - 3 - g.off) * g.h, eta = (k - 3) * g.h;
        const double zx = g.slope[i] * g.fade[k], j = 1 + g.depth[i] * g.fadeDer[k];
        if (!(j > 1e-9 && isfinite(j))) atomicMin(fail, (unsigned long long)q);
        g.coords[2 * q] = x;
        g.coords[2 * q + 1] = eta + g.depth[i] * g.fade[k];
        const double a = -zx / j, b = 1 / j;
        g.a[q] = a, g.b[q] = b, g.jac[q] = j;
        if (i >= 3 && i < g.nx - 3 && k >= 3 && k < g.nz - 3) r.minJ = j, r.maxMetric = sqrt(1 + a * a + b * b);
    }
    sh[threadIdx.x] = r;
    __syncthreads();
    for (int s = GB / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            Cfl2 &x = sh[threadIdx.x];
            const Cfl2 &y = sh[threadIdx.x + s];
            x.minJ = fmin(x.minJ, y.minJ), x.maxMetric = fmax(x.maxMetric, y.maxMetric);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = sh[0];
}

} // namespace

// CurvilinearSolver2D.cpp:247-266: the session's budget against 80 % of the free host memory and 85 % of the free
// memory of the device, with the reference's message.
void preflightMemory2D(const swl_params2d &p, int device) {
    const MemoryBudget2D budget = memoryBudget2D(p, 2);
    const auto require_memory = [](double bytes, double free, double fraction, const char *name) {
        if (bytes < free * fraction) return;
        constexpr double gib = 1024. * 1024 * 1024;
        std::ostringstream message;
        message.imbue(std::locale::classic());
        message << "Insufficient " << name << " for a transactional 2D rebuild. " << std::fixed << std::setprecision(2)
                << "Additional budget: " << bytes / gib << " GiB; available: " << free / gib
                << " GiB; required free with reserve: " << bytes / fraction / gib << " GiB.";
        throw std::runtime_error(message.str());
    };
    const long pages = sysconf(_SC_AVPHYS_PAGES), page = sysconf(_SC_PAGESIZE);
    if (pages > 0 && page > 0) require_memory(budget.host, (double)pages * (double)page, .8, "RAM");
    DeviceScope scope(device);
    size_t freeBytes = 0, total = 0;
    check(hipMemGetInfo(&freeBytes, &total), "memory query");
    require_memory(budget.device, (double)freeBytes, .85, "VRAM");
}

// Fills m.coordinates, m.a, m.b, m.jac and m.minJacobian, m.maxMetric on `device` from the profiles already in
// m (surfaceDepth, surfaceSlope, fade, fadeDerivative).  Returns the linear index of the first node with a
// nonpositive or nonfinite Jacobian, or -1 (the host raises the reference's message when it reaches it).
long long buildGeometry2DOnDevice(Model2D &m, int device) {
    DeviceScope scope(device);
    const hipStream_t st = setup_stream(); // set-up work waits for this thread's own stream only
    const size_t N = m.count;
    SetupBuffer coords, a, b, jac, depth, slope, fade, fadeDer, fail, parts;
    coords.allocate(2 * N * sizeof(double)), a.allocate(N * sizeof(double)), b.allocate(N * sizeof(double));
    jac.allocate(N * sizeof(double));
    depth.upload(m.surfaceDepth.data(), m.surfaceDepth.size() * sizeof(double));
    slope.upload(m.surfaceSlope.data(), m.surfaceSlope.size() * sizeof(double));
    fade.upload(m.fade.data(), m.fade.size() * sizeof(double));
    fadeDer.upload(m.fadeDerivative.data(), m.fadeDerivative.size() * sizeof(double));
    const unsigned long long none = ~0ull;
    fail.upload(&none, sizeof none);
    require(m.surfaceDepth.size() == (size_t)m.nx && m.fade.size() == (size_t)m.nz, "2D geometry profiles do not match the grid");
    const unsigned nb = (unsigned)((N + GB - 1) / GB);
    parts.allocate((size_t)nb * sizeof(Cfl2));
    Geo2 g{m.nx, m.nz, m.offsetX, m.h, depth.as<double>(), slope.as<double>(), fade.as<double>(), fadeDer.as<double>(),
           coords.as<double>(), a.as<double>(), b.as<double>(), jac.as<double>()};
    geometry2d_kernel<<<nb, GB, 0, st>>>(g, fail.as<unsigned long long>(), parts.as<Cfl2>());
    check(hipGetLastError(), "2D geometry launch");
    unsigned long long first = none;
    check(hipMemcpyAsync(&first, fail.ptr, sizeof first, hipMemcpyDeviceToHost, st), "2D geometry");
    check(hipStreamSynchronize(st), "2D geometry");
    std::vector<Cfl2> h(nb);
    m.coordinates.resize(2 * N), m.a.resize(N), m.b.resize(N), m.jac.resize(N);
    check(hipMemcpyAsync(h.data(), parts.ptr, h.size() * sizeof(Cfl2), hipMemcpyDeviceToHost, st), "2D CFL");
    check(hipMemcpyAsync(m.coordinates.data(), coords.ptr, 2 * N * sizeof(double), hipMemcpyDeviceToHost, st), "2D geometry");
    check(hipMemcpyAsync(m.a.data(), a.ptr, N * sizeof(double), hipMemcpyDeviceToHost, st), "2D geometry");
    check(hipMemcpyAsync(m.b.data(), b.ptr, N * sizeof(double), hipMemcpyDeviceToHost, st), "2D geometry");
    check(hipMemcpyAsync(m.jac.data(), jac.ptr, N * sizeof(double), hipMemcpyDeviceToHost, st), "2D geometry");
    check(hipStreamSynchronize(st), "2D geometry");
    Cfl2 r{DBL_MAX, 0.0};
    for (const auto &x : h) r.minJ = std::min(r.minJ, x.minJ), r.maxMetric = std::max(r.maxMetric, x.maxMetric);
    m.minJacobian = r.minJ, m.maxMetric = r.maxMetric;
    return first == none ? -1 : (long long)first;
}

} // namespace swl
