// HipEngine3D.hpp — the reference-side 3D engine adapter over libswl's C-ABI (include/swl.h).
//
// Drop-in for Seis3D::Curvilinear::CudaEngine3D (src/three_d/CudaEngine3D.hpp:7-24): same constructor, same
// methods, exceptions "HIP <action>: <message>" on failure.  The façade's call sites (CurvilinearSolver.cpp:430-434
// create, :460 step, :482 rollback, :505 resetFields, :216 diagnostics, :490 sample, :519 slices, :564 download,
// :589 residentMemory) only change the type they construct.  It lives beside CudaEngine3D.hpp in src/three_d/ and
// needs that directory and include/ on the include path, plus include/swl.h of this library; link with libswl.so.
//
// The mesh is the reference's FP64 Point::Mesh.  statePrecision Float32 keeps the RK state in FP32 with the
// surface closure, projection and energy in FP64 (the reference's 3D path is FP64 only).  The adapter also
// passes two hints the engine checks against the arrays before using them: the first plane from which the
// metrics are the identity (scanned here), and that the materials may depend on depth only.
#pragma once
#include "PointKernels3D.hpp"
#include "Seis3D/CurvilinearSolver.hpp"
#include "swl.h"

#include <algorithm>
#include <array>
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace Seis3D::Curvilinear {

class HipEngine3D {
  public:
    HipEngine3D(const Point::Mesh &g, const std::vector<ReceiverTrace3D> &receivers, int device, bool fp32State = false)
        : nx_(g.nx), ny_(g.ny), nz_(g.nz), count_(g.count),
          scale_(g.sourceAmplitude / (g.density * g.h * g.h * g.h)) { // the forcing scale of CudaEngine3D::step
        swl_mesh3d_desc d{};
        d.precision = fp32State ? 1 : 0;
        d.nx = g.nx, d.ny = g.ny, d.nz = g.nz, d.h = g.h;
        d.metrics = g.metrics, d.jacobian = g.jacobian, d.normals = g.normals;
        // per-node materials; a homogeneous mesh (no material arrays) repeats its constants
        std::vector<double> lambda, mu, density;
        const auto expand = [&](const double *a, double c, std::vector<double> &v) -> const double * {
            if (a) return a;
            v.assign(g.count, c);
            return v.data();
        };
        d.lambda = expand(g.materialLambda, g.lambda, lambda);
        d.mu = expand(g.materialMu, g.mu, mu);
        d.density = expand(g.materialDensity, g.density, density);
        d.density_global = g.density; // PointKernels3D.hpp:107 scales the source by the global density
        d.damping = g.damping, d.source = g.source;
        for (int a = 0; a < 3; ++a) d.source_gradient[a] = g.sourceGradient[a], d.source_direction[a] = g.sourceDirection[a];
        d.flat_from = identityFrom(g);
        d.material_by_depth = 1; // checked by the engine, which streams per-node arrays when it does not hold
        std::vector<std::int64_t> index;
        index.reserve(receivers.size());
        for (const auto &r : receivers) index.push_back(static_cast<std::int64_t>(r.gridIndex));
        ok(swl_engine3d_create(&d, index.data(), static_cast<int>(index.size()), device, &engine_), "engine creation");
    }
    ~HipEngine3D() {
        if (engine_) swl_engine3d_destroy(engine_);
    }
    HipEngine3D(const HipEngine3D &) = delete;
    HipEngine3D &operator=(const HipEngine3D &) = delete;

    void step(double time, double dt, double frequency, int step) {
        ok(swl_engine3d_step(engine_, time, dt, frequency, scale_, step), "step");
    }
    void rollback() { ok(swl_engine3d_rollback(engine_), "rollback"); }
    void resetFields() { ok(swl_engine3d_reset(engine_), "reset"); }
    // fills what CudaEngine3D::diagnostics fills (CudaEngine3D.cu:127-133)
    void diagnostics(Diagnostics3D &d) {
        double o[10];
        ok(swl_engine3d_diagnostics(engine_, o), "diagnostics");
        d.elasticEnergy = o[0], d.maxAbsVelocity = o[1], d.maxTractionRateRelative = o[2], d.allFinite = o[3] != 0;
        for (int c = 0; c < 5; ++c) d.displayPeaks[c] = o[4 + c]; // Vz, Vx, Vy, |V|, pressure
        d.peakEnergy = std::max(d.peakEnergy, d.elasticEnergy);
    }
    // appends Vz to samples and Vx, Vy to samplesVx, samplesVy, as CudaEngine3D::sample
    void sample(std::vector<ReceiverTrace3D> &receivers) {
        values_.resize(3 * receivers.size());
        if (!receivers.empty()) ok(swl_engine3d_sample(engine_, values_.data()), "receiver readback");
        for (std::size_t r = 0; r < receivers.size(); ++r) {
            receivers[r].samples.push_back(values_[3 * r + 2]);
            receivers[r].samplesVx.push_back(values_[3 * r]);
            receivers[r].samplesVy.push_back(values_[3 * r + 1]);
        }
    }
    // XZ at y, YZ at x, XY at z of the requested component over the physical nodes
    void slices(const DisplayRequest3D &request, DisplayFrame3D &frame) {
        const int px = nx_ - 6, py = ny_ - 6, pz = nz_ - 6;
        const int widths[3] = {px, py, px}, heights[3] = {pz, pz, py};
        std::vector<float> all(static_cast<std::size_t>(px) * pz + static_cast<std::size_t>(py) * pz +
                               static_cast<std::size_t>(px) * py);
        ok(swl_engine3d_slices(engine_, request.x, request.y, request.z, request.component, all.data()), "slices");
        std::size_t at = 0;
        frame.bytes = 0;
        for (int p = 0; p < 3; ++p) {
            auto &plane = frame.planes[p];
            const std::size_t n = static_cast<std::size_t>(widths[p]) * heights[p];
            plane.width = widths[p], plane.height = heights[p];
            plane.values.assign(all.begin() + static_cast<std::ptrdiff_t>(at), all.begin() + static_cast<std::ptrdiff_t>(at + n));
            at += n;
            frame.bytes += n * sizeof(float);
        }
    }
    void download(std::array<std::vector<double>, 9> &fields) const {
        for (int f = 0; f < 9; ++f) {
            fields[f].resize(count_);
            ok(swl_engine3d_download(engine_, f, fields[f].data()), "volume download");
        }
    }
    BackendMemory3D residentMemory() const noexcept {
        BackendMemory3D m;
        if (!engine_) return m;
        m.hostAvailableBytes = sizeof(*this) + values_.capacity() * sizeof(double);
        m.deviceAvailableBytes = swl_engine3d_resident_bytes(engine_);
        return m;
    }

  private:
    swl_engine3d *engine_ = nullptr;
    int nx_ = 0, ny_ = 0, nz_ = 0;
    std::size_t count_ = 0;
    double scale_ = 0;
    std::vector<double> values_;

    static void ok(int status, const char *action) {
        if (status) throw std::runtime_error(std::string("HIP ") + action + ": " + swl_last_error());
    }
    // first padded plane k from which every physical node below has identity metrics and J = 1 (nz when none)
    static int identityFrom(const Point::Mesh &g) {
        int from = g.nz;
        for (int k = g.nz - 4; k >= 3; --k) {
            for (int j = 3; j < g.ny - 3; ++j)
                for (int i = 3; i < g.nx - 3; ++i) {
                    const std::size_t q = Point::index(g, i, j, k);
                    if (g.jacobian[q] != 1.0) return from;
                    for (int a = 0; a < 9; ++a)
                        if (g.metrics[9 * q + a] != (a % 4 == 0 ? 1.0 : 0.0)) return from;
                }
            from = k;
        }
        return from;
    }
};

} // namespace Seis3D::Curvilinear
