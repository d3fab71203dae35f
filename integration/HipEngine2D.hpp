// HipEngine2D.hpp — the reference-side 2D engine adapter over libswl's C-ABI (include/swl.h).
//
// Drop-in for Seis2D::Curvilinear::CudaEngine (src/two_d/CudaEngine2D.hpp:5-24): same constructors, same methods,
// same exceptions on failure ("HIP <action>: <message>", as CudaSupport.cuh:6-9 throws "CUDA ..."), so the façade's
// call sites (CurvilinearSolver2D.cpp:515-519 create, :557-562 step, :580-581 rollback, :614-615 reset, :116
// diagnostics, :126-127 sample, :684-686 frame, :703-704 snapshot, :718 probes, :530 residentBytes) only change
// the type they construct.  It lives beside CudaEngine2D.hpp in src/two_d/ and needs that directory and include/
// on the include path, plus include/swl.h of this library; link with libswl.so.
//
// Two ways in, as the reference has two CudaEngine constructors:
//  * HipEngine(const Point::Mesh&, probes, device, statePrecision): the FP64 mesh.  With
//    statePrecision == Precision::Float32 the engine keeps its RK state in FP32 but the metrics, the surface closure
//    and the projection rows in FP64, from this mesh; the façade builds that FP64 mesh for a HIP FP32 session by
//    skipping its FP32 conversion of the model arrays (CurvilinearSolver2D.cpp:498-503) when the backend is HIP.
//  * HipEngine(const Point::MeshT<float>&, probes, device): the reference's own FP32 mesh, every array already
//    rounded to FP32 (CurvilinearSolver2D.cpp:500-513).  The arrays are widened back to FP64 for the descriptor, so
//    the FP64 closure rows see the FP32-rounded metrics: the reference's FP32 model, stepped with FP32 state.
//
// Receivers: the façade passes `s.probes`, whose x already includes the external-absorber offset
// (CurvilinearSolver2D.cpp:495-496); the storage index is Point::index(mesh, r.x + 3, r.z + 3) as in
// CudaEngine2D.cu:139, with no further offset.
#pragma once
#include "PointKernels2D.hpp"
#include "Seis2D/CurvilinearSolver2D.hpp"
#include "swl.h"

#include <array>
#include <cstddef>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

namespace Seis2D::Curvilinear {

class HipEngine {
  public:
    HipEngine(const Point::Mesh &mesh, const std::vector<Receiver> &probes, int device,
              Precision statePrecision = Precision::Float64) {
        create(mesh, probes, device, statePrecision == Precision::Float32 ? 1 : 0);
    }
    HipEngine(const Point::MeshT<float> &mesh, const std::vector<Receiver> &probes, int device) {
        create(mesh, probes, device, 1);
    }
    ~HipEngine() {
        if (engine_) swl_engine2d_destroy(engine_);
    }
    HipEngine(const HipEngine &) = delete;
    HipEngine &operator=(const HipEngine &) = delete;

    // CudaEngine::step(time, dt, frequency, scale, stepNumber): one RK4 step, stages at t, t+dt/2, t+dt/2, t+dt
    void step(double time, double dt, double frequency, double scale, int stepNumber) {
        ok(swl_engine2d_step(engine_, time, dt, frequency, scale, stepNumber), "step");
    }
    void rollback() { ok(swl_engine2d_rollback(engine_), "rollback"); }
    void reset() { ok(swl_engine2d_reset(engine_), "reset"); }
    // fills what CudaEngine fills: energy, max velocity, traction ratio, finiteness, kernel registers / local bytes
    void diagnostics(Diagnostics &d) {
        swl_diag2d g{};
        ok(swl_engine2d_diagnostics(engine_, &g), "diagnostics");
        d.energy = g.energy, d.maxVelocity = g.max_velocity, d.tractionRelative = g.traction_relative;
        d.finite = g.finite != 0;
        d.rhsRegisters = g.rhs_registers, d.rhsLocalBytes = static_cast<std::size_t>(g.rhs_local_bytes);
    }
    // appends the raw Vx, Vz of every probe, as CudaEngine::sample
    void sample(std::vector<Receiver> &receivers) {
        samples_.resize(2 * receivers.size());
        if (!receivers.empty()) ok(swl_engine2d_sample(engine_, samples_.data()), "receiver readback");
        for (std::size_t r = 0; r < receivers.size(); ++r) {
            receivers[r].vx.push_back(samples_[2 * r]);
            receivers[r].vz.push_back(samples_[2 * r + 1]);
        }
    }
    // component 0 Vz, 1 Vx, 2 |V|, 3 pressure over nx x nz physical nodes starting at column offsetX
    void frame(int component, std::vector<float> &values, int nx, int nz, int offsetX) const {
        values.resize(static_cast<std::size_t>(nx) * nz);
        ok(swl_engine2d_frame(engine_, component, values.data(), nx, nz, offsetX), "display readback");
    }
    // Vx, Vz, Sxx, Szz, Sxz (and Syy, the GSLS memories, the CFS-PML memories) in padded storage, promoted to double
    Snapshot snapshot(bool includeMemories = false) const {
        Snapshot s;
        for (int f = 0; f < 5; ++f) s.fields[f] = download(f);
        if (mechanisms_) {
            s.syy = download(5);
            if (includeMemories)
                for (int f = 6; f < fieldCount_; ++f) s.memory.push_back(download(f));
        }
        if (includeMemories && cpmlCount_) {
            s.boundaryMemory.resize(cpmlCount_);
            ok(swl_engine2d_pml(engine_, s.boundaryMemory.data()), "PML snapshot");
        }
        return s;
    }
    void probes(const std::vector<std::size_t> &indices, std::vector<std::array<double, 5>> &out) const {
        out.resize(indices.size());
        if (indices.empty()) return;
        std::vector<std::int64_t> q(indices.begin(), indices.end());
        ok(swl_engine2d_probes(engine_, q.data(), static_cast<int>(q.size()), out.front().data()), "probe readback");
    }
    std::size_t residentBytes() const { return static_cast<std::size_t>(swl_engine2d_resident_bytes(engine_)); }

  private:
    swl_engine2d *engine_ = nullptr;
    int nx_ = 0, nz_ = 0, mechanisms_ = 0, fieldCount_ = 5;
    std::size_t cpmlCount_ = 0;
    std::vector<double> samples_;

    static void ok(int status, const char *action) {
        if (status) throw std::runtime_error(std::string("HIP ") + action + ": " + swl_last_error());
    }
    std::vector<double> download(int field) const {
        std::vector<double> v(static_cast<std::size_t>(nx_) * nz_);
        ok(swl_engine2d_snapshot(engine_, field, v.data()), "snapshot download");
        return v;
    }

    template <class Real>
    void create(const Point::MeshT<Real> &m, const std::vector<Receiver> &probes, int device, int precision) {
        const std::size_t n = static_cast<std::size_t>(m.nx) * m.nz;
        // FP64 views of the mesh arrays: the FP64 mesh's own arrays, or widened copies of an FP32 mesh's
        std::vector<std::vector<double>> widened;
        const auto view = [&](const Real *a, std::size_t count) -> const double * {
            if (!a) return nullptr;
            if constexpr (std::is_same_v<Real, double>) {
                return a;
            } else {
                widened.emplace_back(a, a + count);
                return widened.back().data();
            }
        };
        widened.reserve(24);
        swl_mesh2d_desc d{};
        d.precision = precision;
        d.nx = m.nx, d.nz = m.nz, d.h = static_cast<double>(m.h);
        d.a = view(m.a, n), d.b = view(m.b, n), d.jac = view(m.jac, n), d.normals = view(m.normals, 2 * n);
        d.lambda = view(m.lambda, n), d.mu = view(m.mu, n), d.rho = view(m.rho, n);
        d.source = view(m.source, n), d.damping = view(m.damping, n);
        for (int a = 0; a < 3; ++a) d.source_direction[a] = static_cast<double>(m.sourceDirection[a]);
        for (int a = 0; a < 2; ++a) d.source_gradient[a] = view(m.sourceGradient[a], n);
        d.mechanisms = m.mechanisms, d.field_count = m.fieldCount;
        if (m.cpmlCount) {
            d.cpml_x = view(m.cpmlX, 3 * static_cast<std::size_t>(m.nx - 6));
            d.cpml_z = view(m.cpmlZ, 3 * static_cast<std::size_t>(m.nz - 6));
            d.cpml_width_x = m.cpmlWidthX, d.cpml_width_z = m.cpmlWidthZ;
            d.cpml_count = static_cast<std::int64_t>(m.cpmlCount);
        }
        for (int l = 0; l < m.mechanisms; ++l) {
            d.rates[l] = static_cast<double>(m.rates[l]);
            d.delta_p[l] = view(m.deltaP[l], n), d.delta_s[l] = view(m.deltaS[l], n);
        }
        std::vector<std::int64_t> index; // CudaEngine2D.cu:139
        index.reserve(probes.size());
        for (const auto &r : probes) index.push_back(static_cast<std::int64_t>(Point::index(m, r.x + 3, r.z + 3)));
        ok(swl_engine2d_create(&d, index.data(), static_cast<int>(index.size()), device, &engine_), "engine creation");
        nx_ = m.nx, nz_ = m.nz, mechanisms_ = m.mechanisms, fieldCount_ = m.fieldCount, cpmlCount_ = m.cpmlCount;
    }
};

} // namespace Seis2D::Curvilinear
