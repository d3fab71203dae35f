// sim3d.cpp — 3D session façade over the HIP engine (ISimulationBackend3D,
// include/Seis3D/CurvilinearSolver.hpp:165-185; behaviour of CpuSimulation3D,
// src/three_d/CurvilinearSolver.cpp:301-592).  initialize() is transactional; advance() keeps the
// reference's per-step semantics (inspection, source-window peak, fault rule with rollback and a
// latched fault, receiver samples every step with t = 0 included) while the engine runs chunks of
// steps on the device.
#include "session3d.hpp"
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>


extern "C" {
int swl_sim3d_create(const swl_params3d *p, int backend, int device, int precision, swl_sim3d **out) {
    return swl::guarded([&] {
        swl::require(p && out, "invalid 3D session arguments");
        swl::require(backend == 2, "Only the HIP backend (2) is provided by this library; the CPU path is the reference's own.");
        swl::require(precision == 0 || precision == 1, "3D precision must be 0 (FP64) or 1 (FP32).");
        const auto begin = std::chrono::steady_clock::now();
        auto s = std::make_unique<swl_sim3d>();
        s->model = swl::buildModel3D(*p, device, precision); // FP64 volume geometry and CFL on the device
        s->device = device, s->precision = precision;
        const auto desc = swl::meshDescriptor3D(s->model, precision);
        if (swl_engine3d_create(&desc, s->model.receiverIndex.data(), static_cast<int>(s->model.receiverIndex.size()), device,
                                &s->engine))
            throw std::runtime_error("Backend initialization failed: " + swl::last_error());
        s->clearRecords();
        s->initializationMs = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - begin).count();
        *out = s.release();
    });
}
int swl_memory_estimate3d(const swl_params3d *p, int precision, double *out) {
    return swl::guarded([&] {
        swl::require(p && out && (precision == 0 || precision == 1), "invalid memory estimate arguments");
        const auto m = swl::estimateModelMemory3D(*p);
        out[0] = (double)m.points, out[1] = m.cpuMiB, out[2] = m.deviceMiB, out[3] = m.cudaHostMiB;
        out[4] = m.withinBudget ? 1 : 0, out[5] = m.points ? swl::hipDeviceBytes3D(*p, precision) : 0;
    });
}
int swl_sim3d_plan(const swl_params3d *p, swl_sim3d **out) {
    return swl::guarded([&] {
        swl::require(p && out, "invalid 3D session arguments");
        auto s = std::make_unique<swl_sim3d>();
        s->model = swl::buildModel3D(*p);
        s->clearRecords();
        *out = s.release();
    });
}
void swl_sim3d_destroy(swl_sim3d *s) { delete s; }
int swl_sim3d_advance(swl_sim3d *s, int count) {
    return swl::guarded([&] {
        swl::require(s->engine != nullptr, "This session was planned without a device; create it on the HIP backend to step.");
        s->advance(count);
    });
}
int swl_sim3d_reset(swl_sim3d *s) {
    return swl::guarded([&] {
        if (s->engine && swl_engine3d_reset(s->engine)) throw std::runtime_error(swl::last_error());
        s->step = 0, s->sourcePeak = 0, s->peak = 0, s->fault.clear();
        s->energy = s->maxVelocity = s->traction = 0, s->finite = true;
        std::fill(s->display, s->display + 5, 0.0);
        s->totalStepMs = 0;
        s->clearRecords();
    });
}
int swl_sim3d_step(const swl_sim3d *s) { return s->step; }
int swl_sim3d_override_source_peak(swl_sim3d *s, double peak) {
    return swl::guarded([&] {
        swl::require(std::isfinite(peak) && peak >= 0, "source-window peak energy must be finite and non-negative");
        s->sourcePeak = peak;
    });
}
int swl_sim3d_total_steps(const swl_sim3d *s) { return s->model.total; }
double swl_sim3d_dt(const swl_sim3d *s) { return s->model.dt; }
int swl_sim3d_layout(const swl_sim3d *s, int64_t *out) {
    out[0] = s->model.nx, out[1] = s->model.ny, out[2] = s->model.nz;
    out[3] = static_cast<int64_t>(s->model.receiverIndex.size());
    out[4] = s->vz.empty() ? 0 : static_cast<int64_t>(s->vz.front().size());
    return 0;
}
int swl_sim3d_diagnostics(const swl_sim3d *s, double *o) {
    const auto &m = s->model;
    const double v[21] = {s->energy,       s->peak,        s->sourcePeak,    s->maxVelocity, s->traction,
                          s->finite ? 1.0 : 0.0, m.minJacobian, m.maxMetricNorm, m.minLength, m.maxEdge,
                          m.waveDt,        m.dampingDt,    m.sourceDt,       s->display[0],  s->display[1],
                          s->display[2],   s->display[3],  s->display[4],    s->initializationMs, s->totalStepMs,
                          s->engine ? static_cast<double>(swl_engine3d_resident_bytes(s->engine)) : 0.0};
    std::memcpy(o, v, sizeof v);
    return 0;
}
int swl_sim3d_receivers(const swl_sim3d *s, double *vx, double *vy, double *vz, double *position, int64_t *index) {
    const size_t n = s->vz.empty() ? 0 : s->vz.front().size();
    for (size_t r = 0; r < s->vz.size(); ++r) {
        if (vx) std::copy(s->vx[r].begin(), s->vx[r].end(), vx + r * n);
        if (vy) std::copy(s->vy[r].begin(), s->vy[r].end(), vy + r * n);
        if (vz) std::copy(s->vz[r].begin(), s->vz[r].end(), vz + r * n);
        if (position) std::copy(&s->model.receiverPosition[3 * r], &s->model.receiverPosition[3 * r] + 3, position + 3 * r);
        if (index) index[r] = s->model.receiverIndex[r];
    }
    return 0;
}
int swl_sim3d_snapshot(const swl_sim3d *s, int field, double *out) {
    return swl::guarded([&] {
        swl::require(s->engine != nullptr, "This session was planned without a device.");
        if (swl_engine3d_download(s->engine, field, out)) throw std::runtime_error(swl::last_error());
    });
}
int swl_sim3d_model(const swl_sim3d *s, int which, double *out) {
    return swl::guarded([&] {
        const auto &m = s->model;
        const auto put = [&](const std::vector<double> &v) { std::copy(v.begin(), v.end(), out); };
        switch (which) {
        case 0: put(m.metrics); break;
        case 1: put(m.jacobian); break;
        case 2: put(m.normals); break;
        case 3: put(m.coordinates); break;
        case 4: put(m.lambdaNode); break;
        case 5: put(m.muNode); break;
        case 6: put(m.densityNode); break;
        case 7: put(m.damping); break;
        case 8: put(m.source); break;
        case 9:
        case 10:
        case 11:
            if (m.gradient[which - 9].empty()) std::fill(out, out + m.count, 0.0);
            else put(m.gradient[which - 9]);
            break;
        default: throw std::runtime_error("unknown 3D model array");
        }
    });
}
int swl_sim3d_read_display_frame(const swl_sim3d *s, int x, int y, int z, int component, float *out) {
    return swl::guarded([&] {
        swl::require(s->engine != nullptr, "Initialize a model before requesting slices.");
        if (!s->finite) throw std::runtime_error("Field is outside the display buffer range.");
        if (swl_engine3d_slices(s->engine, x, y, z, component, out)) throw std::runtime_error(swl::last_error());
    });
}
int swl_sim3d_profile(swl_sim3d *s, int enable) {
    return swl::guarded([&] {
        swl::require(s->engine != nullptr, "no device session");
        if (swl_engine3d_profile(s->engine, enable)) throw std::runtime_error(swl::last_error());
    });
}
int swl_sim3d_set_option(swl_sim3d *s, const char *name, int value) {
    return swl::guarded([&] {
        swl::require(s->engine != nullptr, "no device session");
        if (swl_engine3d_set_option(s->engine, name, value)) throw std::runtime_error(swl::last_error());
    });
}
int swl_sim3d_kernel_timing(const swl_sim3d *s, double *out) {
    return swl::guarded([&] {
        swl::require(s->engine != nullptr, "no device session");
        if (swl_engine3d_kernel_timing(s->engine, out)) throw std::runtime_error(swl::last_error());
    });
}
} // extern "C"
