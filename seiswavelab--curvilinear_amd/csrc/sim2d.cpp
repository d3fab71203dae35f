// sim2d.cpp — 2D session façade over the HIP engine (Seis2D::Curvilinear::Simulation,
// include/Seis2D/CurvilinearSolver2D.hpp:93-123; behaviour of src/two_d/CurvilinearSolver2D.cpp).
// initialize() is transactional (a failed build leaves no session), advance() keeps the
// reference's per-step semantics (inspection, fault rule with rollback and latched fault,
// receiver lerp resampling) while the engine runs chunks of steps on the device.
#include "session2d.hpp"
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <locale>
#include <memory>

namespace swl {
std::string &last_error() {
    thread_local std::string message;
    return message;
}
std::mutex &device_setup_mutex() {
    static std::mutex m;
    return m;
}
} // namespace swl

namespace {
int64_t layoutValue(const swl_sim2d &s, int i) {
    const auto &m = s.model;
    switch (i) {
    case 0: return m.nx;
    case 1: return m.nz;
    case 2: return m.offsetX;
    case 3: return std::min(m.p.sponge_x, m.p.nx / 3);
    case 4: return std::min(m.p.sponge_z, m.p.nz / 3);
    case 5: return m.fields;
    case 6: return m.cpmlCount;
    case 7: return static_cast<int64_t>(m.receiverX.size());
    case 8: return static_cast<int64_t>(s.vx.empty() ? 0 : s.vx.front().size());
    default: return m.mechanisms;
    }
}
} // namespace

extern "C" {
const char *swl_last_error(void) { return swl::last_error().c_str(); }

int swl_sim2d_create(const swl_params2d *p, int backend, int device, swl_sim2d **out) {
    return swl::guarded([&] {
        swl::require(p && out, "invalid 2D session arguments");
        swl::require(backend == 2, "Only the HIP backend (2) is provided by this library; the CPU path is the reference's own.");
        const auto begin = std::chrono::steady_clock::now();
        auto s = std::make_unique<swl_sim2d>();
        s->model = swl::buildModel2D(*p, device);
        s->device = device;
        const auto desc = swl::meshDescriptor(s->model);
        if (swl_engine2d_create(&desc, s->model.probeIndex.data(), static_cast<int>(s->model.probeIndex.size()), device,
                                &s->engine))
            throw std::runtime_error(swl::last_error());
        s->clearRecords();
        s->initializationMs = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - begin).count();
        *out = s.release();
    });
}
int swl_sim2d_plan(const swl_params2d *p, swl_sim2d **out) {
    return swl::guarded([&] {
        swl::require(p && out, "invalid 2D session arguments");
        auto s = std::make_unique<swl_sim2d>();
        s->model = swl::buildModel2D(*p);
        s->clearRecords();
        *out = s.release();
    });
}
void swl_sim2d_destroy(swl_sim2d *s) { delete s; }
int swl_sim2d_advance(swl_sim2d *s, int count) {
    return swl::guarded([&] {
        swl::require(s->engine != nullptr, "This session was planned without a device; create it on the HIP backend to step.");
        s->advance(count);
    });
}
int swl_sim2d_reset(swl_sim2d *s) {
    return swl::guarded([&] {
        if (s->engine && swl_engine2d_reset(s->engine)) throw std::runtime_error(swl::last_error());
        s->step = 0;
        s->fault.clear();
        s->energy = s->peak = s->maxVelocity = s->traction = s->totalStepMs = s->lastStepMs = 0;
        s->finite = true;
        s->clearRecords();
    });
}
int swl_sim2d_step(const swl_sim2d *s) { return s->step; }
int swl_sim2d_parameters(const swl_sim2d *s, swl_params2d *out) {
    *out = s->model.logical;
    out->node_vp = out->node_vs = out->node_density = nullptr;                             // not echoed
    out->surface_xz = s->model.surfaceSamples.empty() ? nullptr : s->model.surfaceSamples.data(); // the session's copy
    return 0;
}
int swl_sim2d_override_peak_energy(swl_sim2d *s, double peak) {
    return swl::guarded([&] {
        swl::require(std::isfinite(peak) && peak >= 0, "peak energy must be finite and non-negative");
        s->peak = peak;
    });
}
int swl_sim2d_total_steps(const swl_sim2d *s) { return s->model.total; }
double swl_sim2d_dt(const swl_sim2d *s) { return s->model.dt; }
double swl_sim2d_record_dt(const swl_sim2d *s) { return s->recordDt(); }
int swl_sim2d_layout(const swl_sim2d *s, int64_t *out) {
    for (int i = 0; i < 10; ++i) out[i] = layoutValue(*s, i);
    return 0;
}
int swl_sim2d_diagnostics(const swl_sim2d *s, double *o) {
    const auto &m = s->model;
    const double v[] = {s->energy,     s->peak,      s->maxVelocity, s->traction,    s->finite ? 1. : 0., m.waveDt,
                        m.dampingDt,   m.sourceDt,   m.cpmlDt,       m.relaxationDt, m.ppw,              m.maxQError,
                        m.maxMetric,   m.minJacobian, m.maxEdge,     m.flatten,      m.sourceWeightSum,  m.sourceRmsRadius,
                        s->initializationMs, s->totalStepMs, s->lastStepMs,
                        s->engine ? static_cast<double>(swl_engine2d_resident_bytes(s->engine)) : 0.};
    std::memcpy(o, v, sizeof v);
    return 0;
}
int swl_sim2d_receivers(const swl_sim2d *s, double *vx, double *vz, double *position, int32_t *xz) {
    const size_t len = s->vx.empty() ? 0 : s->vx.front().size();
    for (size_t r = 0; r < s->vx.size(); ++r) {
        if (vx) std::copy(s->vx[r].begin(), s->vx[r].end(), vx + r * len);
        if (vz) std::copy(s->vz[r].begin(), s->vz[r].end(), vz + r * len);
        if (position) position[2 * r] = s->model.receiverPosition[2 * r], position[2 * r + 1] = s->model.receiverPosition[2 * r + 1];
        if (xz) xz[2 * r] = s->model.receiverX[r], xz[2 * r + 1] = s->model.receiverZ[r];
    }
    return 0;
}
int swl_sim2d_probe_indices(const swl_sim2d *s, int64_t *out) {
    std::copy(s->model.probeIndex.begin(), s->model.probeIndex.end(), out);
    return 0;
}
int swl_sim2d_snapshot(const swl_sim2d *s, int field, double *out) {
    return swl_engine2d_snapshot(s->engine, field, out);
}
int swl_sim2d_pml(const swl_sim2d *s, double *out) { return swl_engine2d_pml(s->engine, out); }
int swl_sim2d_profile(swl_sim2d *s, int enable) {
    if (!s->engine) return swl::guarded([] { throw std::runtime_error("no engine"); });
    return swl_engine2d_profile(s->engine, enable);
}
int swl_sim2d_kernel_timing(const swl_sim2d *s, double *out) {
    if (!s->engine) return swl::guarded([] { throw std::runtime_error("no engine"); });
    return swl_engine2d_kernel_timing(s->engine, out);
}
int swl_sim2d_set_option(swl_sim2d *s, const char *name, int value) {
    if (!s->engine) return swl::guarded([] { throw std::runtime_error("no engine"); });
    return swl_engine2d_set_option(s->engine, name, value);
}
int swl_sim2d_model(const swl_sim2d *s, int which, double *out) {
    return swl::guarded([&] {
        const auto &m = s->model;
        const std::vector<double> *v = nullptr;
        switch (which) {
        case 0: v = &m.a; break;
        case 1: v = &m.b; break;
        case 2: v = &m.jac; break;
        case 3: v = &m.lambda; break;
        case 4: v = &m.mu; break;
        case 5: v = &m.rho; break;
        case 6: v = &m.source; break;
        case 7: v = &m.damping; break;
        case 8: v = &m.gradX; break;
        case 9: v = &m.gradZ; break;
        case 10: v = &m.normals; break;
        case 11: v = &m.coordinates; break;
        case 12: v = &m.vp; break;
        case 13: v = &m.vs; break;
        case 14: v = &m.cpmlX; break;
        case 15: v = &m.cpmlZ; break;
        default:
            if (which >= 16 && which < 19) v = &m.deltaP[which - 16];
            else if (which >= 19 && which < 22) v = &m.deltaS[which - 19];
        }
        swl::require(v && !v->empty(), "model array not present in this session");
        std::copy(v->begin(), v->end(), out);
    });
}
int swl_sim2d_read_frame(const swl_sim2d *s, int component, float *out, int include_absorber) {
    const auto &m = s->model;
    const int nx = include_absorber ? m.p.nx : m.logical.nx, nz = include_absorber ? m.p.nz : m.logical.nz;
    return swl_engine2d_frame(s->engine, component, out, nx, nz, include_absorber ? 0 : m.offsetX);
}
int swl_sim2d_read_probes(const swl_sim2d *s, const int32_t *xz, int count, double *out) {
    return swl::guarded([&] {
        swl::require(count >= 0 && count <= 1000000, "Too many probe nodes.");
        std::vector<int64_t> idx(count);
        for (int n = 0; n < count; ++n) {
            swl::require(xz[2 * n] >= 0 && xz[2 * n] < s->model.logical.nx && xz[2 * n + 1] >= 0 &&
                             xz[2 * n + 1] < s->model.logical.nz,
                         "Probe outside logical model.");
            idx[n] = static_cast<int64_t>(xz[2 * n] + 3 + s->model.offsetX) + static_cast<int64_t>(s->model.nx) * (xz[2 * n + 1] + 3);
        }
        if (count && swl_engine2d_probes(s->engine, idx.data(), count, out)) throw std::runtime_error(swl::last_error());
    });
}
int swl_sim2d_write_receivers(const swl_sim2d *s, const char *path) {
    return swl::guarded([&] {
        // receiver CSV with provenance header (CurvilinearSolver2D.cpp:811-854)
        swl::require(s != nullptr, "Initialize a simulation before exporting traces.");
        const std::filesystem::path file(path ? path : "");
        swl::require(!file.empty() && !file.filename().empty(), "Choose a receiver CSV file name.");
        if (!file.parent_path().empty()) std::filesystem::create_directories(file.parent_path());
        std::ofstream out(file);
        swl::require(static_cast<bool>(out), "Cannot create 2D receiver CSV.");
        out.imbue(std::locale::classic());
        const auto &m = s->model;
        const auto &p = m.logical;
        const char *prec = p.precision == 1 ? "FP32" : "FP64";
        swl_device_info info{};
        swl_device_query(s->device, &info);
        out << std::setprecision(17) << "# P-SV plane strain; x,z in m; velocity in m/s; t=0 included\n# dt=" << s->recordDt()
            << "\n# solver_dt=" << m.dt << "\n# backend=HIP " << info.name;
        out << "\n# calculation_precision=" << prec << "; model_precision=" << prec
            << "; record_storage=Float64; values_sampled_from=" << prec;
        out << "\n# external_absorption=" << p.external_absorption << "; computational_NXZ=" << m.nx - 6 << ',' << m.nz - 6;
        out << "\n# absorption_enabled=" << p.absorbing << "; absorber=" << (p.absorber == 1 ? "CFS-PML_ADE" : "sponge")
            << "; effective_width_XZ=" << std::min(p.sponge_x, p.nx / 3) << ',' << std::min(p.sponge_z, p.nz / 3)
            << "; PML_target_R=" << p.pml_reflection << "; kappa=" << p.pml_kappa_max
            << "; PML_Hz=" << (p.pml_frequency > 0 ? p.pml_frequency : p.frequency);
        out << "\n# GSLS=" << p.attenuation << "; mechanisms=" << p.mechanisms << "; reference_Hz=" << p.q_reference_frequency
            << "; fit_band_Hz=" << p.q_min_frequency << ':' << p.q_max_frequency;
        if (p.layered)
            for (int i = 0; i < p.n_layers; ++i) out << "\n# layer_" << i << "_Qp_Qs=" << p.layer_qp[i] << ',' << p.layer_qs[i];
        else
            out << "; Qp=" << p.qp << "; Qs=" << p.qs;
        out << "\n# source_kind=" << p.source_kind
            << " (0=directional_force,1=isotropic_moment); force_per_length_N_per_m=" << p.force_per_length
            << "; diagonal_moment_per_length_N=" << p.moment_per_length << "; direction_XZ=" << m.direction[0] << ','
            << m.direction[2] << "; Ricker_f0_Hz=" << p.frequency << "\nreceiver,x_m,z_m,time_s,vx_m_s,vz_m_s\n";
        for (size_t r = 0; r < s->vx.size(); ++r)
            for (size_t t = 0; t < s->vx[r].size(); ++t)
                out << r << ',' << m.receiverPosition[2 * r] << ',' << m.receiverPosition[2 * r + 1] << ','
                    << t * s->recordDt() << ',' << s->vx[r][t] << ',' << s->vz[r][t] << '\n';
        swl::require(static_cast<bool>(out), "2D receiver write failed.");
    });
}
int64_t swl_pml_slot2d(int nx, int nz, int wx, int wz, int64_t q, int axis) {
    const int i = static_cast<int>(q % nx) - 3, k = static_cast<int>(q / nx) - 3, nxc = nx - 6, nzc = nz - 6;
    if (axis == 0) {
        const int j = i < wx ? i : i >= nxc - wx ? wx + i - (nxc - wx) : -1;
        return j < 0 ? -1 : 4 * (static_cast<int64_t>(k) * 2 * wx + j);
    }
    return k < nzc - wz ? -1 : 4 * (static_cast<int64_t>(2 * wx) * nzc + static_cast<int64_t>(k - (nzc - wz)) * nxc + i);
}
int swl_memory_budget2d(const swl_params2d *p, int backend, double *out) {
    return swl::guarded([&] {
        swl::require(p && out && backend >= 0 && backend <= 2, "invalid memory budget arguments");
        const auto b = swl::memoryBudget2D(*p, backend);
        out[0] = b.host, out[1] = b.device;
    });
}
size_t swl_params2d_size(void) { return sizeof(swl_params2d); }
size_t swl_params3d_size(void) { return sizeof(swl_params3d); }

int swl_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}
int swl_device_query(int device, swl_device_info *out) {
    return swl::guarded([&] {
        std::memset(out, 0, sizeof *out);
        out->index = device;
        hipDeviceProp_t prop{};
        swl::check(hipGetDeviceProperties(&prop, device), "device properties");
        std::strncpy(out->name, prop.name, sizeof out->name - 1);
        if (!out->name[0] && hipDeviceGetName(out->name, sizeof out->name - 1, device) != hipSuccess) out->name[0] = 0;
        std::strncpy(out->arch, prop.gcnArchName, sizeof out->arch - 1);
        out->total_bytes = prop.totalGlobalMem;
        out->compute_units = prop.multiProcessorCount;
        if (!out->name[0]) { // no marketing name on this system: name the device by its ISA and size
            const std::string isa(prop.gcnArchName, std::strcspn(prop.gcnArchName, ":"));
            std::snprintf(out->name, sizeof out->name, "AMD %s, %d CUs, %.0f GiB", isa.c_str(), prop.multiProcessorCount,
                          prop.totalGlobalMem / 1073741824.0);
        }
        swl::DeviceScope scope(device);
        size_t freeBytes = 0, total = 0;
        swl::check(hipMemGetInfo(&freeBytes, &total), "memory query");
        out->free_bytes = freeBytes;
        out->available = 1;
    });
}
}
