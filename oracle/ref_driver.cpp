// ref_driver.cpp — C-ABI shim over the reference's own public API, compiled together
// with the reference's CPU sources where they lie (see oracle/Makefile, target _ref).
// TEST INFRASTRUCTURE ONLY: it pins the oracle restatement against the real reference
// (include/Seis2D/CurvilinearSolver2D.hpp:93-123, include/Seis3D/CurvilinearSolver.hpp:165-221).
// It contains no reference code; it only fills the reference's parameter structs and
// copies results out through the same function shapes as oracle.h (prefix ref2d_/ref3d_).
#include "oracle.h"
#include "Seis2D/CurvilinearSolver2D.hpp"
#include "Seis3D/CurvilinearSolver.hpp"
#include "two_d/CfsPml2D.hpp"
#include "two_d/PointKernels2D.hpp"
#include <algorithm>
#include <cstring>
#include <memory>
#include <string>

namespace S2 = Seis2D::Curvilinear;
namespace S3 = Seis3D::Curvilinear;

namespace {
void copyError(const std::string &e, char *error, int len) {
    if (error && len > 0) {
        std::strncpy(error, e.c_str(), static_cast<size_t>(len) - 1);
        error[len - 1] = 0;
    }
}
S2::Parameters to2d(const orc_params2d &o) {
    S2::Parameters p;
    p.precision = o.precision ? S2::Precision::Float32 : S2::Precision::Float64;
    p.nx = o.nx, p.nz = o.nz, p.spacing = o.spacing, p.vp = o.vp, p.vs = o.vs, p.density = o.density;
    p.relief = o.relief, p.wavelength = o.wavelength, p.flattenDepth = o.flatten_depth, p.frequency = o.frequency;
    p.duration = o.duration, p.cfl = o.cfl, p.forcePerLength = o.force_per_length;
    p.sourceX = o.source_x, p.sourceZ = o.source_z, p.maxSteps = o.max_steps;
    p.source.kind = o.source_kind ? SeisWave::SourceKind::IsotropicMoment : SeisWave::SourceKind::DirectionalForce;
    p.source.direction = {o.source_direction[0], o.source_direction[1], o.source_direction[2]};
    p.momentPerLength = o.moment_per_length, p.sourceSigmaM = o.source_sigma_m, p.sourceRadiusM = o.source_radius_m;
    p.absorbing = o.absorbing, p.spongeX = o.sponge_x, p.spongeZ = o.sponge_z, p.spongeStrength = o.sponge_strength;
    p.absorber = o.absorber ? S2::Absorber::CfsPml : S2::Absorber::Sponge;
    p.cpml.targetReflection = o.pml_reflection, p.cpml.kappaMax = o.pml_kappa_max, p.cpml.frequency = o.pml_frequency;
    p.externalAbsorption = o.external_absorption;
    p.receiverStart = o.receiver_start, p.receiverCount = o.receiver_count, p.receiverStride = o.receiver_stride;
    p.receiverDepth = o.receiver_depth, p.sampleIntervalMs = o.sample_interval_ms;
    p.autoSurfaceReceivers = false;
    p.layered = o.layered;
    p.layers.clear();
    for (int l = 0; l < o.n_layers; ++l)
        p.layers.push_back({o.layer_thickness[l], o.layer_vp[l], o.layer_vs[l], o.layer_density[l], o.layer_qp[l], o.layer_qs[l]});
    if (!o.layered && p.layers.empty()) p.layers = S2::Parameters{}.layers;
    p.attenuation.enabled = o.attenuation, p.attenuation.mechanisms = o.mechanisms;
    p.attenuation.minFrequency = o.q_min_frequency, p.attenuation.maxFrequency = o.q_max_frequency;
    p.attenuation.referenceFrequency = o.q_reference_frequency, p.attenuation.qp = o.qp, p.attenuation.qs = o.qs;
    p.importedSurface = o.imported_surface, p.surfaceDatum = o.surface_datum;
    p.surfaceSamples.clear();
    for (int i = 0; i < o.n_surface; ++i) p.surfaceSamples.push_back({o.surface_xz[2 * i], o.surface_xz[2 * i + 1]});
    return p;
}
} // namespace

extern "C" {
void *ref2d_create(const orc_params2d *o, char *error, int len) {
    if (o->node_vp) {
        copyError("per-node models are not expressible in the reference 2D API", error, len);
        return nullptr;
    }
    auto s = std::make_unique<S2::Simulation>();
    std::string e;
    S3::ComputeSettings3D settings;
    settings.cpuThreads = o->threads;
    if (!s->initialize(to2d(*o), e, S3::BackendKind3D::Cpu, settings)) {
        copyError(e, error, len);
        return nullptr;
    }
    return s.release();
}
void ref2d_destroy(void *h) { delete static_cast<S2::Simulation *>(h); }
int ref2d_advance(void *h, int count, char *error, int len) {
    std::string e;
    const bool ok = static_cast<S2::Simulation *>(h)->advance(count, e);
    copyError(e, error, len);
    return ok;
}
int ref2d_reset(void *h) {
    std::string e;
    return static_cast<S2::Simulation *>(h)->reset(e);
}
int ref2d_step(void *h) { return static_cast<S2::Simulation *>(h)->step(); }
int ref2d_total(void *h) { return static_cast<S2::Simulation *>(h)->totalSteps(); }
double ref2d_dt(void *h) { return static_cast<S2::Simulation *>(h)->dt(); }
double ref2d_record_dt(void *h) { return static_cast<S2::Simulation *>(h)->recordDt(); }
void ref2d_layout(void *h, int64_t *o) {
    auto &s = *static_cast<S2::Simulation *>(h);
    const auto d = S2::Detail::domainLayout(s.parameters());
    o[0] = s.grid().nx, o[1] = s.grid().nz, o[2] = s.grid().modelOffsetX, o[3] = d.widthX, o[4] = d.widthZ;
    o[5] = s.diagnostics().stateFields, o[6] = static_cast<int64_t>(s.diagnostics().cpmlScalars);
    o[7] = static_cast<int64_t>(s.receivers().size()), o[8] = static_cast<int64_t>(s.receivers().front().vx.size());
    o[9] = s.parameters().attenuation.enabled ? s.parameters().attenuation.mechanisms : 0;
}
void ref2d_diagnostics(void *h, double *o) {
    auto &s = *static_cast<S2::Simulation *>(h);
    const auto &d = s.diagnostics();
    const auto &g = s.grid();
    const double v[] = {d.energy,     d.peakEnergy,  d.maxVelocity, d.tractionRelative, double(d.finite),
                        d.waveDt,     d.dampingDt,   d.sourceDt,    d.cpmlDt,           d.relaxationDt,
                        d.ppw,        d.maxRelativeQError, g.maxMetric, g.minJacobian,   g.maxEdge,
                        g.flattenDepth, d.sourceWeightSum, d.sourceRmsRadiusM};
    std::memcpy(o, v, sizeof v);
}
void ref2d_receivers(void *h, double *vx, double *vz, double *pos, int32_t *xz) {
    auto &s = *static_cast<S2::Simulation *>(h);
    const auto &rs = s.receivers();
    const size_t len = rs.front().vx.size();
    for (size_t r = 0; r < rs.size(); ++r) {
        if (vx) std::copy(rs[r].vx.begin(), rs[r].vx.end(), vx + r * len);
        if (vz) std::copy(rs[r].vz.begin(), rs[r].vz.end(), vz + r * len);
        if (pos) pos[2 * r] = rs[r].position[0], pos[2 * r + 1] = rs[r].position[1];
        if (xz) xz[2 * r] = rs[r].x, xz[2 * r + 1] = rs[r].z;
    }
}
void ref2d_probe_indices(void *h, int64_t *out) {
    auto &s = *static_cast<S2::Simulation *>(h);
    const auto &rs = s.receivers();
    for (size_t r = 0; r < rs.size(); ++r) out[r] = static_cast<int64_t>(s.grid().modelIndex(rs[r].x, rs[r].z));
}
void ref2d_field(void *h, int f, double *out) {
    const auto snap = static_cast<S2::Simulation *>(h)->snapshot(true);
    const std::vector<double> *v = f < 5 ? &snap.fields[f] : f == 5 ? &snap.syy : &snap.memory[f - 6];
    std::copy(v->begin(), v->end(), out);
}
void ref2d_pml(void *h, double *out) {
    const auto snap = static_cast<S2::Simulation *>(h)->snapshot(true);
    std::copy(snap.boundaryMemory.begin(), snap.boundaryMemory.end(), out);
}
int ref2d_model(void *h, int which, double *out) {
    const auto &g = static_cast<S2::Simulation *>(h)->grid();
    const auto put = [&](const auto &a) {
        const auto v = a.toDouble();
        std::copy(v.begin(), v.end(), out);
        return 1;
    };
    switch (which) {
    case 0: return put(g.metricA);
    case 1: return put(g.metricB);
    case 2: return put(g.jacobian);
    case 5: return put(g.density);
    case 10: return put(g.normals);
    case 11: return put(g.coordinates);
    case 12: return put(g.vp);
    case 13: return put(g.vs);
    default: return 0;
    }
}
int64_t ref2d_pml_slot(int nx_pad, int nz_pad, int width_x, int width_z, int64_t q, int axis) {
    S2::Point::MeshT<double> m;
    m.nx = nx_pad, m.nz = nz_pad, m.cpmlWidthX = width_x, m.cpmlWidthZ = width_z;
    m.cpmlCount = 4 * (2 * static_cast<size_t>(width_x) * (nz_pad - 6) + static_cast<size_t>(width_z) * (nx_pad - 6));
    return S2::Point::pmlSlot(m, static_cast<size_t>(q), axis);
}
} // extern "C"
