// ref_driver.cpp — C-ABI shim over the reference's own public API, compiled together
// with the reference's CPU sources where they lie (see oracle/Makefile, target _ref).
// TEST INFRASTRUCTURE ONLY: it pins the oracle restatement against the real reference
// (include/Seis2D/CurvilinearSolver2D.hpp:93-123, include/Seis3D/CurvilinearSolver.hpp:165-221).
// It contains no reference code; it only fills the reference's parameter structs and
// copies results out through the same function shapes as oracle.h (prefix ref2d_/ref3d_).
#include "oracle.h"
#include "Seis2D/CurvilinearSolver2D.hpp"
#include "Seis3D/AbsorbingBoundary3D.hpp"
#include "Seis3D/CurvilinearSolver.hpp"
#include "Seis3D/ModelPresets3D.hpp"
#include "two_d/CfsPml2D.hpp"
#include "two_d/MemoryBudget2D.hpp"
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
    p.autoSurfaceReceivers = o.auto_surface_receivers;
    p.receiverTargetCount = o.receiver_target_count;
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
S3::SimulationParameters to3d(const orc_params3d *o) {
    S3::SimulationParameters p;
    p.nx = o->nx, p.ny = o->ny, p.nz = o->nz, p.spacing = o->spacing, p.vp = o->vp, p.vs = o->vs, p.density = o->density;
    p.layeredMaterial = o->layered != 0;
    p.layerDepthMode = static_cast<S3::LayerDepthMode3D>(o->layer_depth_mode);
    p.layers.clear();
    for (int l = 0; l < o->n_layers; ++l)
        p.layers.push_back({"layer", o->layer_thickness[l], o->layer_vp[l], o->layer_vs[l], o->layer_density[l]});
    p.dominantFrequency = o->dominant_frequency;
    p.sourceX = o->source_x, p.sourceY = o->source_y, p.sourceZ = o->source_z;
    p.source.kind = o->source_kind ? SeisWave::SourceKind::IsotropicMoment : SeisWave::SourceKind::DirectionalForce;
    p.source.direction = {o->source_direction[0], o->source_direction[1], o->source_direction[2]};
    p.sourceForce = o->source_force, p.sourceMoment = o->source_moment;
    p.duration = o->duration, p.maxSteps = o->max_steps, p.cflSafety = o->cfl;
    p.topographyAmplitude = o->topography_amplitude, p.topographyWavelength = o->topography_wavelength;
    p.terrainTransitionDepth = o->terrain_transition_depth;
    if (o->dem_nx > 0 && o->dem_elevations) {
        auto &g = p.surfaceGrid;
        g.nx = o->dem_nx, g.ny = o->dem_ny, g.spacing = o->dem_spacing, g.originX = o->dem_origin_x;
        g.originY = o->dem_origin_y, g.datum = o->dem_datum;
        g.elevations.assign(o->dem_elevations, o->dem_elevations + static_cast<size_t>(o->dem_nx) * o->dem_ny);
    }
    auto &r = p.receiverArray;
    r.kind = static_cast<S3::ReceiverArrayKind3D>(o->receiver_kind);
    r.startX = o->receiver_start_x, r.startY = o->receiver_start_y, r.z = o->receiver_z;
    r.countX = o->receiver_count_x, r.countY = o->receiver_count_y, r.strideX = o->receiver_stride_x;
    r.strideY = o->receiver_stride_y;
    p.absorbingBoundary.enabled = o->absorbing != 0;
    p.absorbingBoundary.cells = {o->absorbing_cells[0], o->absorbing_cells[1], o->absorbing_cells[2]};
    p.absorbingBoundary.strength = o->absorbing_strength;
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

// ---- 3D: CpuSimulation3D on the CPU backend ----
void *ref3d_create(const orc_params3d *o, char *error, int len) {
    const S3::SimulationParameters p = to3d(o);
    auto s = std::make_unique<S3::CpuSimulation3D>(S3::BackendKind3D::Cpu);
    std::string e;
    if (!s->initialize(p, e, S3::ComputeSettings3D{o->threads, -1})) {
        copyError(e, error, len);
        return nullptr;
    }
    copyError("", error, len);
    return s.release();
}
void ref3d_destroy(void *h) { delete static_cast<S3::CpuSimulation3D *>(h); }
int ref3d_advance(void *h, int count, char *error, int len) {
    std::string e;
    const bool ok = static_cast<S3::CpuSimulation3D *>(h)->advance(count, e);
    copyError(e, error, len);
    return ok ? 1 : 0;
}
int ref3d_step(void *h) { return static_cast<S3::CpuSimulation3D *>(h)->currentStep(); }
int ref3d_total(void *h) { return static_cast<S3::CpuSimulation3D *>(h)->totalSteps(); }
double ref3d_dt(void *h) { return static_cast<S3::CpuSimulation3D *>(h)->dt(); }
void ref3d_layout(void *h, int64_t *o) {
    auto &s = *static_cast<S3::CpuSimulation3D *>(h);
    const auto &g = s.grid();
    o[0] = g.nx, o[1] = g.ny, o[2] = g.nz, o[3] = static_cast<int64_t>(s.receivers().size());
    o[4] = s.receivers().empty() ? 0 : static_cast<int64_t>(s.receivers()[0].samples.size());
}
// same order as orc3d_diagnostics; the source-window peak is private to the reference (0 here)
void ref3d_diagnostics(void *h, double *o) {
    const auto &d = static_cast<S3::CpuSimulation3D *>(h)->diagnostics();
    const double v[18] = {d.elasticEnergy,  d.peakEnergy,     0.0,            d.maxAbsVelocity, d.maxTractionRateRelative,
                          d.allFinite ? 1.0 : 0.0, d.minJacobian, d.maxMetricNorm, d.minCellLength, d.maxCellEdge,
                          d.waveDtLimit,    d.dampingDtLimit, d.sourceDtLimit, d.displayPeaks[0], d.displayPeaks[1],
                          d.displayPeaks[2], d.displayPeaks[3], d.displayPeaks[4]};
    std::copy(v, v + 18, o);
}
void ref3d_receivers(void *h, double *vx, double *vy, double *vz, double *position, int64_t *index) {
    const auto &rs = static_cast<S3::CpuSimulation3D *>(h)->receivers();
    for (size_t r = 0; r < rs.size(); ++r) {
        const size_t n = rs[r].samples.size();
        if (vx) std::copy(rs[r].samplesVx.begin(), rs[r].samplesVx.end(), vx + r * n);
        if (vy) std::copy(rs[r].samplesVy.begin(), rs[r].samplesVy.end(), vy + r * n);
        if (vz) std::copy(rs[r].samples.begin(), rs[r].samples.end(), vz + r * n);
        if (position) std::copy(rs[r].position.begin(), rs[r].position.end(), position + 3 * r);
        if (index) index[r] = static_cast<int64_t>(rs[r].gridIndex);
    }
}
void ref3d_field(void *h, int f, double *out) {
    const auto &v = static_cast<S3::CpuSimulation3D *>(h)->field(static_cast<S3::Field3D>(f));
    std::copy(v.begin(), v.end(), out);
}
// 0 metrics (9N), 1 jacobian, 2 normals (3N), 3 coordinates (3N), 7 sponge rate of every physical
// node (reference absorbingRate3D), 12..14 layered Vp, Vs, density of the export snapshot.
int ref3d_model(void *h, int which, double *out) {
    auto &s = *static_cast<S3::CpuSimulation3D *>(h);
    const auto &g = s.grid();
    const size_t n = g.coordinates.size();
    switch (which) {
    case 0: for (size_t q = 0; q < n; ++q) std::copy(g.inverseMetrics[q].begin(), g.inverseMetrics[q].end(), out + 9 * q); return 1;
    case 1: std::copy(g.jacobian.begin(), g.jacobian.end(), out); return 1;
    case 2: for (size_t q = 0; q < n; ++q) std::copy(g.surfaceNormal[q].begin(), g.surfaceNormal[q].end(), out + 3 * q); return 1;
    case 3: for (size_t q = 0; q < n; ++q) std::copy(g.coordinates[q].begin(), g.coordinates[q].end(), out + 3 * q); return 1;
    case 7: {
        const auto &p = s.parameters();
        double maxVp = p.vp;
        if (p.layeredMaterial) {
            maxVp = 0;
            for (const auto &l : p.layers) maxVp = std::max(maxVp, l.vp);
        }
        std::fill(out, out + n, 0.0);
        for (int k = g.ghost; k < g.nz - g.ghost; ++k)
            for (int j = g.ghost; j < g.ny - g.ghost; ++j)
                for (int i = g.ghost; i < g.nx - g.ghost; ++i)
                    out[g.index(i, j, k)] = S3::absorbingRate3D(p, i - g.ghost, j - g.ghost, k - g.ghost, maxVp);
        return 1;
    }
    case 12:
    case 13:
    case 14: {
        const auto snap = s.snapshot();
        const auto &v = which == 12 ? snap.materialVp : which == 13 ? snap.materialVs : snap.materialDensity;
        if (v.size() != n) return 0;
        std::copy(v.begin(), v.end(), out);
        return 1;
    }
    default: return 0;
    }
}
// Detail::memoryBudget (src/two_d/MemoryBudget2D.hpp:12-38): out[0] host bytes, out[1] device bytes
void ref2d_memory_budget(const orc_params2d *o, int cuda, double *out) {
    const auto b = S2::Detail::memoryBudget(to2d(*o), cuda ? S3::BackendKind3D::Cuda : S3::BackendKind3D::Cpu);
    out[0] = b.hostBytes, out[1] = b.deviceBytes;
}
// the session's receiver line after initialize (automatic surface layout applied): start, count, stride, depth, target
void ref2d_receiver_line(void *h, int *out) {
    const auto &p = static_cast<S2::Simulation *>(h)->parameters();
    out[0] = p.receiverStart, out[1] = p.receiverCount, out[2] = p.receiverStride, out[3] = p.receiverDepth;
    out[4] = p.receiverTargetCount;
}
// estimateModelMemory3D (include/Seis3D/ModelPresets3D.hpp:38-49): points, cpu, device, cuda host MiB, within budget
void ref3d_memory_estimate(const orc_params3d *o, double *out) {
    const auto m = S3::estimateModelMemory3D(to3d(o));
    out[0] = (double)m.points, out[1] = m.cpuMiB, out[2] = m.deviceMiB, out[3] = m.cudaHostMiB, out[4] = m.withinBudget;
}
} // extern "C"

// ---- SEG-Y: the reference's own writer and reader (include/SeisCommon/SegyIo.hpp:5-19) over plain arrays
#include "SeisCommon/SegyIo.hpp"
extern "C" {
// traces: n positions / sources (3 each), start times, samples (n x count, row per trace)
int ref_segy_write(const char *path, const double *pos, const double *src, const double *start, const double *samples,
                   int n, int64_t count, double dt, const char *component, const char *units, int profile,
                   double gain, double interval, char *error, int len) {
    SeisWave::RecordDataView view;
    view.dt = dt, view.units = units, view.components = {component};
    for (int r = 0; r < n; ++r) {
        SeisWave::TraceView t;
        for (int c = 0; c < 3; ++c) t.position[c] = pos[3 * r + c], t.source[c] = src[3 * r + c];
        t.startTime = start[r];
        t.channels[0] = std::span<const double>(samples + (size_t)r * count, (size_t)count);
        view.traces.push_back(t);
    }
    SeisWave::SegyExportOptions o;
    o.profile = static_cast<SeisWave::SegyExportProfile>(profile);
    o.sampleGain = gain, o.sampleIntervalSeconds = interval;
    std::string e;
    const bool ok = SeisWave::writeSegy(view, 0, path, e, o);
    copyError(e, error, len);
    return ok;
}
void *ref_segy_read(const char *path, char *error, int len) {
    auto r = std::make_unique<SeisWave::ImportedRecords>();
    std::string e;
    if (!SeisWave::readSegy(path, *r, e)) {
        copyError(e, error, len);
        return nullptr;
    }
    return r.release();
}
void ref_segy_free(void *h) { delete static_cast<SeisWave::ImportedRecords *>(h); }
int ref_segy_info(void *h, int *n, double *dt, char *component, char *units, char *description, int cap) {
    const auto &r = *static_cast<SeisWave::ImportedRecords *>(h);
    *n = (int)r.traces.size(), *dt = r.dt;
    copyError(r.component, component, cap), copyError(r.units, units, cap), copyError(r.description, description, cap);
    return 1;
}
int ref_segy_trace(void *h, int i, double *pos, double *src, double *start, int64_t *count, double *samples) {
    const auto &t = static_cast<SeisWave::ImportedRecords *>(h)->traces[i];
    for (int c = 0; c < 3; ++c) pos[c] = t.position[c], src[c] = t.source[c];
    *start = t.startTime, *count = (int64_t)t.samples.size();
    if (samples) std::copy(t.samples.begin(), t.samples.end(), samples);
    return 1;
}
} // extern "C"

// ---- VTS: the reference's own model document writer/reader and wavefield snapshot writer
//      (include/Seis2D/ModelIo2D.hpp:8-13, include/Seis2D/CurvilinearSolver2D.hpp:124-140)
#include "Seis2D/ModelIo2D.hpp"
extern "C" {
int ref2d_write_model_vts(void *h, const char *path, int threads, int device, char *error, int len) {
    S3::ComputeSettings3D settings;
    settings.cpuThreads = threads, settings.cudaDevice = device;
    std::string e;
    const bool ok = S2::writeModelVts(*static_cast<S2::Simulation *>(h), path, settings, e);
    copyError(e, error, len);
    return ok;
}
// Fills *out (surface samples copied to `surface`, at most `capacity` pairs) and info[4] = backend kind,
// compute.cpuThreads, compute.cudaDevice, referenceDt.
int ref_model_vts_read(const char *path, orc_params2d *out, double *surface, int capacity, double *info, char *error,
                       int len) {
    S2::ModelDocument d;
    std::string e;
    if (!S2::readModelVts(path, d, e)) {
        copyError(e, error, len);
        return 0;
    }
    const auto &p = d.parameters;
    orc_params2d o{};
    o.precision = p.precision == S2::Precision::Float32;
    o.nx = p.nx, o.nz = p.nz, o.spacing = p.spacing, o.vp = p.vp, o.vs = p.vs, o.density = p.density;
    o.relief = p.relief, o.wavelength = p.wavelength, o.flatten_depth = p.flattenDepth, o.frequency = p.frequency;
    o.duration = p.duration, o.cfl = p.cfl, o.force_per_length = p.forcePerLength;
    o.source_x = p.sourceX, o.source_z = p.sourceZ, o.max_steps = p.maxSteps;
    o.source_kind = p.source.kind == SeisWave::SourceKind::IsotropicMoment;
    for (int c = 0; c < 3; ++c) o.source_direction[c] = p.source.direction[c];
    o.moment_per_length = p.momentPerLength, o.source_sigma_m = p.sourceSigmaM, o.source_radius_m = p.sourceRadiusM;
    o.absorbing = p.absorbing, o.sponge_x = p.spongeX, o.sponge_z = p.spongeZ, o.sponge_strength = p.spongeStrength;
    o.absorber = p.absorber == S2::Absorber::CfsPml;
    o.pml_reflection = p.cpml.targetReflection, o.pml_kappa_max = p.cpml.kappaMax, o.pml_frequency = p.cpml.frequency;
    o.external_absorption = p.externalAbsorption;
    o.receiver_start = p.receiverStart, o.receiver_count = p.receiverCount, o.receiver_stride = p.receiverStride;
    o.receiver_depth = p.receiverDepth, o.sample_interval_ms = p.sampleIntervalMs;
    o.auto_surface_receivers = p.autoSurfaceReceivers, o.receiver_target_count = p.receiverTargetCount;
    o.layered = p.layered, o.n_layers = static_cast<int>(p.layers.size());
    for (int l = 0; l < o.n_layers && l < ORC_MAX_LAYERS; ++l) {
        const auto &L = p.layers[l];
        o.layer_thickness[l] = L.thickness, o.layer_vp[l] = L.vp, o.layer_vs[l] = L.vs, o.layer_density[l] = L.density;
        o.layer_qp[l] = L.qp, o.layer_qs[l] = L.qs;
    }
    const auto &a = p.attenuation;
    o.attenuation = a.enabled, o.mechanisms = a.mechanisms, o.q_min_frequency = a.minFrequency;
    o.q_max_frequency = a.maxFrequency, o.q_reference_frequency = a.referenceFrequency, o.qp = a.qp, o.qs = a.qs;
    o.imported_surface = p.importedSurface, o.surface_datum = p.surfaceDatum;
    o.n_surface = static_cast<int>(p.surfaceSamples.size());
    for (int i = 0; i < o.n_surface && i < capacity; ++i) surface[2 * i] = p.surfaceSamples[i][0], surface[2 * i + 1] = p.surfaceSamples[i][1];
    o.threads = d.compute.cpuThreads;
    *out = o;
    info[0] = static_cast<int>(d.backend), info[1] = d.compute.cpuThreads, info[2] = d.compute.cudaDevice;
    info[3] = d.referenceDt;
    return 1;
}
// components[7]: SnapshotExportOptions::components
int ref2d_write_snapshot(void *h, const char *path, const int *components, int include_material, char *error, int len) {
    S2::SnapshotExportOptions o;
    for (int c = 0; c < 7; ++c) o.components[c] = components[c] != 0;
    o.includeMaterial = include_material != 0;
    std::string e;
    const bool ok = S2::writeSnapshot(*static_cast<S2::Simulation *>(h), path, e, o);
    copyError(e, error, len);
    return ok;
}
// writeReceivers (include/Seis2D/CurvilinearSolver2D.hpp:140): the receiver CSV with its provenance header
int ref2d_write_receivers(void *h, const char *path, char *error, int len) {
    std::string e;
    const bool ok = S2::writeReceivers(*static_cast<S2::Simulation *>(h), path, e);
    copyError(e, error, len);
    return ok;
}
} // extern "C"

// ---- 3D receiver CSV and wavefield VTS: the reference's own writers (include/Seis3D/ReceiverIo.hpp,
//      include/Seis3D/VtsWriter.hpp) on a reference CPU session
#include "Seis3D/ReceiverIo.hpp"
#include "Seis3D/VtsWriter.hpp"
extern "C" {
int ref3d_write_receivers_csv(void *h, const char *path, const int *components, char *error, int len) {
    std::string e;
    const std::array<bool, 3> sel{components[0] != 0, components[1] != 0, components[2] != 0};
    const bool ok = S3::writeReceiversCsv(*static_cast<S3::CpuSimulation3D *>(h), path, sel, e);
    copyError(e, error, len);
    return ok;
}
int ref3d_write_receivers_vz_csv(void *h, const char *path, int all, char *error, int len) {
    std::string e;
    const bool ok = S3::writeReceiversCsv(*static_cast<S3::CpuSimulation3D *>(h), path, e, all != 0);
    copyError(e, error, len);
    return ok;
}
int ref3d_write_vts(void *h, const char *path, const int *fields, int include_material, int include_vector, char *error,
                    int len) {
    S3::VtsExportOptions3D o;
    for (int n = 0; n < 11; ++n) o.fields[n] = fields[n] != 0;
    o.includeMaterial = include_material != 0, o.includeVelocityVector = include_vector != 0;
    std::string e;
    const bool ok = S3::writeVts(static_cast<S3::CpuSimulation3D *>(h)->snapshot(), path, o, e);
    copyError(e, error, len);
    return ok;
}
} // extern "C"
