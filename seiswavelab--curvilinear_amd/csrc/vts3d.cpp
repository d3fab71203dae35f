// vts3d.cpp — 3D receiver CSV and wavefield VTS export.
//
// writeReceiversCsv (include/Seis3D/ReceiverIo.hpp:5-10, src/three_d/ReceiverIo.cpp:9-51): provenance header,
// the source's coordinates and grid indices, receiver positions, then one row per accepted step (t = 0 included)
// with the selected Vx / Vy / Vz of every receiver.
// writeVts (include/Seis3D/VtsWriter.hpp:11-24, src/three_d/VtsWriter.cpp:17-117): a VTK StructuredGrid of the
// physical nodes with the selected fields of the current state (downloaded from the device), the materials
// (per node for layered models, with layer ids), the velocity vector and the curvilinear point coordinates;
// written to `path`.tmp-<ticks> and renamed.
//
// Numbers are the "%.17g" form of the C locale (std::to_chars, general format, precision 17), which is what the
// reference's setprecision(17) streams print; rows of nodes are formatted in parallel and written in order.
#include "session3d.hpp"
#include <charconv>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <locale>
#include <string>
#include <system_error>
#include <vector>

namespace {
using swl::require;

constexpr int kNumberChars = 25; // "-1.2345678901234567e-308"
inline char *put(char *p, double v) { return std::to_chars(p, p + kNumberChars, v, std::chars_format::general, 17).ptr; }
inline char *put(char *p, int v) { return std::to_chars(p, p + 12, v).ptr; }
inline char *put(char *p, const char *s) {
    while (*s) *p++ = *s++;
    return p;
}

// Formats `count` items in parallel: item n is written by emit(n, p) (at most `width` characters) followed by
// `sep` and, after every `perLine`-th item, by `brk`; the text goes out in order.
template <class Emit>
void parallelText(std::ofstream &out, size_t count, int width, const char *sep, int perLine, const char *brk,
                  const Emit &emit) {
    constexpr size_t block = 1 << 16;
    std::vector<std::string> text((count + block - 1) / block);
#pragma omp parallel for schedule(dynamic, 1)
    for (long long b = 0; b < (long long)text.size(); ++b) {
        const size_t n0 = (size_t)b * block, n1 = std::min(count, n0 + block);
        std::string &s = text[b];
        s.resize((n1 - n0) * (width + 16));
        char *p = s.data();
        for (size_t n = n0; n < n1; ++n) {
            p = put(emit(n, p), sep);
            if ((n + 1) % perLine == 0) p = put(p, brk);
        }
        s.resize((size_t)(p - s.data()));
    }
    for (const auto &s : text) out.write(s.data(), (std::streamsize)s.size());
}

// SeisWave::unitSourceDirection (include/SeisCommon/SourceMechanism.hpp:32-38)
std::array<double, 3> unitDirection(const swl_params3d &p) {
    if (p.source_kind == 1) return {0, 0, 0};
    const double n = std::hypot(p.source_direction[0], p.source_direction[1], p.source_direction[2]);
    return {p.source_direction[0] / n, p.source_direction[1] / n, p.source_direction[2] / n};
}

void writeCsv(const swl_sim3d *s, const std::filesystem::path &path, const int32_t *selected, bool legacyVzHeader) {
    bool any = false;
    for (int c = 0; c < 3; ++c) any = any || selected[c] != 0;
    require(any, "Select at least one receiver component.");
    require(s != nullptr, "Initialize a simulation before exporting traces.");
    const char *names[] = {"Vx", "Vy", "Vz"};
    std::error_code ec;
    if (!path.parent_path().empty()) std::filesystem::create_directories(path.parent_path(), ec);
    if (ec) throw std::runtime_error(ec.message());
    std::ofstream out(path);
    require(static_cast<bool>(out), "Could not open receiver CSV.");
    out.imbue(std::locale::classic());
    out << std::setprecision(17) << "# ";
    bool comma = false;
    for (int c = 0; c < 3; ++c)
        if (selected[c]) {
            if (comma) out << ',';
            out << names[c];
            comma = true;
        }
    out << " in m/s; time in seconds; coordinates in metres; depth positive down\n";
    const auto &m = s->model;
    const auto &p = m.p;
    const auto d = unitDirection(p);
    out << "# source_kind=" << p.source_kind << " (0=directional_force,1=isotropic_moment); force_N=" << p.source_force
        << "; diagonal_moment_Nm=" << p.source_moment << "; direction=" << d[0] << ',' << d[1] << ',' << d[2]
        << "; Ricker_f0_Hz=" << p.dominant_frequency << '\n';
    const auto &src = m.sourceNode;
    const double *xyz = &m.coordinates[3 * m.index(src[0] + 3, src[1] + 3, src[2] + 3)];
    out << "# source XYZ (m): " << xyz[0] << ',' << xyz[1] << ',' << xyz[2] << "; grid IJK (zero-based): " << src[0] << ','
        << src[1] << ',' << src[2] << '\n';
    const size_t nrec = m.receiverIndex.size();
    for (size_t r = 0; r < nrec; ++r)
        out << "# receiver " << r + 1 << ": " << m.receiverPosition[3 * r] << ',' << m.receiverPosition[3 * r + 1] << ','
            << m.receiverPosition[3 * r + 2] << '\n';
    out << "time_s";
    for (size_t r = 0; r < nrec; ++r)
        for (int c = 0; c < 3; ++c)
            if (selected[c]) {
                out << ",receiver_" << r + 1;
                if (!legacyVzHeader) out << '_' << names[c];
            }
    out << '\n';
    const std::vector<std::vector<double>> *traces[] = {&s->vx, &s->vy, &s->vz};
    for (int sample = 0; sample <= s->step; ++sample) {
        out << sample * m.dt;
        for (size_t r = 0; r < nrec; ++r)
            for (int c = 0; c < 3; ++c)
                if (selected[c]) out << ',' << (*traces[c])[r].at(static_cast<size_t>(sample));
        out << '\n';
    }
    require(out.good(), "Receiver CSV write failed.");
}

// VtsExportOptions3D: fields Vx, Vy, Vz, |V|, mean normal stress, Sxx, Syy, Szz, Sxy, Sxz, Syz.
void writeVts(const swl_sim3d *s, const std::filesystem::path &path, const int32_t *fields, bool includeMaterial,
              bool includeVelocityVector) {
    bool any = false;
    for (int n = 0; n < 11; ++n) any = any || fields[n] != 0;
    require(includeVelocityVector || any, "Select at least one wavefield component.");
    require(s != nullptr, "Uninitialized session");
    require(s->engine != nullptr, "This session was planned without a device.");
    const char *scalarNames[] = {"Vx", "Vy", "Vz", "VelocityMagnitude", "MeanNormalStress", "Sxx", "Syy", "Szz", "Sxy", "Sxz", "Syz"};
    const char *active = nullptr;
    for (int n = 0; n < 11; ++n)
        if (fields[n]) {
            active = scalarNames[n];
            break;
        }
    const auto &m = s->model;
    const auto &p = m.p;
    const int g = 3, nx = m.nx - 2 * g, ny = m.ny - 2 * g, nz = m.nz - 2 * g;
    require(nx >= 1 && ny >= 1 && nz >= 1, "Cannot export an empty grid.");
    require(s->finite, "Refusing to export non-finite simulation fields.");
    std::vector<std::vector<double>> f(9, std::vector<double>(m.count));
    for (int c = 0; c < 9; ++c)
        if (swl_engine3d_download(s->engine, c, f[c].data())) throw std::runtime_error(swl::last_error());
    std::error_code ec;
    if (!path.parent_path().empty()) std::filesystem::create_directories(path.parent_path(), ec);
    if (ec) throw std::runtime_error("Could not create export folder: " + ec.message());
    auto pending = path;
    pending += ".tmp-" + std::to_string(std::chrono::steady_clock::now().time_since_epoch().count());
    struct Pending {
        std::filesystem::path path;
        ~Pending() {
            std::error_code ignored;
            std::filesystem::remove(path, ignored);
        }
    } cleanup{pending};
    std::ofstream out(pending, std::ios::binary);
    out.imbue(std::locale::classic());
    require(static_cast<bool>(out), ("Could not open VTS output file: " + path.string()).c_str());
    out << std::setprecision(17);
    const auto d = unitDirection(p);
    const auto &src = m.sourceNode;
    const double *sxyz = &m.coordinates[3 * m.index(src[0] + g, src[1] + g, src[2] + g)];
    const auto scalar = [&](const char *type, const char *name, auto value) {
        out << "<DataArray type=\"" << type << "\" Name=\"" << name << "\" NumberOfTuples=\"1\" format=\"ascii\">" << value
            << "</DataArray>";
    };
    out << "<?xml version=\"1.0\"?>\n<VTKFile type=\"StructuredGrid\" version=\"1.0\" byte_order=\"LittleEndian\">\n"
        << "  <StructuredGrid WholeExtent=\"0 " << nx - 1 << " 0 " << ny - 1 << " 0 " << nz - 1 << "\">\n    <FieldData>";
    scalar("Float64", "TIME_VALUE", s->step * m.dt);
    scalar("Float64", "SimulationDt", m.dt);
    scalar("Int32", "SimulationStep", s->step);
    scalar("Float64", "SourceX", sxyz[0]), scalar("Float64", "SourceY", sxyz[1]), scalar("Float64", "SourceZ", sxyz[2]);
    scalar("Int32", "SourceKind", p.source_kind);
    scalar("Float64", "SourceForceN", p.source_force), scalar("Float64", "SourceMomentNm", p.source_moment);
    scalar("Float64", "SourceFrequencyHz", p.dominant_frequency);
    scalar("Float64", "SourceDirectionX", d[0]), scalar("Float64", "SourceDirectionY", d[1]);
    scalar("Float64", "SourceDirectionZ", d[2]);
    scalar("Float64", "GridSpacing", m.h);
    scalar("Float64", "TerrainTransitionDepth", m.transition);
    out << "</FieldData>\n    <Piece Extent=\"0 " << nx - 1 << " 0 " << ny - 1 << " 0 " << nz - 1 << "\">\n      <PointData";
    if (active) out << " Scalars=\"" << active << "\"";
    if (includeVelocityVector) out << " Vectors=\"Velocity\"";
    out << ">\n";
    const size_t nodes = (size_t)nx * ny * nz;
    const auto node = [&](size_t n) { // physical node n (x fastest, then y, then z) -> padded index
        const int i = (int)(n % nx), j = (int)((n / nx) % ny), k = (int)(n / ((size_t)nx * ny));
        return m.index(i + g, j + g, k + g);
    };
    const auto array = [&](const char *name, const auto &value) {
        out << "        <DataArray type=\"Float64\" Name=\"" << name << "\" format=\"ascii\">\n          ";
        parallelText(out, nodes, kNumberChars, " ", 6, "\n          ", [&](size_t n, char *c) { return put(c, value(node(n))); });
        out << "\n        </DataArray>\n";
    };
    if (fields[0]) array("Vx", [&](size_t q) { return f[0][q]; });
    if (fields[1]) array("Vy", [&](size_t q) { return f[1][q]; });
    if (fields[2]) array("Vz", [&](size_t q) { return f[2][q]; });
    if (includeMaterial) {
        if (p.layered) { // per node: Vp, Vs from the moduli (CurvilinearSolver.cpp:656), layer ids
            array("Vp", [&](size_t q) { return std::sqrt((m.lambdaNode[q] + 2 * m.muNode[q]) / m.densityNode[q]); });
            array("Vs", [&](size_t q) { return std::sqrt(m.muNode[q] / m.densityNode[q]); });
            array("Density", [&](size_t q) { return m.densityNode[q]; });
            array("LayerId", [&](size_t q) { return swl::layerOfNode3D(m, q); });
        } else {
            array("Vp", [&](size_t) { return p.vp; });
            array("Vs", [&](size_t) { return p.vs; });
            array("Density", [&](size_t) { return p.density; });
        }
    }
    if (fields[3]) array("VelocityMagnitude", [&](size_t q) { return std::hypot(f[0][q], f[1][q], f[2][q]); });
    if (fields[4]) array("MeanNormalStress", [&](size_t q) { return -(f[3][q] + f[4][q] + f[5][q]) / 3.; });
    const char *stress[] = {"Sxx", "Syy", "Szz", "Sxy", "Sxz", "Syz"};
    for (int n = 0; n < 6; ++n)
        if (fields[n + 5]) array(stress[n], [&](size_t q) { return f[3 + n][q]; });
    if (includeVelocityVector) {
        out << "        <DataArray type=\"Float64\" Name=\"Velocity\" NumberOfComponents=\"3\" format=\"ascii\">\n          ";
        parallelText(out, nodes, 3 * kNumberChars + 2, " ", 2, "\n          ", [&](size_t n, char *c) {
            const size_t q = node(n);
            return put(put(put(put(put(c, f[0][q]), " "), f[1][q]), " "), f[2][q]);
        });
        out << "\n        </DataArray>\n";
    }
    out << "      </PointData>\n      <CellData/>\n      <Points>\n"
        << "        <DataArray type=\"Float64\" NumberOfComponents=\"3\" format=\"ascii\">\n          ";
    parallelText(out, nodes, 3 * kNumberChars + 2, " ", 2, "\n          ", [&](size_t n, char *c) {
        const double *x = &m.coordinates[3 * node(n)];
        return put(put(put(put(put(c, x[0]), " "), x[1]), " "), x[2]);
    });
    out << "\n        </DataArray>\n      </Points>\n    </Piece>\n  </StructuredGrid>\n</VTKFile>\n";
    require(out.good(), ("Failed while writing VTS output: " + path.string()).c_str());
    out.close();
    require(static_cast<bool>(out), "Failed closing VTS output.");
    std::filesystem::rename(pending, path, ec); // same-directory replace (include/SeisCommon/AtomicFile.hpp:14-30)
    if (ec) throw std::runtime_error("Could not replace VTS output: " + ec.message());
}
} // namespace

extern "C" {
int swl_sim3d_write_receivers_csv(const swl_sim3d *s, const char *path, const int32_t *components) {
    return swl::guarded([&] {
        require(path != nullptr && components != nullptr, "invalid receiver export arguments");
        writeCsv(s, path, components, false);
    });
}
int swl_sim3d_write_receivers_vz_csv(const swl_sim3d *s, const char *path, int all_velocity_components) {
    return swl::guarded([&] {
        require(path != nullptr, "invalid receiver export arguments");
        const int32_t selected[3] = {all_velocity_components != 0, all_velocity_components != 0, 1};
        writeCsv(s, path, selected, all_velocity_components == 0);
    });
}
int swl_sim3d_write_vts(const swl_sim3d *s, const char *path, const int32_t *fields, int include_material,
                        int include_velocity_vector) {
    return swl::guarded([&] {
        require(path != nullptr && fields != nullptr, "invalid VTS export arguments");
        writeVts(s, path, fields, include_material != 0, include_velocity_vector != 0);
    });
}
} // extern "C"
