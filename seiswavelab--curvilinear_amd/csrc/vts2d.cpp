// vts2d.cpp — 2D model documents and wavefield snapshots as ASCII VTK StructuredGrid (.vts) files.
//
// Model document: writeModelVts / readModelVts (include/Seis2D/ModelIo2D.hpp:8-13, src/two_d/ModelIo2D.cpp:21-254).
// FieldData holds the versioned parameter record (one array per name, in byte order of the names); PointData
// holds Vp, Vs, Density, Jacobian, MetricA, MetricB and Points the X-0-Z coordinates of the logical nodes
// (X fastest, Z positive down).  An FNV-1a 64-bit checksum over every name and value seals the file
// (include/SeisCommon/VtsModelIo.hpp:18-31).  A document is a model, not a wavefield checkpoint: loading it
// rebuilds the session at t = 0.  The reader is bounded (tags <= 8 KiB, <= 80 metadata arrays of
// <= 2e6 values) and streams the point arrays through the checksum without keeping them.
//
// Wavefield snapshot: writeSnapshot (src/two_d/CurvilinearSolver2D.cpp:725-810): provenance scalars and the
// selected components of the state downloaded from the device, plus derived speed and plane-strain pressure.
//
// Numbers are written as the shortest "%.17g" form of the C locale (std::to_chars, general format,
// precision 17), which round-trips every double.  Rows are formatted in parallel and written in order.
#include "model2d.hpp"
#include "session2d.hpp"
#include <charconv>
#include <chrono>
#include <clocale>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <locale.h>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace {
using swl::require;
using Arrays = std::map<std::string, std::vector<double>>;

void require(bool ok, const std::string &message) {
    if (!ok) throw std::runtime_error(message);
}

// FNV-1a, 64-bit (Fowler-Noll-Vo): a name contributes its bytes and a terminating zero byte; a number its
// eight IEEE-754 bytes, least significant first, with -0 folded onto +0 (VtsModelIo.hpp:18-31).
struct Fnv1a {
    uint64_t value = 0xcbf29ce484222325ULL;
    void byte(unsigned char b) { value = (value ^ b) * 0x100000001b3ULL; }
    void name(const std::string &s) {
        for (unsigned char c : s) byte(c);
        byte(0);
    }
    void number(double d) {
        const double canonical = d == 0 ? 0.0 : d;
        uint64_t bits;
        std::memcpy(&bits, &canonical, sizeof bits);
        for (int shift = 0; shift < 64; shift += 8) byte(static_cast<unsigned char>(bits >> shift));
    }
};

constexpr int kNumberChars = 25; // "-1.2345678901234567e-308"
inline char *put(char *p, double v) { return std::to_chars(p, p + kNumberChars, v, std::chars_format::general, 17).ptr; }
inline char *put(char *p, const char *s) {
    while (*s) *p++ = *s++;
    return p;
}

// Buffered output that fails loudly; text is produced with put() into reserve()d space.
class Output {
    std::FILE *f_ = nullptr;
    std::string buf_;
    bool ok_ = true;

public:
    explicit Output(const std::string &path) : f_(std::fopen(path.c_str(), "wb")) { buf_.reserve(1 << 20); }
    ~Output() {
        if (f_) std::fclose(f_);
    }
    bool opened() const { return f_ != nullptr; }
    void text(const std::string &s) { raw(s.data(), s.size()); }
    void text(const char *s) { raw(s, std::strlen(s)); }
    void number(double v) {
        char b[32];
        raw(b, static_cast<size_t>(put(b, v) - b));
    }
    void integer(long long v) { text(std::to_string(v)); }
    void raw(const char *s, size_t n) {
        buf_.append(s, n);
        if (buf_.size() >= (1u << 20)) flush();
    }
    void flush() {
        if (f_ && !buf_.empty() && std::fwrite(buf_.data(), 1, buf_.size(), f_) != buf_.size()) ok_ = false;
        buf_.clear();
    }
    bool close() {
        flush();
        if (f_ && std::fclose(f_) != 0) ok_ = false;
        f_ = nullptr;
        return ok_;
    }
    // nz rows of nx nodes; emit(i, k, p) writes one node's text at p (at most `width` chars) and returns the end.
    template <class Emit> void rows(int nx, int nz, int width, const Emit &emit) {
        flush();
        constexpr int block = 64;
        std::vector<std::string> text(block);
        for (int k0 = 0; k0 < nz; k0 += block) {
            const int kn = std::min(block, nz - k0);
#pragma omp parallel for schedule(dynamic, 1)
            for (int r = 0; r < kn; ++r) {
                auto &s = text[r];
                s.resize(static_cast<size_t>(nx) * width);
                char *p = s.data();
                for (int i = 0; i < nx; ++i) p = emit(i, k0 + r, p);
                s.resize(static_cast<size_t>(p - s.data()));
            }
            for (int r = 0; r < kn; ++r)
                if (f_ && std::fwrite(text[r].data(), 1, text[r].size(), f_) != text[r].size()) ok_ = false;
        }
    }
};

// Buffered input with one character of lookahead.
class Input {
    std::FILE *f_ = nullptr;
    std::vector<char> buf_ = std::vector<char>(1 << 20);
    size_t pos_ = 0, len_ = 0;

public:
    explicit Input(const std::string &path) : f_(std::fopen(path.c_str(), "rb")) {}
    ~Input() {
        if (f_) std::fclose(f_);
    }
    bool opened() const { return f_ != nullptr; }
    int peek() {
        if (pos_ == len_) {
            len_ = f_ ? std::fread(buf_.data(), 1, buf_.size(), f_) : 0, pos_ = 0;
            if (!len_) return EOF;
        }
        return static_cast<unsigned char>(buf_[pos_]);
    }
    int get() {
        const int c = peek();
        if (c != EOF) ++pos_;
        return c;
    }
    void skipSpace() { // std::ws in the C locale
        for (int c = peek(); c == ' ' || c == '\t' || c == '\n' || c == '\v' || c == '\f' || c == '\r'; c = peek()) get();
    }
};

// --- VTS grammar (VtsModelIo.hpp:33-86) ----------------------------------------------------------------
// Value of `name="..."` (or '...'): the first occurrence of name= anywhere in the tag.
std::string attribute(const std::string &tag, const std::string &name) {
    const auto pos = tag.find(name + "=");
    require(pos != std::string::npos, "Missing VTS attribute: " + name);
    const auto start = pos + name.size() + 1;
    require(start < tag.size() && (tag[start] == '"' || tag[start] == '\''), "Invalid VTS attribute.");
    const auto end = tag.find(tag[start], start + 1);
    require(end != std::string::npos, "Unterminated VTS attribute.");
    return tag.substr(start + 1, end - start - 1);
}
// Next "<...>" after white space; at most 8192 characters after '<'.
std::string tag(Input &in) {
    in.skipSpace();
    require(in.get() == '<', "Expected VTS XML tag (truncated or unsupported file).");
    std::string s{"<"};
    for (int i = 0; i < 8192; ++i) {
        const int c = in.get();
        require(c != EOF, "Truncated VTS XML tag.");
        s.push_back(static_cast<char>(c));
        if (c == '>') return s;
    }
    throw std::runtime_error("Oversized VTS XML tag.");
}
void expected(Input &in, const char *name) { require(tag(in) == name, std::string("Expected ") + name); }
int integer(double value, int low, int high) {
    require(std::isfinite(value) && value >= low && value <= high && value == std::floor(value),
            "Invalid integer model parameter.");
    return static_cast<int>(value);
}

locale_t cLocale() {
    static const locale_t c = newlocale(LC_ALL_MASK, "C", nullptr);
    return c;
}
// One formatted double as an istream extracts it in the C locale: optional sign, digits with at most one
// decimal point, then an optional exponent (only after a mantissa digit) with optional sign; the collected
// text must convert completely and not overflow.
bool number(Input &in, double &v, std::string &s) {
    in.skipSpace();
    s.clear();
    int c = in.peek();
    if (c == '+' || c == '-') s.push_back(static_cast<char>(in.get())), c = in.peek();
    bool mantissa = false, dec = false, sci = false;
    for (; c == '0'; c = (in.get(), in.peek()))
        if (!mantissa) s.push_back('0'), mantissa = true;
    while (c != EOF) {
        if (c >= '0' && c <= '9') s.push_back(static_cast<char>(c)), mantissa = true;
        else if (c == '.' && !dec && !sci) s.push_back('.'), dec = true;
        else if ((c == 'e' || c == 'E') && !sci && mantissa) {
            s.push_back('e'), sci = true;
            in.get();
            c = in.peek();
            if (c != '+' && c != '-') continue;
            s.push_back(static_cast<char>(c));
        } else break;
        in.get();
        c = in.peek();
    }
    if (s.empty()) return false;
    char *end = nullptr;
    v = strtod_l(s.c_str(), &end, cLocale());
    return end == s.c_str() + s.size() && !std::isinf(v);
}
// DataArray body of `count` values followed by </DataArray>; hashed (unless hash is null) and optionally kept.
std::vector<double> readArray(Input &in, const std::string &header, Fnv1a *hash, size_t count, bool retain,
                              bool allowFloat32 = false) {
    const bool fp32 = attribute(header, "type") == "Float32";
    require(attribute(header, "format") == "ascii" && (attribute(header, "type") == "Float64" || (allowFloat32 && fp32)),
            "Model import requires supported SeisWave ASCII floating point VTS arrays.");
    const auto name = attribute(header, "Name");
    if (hash) hash->name(name);
    std::vector<double> values;
    if (retain) values.reserve(count);
    std::string scratch;
    for (size_t n = 0; n < count; ++n) {
        double v{};
        require(number(in, v, scratch) && std::isfinite(v), "Missing or nonfinite VTS data: " + name);
        if (fp32) {
            v = static_cast<float>(v);
            require(std::isfinite(v), "Float32 VTS value overflows: " + name);
        }
        if (hash) hash->number(v);
        if (retain) values.push_back(v);
    }
    expected(in, "</DataArray>");
    return values;
}

// --- the parameter record (ModelIo2D.cpp:21-61) ---------------------------------------------------------
const char *const kParameterNames[] = {
    "nx", "nz", "spacing", "vp", "vs", "density", "relief", "wavelength", "flattenDepth", "frequency",
    "duration", "cfl", "forcePerLength", "momentPerLength", "sourceX", "sourceZ", "maxSteps", "absorbing",
    "spongeX", "spongeZ", "spongeStrength", "receiverStart", "receiverCount", "receiverStride",
    "receiverDepth", "sampleIntervalMs", "layered", "importedSurface", "surfaceDatum"};
double parameterValue(const swl_params2d &p, int n) {
    const double v[] = {double(p.nx), double(p.nz), p.spacing, p.vp, p.vs, p.density, p.relief, p.wavelength,
                        p.flatten_depth, p.frequency, p.duration, p.cfl, p.force_per_length, p.moment_per_length,
                        double(p.source_x), double(p.source_z), double(p.max_steps), double(p.absorbing != 0),
                        double(p.sponge_x), double(p.sponge_z), p.sponge_strength, double(p.receiver_start),
                        double(p.receiver_count), double(p.receiver_stride), double(p.receiver_depth),
                        p.sample_interval_ms, double(p.layered != 0), double(p.imported_surface != 0), p.surface_datum};
    return v[n];
}
bool defaultPml(const swl_params2d &p) { return p.pml_reflection == 1e-4 && p.pml_kappa_max == 3 && p.pml_frequency == 0; }

// Schema version: 2 base record, 3 CFS-PML settings, 4 external absorption, 5 precision, 6 physical source
// kernel, 7 surface receiver policy; each optional group appears from the first version that needs it.
Arrays metadata(const swl_sim2d &s, int cpuThreads, int cudaDevice) {
    const auto &m = s.model;
    const auto &p = m.logical;
    Arrays a;
    const auto add = [&](const char *name, double v) { a[name] = {v}; };
    const bool pml = p.absorber != 0 || !defaultPml(p), fp32 = p.precision == 1, physical = p.source_sigma_m > 0;
    const bool survey = p.auto_surface_receivers || p.receiver_target_count != 0, external = p.external_absorption != 0;
    add("SeisWaveModel2DVersion", survey ? 7 : physical ? 6 : fp32 ? 5 : external ? 4 : pml ? 3 : 2);
    if (survey) a["SurfaceReceiverLayout"] = {double(p.auto_surface_receivers != 0), double(p.receiver_target_count)};
    if (fp32 || physical || survey) add("CalculationPrecision", p.precision);
    if (pml || external || fp32 || physical || survey)
        a["CfsPmlSettings"] = {double(p.absorber), p.pml_reflection, p.pml_kappa_max, p.pml_frequency};
    if (external || fp32 || physical || survey) add("ExternalAbsorption", external);
    if (physical || survey) a["PhysicalSourceKernelM"] = {p.source_sigma_m, p.source_radius_m};
    a["GSLSSettings"] = {double(p.attenuation != 0), double(p.mechanisms), p.q_min_frequency, p.q_max_frequency,
                         p.q_reference_frequency, p.qp, p.qs};
    auto &q = a["LayerQpQs"];
    auto &layers = a["LayersThicknessVpVsDensity"];
    for (int l = 0; l < p.n_layers; ++l) {
        q.insert(q.end(), {p.layer_qp[l], p.layer_qs[l]});
        layers.insert(layers.end(), {p.layer_thickness[l], p.layer_vp[l], p.layer_vs[l], p.layer_density[l]});
    }
    for (int n = 0; n < 29; ++n) a[std::string("Parameter_") + kParameterNames[n]] = {parameterValue(p, n)};
    add("SourceKind", p.source_kind);
    a["SourceDirection"] = {p.source_direction[0], p.source_direction[1], p.source_direction[2]};
    auto &samples = a["SurfaceSamplesXElevation"];
    samples.assign(p.surface_xz, p.surface_xz + 2 * static_cast<size_t>(std::max(p.n_surface, 0)));
    add("BackendKind", s.engine ? 1 : 0); // 1: the accelerator backend (BackendKind3D::Cuda's slot), 0: CPU
    add("CpuThreadsRequested", cpuThreads);
    add("CudaDeviceRequested", cudaDevice);
    add("ReferenceDtSeconds", m.dt);
    add("RecordIntervalSeconds", s.recordDt());
    add("EffectiveTotalSteps", m.total);
    add("EffectiveFlattenDepthM", m.flatten);
    add("CoordinatesZPositiveDown", 1);
    return a;
}
int components(const std::string &name) {
    if (name == "SourceDirection") return 3;
    if (name == "LayersThicknessVpVsDensity") return 4;
    return name == "SurfaceSamplesXElevation" || name == "ModelChecksum" ? 2 : 1;
}

const char *const kPointNames[] = {"Vp", "Vs", "Density", "Jacobian", "MetricA", "MetricB"};
// Node values as the session's grid holds them: an FP32 session keeps coordinates and the point arrays in
// single precision (CurvilinearSolver2D.cpp:498-499), so they are written and checksummed rounded to float.
struct PointArrays {
    const swl::Model2D &m;
    const std::vector<double> *v[6];
    bool fp32;
    explicit PointArrays(const swl::Model2D &model)
        : m(model), v{&model.vp, &model.vs, &model.rho, &model.jac, &model.a, &model.b}, fp32(model.logical.precision == 1) {}
    size_t index(int i, int k) const { return static_cast<size_t>(i + 3 + m.offsetX) + static_cast<size_t>(m.nx) * (k + 3); }
    double stored(double x) const { return fp32 ? double(static_cast<float>(x)) : x; }
    double value(int f, size_t q) const { return stored((*v[f])[q]); }
    double coordinate(size_t q, int c) const { return stored(m.coordinates[2 * q + c]); }
};

std::string extent(int nx, int nz) { return "0 " + std::to_string(nx - 1) + " 0 0 0 " + std::to_string(nz - 1); }

// --- reader ---------------------------------------------------------------------------------------------
swl_params2d defaultParameters() { // Seis2D::Curvilinear::Parameters{} (CurvilinearSolver2D.hpp:29-57)
    swl_params2d p{};
    p.precision = 0, p.nx = 200, p.nz = 200, p.spacing = 2.5, p.vp = 3000, p.vs = 1700, p.density = 2200;
    p.relief = 10, p.wavelength = 400, p.flatten_depth = 0, p.frequency = 30, p.duration = 1, p.cfl = .8;
    p.force_per_length = 2e5, p.source_x = -1, p.source_z = 40, p.max_steps = 5000;
    p.source_kind = 0, p.source_direction[0] = 0, p.source_direction[1] = 0, p.source_direction[2] = 1;
    p.moment_per_length = 2e6, p.absorbing = 1, p.sponge_x = p.sponge_z = 16, p.sponge_strength = 3, p.absorber = 0;
    p.pml_reflection = 1e-4, p.pml_kappa_max = 3, p.pml_frequency = 0;
    p.receiver_start = 2, p.receiver_count = 96, p.receiver_stride = 2, p.receiver_depth = 1;
    p.attenuation = 0, p.mechanisms = 3, p.q_min_frequency = 5, p.q_max_frequency = 100, p.q_reference_frequency = 30;
    p.qp = 100, p.qs = 60;
    const double layers[3][4] = {{20, 1500, 700, 2050}, {50, 2500, 1300, 2250}, {0, 4000, 2300, 2650}};
    p.n_layers = 3;
    for (int l = 0; l < 3; ++l) {
        p.layer_thickness[l] = layers[l][0], p.layer_vp[l] = layers[l][1], p.layer_vs[l] = layers[l][2];
        p.layer_density[l] = layers[l][3], p.layer_qp[l] = 100, p.layer_qs[l] = 60;
    }
    return p;
}
const std::vector<double> &at(const Arrays &a, const char *key) {
    const auto it = a.find(key);
    if (it == a.end()) throw std::out_of_range("map::at"); // what std::map::at reports
    return it->second;
}
} // namespace

struct swl_model_doc2d {
    swl_params2d p{};
    std::vector<double> surface;
    int version = 0, backend = 0, cpuThreads = 0, cudaDevice = -1;
    double referenceDt = 0;
};

namespace {
std::unique_ptr<swl_model_doc2d> readDocument(const std::string &path) {
    Input in(path);
    require(in.opened(), "Cannot open model VTS file.");
    auto head = tag(in);
    if (head.rfind("<?xml", 0) == 0) head = tag(in);
    require(head.rfind("<VTKFile ", 0) == 0 && attribute(head, "type") == "StructuredGrid",
            "Expected a SeisWave 2D StructuredGrid model VTS.");
    const auto gridHeader = tag(in);
    require(gridHeader.rfind("<StructuredGrid ", 0) == 0, "Missing StructuredGrid.");
    expected(in, "<FieldData>");
    Arrays arrays;
    Fnv1a hash;
    while (true) {
        const auto a = tag(in);
        if (a == "</FieldData>") break;
        require(a.rfind("<DataArray ", 0) == 0 && arrays.size() < 80, "Invalid model metadata.");
        const auto name = attribute(a, "Name");
        require(!arrays.count(name), "Duplicate model metadata: " + name);
        const auto parsedInt = [&](const char *key, int high) {
            const auto text = attribute(a, key);
            size_t end{};
            const double value = std::stod(text, &end);
            require(end == text.size(), "Invalid array size.");
            return integer(value, 0, high);
        };
        const int comps = parsedInt("NumberOfComponents", 4), tuples = parsedInt("NumberOfTuples", 1000000);
        require(comps > 0 && static_cast<size_t>(comps) * tuples <= 2000000, "Oversized model metadata.");
        arrays[name] = readArray(in, a, name == "ModelChecksum" ? nullptr : &hash, static_cast<size_t>(comps) * tuples, true);
    }
    const auto scalar = [&](const std::string &key) {
        const auto it = arrays.find(key);
        require(it != arrays.end() && it->second.size() == 1, "Missing model parameter: " + key);
        return it->second[0];
    };
    auto doc = std::make_unique<swl_model_doc2d>();
    const int version = integer(scalar("SeisWaveModel2DVersion"), 1, 7);
    doc->version = version;
    auto &p = doc->p;
    p = defaultParameters();
    const auto real = [&](const char *name) { return scalar(std::string("Parameter_") + name); };
    const auto whole = [&](const char *name, int low, int high) { return integer(real(name), low, high); };
    p.nx = whole("nx", 8, 16384), p.nz = whole("nz", 8, 16384);
    p.spacing = real("spacing"), p.vp = real("vp"), p.vs = real("vs"), p.density = real("density");
    p.relief = real("relief"), p.wavelength = real("wavelength"), p.flatten_depth = real("flattenDepth");
    p.frequency = real("frequency"), p.duration = real("duration"), p.cfl = real("cfl");
    p.force_per_length = real("forcePerLength"), p.moment_per_length = real("momentPerLength");
    p.source_x = whole("sourceX", -1, p.nx - 1), p.source_z = whole("sourceZ", 0, p.nz - 1);
    p.max_steps = whole("maxSteps", 0, 200000), p.sponge_x = whole("spongeX", 1, 512), p.sponge_z = whole("spongeZ", 1, 512);
    p.sponge_strength = real("spongeStrength");
    p.receiver_start = whole("receiverStart", 0, p.nx - 1), p.receiver_count = whole("receiverCount", 1, 1024);
    p.receiver_stride = whole("receiverStride", 1, 16384), p.receiver_depth = whole("receiverDepth", 0, p.nz - 1);
    p.sample_interval_ms = real("sampleIntervalMs"), p.surface_datum = real("surfaceDatum");
    p.absorbing = whole("absorbing", 0, 1), p.layered = whole("layered", 0, 1);
    p.imported_surface = whole("importedSurface", 0, 1);
    p.source_kind = integer(scalar("SourceKind"), 0, 1);
    const auto &direction = at(arrays, "SourceDirection");
    require(direction.size() == 3, "Invalid source direction.");
    for (int c = 0; c < 3; ++c) p.source_direction[c] = direction[c];
    if (version >= 6) {
        const auto &kernel = at(arrays, "PhysicalSourceKernelM");
        require(kernel.size() == 2 && ((version >= 7 && kernel[0] == 0 && kernel[1] == 0) ||
                                       (kernel[0] > 0 && kernel[1] >= kernel[0] && p.source_kind == 0)),
                "Invalid physical source kernel metadata.");
        p.source_sigma_m = kernel[0], p.source_radius_m = kernel[1];
    }
    if (version >= 7) {
        const auto &layout = at(arrays, "SurfaceReceiverLayout");
        require(layout.size() == 2, "Invalid surface receiver policy.");
        p.auto_surface_receivers = integer(layout[0], 0, 1), p.receiver_target_count = integer(layout[1], 0, 1024);
    }
    const auto &layers = at(arrays, "LayersThicknessVpVsDensity");
    require(layers.size() % 4 == 0 && layers.size() <= 64, "Invalid model layer table.");
    p.n_layers = static_cast<int>(layers.size() / 4);
    for (int l = 0; l < p.n_layers; ++l) {
        p.layer_thickness[l] = layers[4 * l], p.layer_vp[l] = layers[4 * l + 1], p.layer_vs[l] = layers[4 * l + 2];
        p.layer_density[l] = layers[4 * l + 3], p.layer_qp[l] = 100, p.layer_qs[l] = 60;
    }
    if (version >= 2) {
        const auto &q = at(arrays, "GSLSSettings");
        require(q.size() == 7, "Invalid GSLS settings.");
        p.attenuation = integer(q[0], 0, 1), p.mechanisms = integer(q[1], 1, 3);
        p.q_min_frequency = q[2], p.q_max_frequency = q[3], p.q_reference_frequency = q[4], p.qp = q[5], p.qs = q[6];
        swl::validateGsls(p.mechanisms, p.q_min_frequency, p.q_max_frequency, p.q_reference_frequency, p.qp, p.qs);
        const auto &layerQ = at(arrays, "LayerQpQs");
        require(layerQ.size() == 2 * static_cast<size_t>(p.n_layers), "Layer Q table differs from material table.");
        for (int l = 0; l < p.n_layers; ++l) p.layer_qp[l] = layerQ[2 * l], p.layer_qs[l] = layerQ[2 * l + 1];
    }
    if (version >= 5) p.precision = integer(scalar("CalculationPrecision"), 0, 1);
    if (version >= 4) p.external_absorption = integer(scalar("ExternalAbsorption"), 0, 1);
    if (version >= 3) {
        const auto &a = at(arrays, "CfsPmlSettings");
        require(a.size() == 4, "Invalid CFS-PML metadata.");
        p.absorber = integer(a[0], 0, 1), p.pml_reflection = a[1], p.pml_kappa_max = a[2], p.pml_frequency = a[3];
        require(p.pml_reflection >= 1e-8 && p.pml_reflection <= .1 && p.pml_kappa_max >= 1 && p.pml_kappa_max <= 20 &&
                    p.pml_frequency >= 0 && p.pml_frequency <= 1e6,
                "Invalid CFS-PML profile settings.");
    }
    const auto &samples = at(arrays, "SurfaceSamplesXElevation");
    require(samples.size() % 2 == 0, "Invalid model surface samples.");
    doc->surface = samples;
    doc->cpuThreads = integer(scalar("CpuThreadsRequested"), 0, 64);
    doc->cudaDevice = integer(scalar("CudaDeviceRequested"), -1, 100000);
    doc->backend = integer(scalar("BackendKind"), 0, 1);
    doc->referenceDt = scalar("ReferenceDtSeconds");
    require(doc->referenceDt > 0 && scalar("CoordinatesZPositiveDown") == 1, "Invalid model time step / coordinate convention.");
    const auto ext = extent(p.nx, p.nz);
    require(attribute(gridHeader, "WholeExtent") == ext, "Model dimensions disagree with VTS extent.");
    const auto piece = tag(in);
    require(piece.rfind("<Piece ", 0) == 0 && attribute(piece, "Extent") == ext, "Invalid model piece extent.");
    require(tag(in).rfind("<PointData", 0) == 0, "Missing point material data.");
    const size_t count = static_cast<size_t>(p.nx) * p.nz;
    for (const char *name : kPointNames) {
        const auto a = tag(in);
        require(a.rfind("<DataArray ", 0) == 0 && attribute(a, "Name") == name, "Missing model point array.");
        readArray(in, a, &hash, count, false, version >= 5);
    }
    expected(in, "</PointData>");
    expected(in, "<Points>");
    const auto coordinates = tag(in);
    require(attribute(coordinates, "Name") == "Coordinates" && attribute(coordinates, "NumberOfComponents") == "3",
            "Invalid model coordinates.");
    readArray(in, coordinates, &hash, count * 3, false, version >= 5);
    expected(in, "</Points>");
    expected(in, "</Piece>");
    expected(in, "</StructuredGrid>");
    expected(in, "</VTKFile>");
    in.skipSpace();
    require(in.peek() == EOF, "Unexpected trailing VTS content.");
    const auto &digest = at(arrays, "ModelChecksum");
    require(digest.size() == 2 && digest[0] == static_cast<double>(hash.value >> 32) &&
                digest[1] == static_cast<double>(hash.value & 0xffffffffULL),
            "Model checksum mismatch (file changed or damaged).");
    p.n_surface = static_cast<int>(doc->surface.size() / 2);
    p.surface_xz = doc->surface.empty() ? nullptr : doc->surface.data();
    return doc;
}

// Same-directory replacement after a complete write (include/SeisCommon/AtomicFile.hpp:14-30).
void replaceCompleted(const std::filesystem::path &pending, const std::filesystem::path &target) {
    std::error_code error;
    std::filesystem::rename(pending, target, error);
    if (error) throw std::filesystem::filesystem_error("Could not replace completed output file", pending, target, error);
}

void writeDocument(const swl_sim2d *s, const std::string &path, int cpuThreads, int cudaDevice) {
    require(s != nullptr, "Initialize the 2D model before saving.");
    const auto &m = s->model;
    const auto &p = m.logical;
    const PointArrays points(m);
    const char *pointType = points.fp32 ? "Float32" : "Float64";
    auto arrays = metadata(*s, cpuThreads, cudaDevice);
    Fnv1a hash;
    for (const auto &[name, values] : arrays) {
        hash.name(name);
        for (double v : values) {
            require(std::isfinite(v), "Cannot save nonfinite model metadata.");
            hash.number(v);
        }
    }
    for (int f = 0; f < 6; ++f) {
        hash.name(kPointNames[f]);
        for (int k = 0; k < p.nz; ++k)
            for (int i = 0; i < p.nx; ++i) hash.number(points.value(f, points.index(i, k)));
    }
    hash.name("Coordinates");
    for (int k = 0; k < p.nz; ++k)
        for (int i = 0; i < p.nx; ++i) {
            const size_t q = points.index(i, k);
            hash.number(points.coordinate(q, 0)), hash.number(0), hash.number(points.coordinate(q, 1));
        }
    arrays["ModelChecksum"] = {static_cast<double>(hash.value >> 32), static_cast<double>(hash.value & 0xffffffffULL)};

    const std::filesystem::path target(path);
    auto pending = target;
    pending += ".tmp-" + std::to_string(std::chrono::steady_clock::now().time_since_epoch().count());
    try {
        Output out(pending.string());
        require(out.opened(), "Cannot create 2D model VTS file.");
        const auto ext = extent(p.nx, p.nz);
        out.text("<?xml version=\"1.0\"?>\n<VTKFile type=\"StructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\">\n"
                 "<StructuredGrid WholeExtent=\"" + ext + "\">\n<FieldData>\n");
        for (const auto &[name, values] : arrays) {
            const int c = components(name);
            out.text("<DataArray type=\"Float64\" Name=\"" + name + "\" NumberOfComponents=\"" + std::to_string(c) +
                     "\" NumberOfTuples=\"" + std::to_string(values.size() / c) + "\" format=\"ascii\">");
            for (double v : values) {
                require(std::isfinite(v), "Cannot save nonfinite model metadata.");
                out.number(v), out.text(" ");
            }
            out.text("</DataArray>\n");
        }
        out.text("</FieldData>\n<Piece Extent=\"" + ext + "\"><PointData Scalars=\"Vp\">\n");
        for (int f = 0; f < 6; ++f) {
            out.text(std::string("<DataArray type=\"") + pointType + "\" Name=\"" + kPointNames[f] + "\" format=\"ascii\">\n");
            out.rows(p.nx, p.nz, kNumberChars + 1, [&](int i, int k, char *c) {
                c = put(c, points.value(f, points.index(i, k)));
                *c++ = ' ';
                return c;
            });
            out.text("\n</DataArray>\n");
        }
        out.text(std::string("</PointData><Points><DataArray type=\"") + pointType +
                 "\" Name=\"Coordinates\" NumberOfComponents=\"3\" format=\"ascii\">\n");
        out.rows(p.nx, p.nz, 2 * kNumberChars + 4, [&](int i, int k, char *c) {
            const size_t q = points.index(i, k);
            c = put(put(put(c, points.coordinate(q, 0)), " 0 "), points.coordinate(q, 1));
            *c++ = ' ';
            return c;
        });
        out.text("\n</DataArray></Points></Piece></StructuredGrid></VTKFile>\n");
        require(out.close(), "2D model VTS write failed.");
        replaceCompleted(pending, target);
    } catch (...) {
        std::error_code ignored;
        std::filesystem::remove(pending, ignored);
        throw;
    }
}

// writeSnapshot (CurvilinearSolver2D.cpp:725-810).  components: Vx, Vz, |V|, pressure, Sxx, Szz, Sxz.
void writeWavefield(const swl_sim2d *s, const std::string &path, const int32_t *selected, bool includeMaterial) {
    bool any = false;
    for (int c = 0; c < 7; ++c) any = any || selected[c] != 0;
    require(any, "Select at least one wavefield export component.");
    require(s != nullptr, "No 2D session.");
    require(s->engine != nullptr, "This session was planned without a device; create it on the HIP backend to step.");
    const auto &m = s->model;
    const auto &p = m.logical;
    const bool attenuating = m.mechanisms > 0;
    std::vector<std::vector<double>> fields(attenuating ? 6 : 5, std::vector<double>(m.count));
    for (size_t f = 0; f < fields.size(); ++f)
        if (swl_engine2d_snapshot(s->engine, static_cast<int>(f), fields[f].data())) throw std::runtime_error(swl::last_error());
    const PointArrays points(m);
    const bool fp32 = points.fp32;
    const char *pointType = fp32 ? "Float32" : "Float64";
    Output out(path);
    require(out.opened(), "Cannot create 2D VTS file.");
    const auto ext = extent(p.nx, p.nz);
    out.text("<?xml version=\"1.0\"?>\n<VTKFile type=\"StructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\">"
             "<StructuredGrid WholeExtent=\"" + ext + "\">\n<FieldData>");
    const auto scalar = [&](const char *name, double v) {
        out.text(std::string("<DataArray type=\"Float64\" Name=\"") + name + "\" NumberOfTuples=\"1\" format=\"ascii\">");
        out.number(v), out.text("</DataArray>");
    };
    scalar("CalculationPrecisionBits", fp32 ? 32 : 64);
    scalar("ExternalAbsorption", p.external_absorption != 0);
    scalar("ComputationalNodesX", m.nx - 6), scalar("ComputationalNodesZ", m.nz - 6);
    scalar("LogicalModelOffsetXCells", m.offsetX);
    scalar("AbsorptionEnabled", p.absorbing != 0), scalar("AbsorberKind", p.absorber);
    scalar("AbsorberWidthXCells", std::min(p.sponge_x, p.nx / 3)), scalar("AbsorberWidthZCells", std::min(p.sponge_z, p.nz / 3));
    scalar("CfsPmlTargetReflection", p.pml_reflection), scalar("CfsPmlKappaMax", p.pml_kappa_max);
    scalar("CfsPmlFrequencyHz", p.pml_frequency > 0 ? p.pml_frequency : p.frequency);
    scalar("SourceKind", p.source_kind);
    scalar("SourceForcePerLengthNPerM", p.force_per_length);
    scalar("SourceMomentPerLengthN", p.moment_per_length);
    scalar("SourceFrequencyHz", p.frequency);
    scalar("SourceDirectionX", m.direction[0]), scalar("SourceDirectionZ", m.direction[2]);
    scalar("SimulationDt", m.dt);
    scalar("RecordIntervalSeconds", s->recordDt());
    scalar("SimulationStep", s->step);
    scalar("TIME_VALUE", s->step * m.dt);
    scalar("AttenuationGSLS", p.attenuation != 0), scalar("QReferenceFrequencyHz", p.q_reference_frequency);
    scalar("QBandMinHz", p.q_min_frequency), scalar("QBandMaxHz", p.q_max_frequency);
    scalar("RelaxationMechanisms", p.attenuation ? p.mechanisms : 0);
    const char *names[] = {"Vx", "Vz", "VelocityMagnitude", "Pressure", "Sxx", "Szz", "Sxz", "Vp", "Vs", "Density"};
    int first = 0;
    while (!selected[first]) ++first;
    out.text(std::string("</FieldData><Piece Extent=\"") + ext + "\"><PointData Scalars=\"" + names[first] + "\">\n");
    const auto grid = [&](int f, size_t q) { return points.value(f, q); };
    for (int f = 0; f < 10; ++f) {
        if (f < 7 ? !selected[f] : !includeMaterial) continue;
        out.text(std::string("<DataArray type=\"") + pointType + "\" Name=\"" + names[f] + "\" format=\"ascii\">\n");
        out.rows(p.nx, p.nz, kNumberChars + 1, [&](int i, int k, char *c) {
            const size_t q = points.index(i, k);
            double value;
            if (f < 2) value = fields[f][q];
            else if (f == 2) value = std::sqrt(fields[0][q] * fields[0][q] + fields[1][q] * fields[1][q]);
            else if (f == 3) {
                // plane strain: Syy = lambda (Sxx + Szz) / (2 (lambda + mu)), pressure = -tr(S) / 3
                const double mu = grid(2, q) * grid(1, q) * grid(1, q);
                const double lambda = grid(2, q) * grid(0, q) * grid(0, q) - 2 * mu;
                value = p.attenuation ? -(fields[2][q] + fields[3][q] + fields[5][q]) / 3
                                      : -(1 + lambda / (2 * (lambda + mu))) * (fields[2][q] + fields[3][q]) / 3;
            } else if (f < 7) value = fields[f - 2][q];
            else value = grid(f - 7, q);
            c = put(c, fp32 ? double(static_cast<float>(value)) : value);
            *c++ = ' ';
            return c;
        });
        out.text("\n</DataArray>\n");
    }
    if (p.attenuation) {
        const auto array = [&](const char *name, const auto &value) {
            out.text(std::string("<DataArray type=\"") + pointType + "\" Name=\"" + name + "\" format=\"ascii\">\n");
            out.rows(p.nx, p.nz, kNumberChars + 1, [&](int i, int k, char *c) {
                c = put(c, value(points.index(i, k)));
                *c++ = ' ';
                return c;
            });
            out.text("\n</DataArray>\n");
        };
        if (includeMaterial) {
            array("Qp", [&](size_t q) { return points.stored(m.qp(q)); });
            array("Qs", [&](size_t q) { return points.stored(m.qs(q)); });
        }
        if (selected[3] || selected[4] || selected[5]) array("Syy", [&](size_t q) { return fields[5][q]; });
    }
    out.text(std::string("</PointData><Points><DataArray type=\"") + pointType + "\" NumberOfComponents=\"3\" format=\"ascii\">\n");
    out.rows(p.nx, p.nz, 2 * kNumberChars + 4, [&](int i, int k, char *c) {
        const size_t q = points.index(i, k);
        c = put(put(put(c, points.coordinate(q, 0)), " 0 "), points.coordinate(q, 1));
        *c++ = ' ';
        return c;
    });
    out.text("</DataArray></Points></Piece></StructuredGrid></VTKFile>\n");
    require(out.close(), "2D VTS write failed.");
}
} // namespace

extern "C" {
int swl_sim2d_write_model_vts(const swl_sim2d *s, const char *path, int cpu_threads, int cuda_device) {
    return swl::guarded([&] {
        require(path != nullptr, "Choose a model file name.");
        writeDocument(s, path, cpu_threads, cuda_device);
    });
}
int swl_model_vts_read(const char *path, swl_model_doc2d **out) {
    return swl::guarded([&] {
        require(path != nullptr && out != nullptr, "invalid model document arguments");
        *out = readDocument(path).release();
    });
}
void swl_model_doc2d_destroy(swl_model_doc2d *d) { delete d; }
int swl_model_doc2d_params(const swl_model_doc2d *d, swl_params2d *out) {
    return swl::guarded([&] {
        require(d != nullptr && out != nullptr, "invalid model document arguments");
        *out = d->p;
    });
}
int swl_model_doc2d_info(const swl_model_doc2d *d, double *out) {
    return swl::guarded([&] {
        require(d != nullptr && out != nullptr, "invalid model document arguments");
        const double v[] = {double(d->backend), double(d->cpuThreads), double(d->cudaDevice), d->referenceDt, double(d->version)};
        std::memcpy(out, v, sizeof v);
    });
}
int swl_sim2d_write_snapshot(const swl_sim2d *s, const char *path, const int32_t *components, int include_material) {
    return swl::guarded([&] {
        require(path != nullptr && components != nullptr, "invalid snapshot arguments");
        writeWavefield(s, path, components, include_material != 0);
    });
}
} // extern "C"
