// segy.cpp — receiver records as SEG-Y rev 2 (writer) and rev 0/1/2 (reader).
//
// Byte-compatible with the reference's SeisWave::writeSegy / readSegy (include/SeisCommon/SegyIo.hpp:5-19,
// src/common/SegyIo.cpp): the same four export profiles, the same textual / binary / trace header fields,
// the same IBM-float rounding, the same reader limits and error messages, so a file written here is the
// file the reference writes for the same records and either reader decodes it to the same values.
//   writer: SegyIo.cpp:93-290    reader: SegyIo.cpp:292-450    CP037 textual header: SegyIo.cpp:70-78
#include "session2d.hpp"

#include <algorithm>
#include <bit>
#include <cmath>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <limits>
#include <sstream>
#include <string>
#include <vector>

namespace swl {
namespace {

using Bytes = std::vector<unsigned char>;

// big-endian (or little-endian when `le`) unsigned field of `n` bytes
uint64_t field(const Bytes &b, size_t at, int n, bool le = false) {
    uint64_t v = 0;
    for (int i = 0; i < n; ++i) v = v << 8 | b[at + (le ? n - 1 - i : i)];
    return v;
}
int16_t field16(const Bytes &b, size_t at, bool le) { return std::bit_cast<int16_t>(static_cast<uint16_t>(field(b, at, 2, le))); }
int32_t field32(const Bytes &b, size_t at, bool le) { return std::bit_cast<int32_t>(static_cast<uint32_t>(field(b, at, 4, le))); }
void store(Bytes &b, size_t at, uint64_t v, int n) { // big-endian
    for (int i = n - 1; i >= 0; --i, v >>= 8) b[at + i] = static_cast<unsigned char>(v & 0xff);
}

// ASCII 0..127 -> EBCDIC code page 037 (IBM CP037), the encoding of a traditional SEG-Y textual header
constexpr unsigned char kCp037[128] = {
    /* 0x00 */ 0x00, 0x01, 0x02, 0x03, 0x37, 0x2d, 0x2e, 0x2f, 0x16, 0x05, 0x25, 0x0b, 0x0c, 0x0d, 0x0e, 0x0f,
    /* 0x10 */ 0x10, 0x11, 0x12, 0x13, 0x3c, 0x3d, 0x32, 0x26, 0x18, 0x19, 0x3f, 0x27, 0x1c, 0x1d, 0x1e, 0x1f,
    /* ' '  */ 0x40, 0x5a, 0x7f, 0x7b, 0x5b, 0x6c, 0x50, 0x7d, 0x4d, 0x5d, 0x5c, 0x4e, 0x6b, 0x60, 0x4b, 0x61,
    /* '0'  */ 0xf0, 0xf1, 0xf2, 0xf3, 0xf4, 0xf5, 0xf6, 0xf7, 0xf8, 0xf9, 0x7a, 0x5e, 0x4c, 0x7e, 0x6e, 0x6f,
    /* '@'  */ 0x7c, 0xc1, 0xc2, 0xc3, 0xc4, 0xc5, 0xc6, 0xc7, 0xc8, 0xc9, 0xd1, 0xd2, 0xd3, 0xd4, 0xd5, 0xd6,
    /* 'P'  */ 0xd7, 0xd8, 0xd9, 0xe2, 0xe3, 0xe4, 0xe5, 0xe6, 0xe7, 0xe8, 0xe9, 0xba, 0xe0, 0xbb, 0xb0, 0x6d,
    /* '`'  */ 0x79, 0x81, 0x82, 0x83, 0x84, 0x85, 0x86, 0x87, 0x88, 0x89, 0x91, 0x92, 0x93, 0x94, 0x95, 0x96,
    /* 'p'  */ 0x97, 0x98, 0x99, 0xa2, 0xa3, 0xa4, 0xa5, 0xa6, 0xa7, 0xa8, 0xa9, 0xc0, 0x4f, 0xd0, 0xa1, 0x07};

// IBM System/360 single precision: sign, 7-bit excess-64 base-16 exponent, 24-bit fraction 0.F;
// the fraction is rounded to nearest (ties away from zero) and renormalised when it carries out.
uint32_t to_ibm(double v) {
    if (v == 0) return 0;
    const uint32_t sign = std::signbit(v) ? 0x80000000u : 0u;
    int e2 = 0;
    std::frexp(std::fabs(v), &e2);                             // |v| = m 2^e2, m in [0.5, 1)
    int e16 = static_cast<int>(std::ceil(e2 / 4.0));           // |v| < 16^e16
    if (e16 < -64) return 0;                                   // underflow: smaller than any normalised value
    require(e16 <= 63, "SEG-Y IBM32 sample overflow; reduce the export multiplier.");
    auto frac = static_cast<uint32_t>(std::llround(std::ldexp(std::fabs(v), 24 - 4 * e16)));
    if (frac == 0x1000000u) frac = 0x100000u, ++e16;         // rounding carried into the next hex digit
    require(e16 <= 63, "SEG-Y IBM32 rounding overflow.");
    return sign | static_cast<uint32_t>(e16 + 64) << 24 | frac;
}
double from_ibm(uint32_t bits) {
    const uint32_t frac = bits & 0xffffffu;
    double v = frac ? std::ldexp(static_cast<double>(frac), 4 * (static_cast<int>(bits >> 24 & 0x7f) - 64) - 24) : 0;
    return bits & 0x80000000u ? -v : v;
}
int bytes_per_sample(int format) {
    return format == 1 || format == 2 || format == 5 ? 4 : format == 3 ? 2 : format == 6 ? 8 : format == 8 ? 1 : 0;
}
double header_scalar(int16_t s) { return s > 0 ? s : s < 0 ? -1.0 / s : 1.0; }
int32_t to_mm(double metres) {
    require(std::isfinite(metres) && std::fabs(metres * 1000) < std::numeric_limits<int32_t>::max(),
            "SEG-Y coordinates exceed the supported millimetre range.");
    return static_cast<int32_t>(std::llround(metres * 1000));
}
std::string g17(double v) {
    std::ostringstream s;
    s.imbue(std::locale::classic());
    s << std::setprecision(17) << v;
    return s.str();
}

struct Record {
    double position[3], source[3], start;
    std::vector<double> samples;
};

} // namespace
} // namespace swl

struct swl_segy_records {
    double dt = 0;
    std::string component = "Amplitude", units = "raw", description;
    std::vector<swl::Record> traces;
};

namespace swl {
namespace {

enum Profile { ExactIeee32 = 0, ExactIeee64 = 1, LegacyIbm32 = 2, CompatibleIeee32 = 3 };

void write_segy(const std::string &path, const swl_segy_trace *traces, int ntraces, int64_t nsrc, double dt,
                const std::string &component, const std::string &units, int profile, double gain, double interval) {
    const bool compatible = profile == LegacyIbm32 || profile == CompatibleIeee32;
    const bool f64 = profile == ExactIeee64, ibm = profile == LegacyIbm32;
    require(profile >= 0 && profile <= 3, "Unknown SEG-Y export profile.");
    require(std::isfinite(gain) && gain > 0, "SEG-Y export multiplier must be finite and positive.");
    require(std::isfinite(interval) && interval >= 0, "Invalid SEG-Y output sampling interval.");
    require(ntraces > 0 && traces != nullptr, "Select an existing raw component before SEG-Y export.");
    require(std::isfinite(dt) && dt > 0, "Invalid source sampling interval.");
    const double odt = interval > 0 ? interval : dt;
    require(odt > 0 && odt * 1e6 <= 65535 && (!compatible || odt * 1e6 >= 1),
            "SEG-Y sample interval must be positive and at most 65535 microseconds.");
    require(!compatible || std::fabs(odt * 1e6 - std::round(odt * 1e6)) < 1e-7,
            "Legacy SEG-Y requires whole-microsecond dt; select 0.5 or 1 ms output sampling.");
    require(nsrc > 0, "SEG-Y export requires nonempty traces.");
    const double n_out = odt == dt ? static_cast<double>(nsrc) : std::floor((nsrc - 1) * dt / odt + 1e-9) + 1;
    require(std::isfinite(n_out) && n_out >= 1 && n_out <= 65535, "SEG-Y export requires 1..65535 output samples per trace.");
    const auto ns = static_cast<size_t>(n_out);
    require(ntraces <= 100000, "Too many SEG-Y traces.");
    for (int r = 0; r < ntraces; ++r) {
        const auto &t = traces[r];
        require(t.samples != nullptr, "SEG-Y export requires equal trace lengths.");
        require(std::isfinite(t.start_time) && std::fabs(t.start_time * 1000) <= 32767 &&
                    std::fabs(t.start_time * 1000 - std::round(t.start_time * 1000)) < 1e-7,
                "SEG-Y export delay must be representable as whole milliseconds.");
        for (int c = 0; c < 3; ++c) to_mm(t.position[c]), to_mm(t.source[c]);
        for (int64_t s = 0; s < nsrc; ++s) {
            const double v = t.samples[s];
            require(std::isfinite(v) && std::isfinite(v * gain) &&
                        (f64 || ibm || std::fabs(v * gain) <= std::numeric_limits<float>::max()),
                    "SEG-Y samples/multiplier overflow the selected precision.");
        }
    }
    // output sample t: the input itself, or the linear interpolation at t*odt (clamped to the record)
    auto value = [&](const swl_segy_trace &tr, size_t t) {
        if (odt == dt) return tr.samples[t] * gain;
        const double u = std::clamp(t * odt / dt, 0.0, static_cast<double>(nsrc - 1));
        const auto i = static_cast<size_t>(u);
        return std::lerp(tr.samples[i], tr.samples[std::min<size_t>(i + 1, nsrc - 1)], u - i) * gain;
    };
    size_t underflow = 0;
    double peak = 0;
    for (int r = 0; r < ntraces; ++r)
        for (size_t t = 0; t < ns; ++t) {
            const double v = value(traces[r], t);
            peak = std::max(peak, std::fabs(v));
            if (ibm) underflow += v != 0 && to_ibm(v) == 0;
            else if (!f64) underflow += v != 0 && static_cast<float>(v) == 0;
        }
    // textual header: 40 lines of 80 ASCII characters ("C nn " prefixes), CP037 for the IBM profile
    Bytes text(3200, ' ');
    auto line = [&](int row, const std::string &body) {
        const std::string s = (row < 10 ? "C " : "C") + std::to_string(row) + " " + body;
        std::copy_n(s.begin(), std::min<size_t>(80, s.size()), text.begin() + (row - 1) * 80);
    };
    line(1, "SEISWAVE LAB - ELASTIC WAVE RECEIVER RECORDS");
    line(2, "COMPONENT=" + component + " UNITS=" + (gain == 1 ? units : "scaled(" + units + ")"));
    line(3, "EXACT DT (SECONDS)=" + g17(odt));
    line(4, std::string("SEG-Y REV ") + (compatible ? "1.0" : "2.0") + ", BIG ENDIAN, " + (ibm ? "IBM32" : f64 ? "IEEE64" : "IEEE32"));
    line(5, compatible ? "DT: BINARY BYTES 3217-3218; TRACE BYTES 117-118 (MICROSECONDS)"
                       : "EXACT DT: BINARY HEADER BYTES 3273-3280 (IEEE64 MICROSECONDS)");
    line(6, "COORDINATES/ELEVATIONS IN METRES; SCALARS=-1000; Z POSITIVE DOWN");
    line(7, "NO DISPLAY AGC OR TRACE NORMALIZATION; FIRST INPUT SAMPLE INCLUDED");
    line(8, "X IS INLINE, Y CROSSLINE; ONE VELOCITY COMPONENT PER FILE");
    line(9, "EXPORT_MULTIPLIER=" + g17(gain) + "; ORIGINAL_VALUE=STORED_VALUE/MULTIPLIER");
    line(10, "FFID=1; CHANNEL=1..N; CDP=1..N (ORDINAL, NOT A BINNED CMP GRID)");
    line(11, "SOURCE_DT_SECONDS=" + g17(dt) + "; RESAMPLING=" + (odt == dt ? "NONE" : "LINEAR"));
    line(12, "OUTPUT_PEAK=" + g17(peak) + "; UNDERFLOWED_SAMPLES=" + std::to_string(underflow));
    line(13, compatible ? "TRACE_ID=1 LIVE SEISMIC; COMPONENT IDENTIFIED ON C 2 / FILENAME"
                        : "TRACE_ID: 12=VERTICAL, 14=INLINE; COMPONENT IDENTIFIED ON C 2");
    line(40, "END TEXTUAL HEADER");
    if (ibm)
        for (auto &c : text) c = kCp037[c < 128 ? c : '?'];
    // binary file header (400 bytes; offsets relative to byte 3201)
    Bytes bin(400, 0);
    const auto us = static_cast<uint16_t>(std::clamp<long long>(std::llround(odt * 1e6), 1, 65535));
    store(bin, 12, std::min<size_t>(65535, ntraces), 2); // data traces per ensemble
    store(bin, 16, us, 2), store(bin, 18, us, 2);       // sample interval (this reel / field)
    store(bin, 20, ns, 2), store(bin, 22, ns, 2);       // samples per trace (this reel / field)
    store(bin, 24, ibm ? 1 : f64 ? 6 : 5, 2);           // sample format code
    store(bin, 26, 1, 2);                               // ensemble fold
    store(bin, 28, 1, 2);                               // trace sorting: as recorded
    store(bin, 54, 1, 2);                               // measurement system: metres
    if (!compatible) {
        store(bin, 60, ntraces, 4);                                 // extended traces per ensemble
        store(bin, 68, ns, 4);                                      // extended samples per trace
        store(bin, 72, std::bit_cast<uint64_t>(odt * 1e6), 8);      // extended interval, IEEE64 microseconds
        store(bin, 80, std::bit_cast<uint64_t>(odt * 1e6), 8);
        store(bin, 88, ns, 4);
        store(bin, 96, 0x01020304u, 4);                             // byte-order marker
    }
    bin[300] = compatible ? 1 : 2; // revision major
    store(bin, 302, 1, 2);         // fixed-length traces
    if (!compatible) {
        store(bin, 312, ntraces, 8); // data traces in the file
        store(bin, 320, 3600, 8);    // byte offset of the first trace
    }
    namespace fs = std::filesystem;
    const fs::path target(path);
    if (!target.parent_path().empty()) fs::create_directories(target.parent_path());
    fs::path pending = target;
    pending += ".partial";
    for (int k = 1; fs::exists(pending); ++k) pending = target, pending += ".partial" + std::to_string(k);
    try {
        std::ofstream out(pending, std::ios::binary);
        require(static_cast<bool>(out), "Cannot create SEG-Y output.");
        auto emit = [&](const Bytes &b) {
            out.write(reinterpret_cast<const char *>(b.data()), static_cast<std::streamsize>(b.size()));
            require(static_cast<bool>(out), "SEG-Y write failed.");
        };
        emit(text), emit(bin);
        const int id = compatible ? 1 : component == "Vz" ? 12 : component == "Vx" ? 14 : 1;
        const int unit = gain != 1 ? 0 : units == "m/s" ? 6 : units == "Pa" ? 1 : 0;
        for (int r = 0; r < ntraces; ++r) {
            const auto &tr = traces[r];
            Bytes h(240, 0);
            store(h, 0, r + 1, 4), store(h, 4, r + 1, 4); // trace sequence in line / in file
            store(h, 8, 1, 4), store(h, 12, r + 1, 4);    // FFID, channel
            store(h, 16, 1, 4), store(h, 20, r + 1, 4), store(h, 24, 1, 4); // source point, CDP, CDP trace
            store(h, 28, id, 2);
            store(h, 30, 1, 2), store(h, 32, 1, 2), store(h, 34, 1, 2);
            const double dx = tr.position[0] - tr.source[0], dy = tr.position[1] - tr.source[1];
            const double offset = std::copysign(std::hypot(dx, dy), dx != 0 ? dx : dy);
            require(std::fabs(offset) <= std::numeric_limits<int32_t>::max(), "SEG-Y offset exceeds Int32 metres.");
            store(h, 36, static_cast<uint32_t>(static_cast<int32_t>(std::llround(offset))), 4);
            store(h, 40, static_cast<uint32_t>(to_mm(-tr.position[2])), 4); // receiver elevation (z down)
            store(h, 48, static_cast<uint32_t>(to_mm(tr.source[2])), 4);    // source depth
            store(h, 68, static_cast<uint16_t>(-1000), 2), store(h, 70, static_cast<uint16_t>(-1000), 2);
            store(h, 72, static_cast<uint32_t>(to_mm(tr.source[0])), 4), store(h, 76, static_cast<uint32_t>(to_mm(tr.source[1])), 4);
            store(h, 80, static_cast<uint32_t>(to_mm(tr.position[0])), 4), store(h, 84, static_cast<uint32_t>(to_mm(tr.position[1])), 4);
            store(h, 88, 1, 2); // coordinate units: length
            store(h, 108, static_cast<uint16_t>(std::llround(tr.start_time * 1000)), 2);
            store(h, 114, ns, 2), store(h, 116, us, 2);
            store(h, 180, static_cast<uint32_t>(to_mm((tr.source[0] + tr.position[0]) * .5)), 4);
            store(h, 184, static_cast<uint32_t>(to_mm((tr.source[1] + tr.position[1]) * .5)), 4);
            store(h, 202, unit, 2);
            emit(h);
            Bytes s(ns * (f64 ? 8 : 4));
            for (size_t t = 0; t < ns; ++t) {
                const double v = value(tr, t);
                if (ibm) store(s, 4 * t, to_ibm(v), 4);
                else if (f64) store(s, 8 * t, std::bit_cast<uint64_t>(v), 8);
                else store(s, 4 * t, std::bit_cast<uint32_t>(static_cast<float>(v)), 4);
            }
            emit(s);
        }
        out.close();
        require(static_cast<bool>(out), "SEG-Y final flush failed.");
        std::error_code ec;
        fs::rename(pending, target, ec); // readers see the old file or the complete new one
        if (ec) throw fs::filesystem_error("Could not replace completed output file", pending, target, ec);
    } catch (...) {
        std::error_code ec;
        fs::remove(pending, ec);
        throw;
    }
}

swl_segy_records read_segy(const std::string &path) {
    namespace fs = std::filesystem;
    std::ifstream in(path, std::ios::binary);
    require(static_cast<bool>(in), "Cannot open SEG-Y input.");
    const uint64_t size = fs::file_size(path);
    require(size >= 3600, "SEG-Y input is shorter than its file headers.");
    auto take = [&](Bytes &b, size_t n) {
        b.resize(n);
        require(static_cast<bool>(in.read(reinterpret_cast<char *>(b.data()), static_cast<std::streamsize>(n))),
                "Truncated SEG-Y header or sample data.");
    };
    Bytes text, bin;
    take(text, 3200), take(bin, 400);
    const uint64_t marker = field(bin, 96, 4);
    require(marker == 0 || marker == 0x01020304u || marker == 0x04030201u,
            "Unsupported SEG-Y byte ordering (pair-swapped files are not supported).");
    bool le = marker == 0x04030201u;
    if (!marker && !bytes_per_sample(static_cast<int>(field(bin, 24, 2))) && bytes_per_sample(static_cast<int>(field(bin, 24, 2, true))))
        le = true; // no marker: trust the byte order in which the format code is a known one
    const int format = static_cast<int>(field(bin, 24, 2, le)), width = bytes_per_sample(format), revision = bin[300];
    require(width > 0, "Unsupported SEG-Y sample format. Supported codes: 1,2,3,5,6,8.");
    require(revision <= 2, "Unsupported SEG-Y revision.");
    require(revision < 2 || field(bin, 306, 4, le) == 0,
            "SEG-Y additional trace headers are not supported; convert to standard 240-byte trace headers.");
    uint64_t at = 3600;
    const int extended = revision ? field16(bin, 304, le) : 0;
    if (revision >= 2 && field(bin, 320, 8, le)) {
        at = field(bin, 320, 8, le);
    } else {
        require(extended >= 0 && extended <= 1024, "SEG-Y needs an explicit, bounded extended textual header count.");
        at += static_cast<uint64_t>(extended) * 3200;
    }
    require(at >= 3600 && at <= size, "SEG-Y first trace offset is outside the file.");
    in.seekg(static_cast<std::streamoff>(at));
    const int32_t trailer = revision >= 2 ? field32(bin, 328, le) : 0;
    require(trailer >= 0, "SEG-Y variable trailer counts are not supported.");
    require(static_cast<uint64_t>(trailer) * 3200 <= size - at, "Invalid SEG-Y data trailer size.");
    const uint64_t end = size - static_cast<uint64_t>(trailer) * 3200;
    uint64_t ns = field(bin, 20, 2, le);
    if (revision >= 2 && field(bin, 68, 4, le)) ns = field(bin, 68, 4, le);
    double interval0 = static_cast<double>(field(bin, 16, 2, le));
    if (revision >= 2 && std::bit_cast<double>(field(bin, 72, 8, le)) != 0) interval0 = std::bit_cast<double>(field(bin, 72, 8, le));
    const bool fixed = revision && field(bin, 302, 2, le) == 1;
    const double unit_length = field(bin, 54, 2, le) == 2 ? .3048 : 1; // feet
    swl_segy_records out;
    // the textual header in ASCII (decoded from CP037 when it starts with an EBCDIC 'C')
    std::string ascii(text.begin(), text.end());
    if (text.front() == kCp037['C'])
        for (size_t n = 0; n < text.size(); ++n) {
            const auto *hit = std::find(std::begin(kCp037), std::end(kCp037), text[n]);
            ascii[n] = hit == std::end(kCp037) ? '?' : static_cast<char>(hit - std::begin(kCp037));
        }
    const bool ours = ascii.find("SEISWAVE LAB") != std::string::npos;
    auto token = [&](const char *key) { // value after `key` on textual line 2, up to the next space
        const std::string row = ascii.substr(80, 80);
        const size_t k = row.find(key);
        if (k == std::string::npos) return std::string();
        const size_t b = k + std::strlen(key);
        return row.substr(b, row.find(' ', b) - b);
    };
    if (ours && !token("UNITS=").empty()) out.units = token("UNITS=");
    uint64_t decoded = 0;
    int id0 = -1, unit0 = -1;
    while (at < end) {
        require(out.traces.size() < 4096, "SEG-Y import exceeds 4096 trace positions.");
        require(end - at >= 240, "Incomplete SEG-Y trace header.");
        Bytes h, s;
        take(h, 240);
        at += 240;
        const uint64_t n = fixed ? ns : (field(h, 114, 2, le) ? field(h, 114, 2, le) : ns);
        require(n > 0 && n <= 1000000 && n <= ((256ULL << 20) - decoded) / sizeof(double),
                "SEG-Y sample count exceeds the import memory budget (256 MiB).");
        double us = fixed && interval0 > 0 ? interval0 : static_cast<double>(field(h, 116, 2, le));
        if (us == 0) us = interval0;
        require(std::isfinite(us) && us >= 1e-6 && us <= 1e9, "SEG-Y sample interval is missing, invalid or outside 1 ps..1000 s.");
        if (out.traces.empty()) out.dt = us * 1e-6;
        else
            require(std::fabs(out.dt - us * 1e-6) <= std::max(1e-12, out.dt * 1e-9),
                    "SEG-Y traces have different sample intervals; resample before import.");
        require(n * width <= end - at, "Truncated SEG-Y trace samples.");
        take(s, static_cast<size_t>(n) * width);
        at += n * width;
        decoded += n * sizeof(double);
        Record tr;
        const double xy = header_scalar(field16(h, 70, le)), zs = header_scalar(field16(h, 68, le));
        require(field(h, 88, 2, le) <= 1, "SEG-Y angular coordinates are not supported in this Cartesian record viewer.");
        tr.position[0] = field32(h, 80, le) * xy * unit_length, tr.position[1] = field32(h, 84, le) * xy * unit_length;
        tr.position[2] = -static_cast<double>(field32(h, 40, le)) * zs * unit_length;
        tr.source[0] = field32(h, 72, le) * xy * unit_length, tr.source[1] = field32(h, 76, le) * xy * unit_length;
        tr.source[2] = (field32(h, 48, le) - static_cast<double>(field32(h, 44, le))) * zs * unit_length;
        tr.start = field16(h, 108, le) * .001 * (revision ? header_scalar(field16(h, 214, le)) : 1);
        const int weight = field16(h, 168, le);
        require(weight >= -1022 && weight <= 1022, "SEG-Y trace weighting is outside the supported finite range.");
        tr.samples.resize(n);
        for (uint64_t t = 0; t < n; ++t) {
            const size_t o = t * width;
            double v;
            switch (format) {
            case 5: v = std::bit_cast<float>(static_cast<uint32_t>(field(s, o, 4, le))); break;
            case 6: v = std::bit_cast<double>(field(s, o, 8, le)); break;
            case 2: v = field32(s, o, le); break;
            case 3: v = field16(s, o, le); break;
            case 8: v = std::bit_cast<int8_t>(s[o]); break;
            default: v = from_ibm(static_cast<uint32_t>(field(s, o, 4, le))); break;
            }
            v = std::ldexp(v, -weight);
            require(std::isfinite(v), "SEG-Y contains nonfinite samples.");
            tr.samples[t] = v;
        }
        const int id = static_cast<int>(field(h, 28, 2, le));
        id0 = id0 < 0 ? id : id0 != id ? 0 : id0;
        const int unit = revision ? field16(h, 202, le) : 0;
        unit0 = unit0 < 0 ? unit : unit0 != unit ? 1000 : unit0;
        out.traces.push_back(std::move(tr));
    }
    require(!out.traces.empty(), "SEG-Y contains no data traces.");
    out.component = id0 == 12 ? "Vz" : id0 == 14 ? "Vx" : "Amplitude";
    if (ours && !token("COMPONENT=").empty()) out.component = token("COMPONENT=");
    static const char *const kUnits[] = {"raw", "Pa", "V", "mV", "A", "m", "m/s", "m/s2", "N", "W"};
    if (unit0 > 0 && unit0 < 10) out.units = kUnits[unit0];
    if (unit0 == 1000) out.units = "mixed";
    out.description = "SEG-Y rev " + std::to_string(revision) + " | format " + std::to_string(format) +
                      (le ? " | little endian" : " | big endian");
    return out;
}

void copy_text(const std::string &s, char *out, int capacity) {
    if (!out || capacity <= 0) return;
    const size_t n = std::min<size_t>(s.size(), static_cast<size_t>(capacity - 1));
    std::memcpy(out, s.data(), n);
    out[n] = 0;
}

} // namespace
} // namespace swl

extern "C" {
int swl_segy_write(const char *path, const swl_segy_trace *traces, int ntraces, int64_t count, double dt,
                   const char *component, const char *units, int profile, double gain, double sample_interval) {
    return swl::guarded([&] {
        swl::require(path && component && units, "invalid SEG-Y arguments");
        swl::write_segy(path, traces, ntraces, count, dt, component, units, profile, gain, sample_interval);
    });
}
int swl_segy_read(const char *path, swl_segy_records **out) {
    return swl::guarded([&] {
        swl::require(path && out, "invalid SEG-Y arguments");
        *out = new swl_segy_records(swl::read_segy(path));
    });
}
void swl_segy_records_destroy(swl_segy_records *r) { delete r; }
int swl_segy_records_info(const swl_segy_records *r, int *ntraces, double *dt, char *component, char *units,
                          char *description, int text_capacity) {
    return swl::guarded([&] {
        swl::require(r != nullptr, "invalid SEG-Y records");
        if (ntraces) *ntraces = static_cast<int>(r->traces.size());
        if (dt) *dt = r->dt;
        swl::copy_text(r->component, component, text_capacity);
        swl::copy_text(r->units, units, text_capacity);
        swl::copy_text(r->description, description, text_capacity);
    });
}
int swl_segy_records_trace(const swl_segy_records *r, int index, double *position, double *source, double *start_time,
                           int64_t *count) {
    return swl::guarded([&] {
        swl::require(r && index >= 0 && index < static_cast<int>(r->traces.size()), "SEG-Y trace index out of range");
        const auto &t = r->traces[index];
        for (int c = 0; c < 3; ++c) {
            if (position) position[c] = t.position[c];
            if (source) source[c] = t.source[c];
        }
        if (start_time) *start_time = t.start;
        if (count) *count = static_cast<int64_t>(t.samples.size());
    });
}
int swl_segy_records_samples(const swl_segy_records *r, int index, double *out) {
    return swl::guarded([&] {
        swl::require(r && index >= 0 && index < static_cast<int>(r->traces.size()) && out, "SEG-Y trace index out of range");
        std::copy(r->traces[index].samples.begin(), r->traces[index].samples.end(), out);
    });
}
// The session's receivers as one component per file (as src/two_d/Experiment2D.cpp:44-71 assembles them):
// trace positions (x, 0, z), the source at the coordinates of the source node, dt = recordDt.
int swl_sim2d_write_segy(const swl_sim2d *s, const char *path, int component, int profile, double gain,
                         double sample_interval) {
    return swl::guarded([&] {
        swl::require(s && path, "invalid SEG-Y arguments");
        swl::require(component == 0 || component == 1, "Select an existing raw component before SEG-Y export.");
        const auto &m = s->model;
        const auto &p = m.logical;
        const int si = (p.source_x < 0 ? p.nx / 2 : p.source_x) + 3 + m.offsetX, sk = p.source_z + 3;
        const size_t q = static_cast<size_t>(si) + static_cast<size_t>(m.nx) * sk;
        const double sx = m.coordinates[2 * q], sz = m.coordinates[2 * q + 1];
        const auto &rows = component == 0 ? s->vx : s->vz;
        const size_t n = rows.size();
        swl::require(n > 0, "Select an existing raw component before SEG-Y export.");
        std::vector<swl_segy_trace> traces(n);
        for (size_t r = 0; r < n; ++r) {
            swl::require(rows[r].size() == rows[0].size(), "SEG-Y export requires equal trace lengths.");
            auto &t = traces[r];
            t.position[0] = m.receiverPosition[2 * r], t.position[1] = 0, t.position[2] = m.receiverPosition[2 * r + 1];
            t.source[0] = sx, t.source[1] = 0, t.source[2] = sz;
            t.start_time = 0;
            t.samples = rows[r].data();
        }
        swl::write_segy(path, traces.data(), static_cast<int>(n), static_cast<int64_t>(rows[0].size()), s->recordDt(),
                        component == 0 ? "Vx" : "Vz", "m/s", profile, gain, sample_interval);
    });
}
} // extern "C"
