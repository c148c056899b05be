// netio.cpp -- the reference's network files: readers and the matching writers.
//
// Files, in a directory given by the caller (the reference reads and writes the current
// directory):
//   sample_geom.csv      line 1 "x0, y0, z0" (fixed, prec_cnt digits), line 2 "lx, ly, lz"
//                        (6 significant digits)             Generate_Network.cpp:6118-6134
//                        read by Read_Network::Read_sample_geometry, Read_Network.cpp:153-199
//   cnt_struct.csv       "n, n" then one "n_points, radius" line per tube (6 significant)
//   cnt_coordinates.csv  one "x, y, z" line per point, tube by tube (fixed, prec_cnt)
//                        Printer::Print_cnt_points_and_structure, Printer.cpp:273-316;
//                        Read_cnt_data_from_csv, Read_Network.cpp:201-344
//   cnt_struct.dat       int32 n, then per tube int32 n_points, float64 radius
//   cnt_coordinates.dat  float64 x, y, z per point, tube by tube
//                        Printer.cpp:344-380; Read_cnt_data_from_dat, Read_Network.cpp:346-462
//   gnp_data.csv         "l, l, t, theta, phi, cx, cy, cz" per plate (l, t and the angles
//                        with 6 significant digits, the centre fixed with prec_gnp digits)
//                        Printer::Print_gnp_data, Printer.cpp:140-167;
//                        Read_gnp_data_from_csv, Read_Network.cpp:464-571
//   gnp_data.dat         int32 n, then per plate float64 l, t, theta, phi, cx, cy, cz
//                        Printer.cpp:205-243; Read_gnp_data_from_dat, Read_Network.cpp:573-678
//
// Text fields are parsed as the reference's stringstreams do: leading white space is
// skipped, the number is read (strtod/strtol, the conversion libstdc++'s extractors use),
// and exactly one character after it is skipped (the ignore() that eats the comma).  Where
// the reference would read a short or malformed file into zeros or stale values, these
// readers stop with an error naming the file and line.  The plate rotation is
// Generate_Network::Get_transformation_matrix (Generate_Network.cpp:3466-3490); the writer
// recovers the angles from it with Printer::Recover_angle (Printer.cpp:170-202).
#include "cicy_netio.h"

#include <cerrno>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

constexpr double kPi = 3.1415926535897932384626433832795;  // Input_Reader.h:23
constexpr double kZero = 1.0e-10;                           // MathMatrix.h:18

struct Err {
    char *buf;
    int64_t len;
    int fail(const char *fmt, ...) {
        if (buf && len > 0) {
            va_list ap;
            va_start(ap, fmt);
            std::vsnprintf(buf, (size_t)len, fmt, ap);
            va_end(ap);
        }
        return 0;
    }
};

std::string path_of(const char *dir, const char *name) {
    std::string p = dir && *dir ? dir : ".";
    if (p.back() != '/') p += '/';
    return p + name;
}

// One text line, read field by field as "ss >> value; ss.ignore();".
struct Fields {
    const char *p;
    bool ok = true;
    explicit Fields(const std::string &line) : p(line.c_str()) {}
    void skip_ws() {
        while (*p == ' ' || *p == '\t' || *p == '\r' || *p == '\n' || *p == '\v' || *p == '\f')
            p++;
    }
    double real() {
        skip_ws();
        char *end = nullptr;
        errno = 0;
        const double v = std::strtod(p, &end);
        if (end == p) {
            ok = false;
            return 0.0;
        }
        p = end;
        if (*p) p++;  // ignore()
        return v;
    }
    long long integer() {
        skip_ws();
        char *end = nullptr;
        const long long v = std::strtoll(p, &end, 10);
        if (end == p) {
            ok = false;
            return 0;
        }
        p = end;
        if (*p) p++;
        return v;
    }
};

bool read_lines(const std::string &path, std::vector<std::string> &lines) {
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    std::string all;
    char buf[1 << 16];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) all.append(buf, n);
    std::fclose(f);
    // getline semantics: a final line without '\n' still counts, an empty tail does not
    size_t a = 0;
    while (a < all.size()) {
        size_t b = all.find('\n', a);
        if (b == std::string::npos) b = all.size();
        lines.emplace_back(all, a, b - a);
        a = b + 1;
    }
    return true;
}

struct Bin {
    std::vector<char> data;
    size_t pos = 0;
    bool load(const std::string &path) {
        FILE *f = std::fopen(path.c_str(), "rb");
        if (!f) return false;
        char buf[1 << 16];
        size_t n;
        while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) data.insert(data.end(), buf, buf + n);
        std::fclose(f);
        return true;
    }
    template <class T>
    bool get(T &v) {
        if (pos + sizeof(T) > data.size()) return false;
        std::memcpy(&v, data.data() + pos, sizeof(T));
        pos += sizeof(T);
        return true;
    }
};

void rotation(double theta, double phi, double *m) {
    m[0] = std::cos(phi) * std::cos(theta);
    m[1] = -std::sin(phi);
    m[2] = std::cos(phi) * std::sin(theta);
    m[3] = std::sin(phi) * std::cos(theta);
    m[4] = std::cos(phi);
    m[5] = std::sin(phi) * std::sin(theta);
    m[6] = -std::sin(theta);
    m[7] = 0.0;
    m[8] = std::cos(theta);
}

double recover_angle(double c, double s) {
    const double two_pi = 2 * kPi;
    if (c > kZero) return s > kZero ? std::acos(c) : std::asin(s) + two_pi;
    return s > kZero ? std::acos(c) : two_pi - std::acos(c);
}

}  // namespace

extern "C" {

int cicy_read_sample_geometry(const char *dir, double *out, char *err, int64_t err_len) {
    Err e{err, err_len};
    if (!out) return e.fail("cicy_read_sample_geometry: null output");
    const std::string path = path_of(dir, "sample_geom.csv");
    std::vector<std::string> lines;
    if (!read_lines(path, lines)) return e.fail("failed to open %s", path.c_str());
    if (lines.size() < 2) return e.fail("%s: expected two lines", path.c_str());
    for (int l = 0; l < 2; l++) {
        Fields f(lines[(size_t)l]);
        for (int k = 0; k < 3; k++) out[3 * l + k] = f.real();
        if (!f.ok) return e.fail("%s: malformed line %d", path.c_str(), l + 1);
    }
    return 1;
}

int cicy_read_cnts(const char *dir, int fmt, int64_t *n_cnts, int64_t *n_points, int64_t *npts,
                   double *radii, double *xyz, char *err, int64_t err_len) {
    Err e{err, err_len};
    if (!n_cnts || !n_points || (fmt != 0 && fmt != 1))
        return e.fail("cicy_read_cnts: null counts or format not 0 (csv) / 1 (dat)");
    const bool fill = npts || radii || xyz;
    if (fill && (!npts || !radii || !xyz))
        return e.fail("cicy_read_cnts: give all three output arrays or none");
    std::vector<int64_t> np;
    std::vector<double> rad;
    if (fmt == 0) {
        const std::string path = path_of(dir, "cnt_struct.csv");
        std::vector<std::string> lines;
        if (!read_lines(path, lines)) return e.fail("failed to open %s", path.c_str());
        if (lines.empty()) return e.fail("%s is empty", path.c_str());
        Fields h(lines[0]);
        const long long n = h.integer(), n_check = h.integer();
        if (!h.ok || n < 0) return e.fail("%s: malformed line 1", path.c_str());
        if (n != n_check)
            return e.fail("%s: the first line gives different numbers of tubes: %lld and %lld",
                          path.c_str(), n, n_check);
        if ((long long)lines.size() < n + 1)
            return e.fail("%s ends before all %lld tubes were read", path.c_str(), n);
        for (long long i = 0; i < n; i++) {
            Fields f(lines[(size_t)i + 1]);
            const long long p = f.integer();
            const double r = f.real();
            if (!f.ok || p < 0) return e.fail("%s: malformed line %lld", path.c_str(), i + 2);
            np.push_back(p);
            rad.push_back(r);
        }
    } else {
        const std::string path = path_of(dir, "cnt_struct.dat");
        Bin b;
        if (!b.load(path)) return e.fail("failed to open %s", path.c_str());
        int32_t n = 0;
        if (!b.get(n) || n < 0) return e.fail("%s: no tube count", path.c_str());
        for (int32_t i = 0; i < n; i++) {
            int32_t p;
            double r;
            if (!b.get(p) || !b.get(r))
                return e.fail("%s ends before all %d tubes were read", path.c_str(), n);
            if (p < 0) return e.fail("%s: tube %d has %d points", path.c_str(), i, p);
            np.push_back(p);
            rad.push_back(r);
        }
    }
    int64_t total = 0;
    for (int64_t p : np) total += p;
    *n_cnts = (int64_t)np.size();
    *n_points = total;
    if (!fill) return 1;
    std::memcpy(npts, np.data(), np.size() * sizeof(int64_t));
    std::memcpy(radii, rad.data(), rad.size() * sizeof(double));
    if (fmt == 0) {
        const std::string path = path_of(dir, "cnt_coordinates.csv");
        std::vector<std::string> lines;
        if (!read_lines(path, lines)) return e.fail("failed to open %s", path.c_str());
        if ((int64_t)lines.size() < total)
            return e.fail("%s ends before all %lld points were read", path.c_str(),
                          (long long)total);
        for (int64_t q = 0; q < total; q++) {
            Fields f(lines[(size_t)q]);
            for (int k = 0; k < 3; k++) xyz[3 * q + k] = f.real();
            if (!f.ok) return e.fail("%s: malformed line %lld", path.c_str(), (long long)q + 1);
        }
    } else {
        const std::string path = path_of(dir, "cnt_coordinates.dat");
        Bin b;
        if (!b.load(path)) return e.fail("failed to open %s", path.c_str());
        for (int64_t q = 0; q < 3 * total; q++)
            if (!b.get(xyz[q]))
                return e.fail("%s ends before all %lld points were read", path.c_str(),
                              (long long)total);
    }
    return 1;
}

int cicy_read_gnps(const char *dir, int fmt, int64_t *n_gnps, double *geo, double *rot,
                   char *err, int64_t err_len) {
    Err e{err, err_len};
    if (!n_gnps || (fmt != 0 && fmt != 1))
        return e.fail("cicy_read_gnps: null count or format not 0 (csv) / 1 (dat)");
    std::vector<double> g;  // l, t, theta, phi, cx, cy, cz per plate
    if (fmt == 0) {
        const std::string path = path_of(dir, "gnp_data.csv");
        std::vector<std::string> lines;
        if (!read_lines(path, lines)) return e.fail("failed to open %s", path.c_str());
        for (size_t i = 0; i < lines.size(); i++) {
            Fields f(lines[i]);
            const double l = f.real();
            (void)f.real();  // the side length is written twice
            const double t = f.real(), theta = f.real(), phi = f.real();
            const double cx = f.real(), cy = f.real(), cz = f.real();
            if (!f.ok) return e.fail("%s: malformed line %zu", path.c_str(), i + 1);
            g.insert(g.end(), {l, t, theta, phi, cx, cy, cz});
        }
    } else {
        const std::string path = path_of(dir, "gnp_data.dat");
        Bin b;
        if (!b.load(path)) return e.fail("failed to open %s", path.c_str());
        int32_t n = 0;
        if (!b.get(n) || n < 0) return e.fail("%s: no plate count", path.c_str());
        for (int32_t i = 0; i < n; i++)
            for (int k = 0; k < 7; k++) {
                double v;
                if (!b.get(v))
                    return e.fail("%s ends before all %d plates were read", path.c_str(), n);
                g.push_back(v);
            }
    }
    const int64_t n = (int64_t)(g.size() / 7);
    const int64_t cap = *n_gnps;
    *n_gnps = n;
    if (!geo && !rot) return 1;
    if (cap < n) return e.fail("cicy_read_gnps: room for %lld plates, %lld in the file",
                               (long long)cap, (long long)n);
    for (int64_t i = 0; i < n; i++) {
        if (geo) std::memcpy(geo + 7 * i, g.data() + 7 * i, 7 * sizeof(double));
        if (rot) rotation(g[(size_t)(7 * i + 2)], g[(size_t)(7 * i + 3)], rot + 9 * i);
    }
    return 1;
}

int cicy_write_network_files(const char *dir, int fmt, int prec_cnt, int prec_gnp,
                             const double *sample, int64_t n_cnts, const int64_t *npts,
                             const double *radii, const double *xyz, int64_t n_gnps,
                             const double *lt, const double *rot, const double *centre,
                             char *err, int64_t err_len) {
    Err e{err, err_len};
    if ((fmt != 0 && fmt != 1) || !sample || n_cnts < 0 || n_gnps < 0 ||
        (n_cnts > 0 && (!npts || !radii || !xyz)) || (n_gnps > 0 && (!lt || !rot || !centre)))
        return e.fail("cicy_write_network_files: invalid arguments");
    // sample_geom.csv is text in both formats (Generate_Network.cpp:6118-6134)
    {
        const std::string path = path_of(dir, "sample_geom.csv");
        FILE *f = std::fopen(path.c_str(), "w");
        if (!f) return e.fail("failed to create %s", path.c_str());
        std::fprintf(f, "%.*f, %.*f, %.*f\n%g, %g, %g\n", prec_cnt, sample[0], prec_cnt,
                     sample[1], prec_cnt, sample[2], sample[3], sample[4], sample[5]);
        std::fclose(f);
    }
    if (n_cnts > 0) {
        const char *ns = fmt ? "cnt_struct.dat" : "cnt_struct.csv";
        const char *nc = fmt ? "cnt_coordinates.dat" : "cnt_coordinates.csv";
        FILE *fs = std::fopen(path_of(dir, ns).c_str(), fmt ? "wb" : "w");
        FILE *fc = std::fopen(path_of(dir, nc).c_str(), fmt ? "wb" : "w");
        if (!fs || !fc) {
            if (fs) std::fclose(fs);
            if (fc) std::fclose(fc);
            return e.fail("failed to create the tube files in %s", dir ? dir : ".");
        }
        if (fmt == 0) {
            std::fprintf(fs, "%lld, %lld\n", (long long)n_cnts, (long long)n_cnts);
        } else {
            const int32_t n = (int32_t)n_cnts;
            std::fwrite(&n, 4, 1, fs);
        }
        int64_t q = 0;
        for (int64_t i = 0; i < n_cnts; i++) {
            if (fmt == 0) {
                for (int64_t j = 0; j < npts[i]; j++, q++)
                    std::fprintf(fc, "%.*f, %.*f, %.*f\n", prec_cnt, xyz[3 * q], prec_cnt,
                                 xyz[3 * q + 1], prec_cnt, xyz[3 * q + 2]);
                std::fprintf(fs, "%d, %g\n", (int)npts[i], radii[i]);
            } else {
                const int32_t p = (int32_t)npts[i];
                std::fwrite(&p, 4, 1, fs);
                std::fwrite(&radii[i], 8, 1, fs);
                std::fwrite(xyz + 3 * q, 8, (size_t)(3 * npts[i]), fc);
                q += npts[i];
            }
        }
        std::fclose(fs);
        std::fclose(fc);
    }
    if (n_gnps > 0) {
        const std::string path = path_of(dir, fmt ? "gnp_data.dat" : "gnp_data.csv");
        FILE *f = std::fopen(path.c_str(), fmt ? "wb" : "w");
        if (!f) return e.fail("failed to create %s", path.c_str());
        if (fmt == 1) {
            const int32_t n = (int32_t)n_gnps;
            std::fwrite(&n, 4, 1, f);
        }
        for (int64_t i = 0; i < n_gnps; i++) {
            const double *m = rot + 9 * i;
            const double phi = recover_angle(m[4], -m[1]);
            const double theta = recover_angle(m[8], -m[6]);
            const double l = lt[2 * i], t = lt[2 * i + 1];
            const double *c = centre + 3 * i;
            if (fmt == 0) {
                std::fprintf(f, "%g, %g, %g, %g, %g, %.*f, %.*f, %.*f\n", l, l, t, theta, phi,
                             prec_gnp, c[0], prec_gnp, c[1], prec_gnp, c[2]);
            } else {
                const double v[7] = {l, t, theta, phi, c[0], c[1], c[2]};
                std::fwrite(v, 8, 7, f);
            }
        }
        std::fclose(f);
    }
    return 1;
}

}  // extern "C"
