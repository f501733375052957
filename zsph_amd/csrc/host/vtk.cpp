// Legacy-VTK particle files.  Writer: the reference's write_vtk
// (single_gpu_sph_src/io/write_vtk.cu:29-208: "zsph data", UNSTRUCTURED_GRID,
// POINTS float, POINT_DATA with SCALARS/VECTORS, big-endian BINARY or ASCII),
// emitting every field requested in the mask.  Reader: replaces the libVTK
// reader of io/read_vkt.cu:3-92 (fields p, rho, vel, acc, type, table;
// missing p/rho/vel/acc default to 0 / rho0 / 0 / 0) and accepts float or
// double arrays, so files written here restart without losing fields.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "zsph.h"

namespace {

inline uint32_t bswap32(uint32_t v) {
    return (v >> 24) | ((v >> 8) & 0xff00u) | ((v << 8) & 0xff0000u) | (v << 24);
}
inline uint64_t bswap64(uint64_t v) {
    return ((uint64_t)bswap32((uint32_t)v) << 32) | bswap32((uint32_t)(v >> 32));
}

struct Out {
    std::ofstream f;
    bool bin;
    void floats(const float *a, int n, int stride, int ncomp) {
        if (bin) {
            std::vector<uint32_t> buf((size_t)n * ncomp);
            for (int i = 0; i < n; ++i)
                for (int c = 0; c < ncomp; ++c) {
                    uint32_t u;
                    std::memcpy(&u, &a[(size_t)i * stride + c], 4);
                    buf[(size_t)i * ncomp + c] = bswap32(u);
                }
            f.write((const char *)buf.data(), (std::streamsize)(buf.size() * 4));
            f << "\n";
        } else {
            char tmp[64];
            for (int i = 0; i < n; ++i) {
                for (int c = 0; c < ncomp; ++c) {
                    std::snprintf(tmp, sizeof(tmp), c ? " %.9g" : "%.9g", a[(size_t)i * stride + c]);
                    f << tmp;
                }
                f << "\n";
            }
        }
    }
    void ints(const int *a, int n) {
        if (bin) {
            std::vector<uint32_t> buf(n);
            for (int i = 0; i < n; ++i) buf[i] = bswap32((uint32_t)a[i]);
            f.write((const char *)buf.data(), (std::streamsize)(buf.size() * 4));
            f << "\n";
        } else {
            for (int i = 0; i < n; ++i) f << a[i] << "\n";
        }
    }
};

struct In {
    std::string s;
    size_t i = 0;
    void ws() {
        while (i < s.size() && (s[i] == ' ' || s[i] == '\n' || s[i] == '\r' || s[i] == '\t')) ++i;
    }
    std::string tok() {
        ws();
        size_t b = i;
        while (i < s.size() && !(s[i] == ' ' || s[i] == '\n' || s[i] == '\r' || s[i] == '\t')) ++i;
        return s.substr(b, i - b);
    }
    std::string line() {
        size_t b = i;
        while (i < s.size() && s[i] != '\n') ++i;
        std::string r = s.substr(b, i - b);
        if (i < s.size()) ++i;
        if (!r.empty() && r.back() == '\r') r.pop_back();
        return r;
    }
    void rest_of_line() {
        while (i < s.size() && s[i] != '\n') ++i;
        if (i < s.size()) ++i;
    }
    // Read `count` values of VTK type `ty` as doubles.  The count is checked
    // against what is left of the file before any memory is set aside.
    bool values(bool bin, const std::string &ty, size_t count, std::vector<double> *out) {
        if (bin) {
            size_t w = (ty == "double" || ty == "long" || ty == "unsigned_long") ? 8 : 4;
            if (ty == "short" || ty == "unsigned_short") w = 2;
            if (ty == "char" || ty == "unsigned_char" || ty == "bit") return false;
            if (count > (s.size() - i) / w) return false;
            out->resize(count);
            for (size_t k = 0; k < count; ++k) {
                const char *p = s.data() + i + w * k;
                double v;
                if (w == 8) {
                    uint64_t u; std::memcpy(&u, p, 8); u = bswap64(u);
                    if (ty == "double") { double d; std::memcpy(&d, &u, 8); v = d; }
                    else if (ty == "long") v = (double)(int64_t)u;
                    else v = (double)u;
                } else if (w == 4) {
                    uint32_t u; std::memcpy(&u, p, 4); u = bswap32(u);
                    if (ty == "float") { float fl; std::memcpy(&fl, &u, 4); v = fl; }
                    else if (ty == "int") v = (double)(int32_t)u;
                    else v = (double)u;
                } else {
                    uint16_t u; std::memcpy(&u, p, 2); u = (uint16_t)((u >> 8) | (u << 8));
                    v = (ty == "short") ? (double)(int16_t)u : (double)u;
                }
                (*out)[k] = v;
            }
            i += w * count;
        } else {
            if (count > (s.size() - i + 1) / 2) return false;  // each value takes a digit and a separator
            out->resize(count);
            for (size_t k = 0; k < count; ++k) {
                std::string t = tok();
                if (t.empty()) return false;
                char *e = nullptr;
                (*out)[k] = std::strtod(t.c_str(), &e);
                if (e == t.c_str()) return false;
            }
        }
        return true;
    }
};

// A count token: a non-negative int, else -1.
int count_of(const std::string &t) {
    char *e = nullptr;
    const long v = std::strtol(t.c_str(), &e, 10);
    return (e == t.c_str() || *e || v < 0 || v > 2147483647L) ? -1 : (int)v;
}
// Conversions of file values that are undefined for out-of-range numbers
// (float: finite values beyond FLT_MAX; int: anything outside int).
bool as_float(double v, float *out) {
    if (std::isfinite(v) && std::fabs(v) > 3.4028234663852886e38) return false;
    *out = (float)v;
    return true;
}
bool as_int(double v, int *out) {
    if (!(v >= -2147483648.0 && v <= 2147483647.0)) return false;
    *out = (int)v;
    return true;
}

}  // namespace

extern "C" int zsph_write_vtk(const char *path, const zsph_host_state *st, int n, int fields,
                              int binary) {
    if (!path || !st || n < 0 || !st->pos) return ZSPH_E_ARG;
    Out o;
    o.bin = binary != 0;
    o.f.open(path, std::ios::out | std::ios::binary);
    if (!o.f) return ZSPH_E_IO;
    o.f << "# vtk DataFile Version 3.0\n";
    o.f << "zsph data\n";
    o.f << (o.bin ? "BINARY\n" : "ASCII\n");
    o.f << "DATASET UNSTRUCTURED_GRID\n";
    o.f << "POINTS " << n << " float\n";
    o.floats(st->pos, n, 3, 3);
    o.f << "POINT_DATA " << n << "\n";
    if ((fields & ZSPH_F_RHOP) && st->rhop) {
        o.f << "SCALARS rho float 1\nLOOKUP_TABLE default\n";
        o.floats(st->rhop, n, 2, 1);
        o.f << "SCALARS p float 1\nLOOKUP_TABLE default\n";
        o.floats(st->rhop + 1, n, 2, 1);
    }
    if ((fields & ZSPH_F_TYPE) && st->type) {
        o.f << "SCALARS type int 1\nLOOKUP_TABLE default\n";
        o.ints(st->type, n);
    }
    if ((fields & ZSPH_F_HASH) && st->hash) {
        o.f << "SCALARS hash int 1\nLOOKUP_TABLE default\n";
        o.ints(st->hash, n);
    }
    if ((fields & ZSPH_F_TABLE) && st->table) {
        o.f << "SCALARS table int 1\nLOOKUP_TABLE default\n";
        o.ints(st->table, n);
    }
    if ((fields & ZSPH_F_VEL) && st->vel) {
        o.f << "VECTORS vel float\n";
        o.floats(st->vel, n, 3, 3);
    }
    if ((fields & ZSPH_F_ACC) && st->acc_drhodt) {
        o.f << "VECTORS acc float\n";
        o.floats(st->acc_drhodt, n, 4, 3);
        o.f << "SCALARS drhodt float 1\nLOOKUP_TABLE default\n";
        o.floats(st->acc_drhodt + 3, n, 4, 1);
    }
    o.f.close();
    return o.f ? ZSPH_OK : ZSPH_E_IO;
}

namespace {

int read_vtk_impl(const char *path, zsph_host_state *st, int n, float rho0, int *n_out) {
    std::ifstream f(path, std::ios::in | std::ios::binary);
    if (!f) return ZSPH_E_IO;
    In in;
    {
        std::stringstream ss;
        ss << f.rdbuf();
        in.s = ss.str();
    }
    std::string l0 = in.line();
    if (l0.find("vtk") == std::string::npos) return ZSPH_E_IO;
    in.line();  // title
    std::string fmt = in.tok();
    const bool bin = (fmt == "BINARY");
    if (!bin && fmt != "ASCII") return ZSPH_E_IO;
    int npts = -1;
    bool have_p = false, have_rho = false, have_vel = false, have_acc = false, have_type = false,
         have_table = false;
    std::vector<double> buf;
    while (true) {
        std::string kw = in.tok();
        if (kw.empty()) break;
        if (kw == "DATASET") {
            in.tok();
        } else if (kw == "POINTS") {
            npts = count_of(in.tok());
            std::string ty = in.tok();
            in.rest_of_line();
            if (npts < 0) return ZSPH_E_IO;
            if (n_out) { *n_out = npts; return ZSPH_OK; }
            if (npts != n) return ZSPH_E_ARG;  // read_vkt.cu:14 assert
            if (!in.values(bin, ty, (size_t)npts * 3, &buf)) return ZSPH_E_IO;
            if (st->pos)
                for (size_t k = 0; k < buf.size(); ++k)
                    if (!as_float(buf[k], &st->pos[k])) return ZSPH_E_IO;
        } else if (kw == "POINT_DATA" || kw == "CELL_DATA") {
            in.tok();
        } else if (kw == "CELLS") {
            const int nc = count_of(in.tok());
            const int sz = count_of(in.tok());
            in.rest_of_line();
            if (nc < 0 || sz < 0 || !in.values(bin, "int", (size_t)sz, &buf)) return ZSPH_E_IO;
        } else if (kw == "CELL_TYPES") {
            const int nc = count_of(in.tok());
            in.rest_of_line();
            if (nc < 0 || !in.values(bin, "int", (size_t)nc, &buf)) return ZSPH_E_IO;
        } else if (kw == "SCALARS" || kw == "VECTORS") {
            // "SCALARS name type [ncomp]" / "VECTORS name type" on one line
            std::istringstream hdr(in.line());
            std::string name, ty;
            int ncomp = (kw == "VECTORS") ? 3 : 1;
            hdr >> name >> ty;
            if (kw == "SCALARS") {
                int nc = 0;
                if (hdr >> nc) ncomp = nc;
                if (in.tok() != "LOOKUP_TABLE") return ZSPH_E_IO;
                in.rest_of_line();
            }
            if (name.empty() || ty.empty() || ncomp < 1) return ZSPH_E_IO;
            if (npts < 0) return ZSPH_E_IO;
            if (!in.values(bin, ty, (size_t)npts * ncomp, &buf)) return ZSPH_E_IO;
            if (name == "p" && st->rhop) {
                for (int k = 0; k < npts; ++k)
                    if (!as_float(buf[(size_t)k * ncomp], &st->rhop[2 * k + 1])) return ZSPH_E_IO;
                have_p = true;
            } else if (name == "rho" && st->rhop) {
                for (int k = 0; k < npts; ++k)
                    if (!as_float(buf[(size_t)k * ncomp], &st->rhop[2 * k])) return ZSPH_E_IO;
                have_rho = true;
            } else if (name == "vel" && st->vel && ncomp == 3) {
                for (size_t k = 0; k < buf.size(); ++k)
                    if (!as_float(buf[k], &st->vel[k])) return ZSPH_E_IO;
                have_vel = true;
            } else if (name == "acc" && st->acc_drhodt && ncomp == 3) {
                for (int k = 0; k < npts; ++k)
                    for (int c = 0; c < 3; ++c)
                        if (!as_float(buf[3 * (size_t)k + c], &st->acc_drhodt[4 * k + c])) return ZSPH_E_IO;
                have_acc = true;
            } else if (name == "type" && st->type) {
                for (int k = 0; k < npts; ++k)
                    if (!as_int(buf[(size_t)k * ncomp], &st->type[k])) return ZSPH_E_IO;
                have_type = true;
            } else if (name == "table" && st->table) {
                for (int k = 0; k < npts; ++k)
                    if (!as_int(buf[(size_t)k * ncomp], &st->table[k])) return ZSPH_E_IO;
                have_table = true;
            }
        } else if (kw == "FIELD") {
            return ZSPH_E_IO;  // FIELD blocks are not written by either writer
        } else {
            return ZSPH_E_IO;
        }
    }
    if (npts < 0) return ZSPH_E_IO;
    if (!have_type && st->type) return ZSPH_E_IO;  // read_vkt.cu:74: type is mandatory
    for (int k = 0; k < npts; ++k) {
        if (st->rhop) {
            if (!have_rho) st->rhop[2 * k] = rho0;
            if (!have_p) st->rhop[2 * k + 1] = 0.0f;
        }
        if (st->vel && !have_vel) st->vel[3 * k] = st->vel[3 * k + 1] = st->vel[3 * k + 2] = 0.0f;
        if (st->acc_drhodt) {
            if (!have_acc) st->acc_drhodt[4 * k] = st->acc_drhodt[4 * k + 1] = st->acc_drhodt[4 * k + 2] = 0.0f;
            st->acc_drhodt[4 * k + 3] = 0.0f;
        }
        if (st->table && !have_table) st->table[k] = 0;
    }
    return ZSPH_OK;
}

}  // namespace

extern "C" int zsph_read_vtk(const char *path, zsph_host_state *st, int n, float rho0) {
    if (!path || !st || n < 0) return ZSPH_E_ARG;
    return read_vtk_impl(path, st, n, rho0, nullptr);
}

extern "C" int zsph_vtk_count(const char *path, int *n_out) {
    if (!path || !n_out) return ZSPH_E_ARG;
    zsph_host_state st{};
    return read_vtk_impl(path, &st, 0, 0.0f, n_out);
}
