// Checkpoint files: the restart state the reference names but never writes.
// Its case file carries start_step / current_step
// (single_gpu_sph_src/io/read_json.cu:32-33) and its loop starts at
// start_step (deltasph/main.cu:58), but write_json
// (io/write_json.cu:3-12) is never called and reading its own VTK output
// back drops every field (io/read_vkt.cu:26-40 against
// io/write_vtk.cu:66-127).  Here a checkpoint is the complete state between
// steps, byte for byte, behind a JSON header; this file is the format only
// (host code), the device side gathers and scatters the arrays.
#include <cerrno>
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "host_internal.hpp"
#include "zsph.h"

uint64_t zsph_fnv1a64(const void *data, size_t bytes, uint64_t h) {
    const unsigned char *p = static_cast<const unsigned char *>(data);
    for (size_t k = 0; k < bytes; ++k) {
        h ^= p[k];
        h *= 0x100000001b3ull;
    }
    return h;
}

namespace {

std::string header_text(const CkptMeta &m, const std::vector<CkptSection> &secs, uint64_t checksum) {
    char buf[256];
    std::string s = "{\"format\": \"zsph-checkpoint\", \"version\": ";
    s += std::to_string(ZSPH_CKPT_VERSION);
    s += ", \"endian\": \"little\"";
    std::snprintf(buf, sizeof(buf), ", \"step\": %lld, \"time\": %.17g, \"t0\": %.17g", m.step, m.time, m.t0);
    s += buf;
    std::snprintf(buf, sizeof(buf), ", \"ptc_num\": %d, \"cells\": %d, \"scenes\": %d, \"dim\": %d", m.ptc_num,
                  m.cells, m.scenes, m.dim);
    s += buf;
    std::snprintf(buf, sizeof(buf), ", \"grid_x\": %d, \"grid_y\": %d, \"grid_z\": %d", m.grid[0], m.grid[1],
                  m.grid[2]);
    s += buf;
    std::snprintf(buf, sizeof(buf), ", \"killed\": %d, \"has_body\": %d, \"has_piston\": %d, \"closure\": %d",
                  m.killed, m.has_body, m.has_piston, m.closure);
    s += buf;
    std::snprintf(buf, sizeof(buf), ", \"params\": \"%016" PRIx64 "\"", m.params);
    s += buf;
    s += ", \"sections\": \"";
    size_t total = 0;
    for (size_t k = 0; k < secs.size(); ++k) {
        if (k) s += ",";
        s += secs[k].name;
    }
    s += "\"";
    for (const CkptSection &c : secs) {
        std::snprintf(buf, sizeof(buf), ", \"%s_bytes\": %zu", c.name.c_str(), c.bytes);
        s += buf;
        total += c.bytes;
    }
    std::snprintf(buf, sizeof(buf), ", \"payload_bytes\": %zu, \"checksum\": \"%016" PRIx64 "\"}", total, checksum);
    s += buf;
    return s;
}

int io_fail(std::string *err, const std::string &what) {
    if (err) *err = what;
    return ZSPH_E_IO;
}

bool parse_hex64(const std::string &s, uint64_t *v) {
    if (s.size() != 16) return false;
    uint64_t r = 0;
    for (char ch : s) {
        int d;
        if (ch >= '0' && ch <= '9') d = ch - '0';
        else if (ch >= 'a' && ch <= 'f') d = ch - 'a' + 10;
        else return false;
        r = (r << 4) | (uint64_t)d;
    }
    *v = r;
    return true;
}

}  // namespace

int zsph_ckpt_write(const char *path, const CkptMeta &m, const std::vector<CkptSection> &secs, std::string *err) {
    if (!path || !*path) return io_fail(err, "checkpoint: empty path");
    for (const CkptSection &c : secs)
        if (c.name.empty() || c.name.find_first_of(",\"\\ ") != std::string::npos || (c.bytes && !c.data))
            return io_fail(err, "checkpoint: bad section '" + c.name + "'");
    std::string head = header_text(m, secs, 0);  // same length as the final text
    if (head.size() + 1 > ZSPH_CKPT_HEADER) return io_fail(err, "checkpoint: header too long");
    const std::string tmp = std::string(path) + ".tmp";
    FILE *f = std::fopen(tmp.c_str(), "wb");
    if (!f) return io_fail(err, "checkpoint: cannot create " + tmp + ": " + std::strerror(errno));
    std::vector<char> pad(ZSPH_CKPT_HEADER, ' ');
    bool ok = std::fwrite(pad.data(), 1, pad.size(), f) == pad.size();
    uint64_t h = 0xcbf29ce484222325ull;
    for (size_t k = 0; ok && k < secs.size(); ++k) {
        h = zsph_fnv1a64(secs[k].data, secs[k].bytes, h);
        ok = std::fwrite(secs[k].data, 1, secs[k].bytes, f) == secs[k].bytes;
    }
    if (ok) {
        head = header_text(m, secs, h);
        std::memcpy(pad.data(), head.data(), head.size());
        pad.back() = '\n';
        ok = std::fseek(f, 0, SEEK_SET) == 0 && std::fwrite(pad.data(), 1, pad.size(), f) == pad.size();
    }
    ok = (std::fflush(f) == 0) && ok;
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) {
        std::remove(tmp.c_str());
        return io_fail(err, std::string("checkpoint: write to ") + tmp + " failed");
    }
    if (std::rename(tmp.c_str(), path) != 0) {
        std::remove(tmp.c_str());
        return io_fail(err, std::string("checkpoint: cannot rename to ") + path + ": " + std::strerror(errno));
    }
    return ZSPH_OK;
}

int zsph_ckpt_read(const char *path, CkptMeta *m, std::vector<CkptSection> *secs, std::vector<char> *payload,
                   std::string *err) {
    if (!path || !m || !secs) return io_fail(err, "checkpoint: bad argument");
    FILE *f = std::fopen(path, "rb");
    if (!f) return io_fail(err, std::string("checkpoint: cannot open ") + path + ": " + std::strerror(errno));
    std::string head(ZSPH_CKPT_HEADER, '\0');
    if (std::fread(&head[0], 1, head.size(), f) != head.size() || head.back() != '\n') {
        std::fclose(f);
        return io_fail(err, std::string(path) + ": not a checkpoint (short or malformed header)");
    }
    std::map<std::string, double> num;
    std::map<std::string, std::string> str;
    if (!zsph_json_scalars(head, &num, &str) || str["format"] != "zsph-checkpoint") {
        std::fclose(f);
        return io_fail(err, std::string(path) + ": not a checkpoint (header is not a zsph-checkpoint object)");
    }
    if (num["version"] != ZSPH_CKPT_VERSION || str["endian"] != "little") {
        std::fclose(f);
        return io_fail(err, std::string(path) + ": unsupported checkpoint version or byte order");
    }
    // every field must be present; integers must be finite, integral and in
    // range before they are converted (a damaged header is refused, never cast)
    bool ok = true;
    auto getd = [&](const char *k) {
        auto it = num.find(k);
        if (it == num.end() || !std::isfinite(it->second)) {
            ok = false;
            return 0.0;
        }
        return it->second;
    };
    auto geti = [&](const char *k, double lo, double hi) -> long long {
        const double v = getd(k);
        if (!(v >= lo && v <= hi && v == std::floor(v))) {
            ok = false;
            return 0;
        }
        return (long long)v;
    };
    constexpr double IMAX = 2147483647.0, BYTES_MAX = 1125899906842624.0;  // 2^31 - 1, 2^50
    m->step = geti("step", 0.0, 9007199254740992.0);
    m->time = getd("time");
    m->t0 = getd("t0");
    m->ptc_num = (int)geti("ptc_num", 0.0, IMAX);
    m->cells = (int)geti("cells", 0.0, IMAX);
    m->scenes = (int)geti("scenes", 1.0, IMAX);
    m->dim = (int)geti("dim", 2.0, 3.0);
    m->grid[0] = (int)geti("grid_x", 0.0, IMAX);
    m->grid[1] = (int)geti("grid_y", 0.0, IMAX);
    m->grid[2] = (int)geti("grid_z", 0.0, IMAX);
    m->killed = (int)geti("killed", 0.0, IMAX);
    m->has_body = (int)geti("has_body", 0.0, 1.0);
    m->has_piston = (int)geti("has_piston", 0.0, 1.0);
    m->closure = (int)geti("closure", 0.0, 1.0);
    const size_t total = (size_t)geti("payload_bytes", 0.0, BYTES_MAX);
    uint64_t checksum = 0;
    ok = ok && parse_hex64(str["params"], &m->params) && parse_hex64(str["checksum"], &checksum);
    secs->clear();
    size_t sum = 0;
    const std::string list = str["sections"];
    for (size_t a = 0; ok && a < list.size() && secs->size() < 64;) {
        size_t b = list.find(',', a);
        if (b == std::string::npos) b = list.size();
        CkptSection c;
        c.name = list.substr(a, b - a);
        ok = !c.name.empty();
        c.bytes = (size_t)geti((c.name + "_bytes").c_str(), 0.0, BYTES_MAX);
        sum += c.bytes;  // at most 64 sections of < 2^50 bytes: no overflow
        secs->push_back(c);
        a = b + 1;
    }
    if (!ok || list.empty() || sum != total) {
        std::fclose(f);
        return io_fail(err, std::string(path) + ": incomplete checkpoint header");
    }
    if (payload) {
        // the payload must be there before memory is set aside for it
        const long here = std::ftell(f);
        if (std::fseek(f, 0, SEEK_END) != 0 || std::ftell(f) - here != (long)total ||
            std::fseek(f, here, SEEK_SET) != 0) {
            std::fclose(f);
            return io_fail(err, std::string(path) + ": payload size differs from its header");
        }
        payload->resize(total);
        const bool got = std::fread(payload->data(), 1, payload->size(), f) == payload->size();
        const bool at_end = got && std::fgetc(f) == EOF;
        std::fclose(f);
        if (!got || !at_end) return io_fail(err, std::string(path) + ": payload size differs from its header");
        if (zsph_fnv1a64(payload->data(), payload->size()) != checksum)
            return io_fail(err, std::string(path) + ": checksum mismatch (corrupt checkpoint)");
    } else {
        std::fclose(f);
    }
    return ZSPH_OK;
}

extern "C" int zsph_checkpoint_inspect(const char *path, int verify, zsph_checkpoint_info *info) {
    if (!path || !info) return ZSPH_E_ARG;
    CkptMeta m;
    std::vector<CkptSection> secs;
    std::vector<char> payload;
    const int r = zsph_ckpt_read(path, &m, &secs, verify ? &payload : nullptr, nullptr);
    if (r) return r;
    info->step = m.step;
    info->time = m.time;
    info->ptc_num = m.ptc_num;
    info->cells = m.cells;
    info->scenes = m.scenes;
    info->dim = m.dim;
    for (int k = 0; k < 3; ++k) info->grid[k] = m.grid[k];
    info->killed = m.killed;
    info->has_body = m.has_body;
    info->has_piston = m.has_piston;
    info->closure = m.closure;
    size_t total = 0;
    for (const CkptSection &c : secs) total += c.bytes;
    info->payload_bytes = (long long)total;
    return ZSPH_OK;
}
