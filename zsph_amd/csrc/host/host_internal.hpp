// Host-side helpers shared by the C ABI translation units (not exported).
#pragma once
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "zsph.h"

int zsph_check_dt(const zsph_cpu_param *p);

// Inverse EOS used to set a hydrostatic initial density:
// gamma == 1: rho = rho0 + p / c^2 (the reference's linear EOS, lib/time_int.cu:39-40);
// otherwise Tait: rho = rho0 (1 + p / B)^(1/gamma), B = rho0 c^2 / gamma.
inline double zsph_host_rho_of_p(double p, double rho0, double c2, double gamma) {
    if (gamma == 1.0) return rho0 + p / c2;
    const double B = rho0 * c2 / gamma;
    return rho0 * std::pow(1.0 + p / B, 1.0 / gamma);
}

// Top-level number / bool and string members of a JSON object (nested
// values are parsed and skipped).  False if the text is not a JSON object.
bool zsph_json_scalars(const std::string &text, std::map<std::string, double> *nums,
                       std::map<std::string, std::string> *strs);

// FNV-1a, 64-bit (offset basis 0xcbf29ce484222325, prime 0x100000001b3),
// continued from h.
uint64_t zsph_fnv1a64(const void *data, size_t bytes, uint64_t h = 0xcbf29ce484222325ull);

// Checkpoint files.  Layout: a JSON object of top-level scalars padded with
// spaces to ZSPH_CKPT_HEADER bytes (last byte '\n'), then the payload: the
// sections listed in "sections" (comma-separated), in that order, section s
// "<s>_bytes" long.  "checksum" is the FNV-1a 64 of the payload (16 hex
// digits); "params" the FNV-1a 64 of the parameter blocks the run was made
// with.  Written to <path>.tmp and renamed, so a partial file never carries
// the final name.
constexpr size_t ZSPH_CKPT_HEADER = 4096;
constexpr int ZSPH_CKPT_VERSION = 1;
struct CkptMeta {
    long long step = 0;
    double time = 0.0, t0 = 0.0;  // simulated time and the start of the current step
    int ptc_num = 0, cells = 0, scenes = 0, dim = 0, grid[3] = {0, 0, 0};
    int killed = 0, has_body = 0, has_piston = 0, closure = 0;
    uint64_t params = 0;
};
struct CkptSection {
    std::string name;
    size_t bytes = 0;
    const void *data = nullptr;  // writing only
};
// 0 or ZSPH_E_IO (*err says why).
int zsph_ckpt_write(const char *path, const CkptMeta &m, const std::vector<CkptSection> &secs, std::string *err);
// Reads the header (and with payload, the payload, checked against the
// checksum).  0, ZSPH_E_IO (unreadable, truncated, malformed) or ZSPH_E_ARG
// (checksum mismatch, unknown version).
int zsph_ckpt_read(const char *path, CkptMeta *m, std::vector<CkptSection> *secs, std::vector<char> *payload,
                   std::string *err);
