// Host-side checks of libzsph's host code, built with AddressSanitizer and
// UndefinedBehaviorSanitizer (`make -C zsph_amd/csrc host-asan`) and run by
// tests/test_sanitizers.py: case generators, parameter set-up, the VTK and
// JSON readers and writers, the rigid-body routines, and the checkpoint
// reader, including truncated and corrupted files and damaged case files.
// Any sanitizer report aborts the run; exit status 0 means every check held.
//   host_checks <scratch directory>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "host_internal.hpp"
#include "zsph.h"

namespace {

int fails = 0;
#define CHECK(c)                                                                     \
    do {                                                                             \
        if (!(c)) {                                                                  \
            std::fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #c); \
            ++fails;                                                                 \
        }                                                                            \
    } while (0)

struct Host {
    std::vector<float> pos, vel, rhop, acc;
    std::vector<int> type, table, hash;
    explicit Host(int n)
        : pos(3 * (size_t)n), vel(3 * (size_t)n), rhop(2 * (size_t)n), acc(4 * (size_t)n), type(n), table(n),
          hash(n) {}
    zsph_host_state st() {
        return {pos.data(), vel.data(), rhop.data(), acc.data(), type.data(), table.data(), hash.data()};
    }
};

std::string slurp(const std::string &path) {
    std::ifstream f(path, std::ios::binary);
    std::stringstream ss;
    ss << f.rdbuf();
    return ss.str();
}
void spit(const std::string &path, const std::string &s) {
    std::ofstream f(path, std::ios::binary);
    f << s;
}

// A small dam break of dimension dim: generated twice (count, then fill),
// parameters derived, then written to VTK and read back.
void dambreak_and_vtk(int dim, const std::string &dir) {
    zsph_dambreak c{};
    c.dim = dim;
    c.dx = dim == 2 ? 0.02f : 0.05f;
    c.lx = 1.0f;
    c.wy = dim == 2 ? 0.0f : 0.3f;
    c.lz = 0.6f;
    c.cw = 0.3f;
    c.ch = 0.3f;
    c.layers = 3;
    zsph_json j;
    zsph_json_defaults(&j);
    j.cfl = 0.2f;
    j.gamma = 7.0f;
    int n = 0;
    CHECK(zsph_case_dambreak(&c, &j, nullptr, &n) == ZSPH_OK && n > 0);
    Host h(n);
    zsph_host_state st = h.st();
    CHECK(zsph_case_dambreak(&c, &j, &st, &n) == ZSPH_OK);
    zsph_cpu_param cp;
    CHECK(zsph_set_cpu_param(&cp, &j) == ZSPH_OK);
    zsph_params gp;
    zsph_set_gpu_param(&gp, &cp);
    CHECK(gp.ptc_num == n && gp.grid_xdim > 0);
    for (int binary = 0; binary < 2; ++binary) {
        const std::string path = dir + "/case" + std::to_string(dim) + (binary ? "b" : "a") + ".vtk";
        CHECK(zsph_write_vtk(path.c_str(), &st, n, ZSPH_F_ALL, binary) == ZSPH_OK);
        int m = -1;
        CHECK(zsph_vtk_count(path.c_str(), &m) == ZSPH_OK && m == n);
        Host g(n);
        zsph_host_state gs = g.st();
        CHECK(zsph_read_vtk(path.c_str(), &gs, n, j.rho0) == ZSPH_OK);
        CHECK(g.type == h.type);
        if (binary) CHECK(g.pos == h.pos);  // binary float round trip is exact
        // damaged copies: every prefix length on a coarse grid and single-byte
        // changes; the reader must return (any code) without a sanitizer report
        const std::string full = slurp(path);
        const std::string bad = dir + "/bad.vtk";
        for (size_t len = 0; len < full.size(); len += 1 + full.size() / 61) {
            spit(bad, full.substr(0, len));
            (void)zsph_read_vtk(bad.c_str(), &gs, n, j.rho0);
            (void)zsph_vtk_count(bad.c_str(), &m);
        }
        for (size_t at = 0; at < full.size(); at += 1 + full.size() / 97) {
            std::string d = full;
            d[at] = (char)(d[at] ^ 0x5a);
            spit(bad, d);
            (void)zsph_read_vtk(bad.c_str(), &gs, n, j.rho0);
        }
    }
    // hostile counts in an otherwise valid header
    const std::string bad = dir + "/bad.vtk";
    Host g(n);
    zsph_host_state gs = g.st();
    spit(bad, "# vtk DataFile Version 3.0\nx\nBINARY\nDATASET UNSTRUCTURED_GRID\nPOINTS 2147483647 float\n");
    CHECK(zsph_read_vtk(bad.c_str(), &gs, n, j.rho0) != ZSPH_OK);
    spit(bad, "# vtk DataFile Version 3.0\nx\nASCII\nDATASET UNSTRUCTURED_GRID\nCELLS 1 -5\n");
    CHECK(zsph_read_vtk(bad.c_str(), &gs, n, j.rho0) != ZSPH_OK);
    spit(bad, "# vtk DataFile Version 3.0\nx\nASCII\nDATASET UNSTRUCTURED_GRID\nPOINTS 1 double\n1e300 0 0\n");
    CHECK(zsph_read_vtk(bad.c_str(), &gs, 1, j.rho0) != ZSPH_OK);
}

// A case file with every key, read back; then damaged and hostile ones.
void json_files(const std::string &dir) {
    const std::string path = dir + "/case.json";
    spit(path,
         "{\"g\": -9.81, \"rho0\": 1000, \"dx\": 0.01, \"h_factor\": 1.3, \"r_factor\": 2, \"eta_factor\": 0.1,\n"
         " \"cs_factor\": 10, \"delta\": 0.1, \"alpha\": 0.02, \"xmin\": 0, \"xmax\": 1, \"ymin\": 0, \"ymax\": 0,\n"
         " \"zmin\": 0, \"zmax\": 0.6, \"grid_layer_factor\": 3, \"grid_size_factor\": 1, \"dt\": 0.0001,\n"
         " \"start_step\": 0, \"current_step\": 0, \"end_step\": 10, \"output_step\": 5, \"ptc_num\": 100,\n"
         " \"water_ptc_num\": 80, \"air_ptc_num\": 0, \"wall_ptc_num\": 20, \"rigid_ptc_num\": 0,\n"
         " \"thread_num\": 2, \"gpu_id\": 0, \"input_path\": \"in\", \"output_path\": \"out\", \"git_hash\": \"x\",\n"
         " \"dim\": 2, \"gamma\": 7, \"diffusion\": \"MC\", \"nested\": {\"a\": [1, 2, {\"b\": null}]}}\n");
    zsph_json j;
    CHECK(zsph_read_json(path.c_str(), &j) == ZSPH_OK);
    CHECK(j.ptc_num == 100 && j.dim == 2 && j.diffusion == ZSPH_DIFF_MC && std::string(j.output_path) == "out");
    const std::string bad = dir + "/bad.json";
    spit(bad, "{\"ptc_num\": 1e20}");
    CHECK(zsph_read_json(bad.c_str(), &j) == ZSPH_E_IO);
    spit(bad, "{\"dx\": 1e300}");
    CHECK(zsph_read_json(bad.c_str(), &j) == ZSPH_E_IO);
    spit(bad, "{\"a\": " + std::string(100000, '[') + std::string(100000, ']') + "}");
    CHECK(zsph_read_json(bad.c_str(), &j) == ZSPH_E_IO);
    const std::string full = slurp(path);
    for (size_t len = 0; len < full.size(); len += 7) {
        spit(bad, full.substr(0, len));
        (void)zsph_read_json(bad.c_str(), &j);
    }
}

void wavetank_and_rigid() {
    zsph_wavetank w{};
    w.dx = 0.04f;
    w.lx = 1.6f;
    w.wy = 0.4f;
    w.lz = 0.6f;
    w.depth = 0.36f;
    const float lo[3] = {0.64f, 0.12f, 0.24f}, hi[3] = {0.88f, 0.28f, 0.44f};
    for (int k = 0; k < 3; ++k) {
        w.box_lo[k] = lo[k];
        w.box_hi[k] = hi[k];
    }
    w.layers = 3;
    w.piston_travel = 0.04f;
    zsph_json j;
    zsph_json_defaults(&j);
    j.cfl = 0.2f;
    int n = 0;
    CHECK(zsph_case_wavetank(&w, &j, nullptr, &n) == ZSPH_OK && n > 0);
    Host h(n);
    zsph_host_state st = h.st();
    CHECK(zsph_case_wavetank(&w, &j, &st, &n) == ZSPH_OK);
    std::vector<float> body;
    for (int i = 0; i < n; ++i)
        if (h.type[i] == ZSPH_BODY)
            for (int c = 0; c < 3; ++c) body.push_back(h.pos[3 * (size_t)i + c]);
    const int nb = (int)body.size() / 3;
    CHECK(nb > 0);
    zsph_rigid r{};
    CHECK(zsph_rigid_mass_properties(body.data(), nb, 5.0, &r) == ZSPH_OK && r.mass == 5.0);
    r.F[2] = -49.05;
    zsph_rigid_acc_alpha(&r);
    for (int s = 0; s < 10; ++s) zsph_rigid_euler_step(&r, body.data(), nb, 1e-3);
    CHECK(std::isfinite(r.c[2]) && r.vel[2] < 0.0);
}

// Checkpoint files through the library's own writer, then truncated and
// damaged copies through the reader: a payload change must be refused,
// every other damage must end in a return code.
void checkpoints(const std::string &dir) {
    std::vector<float> a(4 * 37), b(4 * 37);
    for (size_t k = 0; k < a.size(); ++k) {
        a[k] = 0.5f * (float)k;
        b[k] = -1.0f * (float)k;
    }
    std::vector<int> t(37, 1);
    CkptMeta m;
    m.step = 12;
    m.time = 0.25;
    m.ptc_num = 37;
    m.cells = 9;
    m.scenes = 1;
    m.dim = 3;
    m.grid[0] = m.grid[1] = m.grid[2] = 3;
    std::vector<CkptSection> secs(3);
    secs[0].name = "posr"; secs[0].bytes = a.size() * 4; secs[0].data = a.data();
    secs[1].name = "velp"; secs[1].bytes = b.size() * 4; secs[1].data = b.data();
    secs[2].name = "type"; secs[2].bytes = t.size() * 4; secs[2].data = t.data();
    const std::string path = dir + "/c.bin";
    std::string err;
    CHECK(zsph_ckpt_write(path.c_str(), m, secs, &err) == ZSPH_OK);
    zsph_checkpoint_info info;
    CHECK(zsph_checkpoint_inspect(path.c_str(), 1, &info) == ZSPH_OK && info.step == 12 && info.ptc_num == 37);
    CkptMeta r;
    std::vector<CkptSection> rs;
    std::vector<char> payload;
    CHECK(zsph_ckpt_read(path.c_str(), &r, &rs, &payload, &err) == ZSPH_OK);
    CHECK(rs.size() == 3 && payload.size() == (a.size() + b.size() + t.size()) * 4);
    CHECK(std::memcmp(payload.data(), a.data(), a.size() * 4) == 0);
    const std::string full = slurp(path), bad = dir + "/bad.bin";
    for (size_t len = 0; len < full.size(); len += 53) {
        spit(bad, full.substr(0, len));
        CHECK(zsph_checkpoint_inspect(bad.c_str(), 1, &info) != ZSPH_OK);
    }
    for (size_t at = 0; at < full.size(); at += 11) {
        std::string d = full;
        d[at] = (char)(d[at] ^ 0x21);
        spit(bad, d);
        const int rc = zsph_checkpoint_inspect(bad.c_str(), 1, &info);
        if (at >= ZSPH_CKPT_HEADER) CHECK(rc == ZSPH_E_IO);  // the checksum catches payload changes
    }
    // hostile header values
    const std::string head = full.substr(0, ZSPH_CKPT_HEADER);
    for (const char *key : {"\"payload_bytes\": ", "\"ptc_num\": ", "\"step\": ", "\"posr_bytes\": "}) {
        for (const char *val : {"1e300", "-7", "nan", "3.5", "18446744073709551616"}) {
            std::string h = head;
            const size_t at = h.find(key);
            CHECK(at != std::string::npos);
            const size_t v0 = at + std::strlen(key), v1 = h.find_first_of(",}", v0);
            std::string nh = h.substr(0, v0) + val + h.substr(v1);
            nh = nh.substr(0, ZSPH_CKPT_HEADER - 1);
            nh.resize(ZSPH_CKPT_HEADER - 1, ' ');
            nh += '\n';
            spit(bad, nh + full.substr(ZSPH_CKPT_HEADER));
            CHECK(zsph_checkpoint_inspect(bad.c_str(), 1, &info) != ZSPH_OK);
        }
    }
}

}  // namespace

int main(int argc, char **argv) {
    const std::string dir = argc > 1 ? argv[1] : "/tmp";
    dambreak_and_vtk(2, dir);
    dambreak_and_vtk(3, dir);
    json_files(dir);
    wavetank_and_rigid();
    checkpoints(dir);
    std::printf("host checks: %d failed\n", fails);
    return fails ? 1 : 0;
}
