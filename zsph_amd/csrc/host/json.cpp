// Case-file reader: the key set of the reference's read_json
// (single_gpu_sph_src/io/read_json.cu:5-51, which uses nlohmann/json) plus
// the optional extension keys of zsph_json.  A small recursive-descent JSON
// parser: top-level object members with number / string / bool values are
// kept, nested values are parsed and skipped.
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>

#include "zsph.h"
#include "host_internal.hpp"

namespace {

struct Val {
    enum Kind { NUM, STR, BOOL, OTHER } kind = OTHER;
    double num = 0.0;
    std::string str;
};

struct Parser {
    static constexpr int MAX_DEPTH = 64;  // nesting limit: the recursion stays bounded on any input
    const std::string &s;
    size_t i = 0;
    int depth = 0;
    bool ok = true;
    explicit Parser(const std::string &src) : s(src) {}

    void ws() {
        while (i < s.size() && std::isspace((unsigned char)s[i])) ++i;
    }
    bool eat(char c) {
        ws();
        if (i < s.size() && s[i] == c) { ++i; return true; }
        return false;
    }
    bool string(std::string *out) {
        ws();
        if (i >= s.size() || s[i] != '"') return ok = false;
        ++i;
        std::string r;
        while (i < s.size() && s[i] != '"') {
            char c = s[i++];
            if (c == '\\' && i < s.size()) {
                char e = s[i++];
                switch (e) {
                case 'n': r += '\n'; break;
                case 't': r += '\t'; break;
                case 'r': r += '\r'; break;
                case 'b': r += '\b'; break;
                case 'f': r += '\f'; break;
                case 'u': i = (i + 4 <= s.size()) ? i + 4 : s.size(); r += '?'; break;
                default: r += e; break;
                }
            } else {
                r += c;
            }
        }
        if (i >= s.size()) return ok = false;
        ++i;
        if (out) *out = r;
        return true;
    }
    bool value(Val *v) {
        ws();
        if (i >= s.size()) return ok = false;
        char c = s[i];
        if ((c == '{' || c == '[') && depth >= MAX_DEPTH) return ok = false;
        if (c == '"') {
            v->kind = Val::STR;
            return string(&v->str);
        }
        if (c == '{') {
            ++i;
            v->kind = Val::OTHER;
            if (eat('}')) return true;
            ++depth;
            do {
                if (!string(nullptr)) return false;
                if (!eat(':')) return ok = false;
                Val sub;
                if (!value(&sub)) return false;
            } while (eat(','));
            --depth;
            return eat('}') ? true : (ok = false);
        }
        if (c == '[') {
            ++i;
            v->kind = Val::OTHER;
            if (eat(']')) return true;
            ++depth;
            do {
                Val sub;
                if (!value(&sub)) return false;
            } while (eat(','));
            --depth;
            return eat(']') ? true : (ok = false);
        }
        if (s.compare(i, 4, "true") == 0) { i += 4; v->kind = Val::BOOL; v->num = 1; return true; }
        if (s.compare(i, 5, "false") == 0) { i += 5; v->kind = Val::BOOL; v->num = 0; return true; }
        if (s.compare(i, 4, "null") == 0) { i += 4; v->kind = Val::OTHER; return true; }
        const char *b = s.c_str() + i;
        char *e = nullptr;
        double d = std::strtod(b, &e);
        if (e == b) return ok = false;
        i += (size_t)(e - b);
        v->kind = Val::NUM;
        v->num = d;
        return true;
    }
    bool object(std::map<std::string, Val> *out) {
        if (!eat('{')) return ok = false;
        if (eat('}')) return true;
        do {
            std::string k;
            if (!string(&k)) return false;
            if (!eat(':')) return ok = false;
            Val v;
            if (!value(&v)) return false;
            (*out)[k] = v;
        } while (eat(','));
        return eat('}') ? true : (ok = false);
    }
};

}  // namespace

bool zsph_json_scalars(const std::string &text, std::map<std::string, double> *nums,
                       std::map<std::string, std::string> *strs) {
    std::map<std::string, Val> kv;
    Parser ps(text);
    if (!ps.object(&kv)) return false;
    for (const auto &e : kv) {
        if (e.second.kind == Val::NUM || e.second.kind == Val::BOOL) (*nums)[e.first] = e.second.num;
        else if (e.second.kind == Val::STR) (*strs)[e.first] = e.second.str;
    }
    return true;
}

extern "C" int zsph_read_json(const char *path, zsph_json *j) {
    if (!path || !j) return ZSPH_E_ARG;
    std::ifstream f(path);
    if (!f) return ZSPH_E_IO;
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string src = ss.str();
    std::map<std::string, Val> kv;
    Parser ps(src);
    if (!ps.object(&kv)) return ZSPH_E_IO;

    zsph_json_defaults(j);
    // numbers outside the field's type are refused (their conversion is undefined)
    bool bad = false;
    auto num = [&](const char *k, float *dst) {
        auto it = kv.find(k);
        if (it != kv.end() && (it->second.kind == Val::NUM || it->second.kind == Val::BOOL)) {
            const double v = it->second.num;
            if (std::isfinite(v) && std::fabs(v) > 3.4028234663852886e38) bad = true;
            else *dst = (float)v;
        }
    };
    auto inum = [&](const char *k, int *dst) {
        auto it = kv.find(k);
        if (it != kv.end() && (it->second.kind == Val::NUM || it->second.kind == Val::BOOL)) {
            const double v = it->second.num;
            if (!(v >= -2147483648.0 && v <= 2147483647.0)) bad = true;
            else *dst = (int)v;
        }
    };
    auto str = [&](const char *k, char *dst, size_t cap) {
        auto it = kv.find(k);
        if (it != kv.end() && it->second.kind == Val::STR) {
            std::strncpy(dst, it->second.str.c_str(), cap - 1);
            dst[cap - 1] = 0;
        }
    };
    // io/read_json.cu:12-49
    num("g", &j->g);
    num("rho0", &j->rho0);
    num("dx", &j->dx);
    num("h_factor", &j->h_factor);
    num("r_factor", &j->r_factor);
    num("eta_factor", &j->eta_factor);
    num("cs_factor", &j->cs_factor);
    num("delta", &j->delta);
    num("alpha", &j->alpha);
    num("xmin", &j->xmin); num("xmax", &j->xmax);
    num("ymin", &j->ymin); num("ymax", &j->ymax);
    num("zmin", &j->zmin); num("zmax", &j->zmax);
    num("grid_layer_factor", &j->grid_layer_factor);
    num("grid_size_factor", &j->grid_size_factor);
    num("dt", &j->dt);
    inum("start_step", &j->start_step);
    inum("current_step", &j->current_step);
    inum("end_step", &j->end_step);
    inum("output_step", &j->output_step);
    inum("ptc_num", &j->ptc_num);
    inum("water_ptc_num", &j->water_ptc_num);
    inum("air_ptc_num", &j->air_ptc_num);
    inum("wall_ptc_num", &j->wall_ptc_num);
    inum("rigid_ptc_num", &j->rigid_ptc_num);
    inum("thread_num", &j->thread_num);
    inum("gpu_id", &j->gpu_id);
    str("input_path", j->input_path, sizeof(j->input_path));
    str("output_path", j->output_path, sizeof(j->output_path));
    str("git_hash", j->git_hash, sizeof(j->git_hash));
    // extensions
    inum("dim", &j->dim);
    num("gamma", &j->gamma);
    num("c0", &j->c0);
    num("cfl", &j->cfl);
    auto it = kv.find("diffusion");
    if (it != kv.end()) {
        if (it->second.kind == Val::STR)
            j->diffusion = (it->second.str == "MC" || it->second.str == "mc") ? ZSPH_DIFF_MC
                                                                               : ZSPH_DIFF_ANTUONO;
        else if (it->second.kind == Val::NUM)
            inum("diffusion", &j->diffusion);
    }
    inum("output_fields", &j->output_fields);
    inum("binary", &j->binary);
    return bad ? ZSPH_E_IO : ZSPH_OK;
}
