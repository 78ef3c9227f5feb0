// Case driver and VTU output: the reference's JSON-driven PhysicsEuler run
// (src/models/euler.cu:15-221, apps/main.cu) over the device solver, and its ASCII VTU writer
// (src/fvmfieldsio.cu:5-287), written for this library.
//
// Case file keys (as the reference reads them): "mesh" (.su2 converted to .fvm beside it, or
// .fvm), "init": {"mach": {x,y,z}, "pressure", "temperature"}, "bcs": {<region>: {"type":
// "slip" | "farfield", far field: "mach", "pressure", "temperature"}}, "controls": {"cfl",
// "time": "steady", "tolerance"}, "save": {"file"}.  Extensions: "controls": {"max_iterations",
// "precision": "f64" | "f32", "renumber": bool}.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "gr_internal.hpp"

namespace gr {

// ---- a small JSON reader (objects, arrays, strings, numbers, true/false/null) ---------------

struct Json {
    enum Kind { Null, Bool, Num, Str, Arr, Obj } kind = Null;
    bool b = false;
    double num = 0;
    std::string str;
    std::vector<Json> arr;
    std::vector<std::pair<std::string, Json>> obj;

    const Json *get(const std::string &k) const {
        if (kind != Obj) return nullptr;
        for (const auto &kv : obj)
            if (kv.first == k) return &kv.second;
        return nullptr;
    }
};

class JsonParser {
  public:
    explicit JsonParser(const std::string &s) : s_(s) {}
    Json parse() {
        Json v = value();
        ws();
        require(p_ == s_.size(), "json: trailing characters at offset " + std::to_string(p_), GR_ERR_IO);
        return v;
    }

  private:
    const std::string &s_;
    size_t p_ = 0;
    void ws() {
        while (p_ < s_.size() && std::isspace((unsigned char)s_[p_])) ++p_;
    }
    [[noreturn]] void fail(const char *what) {
        throw Error(GR_ERR_IO, std::string("json: ") + what + " at offset " + std::to_string(p_));
    }
    Json value() {
        ws();
        if (p_ >= s_.size()) fail("unexpected end");
        const char c = s_[p_];
        if (c == '{') return object();
        if (c == '[') return array();
        if (c == '"') {
            Json j;
            j.kind = Json::Str;
            j.str = string();
            return j;
        }
        if (s_.compare(p_, 4, "true") == 0) {
            p_ += 4;
            Json j;
            j.kind = Json::Bool;
            j.b = true;
            return j;
        }
        if (s_.compare(p_, 5, "false") == 0) {
            p_ += 5;
            Json j;
            j.kind = Json::Bool;
            return j;
        }
        if (s_.compare(p_, 4, "null") == 0) {
            p_ += 4;
            return Json();
        }
        return number();
    }
    std::string string() {
        ++p_;   // opening quote
        std::string out;
        while (p_ < s_.size() && s_[p_] != '"') {
            char c = s_[p_++];
            if (c == '\\') {
                if (p_ >= s_.size()) fail("bad escape");
                const char e = s_[p_++];
                switch (e) {
                case 'n': out += '\n'; break;
                case 't': out += '\t'; break;
                case 'r': out += '\r'; break;
                case 'b': out += '\b'; break;
                case 'f': out += '\f'; break;
                case 'u': {
                    if (p_ + 4 > s_.size()) fail("bad unicode escape");
                    const unsigned cp = (unsigned)std::stoul(s_.substr(p_, 4), nullptr, 16);
                    p_ += 4;
                    if (cp < 0x80) out += (char)cp;
                    else if (cp < 0x800) {
                        out += (char)(0xC0 | (cp >> 6));
                        out += (char)(0x80 | (cp & 0x3F));
                    } else {
                        out += (char)(0xE0 | (cp >> 12));
                        out += (char)(0x80 | ((cp >> 6) & 0x3F));
                        out += (char)(0x80 | (cp & 0x3F));
                    }
                    break;
                }
                default: out += e; break;
                }
            } else {
                out += c;
            }
        }
        if (p_ >= s_.size()) fail("unterminated string");
        ++p_;
        return out;
    }
    Json number() {
        const char *b = s_.c_str() + p_;
        char *e = nullptr;
        const double v = std::strtod(b, &e);
        if (e == b) fail("unexpected character");
        p_ += (size_t)(e - b);
        Json j;
        j.kind = Json::Num;
        j.num = v;
        return j;
    }
    Json array() {
        ++p_;
        Json j;
        j.kind = Json::Arr;
        ws();
        if (p_ < s_.size() && s_[p_] == ']') {
            ++p_;
            return j;
        }
        for (;;) {
            j.arr.push_back(value());
            ws();
            if (p_ < s_.size() && s_[p_] == ',') {
                ++p_;
                continue;
            }
            if (p_ < s_.size() && s_[p_] == ']') {
                ++p_;
                return j;
            }
            fail("expected , or ]");
        }
    }
    Json object() {
        ++p_;
        Json j;
        j.kind = Json::Obj;
        ws();
        if (p_ < s_.size() && s_[p_] == '}') {
            ++p_;
            return j;
        }
        for (;;) {
            ws();
            if (p_ >= s_.size() || s_[p_] != '"') fail("expected a key");
            std::string k = string();
            ws();
            if (p_ >= s_.size() || s_[p_] != ':') fail("expected :");
            ++p_;
            j.obj.emplace_back(k, value());
            ws();
            if (p_ < s_.size() && s_[p_] == ',') {
                ++p_;
                continue;
            }
            if (p_ < s_.size() && s_[p_] == '}') {
                ++p_;
                return j;
            }
            fail("expected , or }");
        }
    }
};

double jnum(const Json &j, const char *key, const std::string &where) {
    const Json *v = j.get(key);
    require(v && v->kind == Json::Num, "case file: number \"" + std::string(key) + "\" missing in " + where,
            GR_ERR_IO);
    return v->num;
}

void jvec(const Json &j, const char *key, const std::string &where, double out[3]) {
    const Json *v = j.get(key);
    require(v && v->kind == Json::Obj, "case file: vector \"" + std::string(key) + "\" missing in " + where,
            GR_ERR_IO);
    out[0] = jnum(*v, "x", where + "." + key);
    out[1] = jnum(*v, "y", where + "." + key);
    out[2] = jnum(*v, "z", where + "." + key);
}

Json read_json(const char *path) {
    std::ifstream in(path);
    require(in.good(), std::string("cannot open case file ") + path, GR_ERR_IO);
    std::stringstream ss;
    ss << in.rdbuf();
    const std::string text = ss.str();
    return JsonParser(text).parse();
}

struct CaseBc {
    std::string region;
    int kind = GR_BC_SLIP;
    double mach[3] = {0, 0, 0};
    double p = 1, T = 1;
};

struct CaseSpec {
    std::string mesh, save;
    double mach[3] = {0, 0, 0};
    double p = 1, T = 1, cfl = 0.5, tol = 0;
    int max_iter = 100000, dtype = GR_F64;
    bool renumber = true;
    std::vector<CaseBc> bcs;
};

// Reads and checks every key before any device work.
CaseSpec parse_case(const char *path) {
    const Json js = read_json(path);
    require(js.kind == Json::Obj, "case file: top level must be an object", GR_ERR_IO);
    CaseSpec c;
    const Json *mesh = js.get("mesh");
    require(mesh && mesh->kind == Json::Str, "case file: \"mesh\" missing", GR_ERR_IO);
    c.mesh = mesh->str;
    const Json *init = js.get("init");
    require(init && init->kind == Json::Obj, "case file: \"init\" missing", GR_ERR_IO);
    jvec(*init, "mach", "init", c.mach);
    c.p = jnum(*init, "pressure", "init");
    c.T = jnum(*init, "temperature", "init");
    const Json *bcs = js.get("bcs");
    require(bcs && bcs->kind == Json::Obj, "case file: \"bcs\" missing", GR_ERR_IO);
    for (const auto &kv : bcs->obj) {
        CaseBc b;
        b.region = kv.first;
        const Json *type = kv.second.get("type");
        require(type && type->kind == Json::Str, "case file: bcs." + kv.first + ".type missing", GR_ERR_IO);
        if (type->str == "slip") {
            b.kind = GR_BC_SLIP;
        } else if (type->str == "farfield") {
            b.kind = GR_BC_FARFIELD;
            jvec(kv.second, "mach", "bcs." + kv.first, b.mach);
            b.p = jnum(kv.second, "pressure", "bcs." + kv.first);
            b.T = jnum(kv.second, "temperature", "bcs." + kv.first);
        } else {
            throw Error(GR_ERR_IO, "case file: unknown boundary type \"" + type->str + "\"");
        }
        c.bcs.push_back(b);
    }
    const Json *ctl = js.get("controls");
    require(ctl && ctl->kind == Json::Obj, "case file: \"controls\" missing", GR_ERR_IO);
    c.cfl = jnum(*ctl, "cfl", "controls");
    require(c.cfl > 0, "case file: controls.cfl must be positive", GR_ERR_IO);
    const Json *tm = ctl->get("time");
    require(tm && tm->kind == Json::Str && tm->str == "steady", "case file: controls.time must be \"steady\"",
            GR_ERR_IO);
    c.tol = jnum(*ctl, "tolerance", "controls");
    if (const Json *mi = ctl->get("max_iterations")) {
        require(mi->kind == Json::Num && mi->num >= 1, "case file: controls.max_iterations must be >= 1", GR_ERR_IO);
        c.max_iter = (int)mi->num;
    }
    if (const Json *pr = ctl->get("precision")) {
        require(pr->kind == Json::Str && (pr->str == "f64" || pr->str == "f32"),
                "case file: controls.precision must be \"f64\" or \"f32\"", GR_ERR_IO);
        c.dtype = pr->str == "f32" ? GR_F32 : GR_F64;
    }
    if (const Json *rn = ctl->get("renumber")) {
        require(rn->kind == Json::Bool, "case file: controls.renumber must be true or false", GR_ERR_IO);
        c.renumber = rn->b;
    }
    if (const Json *save = js.get("save")) {
        const Json *file = save->get("file");
        require(file && file->kind == Json::Str, "case file: save.file missing", GR_ERR_IO);
        c.save = file->str;
    }
    return c;
}

}  // namespace gr

using gr::guarded;
using gr::require;

extern "C" {

int gr_write_vtu(const gr_mesh *m, const char *path, uint32_t n_fields, const char *const *names,
                 const uint32_t *n_comp, const double *const *values) {
    return guarded([&] {
        require(m && path, "null argument");
        require(n_fields == 0 || (names && n_comp && values), "null field arrays");
        const gr::HostMesh &h = gr::host_mesh(m);
        const uint32_t N = h.n_cells;
        FILE *f = std::fopen(path, "w");
        require(f != nullptr, std::string("cannot open ") + path + " for writing", GR_ERR_IO);
        std::unique_ptr<FILE, int (*)(FILE *)> guard(f, &std::fclose);
        std::fprintf(f, "<VTKFile type=\"UnstructuredGrid\" version=\"1.0\" byte_order=\"LittleEndian\" "
                        "header_type=\"UInt64\">\n  <UnstructuredGrid>\n");
        std::fprintf(f, "    <Piece NumberOfPoints=\"%u\" NumberOfCells=\"%u\">\n      <CellData>\n", h.n_nodes, N);
        for (uint32_t k = 0; k < n_fields; ++k) {
            require(n_comp[k] == 1 || n_comp[k] == 3 || n_comp[k] == 9, "fields have 1, 3 or 9 components");
            require(names[k] && values[k], "null field name or values");
            std::fprintf(f, "        <DataArray type=\"Float64\" Name=\"%s\" NumberOfComponents=\"%u\" format=\"ascii\">\n",
                         names[k], n_comp[k]);
            for (uint32_t i = 0; i < N; ++i) {
                std::fputs("         ", f);
                for (uint32_t c = 0; c < n_comp[k]; ++c) {
                    const double v = values[k][(size_t)n_comp[k] * i + c];
                    if (std::isnan(v)) std::fputs(" nan", f);
                    else std::fprintf(f, " %.17g", v);
                }
                std::fputc('\n', f);
            }
            std::fputs("        </DataArray>\n", f);
        }
        std::fputs("      </CellData>\n      <Points>\n        <DataArray type=\"Float64\" NumberOfComponents=\"3\" "
                   "format=\"ascii\">\n", f);
        for (uint32_t n = 0; n < h.n_nodes; ++n)
            std::fprintf(f, "          %.17g %.17g %.17g\n", h.nodes[3 * n], h.nodes[3 * n + 1], h.nodes[3 * n + 2]);
        std::fputs("        </DataArray>\n      </Points>\n      <Cells>\n", f);
        // polyhedra (VTK type 42): sorted unique nodes of each cell, then its face stream
        std::fputs("        <DataArray type=\"Int64\" Name=\"connectivity\" format=\"ascii\">\n", f);
        std::vector<uint64_t> offsets(N), face_offsets(N);
        uint64_t off = 0, foff = 0;
        std::vector<uint32_t> nodes;
        for (uint32_t i = 0; i < N; ++i) {
            nodes.clear();
            for (uint32_t q = h.cell_faces_starts[i]; q < h.cell_faces_starts[i + 1]; ++q) {
                const uint32_t fc = h.cell_faces[q];
                for (uint32_t t = h.face_nodes_starts[fc]; t < h.face_nodes_starts[fc + 1]; ++t)
                    nodes.push_back(h.face_nodes[t]);
            }
            std::sort(nodes.begin(), nodes.end());
            nodes.erase(std::unique(nodes.begin(), nodes.end()), nodes.end());
            std::fputs("          ", f);
            for (uint32_t n : nodes) std::fprintf(f, " %u", n);
            std::fputc('\n', f);
            off += nodes.size();
            offsets[i] = off;
        }
        std::fputs("        </DataArray>\n        <DataArray type=\"Int64\" Name=\"offsets\" format=\"ascii\">\n", f);
        for (uint32_t i = 0; i < N; ++i) std::fprintf(f, "          %llu\n", (unsigned long long)offsets[i]);
        std::fputs("        </DataArray>\n        <DataArray type=\"UInt8\" Name=\"types\" format=\"ascii\">\n", f);
        for (uint32_t i = 0; i < N; ++i) std::fputs("          42\n", f);
        std::fputs("        </DataArray>\n        <DataArray type=\"Int64\" Name=\"faces\" format=\"ascii\">\n", f);
        for (uint32_t i = 0; i < N; ++i) {
            const uint32_t s = h.cell_faces_starts[i], e = h.cell_faces_starts[i + 1];
            std::fprintf(f, "          %u", e - s);
            ++foff;
            for (uint32_t q = s; q < e; ++q) {
                const uint32_t fc = h.cell_faces[q];
                const uint32_t a = h.face_nodes_starts[fc], b = h.face_nodes_starts[fc + 1];
                std::fprintf(f, " %u", b - a);
                ++foff;
                for (uint32_t t = a; t < b; ++t) std::fprintf(f, " %u", h.face_nodes[t]);
                foff += b - a;
            }
            std::fputc('\n', f);
            face_offsets[i] = foff;
        }
        std::fputs("        </DataArray>\n        <DataArray type=\"Int64\" Name=\"faceoffsets\" format=\"ascii\">\n", f);
        for (uint32_t i = 0; i < N; ++i) std::fprintf(f, "          %llu\n", (unsigned long long)face_offsets[i]);
        std::fputs("        </DataArray>\n      </Cells>\n    </Piece>\n  </UnstructuredGrid>\n</VTKFile>\n", f);
        require(std::ferror(f) == 0, std::string("write error on ") + path, GR_ERR_IO);
    });
}

int gr_case_check(const char *case_path) {
    return guarded([&] {
        require(case_path != nullptr, "null argument");
        (void)gr::parse_case(case_path);
    });
}

int gr_case_run(const char *case_path, int verbose, gr_solve_result *result) {
    return guarded([&] {
        require(case_path != nullptr, "null argument");
        const gr::CaseSpec c = gr::parse_case(case_path);
        auto log = [&](const std::string &s) {
            if (verbose) std::printf("%s\n", s.c_str());
        };
        auto ok = [](int rc) {
            if (rc != GR_OK) throw gr::Error(rc, gr_last_error());
        };
        std::string mesh_path = c.mesh;
        if (mesh_path.size() > 3 && mesh_path.compare(mesh_path.size() - 3, 3, "su2") == 0) {
            const std::string fvm = mesh_path.substr(0, mesh_path.size() - 3) + "fvm";
            log("Converting mesh to file " + fvm);
            ok(gr_mesh_from_su2(mesh_path.c_str(), fvm.c_str(), 8));
            mesh_path = fvm;
        }
        gr_mesh *m = nullptr;
        ok(gr_mesh_read_fvm(mesh_path.c_str(), 8, &m));
        std::unique_ptr<gr_mesh, void (*)(gr_mesh *)> mguard(m, &gr_mesh_free);
        if (c.renumber) ok(gr_mesh_renumber_rcm(m, nullptr, nullptr));
        ok(gr_mesh_compute_metrics(m));
        const gr::HostMesh &hm = gr::host_mesh(m);
        log("Mesh: " + std::to_string(hm.n_cells) + " cells, " + std::to_string(hm.n_faces) + " faces, " +
            std::to_string(hm.n_ghosts) + " ghosts");
        gr_solver_config cfg{};
        cfg.dtype = c.dtype;
        cfg.cfl = c.cfl;
        cfg.limiter = GR_LIM_MICHALAK;
        cfg.corrected = 0;
        cfg.use_graph = 1;
        gr_solver *s = nullptr;
        ok(gr_solver_create(m, &cfg, &s));
        std::unique_ptr<gr_solver, void (*)(gr_solver *)> sguard(s, &gr_solver_free);
        for (const auto &b : c.bcs) ok(gr_solver_set_bc(s, b.region.c_str(), b.kind, b.mach, b.p, b.T));
        ok(gr_solver_init_freestream(s, c.mach, c.p, c.T));
        gr_solve_result res{};
        ok(gr_solver_iterate(s, c.max_iter, c.tol, 0.0, &res));
        char line[160];
        std::snprintf(line, sizeof line, "Iterations: %d, residual ratio %.6e, %s", res.iterations, res.ratio,
                      res.converged ? "converged" : "not converged");
        log(line);
        if (result) *result = res;
        if (!c.save.empty()) {
            const size_t NG = (size_t)hm.n_cells + hm.n_ghosts;
            std::vector<double> rho(NG), rhou(3 * NG), rhoe(NG);
            ok(gr_solver_get_state(s, rho.data(), rhou.data(), rhoe.data()));
            const char *names[3] = {"rho", "rhou", "rhoe"};
            const uint32_t nc[3] = {1, 3, 1};
            const double *vals[3] = {rho.data(), rhou.data(), rhoe.data()};
            log("Writing Euler solution to file " + c.save);
            ok(gr_write_vtu(m, c.save.c_str(), 3, names, nc, vals));
        }
    });
}

}  // extern "C"
