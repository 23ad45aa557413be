// vp_serve — a request/response front end for the serving form.
//
// Reads one JSON request per line on stdin and writes one JSON response per
// line on stdout.  Requests go to a vp_server (continuous batching on the
// GPU); scenes (PLY files) and environments (Radiance .hdr files) are loaded
// once and stay resident, keyed by path (and sigma multiplier).  Responses
// are written in completion order, not request order.
//
// Request fields (all but "scene" and "output" optional):
//   {"id": "...", "scene": "file.ply", "sigma_multiplier": 7.5,
//    "env": "file.hdr" (default: constant white),
//    "width": 256, "height": 256, "spp": 16, "seed": 42, "max_depth": 128,
//    "rr_depth": 12, "rr_max_survival": 0.99, "phase_g": 0.85, "ris": false,
//    "ris_candidates": 6, "firefly_clamp": 0, "filter_width": 0, "active_cap": 128,
//    "shadow_rr": 0 (shadow-ray roulette threshold, 0 = off),
//    "camera": {"eye": [0,0,-3.5], "target": [0,0,0], "up": [0,1,0], "fov": 40}
//              or {"eye": ..., "target": ..., "up": ..., "ortho_height": 6},
//    "output": "out.exr"}
// Response: {"id": "...", "status": "ok", "latency_ms": 12.5, "output": "out.exr"}
//        or {"id": "...", "status": "error", "message": "..."}
// At end of input, after every request has finished:
//   {"stats": {"completed": n, "samples": s, "p50_ms": .., "p99_ms": .., "busy_ms": ..}}
// Option: --reserve-cus keeps one CU of each XCD out of the render launches
// (vp_server_config.reserve_cus), so a request is answered as soon as its own
// paths finish rather than when the launch it shares frees compute units.
//
// The scene and camera conventions are the reference's (PLY loading as
// src/thesis/host/utils/io/ply.cpp:34-175, cameras as
// include/thesis/host/params/camera.h:40-111, EXR output as
// src/thesis/host/utils/io/exr.cpp:35-105).
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/vp_api.h"
#include "../../include/vp_io.h"

namespace {

// ---- a small JSON reader: objects, arrays, strings, numbers, literals -------
struct Json {
    enum Kind { Null, Bool, Num, Str, Arr, Obj } kind = Null;
    bool b = false;
    double n = 0.0;
    std::string s;
    std::vector<Json> a;
    std::vector<std::pair<std::string, Json>> o;
    const Json* get(const std::string& k) const {
        for (const auto& kv : o)
            if (kv.first == k) return &kv.second;
        return nullptr;
    }
};

struct Parser {
    const char* p;
    const char* end;
    std::string err;
    void ws() {
        while (p < end && (*p == ' ' || *p == '\t' || *p == '\r' || *p == '\n')) ++p;
    }
    bool fail(const char* m) {
        if (err.empty()) err = m;
        return false;
    }
    bool lit(const char* w) {
        const size_t l = std::strlen(w);
        if (static_cast<size_t>(end - p) < l || std::strncmp(p, w, l) != 0) return false;
        p += l;
        return true;
    }
    bool str(std::string& out) {
        if (p >= end || *p != '"') return fail("expected a string");
        ++p;
        while (p < end && *p != '"') {
            char c = *p++;
            if (c == '\\') {
                if (p >= end) return fail("bad escape");
                const char e = *p++;
                switch (e) {
                    case '"': c = '"'; break;
                    case '\\': c = '\\'; break;
                    case '/': c = '/'; break;
                    case 'n': c = '\n'; break;
                    case 't': c = '\t'; break;
                    case 'r': c = '\r'; break;
                    case 'b': c = '\b'; break;
                    case 'f': c = '\f'; break;
                    default: return fail("unsupported escape");
                }
            }
            out.push_back(c);
        }
        if (p >= end) return fail("unterminated string");
        ++p;
        return true;
    }
    bool value(Json& v, int depth) {
        if (depth > 32) return fail("nesting too deep");
        ws();
        if (p >= end) return fail("unexpected end of input");
        if (*p == '{') {
            ++p;
            v.kind = Json::Obj;
            ws();
            if (p < end && *p == '}') return ++p, true;
            while (true) {
                ws();
                std::string k;
                if (!str(k)) return false;
                ws();
                if (p >= end || *p != ':') return fail("expected ':'");
                ++p;
                Json x;
                if (!value(x, depth + 1)) return false;
                v.o.emplace_back(std::move(k), std::move(x));
                ws();
                if (p < end && *p == ',') {
                    ++p;
                    continue;
                }
                if (p < end && *p == '}') return ++p, true;
                return fail("expected ',' or '}'");
            }
        }
        if (*p == '[') {
            ++p;
            v.kind = Json::Arr;
            ws();
            if (p < end && *p == ']') return ++p, true;
            while (true) {
                Json x;
                if (!value(x, depth + 1)) return false;
                v.a.push_back(std::move(x));
                ws();
                if (p < end && *p == ',') {
                    ++p;
                    continue;
                }
                if (p < end && *p == ']') return ++p, true;
                return fail("expected ',' or ']'");
            }
        }
        if (*p == '"') {
            v.kind = Json::Str;
            return str(v.s);
        }
        if (lit("true")) return v.kind = Json::Bool, v.b = true, true;
        if (lit("false")) return v.kind = Json::Bool, v.b = false, true;
        if (lit("null")) return v.kind = Json::Null, true;
        char* e = nullptr;
        const double d = std::strtod(p, &e);
        if (e == p) return fail("unexpected character");
        p = e;
        v.kind = Json::Num;
        v.n = d;
        return true;
    }
};

bool parse_json(const std::string& line, Json& out, std::string& err) {
    Parser ps{line.data(), line.data() + line.size(), {}};
    if (!ps.value(out, 0)) {
        err = ps.err;
        return false;
    }
    ps.ws();
    if (ps.p != ps.end) {
        err = "trailing characters";
        return false;
    }
    if (out.kind != Json::Obj) {
        err = "a request must be a JSON object";
        return false;
    }
    return true;
}

std::string id_text(const Json& x) {
    if (x.kind == Json::Str) return x.s;
    char buf[64];
    std::snprintf(buf, sizeof(buf), "%.17g", x.n);
    return buf;
}

std::string quote(const std::string& s) {
    std::string r = "\"";
    for (char c : s) {
        if (c == '"' || c == '\\') r += '\\', r += c;
        else if (c == '\n') r += "\\n";
        else if (static_cast<unsigned char>(c) < 0x20) r += ' ';
        else r += c;
    }
    return r + "\"";
}

// ---- request decoding -------------------------------------------------------
bool num(const Json& o, const char* k, double& v, std::string& err, double lo, double hi) {
    const Json* x = o.get(k);
    if (!x) return true;
    if (x->kind != Json::Num || !(x->n >= lo && x->n <= hi)) {
        err = std::string("field '") + k + "' must be a number in range";
        return false;
    }
    v = x->n;
    return true;
}
bool vec3(const Json& o, const char* k, float out[3], std::string& err) {
    const Json* x = o.get(k);
    if (!x) return true;
    if (x->kind != Json::Arr || x->a.size() != 3) {
        err = std::string("field '") + k + "' must be an array of 3 numbers";
        return false;
    }
    for (int i = 0; i < 3; ++i) {
        if (x->a[i].kind != Json::Num) {
            err = std::string("field '") + k + "' must be an array of 3 numbers";
            return false;
        }
        out[i] = static_cast<float>(x->a[i].n);
    }
    return true;
}

struct Pending {
    std::string id, output;
    uint64_t ticket = 0;
    std::vector<float> image;
    uint32_t w = 0, h = 0;
};

class Service {
   public:
    ~Service() {
        if (srv_) vp_server_destroy(srv_);
        for (auto& kv : scenes_) vp_scene_destroy(kv.second);
        for (auto& kv : envs_) vp_env_destroy(kv.second);
        if (ctx_) vp_context_destroy(ctx_);
    }
    // Submit one request; returns false (with a message) when it is rejected.
    bool submit(const Json& r, std::unique_ptr<Pending>& out, std::string& err) {
        auto P = std::make_unique<Pending>();
        if (const Json* id = r.get("id")) P->id = id_text(*id);
        const Json* scene = r.get("scene");
        const Json* output = r.get("output");
        if (!scene || scene->kind != Json::Str) return err = "field 'scene' (a PLY path) is required", false;
        if (!output || output->kind != Json::Str) return err = "field 'output' (an EXR path) is required", false;
        P->output = output->s;
        double sigma = 7.5, width = 256, height = 256, spp = 16, seed = 42, max_depth = 128, rr_depth = 12;
        double rr_surv = 0.99, g = 0.85, ris_k = 6, clamp = 0, filt = 0, cap = 128, srr = 0;
        if (!num(r, "sigma_multiplier", sigma, err, 1e-30, 3e38) || !num(r, "width", width, err, 1, 1024) ||
            !num(r, "height", height, err, 1, 1024) || !num(r, "spp", spp, err, 1, 1 << 20) ||
            !num(r, "seed", seed, err, 0, 4294967295.0) || !num(r, "max_depth", max_depth, err, 1, 1 << 20) ||
            !num(r, "rr_depth", rr_depth, err, 0, 1 << 20) || !num(r, "rr_max_survival", rr_surv, err, 0, 1) ||
            !num(r, "phase_g", g, err, -0.999, 0.999) || !num(r, "ris_candidates", ris_k, err, 1, 64) ||
            !num(r, "firefly_clamp", clamp, err, 0, 3e38) || !num(r, "filter_width", filt, err, 0, 3e38) ||
            !num(r, "active_cap", cap, err, 1, 4096) || !num(r, "shadow_rr", srr, err, 0, 3e38))
            return false;
        bool ris = false;
        if (const Json* x = r.get("ris")) ris = x->kind == Json::Bool && x->b;
        float eye[3] = {0.0f, 0.0f, -3.5f}, target[3] = {0.0f, 0.0f, 0.0f}, up[3] = {0.0f, 1.0f, 0.0f};
        double fov = 40.0, ortho = 0.0;
        if (const Json* c = r.get("camera")) {
            if (c->kind != Json::Obj) return err = "field 'camera' must be an object", false;
            if (!vec3(*c, "eye", eye, err) || !vec3(*c, "target", target, err) || !vec3(*c, "up", up, err) ||
                !num(*c, "fov", fov, err, 1e-3, 179.0) || !num(*c, "ortho_height", ortho, err, 0, 3e38))
                return false;
        }
        if (!ensure_context(err)) return false;
        vp_scene* sc = nullptr;
        vp_env* env = nullptr;
        if (!load_scene(scene->s, static_cast<float>(sigma), sc, err)) return false;
        std::string env_path;
        if (const Json* e = r.get("env")) env_path = e->kind == Json::Str ? e->s : "";
        if (!load_env(env_path, env, err)) return false;
        vp_job job{};
        job.scene = sc;
        job.env = env;
        P->w = static_cast<uint32_t>(width);
        P->h = static_cast<uint32_t>(height);
        vp_status st = ortho > 0.0 ? vp_camera_orthographic(P->w, P->h, eye, target, up, static_cast<float>(ortho),
                                                             &job.camera)
                                   : vp_camera_perspective(P->w, P->h, eye, target, up, static_cast<float>(fov),
                                                           &job.camera);
        if (st != VP_OK) return err = vp_last_error(), false;
        vp_default_params(&job.params);
        job.params.width = P->w;
        job.params.height = P->h;
        job.params.spp = static_cast<uint32_t>(spp);
        job.params.seed = static_cast<uint32_t>(seed);
        job.params.max_depth = static_cast<uint32_t>(max_depth);
        job.params.rr_depth = static_cast<uint32_t>(rr_depth);
        job.params.rr_max_survival = static_cast<float>(rr_surv);
        job.params.phase_g = static_cast<float>(g);
        job.params.use_ris = ris ? 1u : 0u;
        job.params.ris_candidates = static_cast<uint32_t>(ris_k);
        job.params.firefly_clamp = static_cast<float>(clamp);
        job.params.filter_stddev = static_cast<float>(filt);
        job.params.active_cap = static_cast<uint32_t>(cap);
        job.params.shadow_rr = static_cast<float>(srr);
        P->image.assign(static_cast<size_t>(P->w) * P->h * 4, 0.0f);
        job.out_rgba = P->image.data();
        if (vp_server_submit(srv_, &job, &P->ticket) != VP_OK) return err = vp_last_error(), false;
        out = std::move(P);
        return true;
    }
    vp_server* server() const { return srv_; }
    void set_reserve_cus(bool on) { reserve_cus_ = on; }

   private:
    bool ensure_context(std::string& err) {
        if (ctx_) return true;
        if (vp_context_create(0, &ctx_) != VP_OK) return err = vp_last_error(), false;
        vp_server_config cfg{uint64_t(16) << 30, 64, 384, reserve_cus_ ? 1u : 0u};
        if (vp_server_create(ctx_, &cfg, &srv_) != VP_OK) return err = vp_last_error(), false;
        return true;
    }
    bool load_scene(const std::string& path, float sigma, vp_scene*& out, std::string& err) {
        const std::string key = path + "#" + std::to_string(sigma);
        auto it = scenes_.find(key);
        if (it != scenes_.end()) return out = it->second, true;
        vp_primitive* prims = nullptr;
        uint32_t n = 0;
        if (vp_io_load_ply(path.c_str(), sigma, nullptr, &prims, &n) != 0) return err = vp_io_last_error(), false;
        // median-split leaf order (host build, once per cached scene)
        const vp_status st = vp_scene_create_ex(ctx_, prims, n, VP_SCENE_BVH_MEDIAN, &out);
        vp_io_free(prims);
        if (st != VP_OK) return err = vp_last_error(), false;
        scenes_[key] = out;
        return true;
    }
    bool load_env(const std::string& path, vp_env*& out, std::string& err) {
        auto it = envs_.find(path);
        if (it != envs_.end()) return out = it->second, true;
        std::vector<float> rgba;
        uint32_t w = 0, h = 0;
        if (path.empty()) {
            w = 64, h = 32;
            rgba.assign(static_cast<size_t>(w) * h * 4, 1.0f);
        } else {
            float* data = nullptr;
            if (vp_io_load_hdr(path.c_str(), &data, &w, &h) != 0) return err = vp_io_last_error(), false;
            rgba.assign(data, data + static_cast<size_t>(w) * h * 4);
            vp_io_free(data);
        }
        if (vp_env_create(ctx_, rgba.data(), w, h, &out) != VP_OK) return err = vp_last_error(), false;
        envs_[path] = out;
        return true;
    }
    vp_context* ctx_ = nullptr;
    vp_server* srv_ = nullptr;
    bool reserve_cus_ = false;
    std::map<std::string, vp_scene*> scenes_;
    std::map<std::string, vp_env*> envs_;
};

}  // namespace

int main(int argc, char** argv) {
    Service svc;
    for (int i = 1; i < argc; ++i) {
        if (std::string(argv[i]) == "--reserve-cus") {
            svc.set_reserve_cus(true);
        } else {
            std::fprintf(stderr, "vp_serve: unknown option %s (known: --reserve-cus)\n", argv[i]);
            return 2;
        }
    }
    std::mutex out_mu, pend_mu;
    std::vector<std::unique_ptr<Pending>> pending;
    bool input_done = false;
    auto emit = [&](const std::string& line) {
        std::lock_guard<std::mutex> lk(out_mu);
        std::fwrite(line.data(), 1, line.size(), stdout);
        std::fputc('\n', stdout);
        std::fflush(stdout);
    };
    // completion thread: poll tickets, write EXRs, answer
    std::thread poller([&] {
        while (true) {
            std::vector<std::unique_ptr<Pending>> done;
            std::vector<double> lat;
            std::vector<vp_status> sts;
            bool finished = false;
            {
                std::lock_guard<std::mutex> lk(pend_mu);
                for (size_t i = 0; i < pending.size();) {
                    double ms = 0.0;
                    const vp_status st = vp_server_poll(svc.server(), pending[i]->ticket, &ms);
                    if (st == VP_ERR_NOT_READY) {
                        ++i;
                        continue;
                    }
                    done.push_back(std::move(pending[i]));
                    lat.push_back(ms);
                    sts.push_back(st);
                    pending.erase(pending.begin() + static_cast<long>(i));
                }
                finished = input_done && pending.empty();
            }
            for (size_t k = 0; k < done.size(); ++k) {
                Pending& P = *done[k];
                if (sts[k] != VP_OK) {
                    emit("{\"id\": " + quote(P.id) + ", \"status\": \"error\", \"message\": \"render failed\"}");
                } else if (vp_io_save_exr(P.output.c_str(), P.image.data(), P.w, P.h) != 0) {
                    emit("{\"id\": " + quote(P.id) + ", \"status\": \"error\", \"message\": " +
                         quote(vp_io_last_error()) + "}");
                } else {
                    char buf[64];
                    std::snprintf(buf, sizeof(buf), "%.3f", lat[k]);
                    emit("{\"id\": " + quote(P.id) + ", \"status\": \"ok\", \"latency_ms\": " + buf +
                         ", \"output\": " + quote(P.output) + "}");
                }
            }
            if (finished) return;
            std::this_thread::sleep_for(std::chrono::milliseconds(1));
        }
    });
    std::string line;
    while (std::getline(std::cin, line)) {
        if (line.find_first_not_of(" \t\r") == std::string::npos) continue;
        Json req;
        std::string err;
        std::string id;
        if (!parse_json(line, req, err)) {
            emit("{\"id\": null, \"status\": \"error\", \"message\": " + quote("bad request: " + err) + "}");
            continue;
        }
        if (const Json* x = req.get("id")) id = id_text(*x);
        std::unique_ptr<Pending> P;
        if (!svc.submit(req, P, err)) {
            emit("{\"id\": " + quote(id) + ", \"status\": \"error\", \"message\": " + quote(err) + "}");
            continue;
        }
        std::lock_guard<std::mutex> lk(pend_mu);
        pending.push_back(std::move(P));
    }
    {
        std::lock_guard<std::mutex> lk(pend_mu);
        input_done = true;
    }
    poller.join();
    if (svc.server()) {
        vp_server_stats st{};
        vp_server_stats_get(svc.server(), &st);
        char buf[320];
        std::snprintf(buf, sizeof(buf),
                      "{\"stats\": {\"completed\": %llu, \"samples\": %llu, \"p50_ms\": %.3f, \"p99_ms\": %.3f, "
                      "\"busy_ms\": %.3f}}",
                      static_cast<unsigned long long>(st.completed), static_cast<unsigned long long>(st.samples),
                      st.p50_ms, st.p99_ms, st.busy_ms);
        emit(buf);
    }
    return 0;
}
