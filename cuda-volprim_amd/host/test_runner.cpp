// test_runner — the reference harness's command line (test/test_runner.cpp:95-434)
// over the MI355X renderer: the same flags and defaults, the validation-scene
// dispatch, multi-view output naming (<output-dir>/<scene>/NNNN.exr).
//
// Additions: --env <file.hdr> overrides the scene's environment map and
// --active-cap sets the active-set capacity.  --denoise is rejected (no AI
// denoiser on this platform).  The --all/--validation/--stress registries of
// geometric_validation.cpp are not rebuilt; the validation scenes are.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <iostream>
#include <string>

#include "app.h"

namespace fs = std::filesystem;

namespace {

struct TestConfig {
    std::string scene_name;
    bool list_scenes = false, run_all = false, run_validation = false, run_stress = false;
    uint32_t spp = 64;
    uint32_t width = 1920, height = 1080;
    std::string output_dir = "test_results";
    std::string output_file;
    float sigma_multiplier = 7.5f;
    bool denoise = false;
    uint32_t seed = 42;
    uint32_t max_depth = 128, rr_depth = 12;
    float rr_max_survival = 0.99f, firefly_clamp = 0.0f, filter_stddev = 0.0f, hg_g = 0.85f;
    bool use_ris = false;
    uint32_t ris_candidates = 6;
    bool measure_caps = false;
    std::string env;
    uint32_t active_cap = 128;
    float shadow_rr = 0.0f;
    float adaptive_threshold = 0.0f;
    uint32_t adaptive_min_samples = 32;
};

void usage() {
    std::printf(
        "Test Runner for Gaussian Volumetric Path Tracer (MI355X)\n"
        "  --scene NAME            run one scene (see --list)\n"
        "  --list                  list scenes\n"
        "  --spp N                 samples per pixel (64)\n"
        "  --sigma-multiplier F    sigma_t scaling factor (7.5)\n"
        "  --seed N                RNG seed (42)\n"
        "  --width N --height N    image size for single-view scenes (1920x1080)\n"
        "  --output FILE           output file (single scene)\n"
        "  --output-dir DIR        output directory (test_results)\n"
        "  --max-depth N           maximum path depth (128)\n"
        "  --phase-g F             Henyey-Greenstein g in (-1,1) (0.85)\n"
        "  --rr-depth N            depth at which Russian roulette begins (12)\n"
        "  --rr-max-survival F     Russian roulette max survival (0.99)\n"
        "  --firefly-clamp F       per-sample luminance clamp, 0 = off (biased when > 0)\n"
        "  --filter-width F        Gaussian pixel-filter stddev in px, 0 = box\n"
        "  --ris                   product-RIS NEE instead of MIS\n"
        "  --ris-candidates N      RIS candidate count K (6)\n"
        "  --measure-caps          measure max hits/ray and point overlap, print suggested caps\n"
        "  --env FILE.hdr          environment map override\n"
        "  --active-cap N          active-set capacity (128)\n"
        "  --shadow-rr F           shadow-ray roulette threshold, 0 = off (unbiased when > 0)\n"
        "  --adaptive-threshold F  adaptive sampling: stop a pixel once its relative sample deviation is\n"
        "                          below F (0 = off, the reference's compile-time default)\n"
        "  --adaptive-min-samples N  samples before a pixel may stop (default 32)\n");
}

bool parse(int argc, char** argv, TestConfig& c, std::string& err) {
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto next = [&](const char* name) -> const char* {
            if (i + 1 >= argc) {
                err = std::string(name) + " needs a value";
                return nullptr;
            }
            return argv[++i];
        };
        auto num_u = [&](const char* name, uint32_t& dst, bool positive) -> bool {
            const char* v = next(name);
            if (!v) return false;
            char* end = nullptr;
            const long x = std::strtol(v, &end, 10);
            if (*end || x < 0 || (positive && x == 0)) {
                err = std::string(name) + ": expected a " + (positive ? "positive" : "non-negative") + " integer";
                return false;
            }
            dst = static_cast<uint32_t>(x);
            return true;
        };
        auto num_f = [&](const char* name, float& dst, float lo, float hi) -> bool {
            const char* v = next(name);
            if (!v) return false;
            char* end = nullptr;
            const float x = std::strtof(v, &end);
            if (*end || !(x >= lo && x <= hi)) {
                err = std::string(name) + ": value out of range";
                return false;
            }
            dst = x;
            return true;
        };
        if (a == "--help" || a == "-h") {
            usage();
            std::exit(0);
        } else if (a == "--scene") {
            const char* v = next("--scene");
            if (!v) return false;
            c.scene_name = v;
        } else if (a == "--list") c.list_scenes = true;
        else if (a == "--all") c.run_all = true;
        else if (a == "--validation") c.run_validation = true;
        else if (a == "--stress") c.run_stress = true;
        else if (a == "--spp") { if (!num_u("--spp", c.spp, true)) return false; }
        else if (a == "--sigma-multiplier") { if (!num_f("--sigma-multiplier", c.sigma_multiplier, 1e-30f, 3e38f)) return false; }
        else if (a == "--seed") { if (!num_u("--seed", c.seed, false)) return false; }
        else if (a == "--width") { if (!num_u("--width", c.width, true)) return false; }
        else if (a == "--height") { if (!num_u("--height", c.height, true)) return false; }
        else if (a == "--output") { const char* v = next("--output"); if (!v) return false; c.output_file = v; }
        else if (a == "--output-dir") { const char* v = next("--output-dir"); if (!v) return false; c.output_dir = v; }
        else if (a == "--denoise") c.denoise = true;
        else if (a == "--max-depth") { if (!num_u("--max-depth", c.max_depth, true)) return false; }
        else if (a == "--phase-g") { if (!num_f("--phase-g", c.hg_g, -0.999f, 0.999f)) return false; }
        else if (a == "--rr-depth") { if (!num_u("--rr-depth", c.rr_depth, false)) return false; }
        else if (a == "--rr-max-survival") { if (!num_f("--rr-max-survival", c.rr_max_survival, 0.0f, 1.0f)) return false; }
        else if (a == "--firefly-clamp") { if (!num_f("--firefly-clamp", c.firefly_clamp, 0.0f, 3e38f)) return false; }
        else if (a == "--filter-width") { if (!num_f("--filter-width", c.filter_stddev, 0.0f, 3e38f)) return false; }
        else if (a == "--ris") c.use_ris = true;
        else if (a == "--ris-candidates") { if (!num_u("--ris-candidates", c.ris_candidates, true)) return false; }
        else if (a == "--measure-caps") c.measure_caps = true;
        else if (a == "--env") { const char* v = next("--env"); if (!v) return false; c.env = v; }
        else if (a == "--active-cap") { if (!num_u("--active-cap", c.active_cap, true)) return false; }
        else if (a == "--shadow-rr") { if (!num_f("--shadow-rr", c.shadow_rr, 0.0f, 3e38f)) return false; }
        else if (a == "--adaptive-threshold") {
            if (!num_f("--adaptive-threshold", c.adaptive_threshold, 0.0f, 3e38f)) return false;
        } else if (a == "--adaptive-min-samples") {
            if (!num_u("--adaptive-min-samples", c.adaptive_min_samples, false)) return false;
        }
        else {
            err = "unknown argument '" + a + "'";
            return false;
        }
    }
    return true;
}

vpapp::Config renderer_config(const TestConfig& t, uint32_t w, uint32_t h, const std::string& out,
                              const std::string& env) {
    vpapp::Config c;  // test_runner.cpp:188-191 fills the renderer's Config the same way
    c.image_width_ = w;
    c.image_height_ = h;
    c.num_samples_per_pixel_ = t.spp;
    c.output_path_ = out;
    c.seed_ = t.seed;
    c.denoise_ = t.denoise;
    c.max_bounces_ = t.max_depth;
    c.rr_depth_ = t.rr_depth;
    c.rr_max_survival_ = t.rr_max_survival;
    c.firefly_clamp_luminance_ = t.firefly_clamp;
    c.pixel_filter_stddev_ = t.filter_stddev;
    c.hg_g_ = t.hg_g;
    c.use_ris_ = t.use_ris;
    c.ris_num_candidates_ = t.ris_candidates;
    c.measure_caps_ = t.measure_caps;
    c.env_map_path_ = env;
    c.active_cap_ = t.active_cap;
    c.shadow_rr_ = t.shadow_rr;
    c.adaptive_threshold_ = t.adaptive_threshold;
    c.adaptive_min_samples_ = t.adaptive_min_samples;
    return c;
}

int run_multiview(const vpapp::Scene& scene, const TestConfig& t) {
    std::printf("Running multi-view test: %s\nDescription: %s\nPrimitives: %zu\nCameras: %zu\nSPP: %u\n",
                scene.name.c_str(), scene.description.c_str(), scene.primitives.size(), scene.cameras.size(), t.spp);
    if (scene.primitives.empty() || scene.cameras.empty()) {
        std::fprintf(stderr, "Multi-view test failed: empty scene\n");
        return 1;
    }
    const fs::path dir = fs::path(t.output_dir) / scene.name;
    fs::create_directories(dir);
    for (size_t k = 0; k < scene.cameras.size(); ++k) {
        char name[32];
        std::snprintf(name, sizeof(name), "%04zu.exr", k);
        const std::string out = (dir / name).string();  // test_runner.cpp:236-239 naming
        const auto& v = scene.cameras[k];
        const std::string env = !t.env.empty() ? t.env : scene.env_map_override.value_or("assets/environment_maps/white_constant.hdr");
        std::vector<vp_primitive> prims = scene.primitives;
        vpapp::Renderer r(renderer_config(t, v.width, v.height, out, env), std::move(prims), v.camera);
        r.render();
        std::printf("  Camera %zu completed -> %s\n", k, out.c_str());
    }
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    TestConfig t;
    std::string err;
    if (!parse(argc, argv, t, err)) {
        std::fprintf(stderr, "Argument parse error: %s\n", err.c_str());
        return 1;
    }
    if (t.list_scenes) {
        std::printf("\nValidation scenes (run via --scene <name>):\n");
        for (const auto& n : vpapp::scene_names()) std::printf("    %s\n", n.c_str());
        return 0;
    }
    if (t.run_all || t.run_validation || t.run_stress) {
        std::fprintf(stderr, "--all/--validation/--stress: the geometric registry scenes are not part of this build; "
                             "run the validation scenes with --scene (see --list)\n");
        return 1;
    }
    if (t.scene_name.empty()) {
        std::fprintf(stderr, "Error: No action specified. Use --scene or --list.\n");
        return 1;
    }
    vpapp::Scene scene;
    bool ok = false;
    if (t.scene_name == "single_gaussian_validation") ok = vpapp::single_gaussian_validation(t.sigma_multiplier, scene, err);
    else if (t.scene_name == "two_gaussian_validation") ok = vpapp::two_gaussian_validation(t.sigma_multiplier, scene, err);
    else if (t.scene_name == "cluster_validation") ok = vpapp::cluster_validation(t.sigma_multiplier, scene, err);
    else if (t.scene_name == "asset_validation") ok = vpapp::asset_validation(t.sigma_multiplier, scene, err);
    else {
        std::fprintf(stderr, "Error: Scene '%s' not found.\nUse --list to see available scenes.\n", t.scene_name.c_str());
        return 1;
    }
    if (!ok) {
        std::fprintf(stderr, "Failed to build %s: %s\n", t.scene_name.c_str(), err.c_str());
        return 1;
    }
    return run_multiview(scene, t);
}
