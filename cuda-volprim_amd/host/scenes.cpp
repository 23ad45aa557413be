// scenes.cpp — the validation scenes of the reference harness, rebuilt on the
// C ABI.  Geometry, masses, cameras and the SG_* environment variables follow
// test/scenes/single_gaussian.cpp:26-313 and test/scenes/cloud_validation.cpp:94-139.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>

#include "app.h"

namespace vpapp {

namespace {

constexpr uint32_t kSgWidth = 256, kSgHeight = 256;
constexpr float kSgScale = 1.0f;
constexpr float kSgOrthoHeight = 6.0f * kSgScale;
constexpr float kSgCameraDistance = 5.0f * kSgScale;
constexpr float kPi = 3.14159265358979323846f;

const char* env_or(const char* k, const char* d) {
    const char* v = std::getenv(k);
    return v ? v : d;
}

// SG_ALBEDO: scalar (grey) or "r,g,b" (single_gaussian.cpp:26-33)
void scatter_albedo3(float a[3]) {
    a[0] = a[1] = a[2] = 0.0f;
    const char* e = std::getenv("SG_ALBEDO");
    if (!e) return;
    float r = 0, g = 0, b = 0;
    const int n = std::sscanf(e, "%f,%f,%f", &r, &g, &b);
    if (n >= 3) {
        a[0] = r, a[1] = g, a[2] = b;
    } else {
        a[0] = a[1] = a[2] = r;
    }
}

std::string env_map_path() {
    return std::string(env_or("SG_ENV", "white_constant")) == "meadow" ? "assets/environment_maps/meadow_2_4k.hdr"
                                                                         : "assets/environment_maps/white_constant.hdr";
}

struct View {
    float o[3], t[3], u[3];
};
View view_config() {
    const std::string v = env_or("SG_VIEW", "posz");
    const float D = kSgCameraDistance;
    View r{{0, 0, -D}, {0, 0, 0}, {0, 1, 0}};
    if (v == "negz") r = View{{0, 0, D}, {0, 0, 0}, {0, 1, 0}};
    else if (v == "posx") r = View{{-D, 0, 0}, {0, 0, 0}, {0, 1, 0}};
    else if (v == "negx") r = View{{D, 0, 0}, {0, 0, 0}, {0, 1, 0}};
    else if (v == "posy") r = View{{0, -D, 0}, {0, 0, 0}, {0, 0, 1}};
    else if (v == "negy") r = View{{0, D, 0}, {0, 0, 0}, {0, 0, 1}};
    else if (v == "sun") {
        const float s[3] = {0.7197f, 0.4530f, 0.5261f};
        r = View{{-D * s[0], -D * s[1], -D * s[2]}, {0, 0, 0}, {0, 1, 0}};
    }
    return r;
}

void zrot_quat(float deg, float q[4]) {
    const float h = 0.5f * deg * kPi / 180.0f;
    q[0] = std::cos(h), q[1] = 0.0f, q[2] = 0.0f, q[3] = std::sin(h);
}

}  // namespace

std::vector<std::string> scene_names() {
    return {"single_gaussian_validation", "two_gaussian_validation", "cluster_validation", "asset_validation"};
}

bool single_gaussian_validation(float sigma_multiplier, Scene& s, std::string&) {
    s = Scene{};
    s.name = "single_gaussian_validation";
    s.description = "Single isotropic absorber Gaussian at origin, ortho view in [-3,3]^2 - closed-form analytic "
                    "verification of escape transmittance";
    float albedo[3];
    scatter_albedo3(albedo);
    const bool transformed = std::string(env_or("SG_TRANSFORMED", "0")) == "1";
    float scale[3] = {kSgScale, kSgScale, kSgScale};
    float q[4] = {1, 0, 0, 0};
    if (transformed) {
        scale[0] = 1.0f, scale[1] = 0.5f, scale[2] = 0.75f;
        zrot_quat(30.0f, q);
        s.description = "Single ANISOTROPIC+ROTATED absorber Gaussian at origin (scale=(1,0.5,0.75), 30 deg about +Z)";
    }
    const float c[3] = {0, 0, 0};
    s.primitives.push_back(make_primitive(c, q, scale, albedo, sigma_multiplier));
    const View v = view_config();
    const bool persp = std::string(env_or("SG_PERSP", "0")) == "1";
    const float fov = static_cast<float>(std::atof(env_or("SG_FOV", "90")));
    const vp_camera cam = persp ? create_perspective(kSgWidth, kSgHeight, v.o, v.t, v.u, fov)
                                : create_orthographic(kSgWidth, kSgHeight, v.o, v.t, v.u, kSgOrthoHeight);
    s.cameras.push_back({cam, kSgWidth, kSgHeight});
    s.env_map_override = env_map_path();
    return true;
}

bool two_gaussian_validation(float sigma_multiplier, Scene& s, std::string&) {
    s = Scene{};
    s.name = "two_gaussian_validation";
    s.description = "Two isotropic absorber Gaussians at distinct positions/depths";
    float albedo[3];
    scatter_albedo3(albedo);
    const float scale[3] = {1, 1, 1}, q[4] = {1, 0, 0, 0};
    const float c0[3] = {-0.5f, 0.0f, -0.5f}, c1[3] = {0.5f, 0.0f, 0.5f};
    s.primitives.push_back(make_primitive(c0, q, scale, albedo, sigma_multiplier));
    s.primitives.push_back(make_primitive(c1, q, scale, albedo, sigma_multiplier));
    const float o[3] = {0, 0, -kSgCameraDistance}, t[3] = {0, 0, 0}, u[3] = {0, 1, 0};
    s.cameras.push_back({create_orthographic(kSgWidth, kSgHeight, o, t, u, kSgOrthoHeight), kSgWidth, kSgHeight});
    s.env_map_override = "assets/environment_maps/white_constant.hdr";
    return true;
}

bool cluster_validation(float, Scene& s, std::string& err) {
    s = Scene{};
    const std::string mode = env_or("SG_CLUSTER_MODE", "n5");
    s.name = "cluster_validation";
    s.description = "Procedural overlap cluster (mode=" + mode + ")";
    float albedo[3];
    scatter_albedo3(albedo);
    float ortho = 6.0f;
    const float id[4] = {1, 0, 0, 0}, one[3] = {1, 1, 1};
    if (mode == "n5") {
        ortho = 8.0f;
        const float P[5][3] = {{0, 0, 0}, {0.5f, 0, 0.3f}, {-0.5f, 0, -0.3f}, {0, 0.5f, 0.4f}, {0, -0.5f, -0.4f}};
        for (const auto& c : P) s.primitives.push_back(make_primitive(c, id, one, albedo, 2.0f));
    } else if (mode == "stress") {
        const int K = std::atoi(env_or("SG_STRESS_K", "80"));
        if (K <= 0) {
            err = "SG_STRESS_K must be positive";
            return false;
        }
        const float Z0 = 1.5f, M = 10.0f / static_cast<float>(K);
        for (int k = 0; k < K; ++k) {
            const float z = K == 1 ? 0.0f : -Z0 + 2.0f * Z0 * static_cast<float>(k) / static_cast<float>(K - 1);
            const float c[3] = {0, 0, z};
            s.primitives.push_back(make_primitive(c, id, one, albedo, M));
        }
    } else if (mode == "traits") {
        ortho = 4.0f;
        struct Spec {
            float c[3], s[3], deg, M;
        };
        const Spec specs[8] = {
            {{0.00f, 0.00f, 0.00f}, {0.50f, 0.30f, 0.40f}, 0.0f, 2.0f},
            {{0.40f, 0.10f, 0.20f}, {0.40f, 0.50f, 0.30f}, 30.0f, 1.5f},
            {{-0.35f, 0.20f, -0.20f}, {0.30f, 0.40f, 0.50f}, -45.0f, 2.5f},
            {{0.10f, -0.40f, 0.30f}, {0.50f, 0.35f, 0.30f}, 60.0f, 1.8f},
            {{-0.20f, -0.30f, -0.30f}, {0.45f, 0.30f, 0.40f}, -20.0f, 2.2f},
            {{0.30f, 0.35f, -0.25f}, {0.35f, 0.45f, 0.30f}, 15.0f, 1.6f},
            {{-0.40f, -0.10f, 0.35f}, {0.40f, 0.40f, 0.50f}, 80.0f, 2.0f},
            {{0.15f, 0.40f, 0.10f}, {0.30f, 0.50f, 0.35f}, -60.0f, 1.9f},
        };
        for (const auto& sp : specs) {
            float q[4];
            zrot_quat(sp.deg, q);
            s.primitives.push_back(make_primitive(sp.c, q, sp.s, albedo, sp.M));
        }
    } else {
        err = "unknown SG_CLUSTER_MODE: " + mode;
        return false;
    }
    const float o[3] = {0, 0, -6.0f}, t[3] = {0, 0, 0}, u[3] = {0, 1, 0};
    s.cameras.push_back({create_orthographic(kSgWidth, kSgHeight, o, t, u, ortho), kSgWidth, kSgHeight});
    s.env_map_override = "assets/environment_maps/white_constant.hdr";
    return true;
}

bool asset_validation(float sigma_multiplier, Scene& s, std::string& err) {
    s = Scene{};
    const std::string ply = env_or("SG_PLY", "assets/models/bunny/bunny_pyr0.ply");
    const float alb = static_cast<float>(std::atof(env_or("SG_ALBEDO", "0")));
    const float over[3] = {alb > 0.0f ? alb : -1.0f, alb > 0.0f ? alb : -1.0f, alb > 0.0f ? alb : -1.0f};
    s.name = "asset_validation";
    s.description = "Generic Gaussian asset: " + ply;
    vp_primitive* prims = nullptr;
    uint32_t n = 0;
    if (vp_io_load_ply(ply.c_str(), sigma_multiplier, over, &prims, &n) != 0) {
        err = "Failed to load asset '" + ply + "': " + vp_io_last_error();
        return false;
    }
    s.primitives.assign(prims, prims + n);
    vp_io_free(prims);
    const float dist = static_cast<float>(std::atof(env_or("SG_DIST", "3.5")));
    const float fov = static_cast<float>(std::atof(env_or("SG_FOV", "40")));
    const uint32_t res = static_cast<uint32_t>(std::atoi(env_or("SG_RES", "512")));
    const std::string view = env_or("SG_VIEW", "negz");
    float o[3] = {0, 0, -dist}, up[3] = {0, 1, 0};
    if (view == "posz") o[0] = 0, o[1] = 0, o[2] = dist;
    else if (view == "posx") o[0] = dist, o[1] = 0, o[2] = 0;
    else if (view == "negx") o[0] = -dist, o[1] = 0, o[2] = 0;
    else if (view == "posy") o[0] = 0, o[1] = dist, o[2] = 0, up[1] = 0, up[2] = 1;
    else if (view == "negy") o[0] = 0, o[1] = -dist, o[2] = 0, up[1] = 0, up[2] = 1;
    else if (view == "diag") o[0] = dist * 0.6f, o[1] = dist * 0.5f, o[2] = -dist * 0.6f;
    const float t[3] = {0, 0, 0};
    s.cameras.push_back({create_perspective(res, res, o, t, up, fov), res, res});
    const std::string sel = env_or("SG_ENV", "");
    s.env_map_override = sel == "meadow"   ? "assets/environment_maps/meadow_2_4k.hdr"
                         : sel == "studio" ? "assets/environment_maps/ferndale_studio_01_4k.hdr"
                         : sel == "sky"    ? "assets/environment_maps/sky.hdr"
                                           : "assets/environment_maps/white_constant.hdr";
    return true;
}

}  // namespace vpapp
