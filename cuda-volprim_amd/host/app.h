// app.h — C++ host facade of the volumetric path tracer, implemented over the
// C ABI (include/vp_api.h, include/vp_io.h).
//
// Mirrors the reference's host surface (kfernandez31/cuda-volprim):
//   vpapp::Config          app::Config (include/thesis/host/app/config.h:14-50)
//   vpapp::Renderer        app::Renderer(const Config&, std::vector<Primitive>&&,
//                          std::optional<Camera>) + render()
//                          (include/thesis/host/app/renderer.h:77-81)
//   vpapp::Scene + factories
//                          test::scenes::MultiViewTestScene and the validation
//                          factories (test/scenes/single_gaussian.cpp:95-313,
//                          test/scenes/cloud_validation.cpp:94-139)
// Host errors are reported the reference's way: a log line and exit(1) from
// the facade; the C ABI underneath only returns codes.
#pragma once

#include <cstdint>
#include <optional>
#include <string>
#include <vector>

#include "../../include/vp_api.h"
#include "../../include/vp_io.h"

namespace vpapp {

struct Config {
    uint32_t image_width = 1920;
    uint32_t image_height = 1080;
    uint32_t num_samples_per_pixel = 64;
    uint32_t seed = 42;
    std::string output_path = "output.exr";
    std::string env_map_path = "assets/environment_maps/white_constant.hdr";
    bool denoise = false;
    uint32_t max_bounces = 128;
    uint32_t rr_depth = 12;
    float rr_max_survival = 0.99f;
    float firefly_clamp_luminance = 0.0f;
    float pixel_filter_stddev = 0.0f;
    float hg_g = 0.85f;
    bool use_ris = false;
    uint32_t ris_num_candidates = 6;
    bool measure_caps = false;
    uint32_t active_cap = 128;
};

struct ViewCamera {
    vp_camera camera;
    uint32_t width, height;
};

// Camera builders of host::params::Camera (include/thesis/host/params/camera.h:127-150).
vp_camera create_perspective(uint32_t w, uint32_t h, const float eye[3], const float target[3], const float up[3],
                             float vfov_deg);
vp_camera create_orthographic(uint32_t w, uint32_t h, const float eye[3], const float target[3], const float up[3],
                              float ortho_height);

struct Scene {
    std::string name;
    std::string description;
    std::vector<vp_primitive> primitives;
    std::vector<ViewCamera> cameras;
    std::optional<std::string> env_map_override;
};

// Validation ladder scenes (the SG_* environment variables of the reference apply).
bool single_gaussian_validation(float sigma_multiplier, Scene& out, std::string& err);
bool two_gaussian_validation(float sigma_multiplier, Scene& out, std::string& err);
bool cluster_validation(float sigma_multiplier, Scene& out, std::string& err);
bool asset_validation(float sigma_multiplier, Scene& out, std::string& err);
// Names the dispatcher knows.
std::vector<std::string> scene_names();

// Environment: a Radiance .hdr file, or the built-in constant white map when
// the path names white_constant.hdr and the file is absent.
bool load_environment(const std::string& path, std::vector<float>& rgba, uint32_t& w, uint32_t& h, std::string& err);

vp_primitive make_primitive(const float center[3], const float quat_wxyz[4], const float scale[3],
                            const float albedo[3], float mass);

class Renderer {
   public:
    Renderer(const Config& config, std::vector<vp_primitive>&& primitives,
             std::optional<vp_camera> camera = std::nullopt);
    ~Renderer();
    Renderer(const Renderer&) = delete;
    Renderer& operator=(const Renderer&) = delete;
    // Render all samples, report overflow / measured caps, write the EXR.
    void render();
    const std::vector<float>& image() const { return image_; }
    vp_render_stats stats() const { return stats_; }

   private:
    Config config_;
    vp_camera camera_{};
    vp_context* ctx_ = nullptr;
    vp_scene* scene_ = nullptr;
    vp_env* env_ = nullptr;
    uint32_t num_primitives_ = 0;
    std::vector<float> image_;
    vp_render_stats stats_{};
};

}  // namespace vpapp
