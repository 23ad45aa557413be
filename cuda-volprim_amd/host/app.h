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

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <optional>
#include <string>
#include <utility>
#include <vector>

#include "../../include/vp_api.h"
#include "../../include/vp_io.h"

namespace vpapp {

namespace fs = std::filesystem;
using uint = unsigned int;

namespace utils {
// Error and Result of the reference's host utilities
// (include/thesis/host/utils/result.h:48-121): a value or an (code, message)
// error, with try_unwrap_or_exit logging the error and exiting with its code.
struct Error {
    int code_;
    std::string msg_;
};
template <typename T>
class Result {
   public:
    Result(T v) : v_(std::move(v)) {}
    Result(Error e) : e_(std::move(e)) {}
    explicit operator bool() const { return v_.has_value(); }
    bool has_value() const { return v_.has_value(); }
    T& operator*() { return *v_; }
    const T& operator*() const { return *v_; }
    T* operator->() { return &*v_; }
    const Error& error() const { return *e_; }

   private:
    std::optional<T> v_;
    std::optional<Error> e_;
};
inline Error make_error(int code, std::string msg) { return Error{code, std::move(msg)}; }
// Logs "Fatal: <msg> (code <code>)" and exits with the error's code.
template <typename T>
T try_unwrap_or_exit(Result<T>&& res) {
    if (!res) {
        std::fprintf(stderr, "[error] Fatal: %s (code %d)\n", res.error().msg_.c_str(), res.error().code_);
        std::exit(res.error().code_);
    }
    return std::move(*res);
}
}  // namespace utils

// app::Config (include/thesis/host/app/config.h:14-50): the same field names,
// types and defaults, so a caller that fills the reference's Config compiles
// against this one.  The OptiX program and module names have no meaning here
// (there is no SBT and no module); they are kept for source compatibility
// and ignored.  active_cap_, shadow_rr_ and the adaptive_ fields are this
// build's: the reference fixes its active-set capacity and adaptive sampling
// at compile time (device/core/constants.cuh:45-48,196-216) and has no
// shadow-ray roulette.
struct Config {
    std::string raygen_function_name_ = "__raygen__rg";
    std::string miss_function_name_ = "__miss__ms";
    std::string anyhit_function_name_ = "__anyhit__ah";
    std::string launch_params_variable_name_ = "launch_params";

    fs::path output_path_ = "output.exr";
    fs::path module_blob_path_ = fs::path("build") / "device_program.optixir";
    fs::path env_map_path_ = fs::path("assets") / "environment_maps" / "meadow_2_4k.hdr";

    size_t num_samples_per_pixel_ = 1;
    size_t image_width_ = 1000;
    size_t image_height_ = 750;

    uint seed_ = 42;
    bool denoise_ = false;

    size_t max_bounces_ = 128;
    size_t rr_depth_ = 12;
    float rr_max_survival_ = 0.99f;
    float firefly_clamp_luminance_ = 0.0f;
    float pixel_filter_stddev_ = 0.0f;
    float hg_g_ = 0.85f;
    bool use_ris_ = false;
    uint32_t ris_num_candidates_ = 6;
    bool measure_caps_ = false;

    uint32_t active_cap_ = 128;
    // NEE shadow-ray roulette threshold (0 = off, the reference's estimator);
    // this build's option, --shadow-rr
    float shadow_rr_ = 0.0f;
    // adaptive sampling (the reference's compile-time ENABLE_ADAPTIVE_SAMPLING,
    // ADAPTIVE_THRESHOLD and ADAPTIVE_MIN_SAMPLES, device/core/constants.cuh:
    // 196-216) as options: --adaptive-threshold (0 = off) and
    // --adaptive-min-samples
    float adaptive_threshold_ = 0.0f;
    uint32_t adaptive_min_samples_ = 32;

    // Config::parse (src/thesis/host/app/config.cpp:11-69): the same options
    // (--output --height --width --samples_per_pixel --raygen --miss --anyhit
    // --launch_params --module_blob --env_map --seed --denoise --max-depth
    // --rr-depth --rr-max-survival --firefly-clamp --filter-width --phase-g
    // --measure-caps, plus --ris --ris-candidates --active-cap --shadow-rr
    // --adaptive-threshold --adaptive-min-samples), the same
    // range checks, an error Result on a bad argument, never an exception.
    [[nodiscard]] static utils::Result<Config> parse(int argc, char* argv[]) noexcept;
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
