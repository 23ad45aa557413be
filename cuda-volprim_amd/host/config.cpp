// config.cpp — vpapp::Config::parse, the reference's command line of the
// renderer (src/thesis/host/app/config.cpp:11-69) without CLI11: the same
// option names, value checks and error reporting through a Result.
#include <cerrno>
#include <cstdlib>
#include <cstring>
#include <string>

#include "app.h"

namespace vpapp {

namespace {

const char* kUsage =
    "Gaussian-mixture volumetric path tracer (MI355X)\n"
    "Image settings:    --output PATH --height N --width N --samples_per_pixel N\n"
    "Entrypoint names:  --raygen S --miss S --anyhit S --launch_params S   (accepted, unused)\n"
    "File paths:        --module_blob PATH (accepted, unused) --env_map PATH\n"
    "Runtime tweaks:    --seed N --denoise\n"
    "Render parameters: --max-depth N --rr-depth N --rr-max-survival X --firefly-clamp X --filter-width X\n"
    "                   --phase-g X --measure-caps --ris --ris-candidates N --active-cap N --shadow-rr X\n"
    "                   --adaptive-threshold X --adaptive-min-samples N\n";

bool to_size(const std::string& v, size_t& out) {
    if (v.empty() || v[0] == '-') return false;
    char* end = nullptr;
    errno = 0;
    const unsigned long long x = std::strtoull(v.c_str(), &end, 10);
    if (errno != 0 || *end != '\0') return false;
    out = static_cast<size_t>(x);
    return true;
}

bool to_float(const std::string& v, float& out) {
    if (v.empty()) return false;
    char* end = nullptr;
    errno = 0;
    const float x = std::strtof(v.c_str(), &end);
    if (errno != 0 || *end != '\0') return false;
    out = x;
    return true;
}

}  // namespace

utils::Result<Config> Config::parse(int argc, char* argv[]) noexcept {
    Config c;
    try {
        for (int i = 1; i < argc; ++i) {
            std::string a = argv[i];
            std::string val;
            const size_t eq = a.find('=');
            bool has_inline = false;
            if (a.rfind("--", 0) == 0 && eq != std::string::npos) {
                val = a.substr(eq + 1);
                a = a.substr(0, eq);
                has_inline = true;
            }
            if (a == "-h" || a == "--help") return utils::make_error(0, kUsage);
            // flags
            if (a == "--denoise" || a == "--measure-caps" || a == "--ris") {
                if (has_inline) return utils::make_error(1, "Argument parse error: " + a + " takes no value");
                (a == "--denoise" ? c.denoise_ : a == "--ris" ? c.use_ris_ : c.measure_caps_) = true;
                continue;
            }
            static const char* const kValued[] = {
                "--output", "--env_map", "--module_blob", "--raygen", "--miss", "--anyhit", "--launch_params",
                "--height", "--width", "--samples_per_pixel", "--seed", "--max-depth", "--rr-depth",
                "--ris-candidates", "--active-cap", "--rr-max-survival", "--firefly-clamp", "--filter-width",
                "--phase-g", "--shadow-rr", "--adaptive-threshold", "--adaptive-min-samples"};
            bool known = false;
            for (const char* k : kValued) known = known || a == k;
            if (!known)
                return utils::make_error(1, "Argument parse error: the following argument was not expected: " + a);
            if (!has_inline) {
                if (i + 1 >= argc) return utils::make_error(1, "Argument parse error: " + a + " requires a value");
                val = argv[++i];
            }
            auto bad = [&](const char* why) {
                return utils::make_error(1, "Argument parse error: " + a + " " + val + ": " + why);
            };
            size_t n = 0;
            float f = 0.0f;
            if (a == "--output") c.output_path_ = val;
            else if (a == "--env_map") c.env_map_path_ = val;
            else if (a == "--module_blob") c.module_blob_path_ = val;
            else if (a == "--raygen") c.raygen_function_name_ = val;
            else if (a == "--miss") c.miss_function_name_ = val;
            else if (a == "--anyhit") c.anyhit_function_name_ = val;
            else if (a == "--launch_params") c.launch_params_variable_name_ = val;
            else if (a == "--height" || a == "--width" || a == "--samples_per_pixel" || a == "--seed" ||
                     a == "--max-depth" || a == "--rr-depth" || a == "--ris-candidates" || a == "--active-cap" ||
                     a == "--adaptive-min-samples") {
                if (!to_size(val, n)) return bad("not a non-negative integer");
                if (a == "--height") c.image_height_ = n;
                else if (a == "--width") c.image_width_ = n;
                else if (a == "--samples_per_pixel") c.num_samples_per_pixel_ = n;
                else if (a == "--seed") c.seed_ = static_cast<uint>(n);
                else if (a == "--max-depth") {
                    if (n == 0) return bad("must be positive");  // CLI::PositiveNumber
                    c.max_bounces_ = n;
                } else if (a == "--rr-depth") c.rr_depth_ = n;
                else if (a == "--ris-candidates") c.ris_num_candidates_ = static_cast<uint32_t>(n);
                else if (a == "--adaptive-min-samples") c.adaptive_min_samples_ = static_cast<uint32_t>(n);
                else c.active_cap_ = static_cast<uint32_t>(n);
            } else if (a == "--rr-max-survival" || a == "--firefly-clamp" || a == "--filter-width" ||
                       a == "--phase-g" || a == "--shadow-rr" || a == "--adaptive-threshold") {
                if (!to_float(val, f)) return bad("not a number");
                if (a == "--rr-max-survival") {
                    if (!(f >= 0.0f && f <= 1.0f)) return bad("not in [0, 1]");  // CLI::Range(0, 1)
                    c.rr_max_survival_ = f;
                } else if (a == "--firefly-clamp") {
                    if (!(f >= 0.0f)) return bad("must be non-negative");  // CLI::NonNegativeNumber
                    c.firefly_clamp_luminance_ = f;
                } else if (a == "--filter-width") {
                    if (!(f >= 0.0f)) return bad("must be non-negative");
                    c.pixel_filter_stddev_ = f;
                } else if (a == "--shadow-rr") {
                    if (!(f >= 0.0f)) return bad("must be non-negative");
                    c.shadow_rr_ = f;
                } else if (a == "--adaptive-threshold") {
                    if (!(f >= 0.0f)) return bad("must be non-negative");
                    c.adaptive_threshold_ = f;
                } else {
                    if (!(f >= -0.999f && f <= 0.999f)) return bad("not in [-0.999, 0.999]");  // CLI::Range
                    c.hg_g_ = f;
                }
            } else {
                return utils::make_error(1, "Argument parse error: the following argument was not expected: " + a);
            }
        }
    } catch (const std::exception& e) {
        return utils::make_error(1, std::string("Argument parsing failed: ") + e.what());
    } catch (...) {
        return utils::make_error(1, "Argument parsing failed: unknown error");
    }
    return c;
}

}  // namespace vpapp
