"""The C++ facade is a drop-in for the reference's app::Config / Config::parse
(include/thesis/host/app/config.h:14-50, src/thesis/host/app/config.cpp:11-69):
a caller written against the reference's field names and types compiles
against cuda-volprim_amd/host/app.h, the defaults are the reference's, and the
command line parses the reference's options with its range checks.  CPU only
(the caller never constructs a Renderer)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "cuda-volprim_amd", "host")

CALLER = r'''
#include <cstdio>
#include <string>
#include <type_traits>
#include "app.h"

namespace app = vpapp;  // the reference's namespace alias in a caller: thesis::host::app

int main(int argc, char* argv[]) {
    app::Config renderer_config;
    // field types of config.h:14-50
    static_assert(std::is_same<decltype(renderer_config.image_width_), size_t>::value, "image_width_");
    static_assert(std::is_same<decltype(renderer_config.image_height_), size_t>::value, "image_height_");
    static_assert(std::is_same<decltype(renderer_config.num_samples_per_pixel_), size_t>::value, "spp");
    static_assert(std::is_same<decltype(renderer_config.max_bounces_), size_t>::value, "max_bounces_");
    static_assert(std::is_same<decltype(renderer_config.rr_depth_), size_t>::value, "rr_depth_");
    static_assert(std::is_same<decltype(renderer_config.seed_), unsigned int>::value, "seed_");
    static_assert(std::is_same<decltype(renderer_config.output_path_), std::filesystem::path>::value, "output");
    static_assert(std::is_same<decltype(renderer_config.env_map_path_), std::filesystem::path>::value, "env");
    static_assert(std::is_same<decltype(renderer_config.hg_g_), float>::value, "hg_g_");
    static_assert(std::is_same<decltype(renderer_config.use_ris_), bool>::value, "use_ris_");
    static_assert(std::is_same<decltype(renderer_config.ris_num_candidates_), uint32_t>::value, "ris");
    // defaults
    std::printf("defaults %zu %zu %zu %u %d %zu %zu %.2f %.1f %.1f %.2f %d %u %d %s %s %s %s %s %s\n",
                renderer_config.image_width_, renderer_config.image_height_, renderer_config.num_samples_per_pixel_,
                renderer_config.seed_, (int)renderer_config.denoise_, renderer_config.max_bounces_,
                renderer_config.rr_depth_, renderer_config.rr_max_survival_, renderer_config.firefly_clamp_luminance_,
                renderer_config.pixel_filter_stddev_, renderer_config.hg_g_, (int)renderer_config.use_ris_,
                renderer_config.ris_num_candidates_, (int)renderer_config.measure_caps_,
                renderer_config.output_path_.string().c_str(), renderer_config.env_map_path_.string().c_str(),
                renderer_config.raygen_function_name_.c_str(), renderer_config.miss_function_name_.c_str(),
                renderer_config.anyhit_function_name_.c_str(), renderer_config.launch_params_variable_name_.c_str());
    // the test_runner's way of filling it (test/test_runner.cpp:188-191)
    renderer_config.image_width_ = 64;
    renderer_config.output_path_ = std::filesystem::path("out") / "0000.exr";
    auto res = app::Config::parse(argc, argv);
    if (!res) {
        std::printf("error %d %s\n", res.error().code_, res.error().msg_.c_str());
        return 3;
    }
    const app::Config& c = *res;
    std::printf("parsed %zu %zu %zu %u %d %zu %zu %.2f %.1f %.1f %.2f %d %u %d %s %s\n", c.image_width_,
                c.image_height_, c.num_samples_per_pixel_, c.seed_, (int)c.denoise_, c.max_bounces_, c.rr_depth_,
                c.rr_max_survival_, c.firefly_clamp_luminance_, c.pixel_filter_stddev_, c.hg_g_, (int)c.use_ris_,
                c.ris_num_candidates_, (int)c.measure_caps_, c.output_path_.string().c_str(),
                c.env_map_path_.string().c_str());
    return 0;
}
'''


@pytest.fixture(scope="module")
def caller(tmp_path_factory):
    d = tmp_path_factory.mktemp("facade")
    src = d / "caller.cpp"
    src.write_text(CALLER)
    exe = d / "caller"
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-I", HOST, str(src), os.path.join(HOST, "config.cpp"),
                    "-o", str(exe)], check=True)
    return str(exe)


def run(exe, *args):
    p = subprocess.run([exe, *args], capture_output=True, text=True)
    return p.returncode, p.stdout.splitlines()


def test_reference_defaults(caller):
    rc, out = run(caller)
    assert rc == 0
    # config.h:15-45: spp 1, 1000x750, seed 42, no denoise, 128 bounces, RR 12,
    # survival 0.99, no clamp, box filter, g 0.85, MIS, 6 candidates, no
    # measuring run, output.exr, the meadow map, the OptiX entry names
    assert out[0] == ("defaults 1000 750 1 42 0 128 12 0.99 0.0 0.0 0.85 0 6 0 output.exr "
                      "assets/environment_maps/meadow_2_4k.hdr __raygen__rg __miss__ms __anyhit__ah launch_params")


def test_parse_reference_options(caller):
    rc, out = run(caller, "--output", "a/b.exr", "--height", "20", "--width=30", "--samples_per_pixel", "7",
                  "--seed", "9", "--denoise", "--max-depth", "40", "--rr-depth", "3", "--rr-max-survival", "0.5",
                  "--firefly-clamp", "4", "--filter-width", "0.5", "--phase-g", "-0.3", "--measure-caps",
                  "--env_map", "x.hdr", "--raygen", "__raygen__x", "--module_blob", "m.optixir", "--ris",
                  "--ris-candidates", "4")
    assert rc == 0, out
    assert out[1] == "parsed 30 20 7 9 1 40 3 0.50 4.0 0.5 -0.30 1 4 1 a/b.exr x.hdr"


@pytest.mark.parametrize("args,msg", [
    (["--max-depth", "0"], "must be positive"),            # CLI::PositiveNumber
    (["--rr-max-survival", "1.5"], "not in [0, 1]"),       # CLI::Range(0, 1)
    (["--phase-g", "1"], "not in [-0.999, 0.999]"),        # CLI::Range(-0.999, 0.999)
    (["--firefly-clamp", "-1"], "must be non-negative"),   # CLI::NonNegativeNumber
    (["--width", "abc"], "not a non-negative integer"),
    (["--height"], "requires a value"),
    (["--bogus"], "not expected"),
])
def test_parse_errors_are_results(caller, args, msg):
    rc, out = run(caller, *args)
    assert rc == 3 and out[-1].startswith("error 1 ") and msg in out[-1], out
