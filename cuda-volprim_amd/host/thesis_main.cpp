// thesis_main.cpp — the renderer's own executable (src/thesis/host/main.cpp:21-40):
// parse the Config from the command line, render the default scene (one
// purple Gaussian at the origin: scale 0.5, albedo (0.5, 0, 0.5), mass 0.5,
// identity rotation) with the default camera, write the EXR.
#include <utility>
#include <vector>

#include "app.h"

int main(int argc, char* argv[]) {
    auto config = vpapp::utils::try_unwrap_or_exit(vpapp::Config::parse(argc, argv));
    const float c[3] = {0.0f, 0.0f, 0.0f}, q[4] = {1.0f, 0.0f, 0.0f, 0.0f}, s[3] = {0.5f, 0.5f, 0.5f},
                a[3] = {0.5f, 0.0f, 0.5f};
    std::vector<vp_primitive> scene{vpapp::make_primitive(c, q, s, a, 0.5f)};
    vpapp::Renderer renderer(config, std::move(scene));
    renderer.render();
    return 0;
}
