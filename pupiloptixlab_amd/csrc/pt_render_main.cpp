// pt_render: headless counterpart of example/path_tracer/main.cpp:5-21.
//
//   pt_render --scene file.xml | --cornell | --terrain [--seed N]
//             [--spp N] [--max-depth N] [--denoise] [--out image.exr]
//
// Builds the system, adds the path-tracing pass (and, with --denoise, the
// denoising pass after it, as main.cpp adds DenoisePass), loads the scene,
// runs one frame per sample and writes the final image as EXR.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "pt_system.h"

int main(int argc, char **argv) {
    std::string xml, out = "pt_render.exr";
    int kind = PT_SCENE_CORNELL;
    uint32_t spp = 16, seed = 0;
    int max_depth = 0;
    bool denoise = false;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto next = [&](const char *what) -> const char * {
            if (i + 1 >= argc) {
                std::fprintf(stderr, "missing value for %s\n", what);
                std::exit(2);
            }
            return argv[++i];
        };
        if (a == "--scene") xml = next("--scene");
        else if (a == "--cornell") kind = PT_SCENE_CORNELL;
        else if (a == "--terrain") kind = PT_SCENE_TERRAIN;
        else if (a == "--door-room") kind = PT_SCENE_DOOR_ROOM;
        else if (a == "--spp") spp = (uint32_t)std::atoi(next("--spp"));
        else if (a == "--seed") seed = (uint32_t)std::atoi(next("--seed"));
        else if (a == "--max-depth") max_depth = std::atoi(next("--max-depth"));
        else if (a == "--out") out = next("--out");
        else if (a == "--denoise") denoise = true;
        else {
            std::fprintf(stderr, "unknown argument %s\n", a.c_str());
            return 2;
        }
    }
    Pupil::System system;
    if (!system.Init(0)) {
        std::fprintf(stderr, "%s\n", system.LastError().c_str());
        return 1;
    }
    Pupil::pt::PTPass pass(&system);
    system.AddPass(&pass);
    Pupil::DenoisePass dn_pass(&system, Pupil::DenoisePass::Config{});
    if (denoise) system.AddPass(&dn_pass);
    bool ok;
    if (!xml.empty()) {
        ok = system.SetScene(xml);
    } else {
        pt_gen_params gp{};
        gp.seed = seed;
        ok = system.SetScene(kind, gp);
    }
    if (!ok) {
        std::fprintf(stderr, "%s\n", system.LastError().c_str());
        system.Destroy();
        return 1;
    }
    if (max_depth > 0) pass.SetMaxDepth(max_depth);
    system.Run(spp);
    const pt_scene_desc *d = system.Scene();
    std::vector<float> img((size_t)d->width * d->height * 4);
    if (!pass.ReadBuffer(PT_BUF_FRAME, img.data(), img.size() * sizeof(float)) ||
        pt_write_exr(out.c_str(), img.data(), d->width, d->height) != PT_OK) {
        std::fprintf(stderr, "output failed: %s\n", pass.LastError().c_str());
        system.Destroy();
        return 1;
    }
    std::printf("%s\n", pass.Inspector().c_str());
    if (denoise) std::printf("%s\n", dn_pass.Inspector().c_str());
    std::printf("wrote %s (%ux%u, %u spp)\n", out.c_str(), d->width, d->height, spp);
    system.Destroy();
    return 0;
}
