// Command-line renderer: the reference's CLI (main.cu:10-97) over the C ABI.
//
//   cavalry -i scene.pbrt [-o out.png] [--spp N] [--integrator path|rlpath]
//           [--bvh-opt-rounds R] [--pfm out.pfm] [--device D]
//
// Prints "Preparation took Xms" and "Rendering took Xms" as the reference's
// Timer does (Utils/Timer.h:98-105), so scripts that scrape those lines keep
// working. Unrecognised options are ignored, as with cxxopts'
// allow_unrecognised_options.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/cavalry.hpp"

namespace {

void usage() {
    std::printf(
        "Cavalry (MI355X) - A GPU Ray Tracer\n"
        "Usage:\n"
        "  cavalry [OPTION...]\n\n"
        "  -i, --input arg           Input file name\n"
        "  -o, --output arg          Override output file name\n"
        "      --spp arg             Override samples per pixel\n"
        "      --integrator arg      Override integrator\n"
        "      --bvh-opt-rounds arg  BVH Optmization Rounds\n"
        "      --pfm arg             Also write the float film as PFM\n"
        "      --device arg          HIP device (default 0)\n"
        "  -h, --help                Print usage\n");
}

double ms_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

int main(int argc, char** argv) {
    std::string input, output, integrator, pfm;
    int spp = 0, rounds = -1, device = 0;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto value = [&](const char* name) -> std::string {
            const std::string pre = std::string(name) + "=";
            if (a.compare(0, pre.size(), pre) == 0) return a.substr(pre.size());
            if (i + 1 >= argc) {
                std::fprintf(stderr, "option %s needs a value\n", name);
                std::exit(2);
            }
            return argv[++i];
        };
        auto is = [&](const char* name) { return a == name || a.compare(0, std::strlen(name) + 1, std::string(name) + "=") == 0; };
        if (a == "-h" || a == "--help") { usage(); return 0; }
        else if (a == "-i" || is("--input")) input = value("--input");
        else if (a == "-o" || is("--output")) output = value("--output");
        else if (is("--spp")) spp = std::atoi(value("--spp").c_str());
        else if (is("--integrator")) integrator = value("--integrator");
        else if (is("--bvh-opt-rounds")) rounds = std::atoi(value("--bvh-opt-rounds").c_str());
        else if (is("--pfm")) pfm = value("--pfm");
        else if (is("--device")) device = std::atoi(value("--device").c_str());
    }
    if (input.empty()) {
        std::fprintf(stderr, "No Input File Found (use --input scene.pbrt)\n");
        return 2;
    }
    try {
        cav::Context ctx(device);
        cav_pbrt_overrides ov{};
        ov.output = output.c_str();
        ov.integrator = integrator.c_str();
        ov.spp = spp;
        ov.bvh_opt_rounds = rounds;
        const auto t0 = std::chrono::steady_clock::now();
        cav::RenderSetup setup = cav::readRenderSetup(ctx, input, ov);
        const double prep = ms_since(t0);
        const auto t1 = std::chrono::steady_clock::now();
        cav::Film& film = setup.renderer.render(*setup.scene);
        const double total = prep + ms_since(t1);
        film.saveToPNG(setup.setup.output);
        if (!pfm.empty()) film.savePFM(pfm);
        const cav_render_stats& st = setup.renderer.integrator->stats;
        std::printf("%d x %d, %d spp, %s, max depth %d, %lld triangles, %d lights, %d BVH rounds\n",
                    setup.setup.width, setup.setup.height, setup.setup.spp,
                    setup.setup.integrator == CAV_INTEGRATOR_RLPATH   ? "rlpath"
                    : setup.setup.integrator == CAV_INTEGRATOR_DIRECT ? "directlighting"
                                                                      : "path",
                    setup.setup.max_depth,
                    (long long)setup.setup.ntris, setup.setup.nlights, setup.setup.bvh_opt_rounds);
        if (setup.setup.warnings) std::printf("%d unsupported feature(s) replaced or ignored\n", setup.setup.warnings);
        std::printf("device time %.3f ms, %.2f Mrays/s (%lld closest + %lld shadow rays)\n", st.ms_total,
                    st.mrays_per_s, (long long)st.closest_rays, (long long)st.shadow_rays);
        std::printf("Preparation took %.0fms\n", prep);
        std::printf("Rendering took %.0fms\n", total);
        std::printf("wrote %s\n", setup.setup.output);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
