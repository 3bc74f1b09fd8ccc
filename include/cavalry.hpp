/* C++ host classes over the C ABI of cavalry.h, with the shape of the
 * reference's host interfaces:
 *   Scene       ~ Scene + Scene::prepareForRender   (Core/Scene.h:60-95)
 *   Film        ~ Film / RenderResult               (Core/Film.h, Core/RenderResult.h)
 *   Integrator  ~ Integrator::render                (Core/Integrator.h:12-21), with
 *                 PathTracingIntegrator / RLPTIntegrator and the factory
 *                 CreateIntegrator::createFromObjectDefinition (Integrators/CreateIntegrator.cu:10-25)
 *   Renderer    ~ Renderer::render                  (Core/Renderer.h:12-19)
 *   readRenderSetup ~ SceneLoading::readRenderSetup (SceneLoading/SceneLoading.cu)
 * Errors throw cav::Error (the reference calls exit()). Header only. */
#pragma once
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "cavalry.h"

namespace cav {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void check(cav_ctx* ctx, int rc, const char* what) {
    if (rc != CAV_OK) {
        const char* m = ctx ? cav_last_error(ctx) : nullptr;
        throw Error(rc, std::string(what) + ": " + (m && m[0] ? m : "error " + std::to_string(rc)));
    }
}

class Context {
public:
    explicit Context(int device = 0) { check(nullptr, cav_ctx_create(device, &ctx_), "cav_ctx_create"); }
    ~Context() { if (ctx_) cav_ctx_destroy(ctx_); }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    cav_ctx* get() const { return ctx_; }

private:
    cav_ctx* ctx_ = nullptr;
};

class Film {
public:
    Film(int width, int height) : width(width), height(height), rgb((size_t)width * height * 3), weight((size_t)width * height) {}
    int width, height;
    std::vector<float> rgb, weight;  // per-pixel sample sums and filter weights

    // Film::readCurrentResult: 8-bit sRGB
    std::vector<uint8_t> readCurrentResult() const {
        std::vector<uint8_t> out((size_t)width * height * 3);
        check(nullptr, cav_film_to_srgb8(rgb.data(), weight.data(), width, height, out.data()), "cav_film_to_srgb8");
        return out;
    }
    // RenderResult::saveToPNG
    void saveToPNG(const std::string& path) const {
        const std::vector<uint8_t> img = readCurrentResult();
        check(nullptr, cav_write_png(path.c_str(), img.data(), width, height), "cav_write_png");
    }
    void savePFM(const std::string& path) const {
        check(nullptr, cav_write_pfm(path.c_str(), rgb.data(), weight.data(), width, height), "cav_write_pfm");
    }
};

class Scene {
public:
    Scene(Context& ctx, const float* tris, int ntris, const cav_mesh* meshes, int nmeshes, int bvhOptimizationRounds,
          int mortonBits = 63)
        : ctx_(ctx.get()) {
        check(ctx_, cav_scene_create(ctx_, tris, ntris, meshes, nmeshes, nullptr, bvhOptimizationRounds, mortonBits, &s_),
              "cav_scene_create");
    }
    Scene(Context& ctx, const std::string& pbrtPath, const cav_pbrt_overrides& ov, cav_render_setup& setup)
        : ctx_(ctx.get()) {
        check(ctx_, cav_scene_create_from_pbrt(ctx_, pbrtPath.c_str(), &ov, &s_, &setup), "cav_scene_create_from_pbrt");
    }
    ~Scene() { if (s_) cav_scene_destroy(s_); }
    Scene(const Scene&) = delete;
    Scene& operator=(const Scene&) = delete;
    cav_scene* get() const { return s_; }
    cav_ctx* ctx() const { return ctx_; }

private:
    cav_ctx* ctx_ = nullptr;
    cav_scene* s_ = nullptr;
};

class Integrator {
public:
    virtual ~Integrator() = default;
    int spp = 1;  // the sampler's samples per pixel
    virtual void render(const Scene& scene, const cav_camera& camera, Film& film) = 0;
    cav_render_stats stats{};

protected:
    void run(const Scene& scene, const cav_camera& camera, Film& film, cav_render_params p) {
        check(scene.ctx(), cav_render(scene.ctx(), scene.get(), &camera, &p, film.rgb.data(), film.weight.data(), &stats),
              "cav_render");
    }
};

class PathTracingIntegrator : public Integrator {
public:
    explicit PathTracingIntegrator(int maxDepth) : maxDepth(maxDepth) {}
    int maxDepth;
    void render(const Scene& scene, const cav_camera& camera, Film& film) override {
        cav_render_params p{};
        p.integrator = CAV_INTEGRATOR_PATH;
        p.spp = spp;
        p.max_depth = maxDepth;
        run(scene, camera, film, p);
    }
};

class RLPTIntegrator : public Integrator {
public:
    explicit RLPTIntegrator(int maxDepth, int qGrid = 32, float alphaScale = 1.f)
        : maxDepth(maxDepth), qGrid(qGrid), alphaScale(alphaScale) {}
    int maxDepth, qGrid;
    float alphaScale;
    void render(const Scene& scene, const cav_camera& camera, Film& film) override {
        cav_render_params p{};
        p.integrator = CAV_INTEGRATOR_RLPATH;
        p.spp = spp;
        p.max_depth = maxDepth;
        p.q_grid = qGrid;
        p.alpha_scale = alphaScale;
        run(scene, camera, film, p);
    }
};

class DirectLightingIntegrator : public Integrator {
public:
    void render(const Scene& scene, const cav_camera& camera, Film& film) override {
        cav_render_params p{};
        p.integrator = CAV_INTEGRATOR_DIRECT;
        p.spp = spp;
        p.max_depth = 1;
        run(scene, camera, film, p);
    }
};

// CreateIntegrator::createFromObjectDefinition, from a parsed setup
inline std::unique_ptr<Integrator> createIntegrator(const cav_render_setup& s) {
    std::unique_ptr<Integrator> r;
    if (s.integrator == CAV_INTEGRATOR_RLPATH) r = std::make_unique<RLPTIntegrator>(s.max_depth);
    else if (s.integrator == CAV_INTEGRATOR_DIRECT) r = std::make_unique<DirectLightingIntegrator>();
    else r = std::make_unique<PathTracingIntegrator>(s.max_depth);
    r->spp = s.spp;
    return r;
}

class Renderer {
public:
    std::unique_ptr<Integrator> integrator;
    cav_camera camera{};
    std::unique_ptr<Film> film;
    // Renderer::render: the integrator fills the film
    Film& render(const Scene& scene) {
        integrator->render(scene, camera, *film);
        return *film;
    }
};

// readRenderSetup + prepareForRender: the scene on the device and a renderer
// configured from the file and the overrides
struct RenderSetup {
    cav_render_setup setup{};
    std::unique_ptr<Scene> scene;
    Renderer renderer;
};

inline RenderSetup readRenderSetup(Context& ctx, const std::string& path, const cav_pbrt_overrides& ov) {
    RenderSetup r;
    r.scene = std::make_unique<Scene>(ctx, path, ov, r.setup);
    r.renderer.integrator = createIntegrator(r.setup);
    r.renderer.camera = r.setup.camera;
    r.renderer.film = std::make_unique<Film>(r.setup.width, r.setup.height);
    return r;
}

}  // namespace cav
