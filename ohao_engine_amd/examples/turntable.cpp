// Turntable driver over the C++ façade: a model in a 10 m room (or under an
// environment map), a camera orbiting it, `spp` samples per frame, one PNG per
// frame. The call pattern is the reference's examples/turntable.cpp:74-231:
// setScene, per frame camera setPosition/setRotation/setFov,
// resetAccumulation, render() x (spp + 3), getPixelSpan.
//
//   turntable <model.obj|sphere> <cornell|dark|env|mirror> [spp=64] [frames=120]
//             [--env=sky.hdr] [--size=WxH] [--out=dir] [--denoise=learned|atrous]
#include <sys/stat.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "ohao/renderer.hpp"

using namespace ohao;

namespace {

void wall(Scene& s, const char* name, vec3 a, vec3 b, vec3 c, vec3 d, vec3 n, vec3 color, float rough = 0.95f,
          float metal = 0.0f) {
    auto act = s.createActor(name);
    auto mc = act->addComponent<MeshComponent>();
    mc->setModel(makeQuad(a, b, c, d, n));
    Material& m = act->addComponent<MaterialComponent>()->getMaterial();
    m.baseColor = color;
    m.roughness = rough;
    m.metallic = metal;
}

// Built-in model: a UV sphere with a smooth radial bump pattern.
std::shared_ptr<Model> bumpy_sphere(int seg) {
    auto m = makeSphere(1.0f, seg);
    for (Vertex& v : m->vertices) {
        const vec3 p = v.position;
        const float r = 1.0f + 0.08f * std::sin(7.0f * p.x) * std::sin(5.0f * p.y + 1.0f) * std::sin(6.0f * p.z);
        v.position = p * r;
    }
    // recompute smooth normals
    std::vector<vec3> acc(m->vertices.size(), vec3(0.0f));
    for (size_t i = 0; i + 2 < m->indices.size(); i += 3) {
        const uint32_t a = m->indices[i], b = m->indices[i + 1], c = m->indices[i + 2];
        const vec3 n = cross(m->vertices[b].position - m->vertices[a].position,
                             m->vertices[c].position - m->vertices[a].position);
        acc[a] = acc[a] + n; acc[b] = acc[b] + n; acc[c] = acc[c] + n;
    }
    for (size_t i = 0; i < acc.size(); ++i)
        if (length(acc[i]) > 0) m->vertices[i].normal = normalize(acc[i]);
    return m;
}

std::vector<float> sky(uint32_t w, uint32_t h) {
    std::vector<float> px((size_t)w * h * 4, 1.0f);
    for (uint32_t y = 0; y < h; ++y) {
        const float t = 1.0f - (float)y / (float)(h - 1);   // 1 at the zenith row
        for (uint32_t x = 0; x < w; ++x) {
            float* p = &px[((size_t)y * w + x) * 4];
            p[0] = 0.2f + 0.5f * t;
            p[1] = 0.25f + 0.55f * t;
            p[2] = 0.35f + 0.65f * t;
        }
    }
    return px;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 3) {
        std::printf("Usage: turntable <model.obj|sphere> <cornell|dark|env|mirror> [spp] [frames] "
                    "[--env=file.hdr] [--size=WxH] [--out=dir] [--denoise=learned|atrous] [--spin|--spin-rebuild]\n");
        return 1;
    }
    const std::string model_path = argv[1], mode = argv[2];
    int spp = 64, frames = 120, pos = 0;
    uint32_t W = 1280, H = 720;
    std::string env_path, out_dir = "renders/turntable";
    DenoiseMode dmode = DenoiseMode::None;
    bool spin = false;      // move the model (refit, no rebuild) instead of orbiting the camera
    bool rebuild = false;   // --spin-rebuild: the same motion through a full scene rebuild
    for (int i = 3; i < argc; ++i) {
        const std::string a = argv[i];
        if (a.rfind("--env=", 0) == 0) env_path = a.substr(6);
        else if (a.rfind("--size=", 0) == 0) std::sscanf(a.c_str() + 7, "%ux%u", &W, &H);
        else if (a.rfind("--out=", 0) == 0) out_dir = a.substr(6);
        else if (a.rfind("--denoise=", 0) == 0) dmode = parseDenoiseMode(a.substr(10));
        else if (a == "--spin") spin = true;
        else if (a == "--spin-rebuild") spin = rebuild = true;
        else if (pos == 0) { spp = std::max(1, std::atoi(a.c_str())); ++pos; }
        else if (pos == 1) { frames = std::max(1, std::atoi(a.c_str())); ++pos; }
    }
    std::printf("ohao turntable: %s mode, %d spp, %d frames, %ux%u\n", mode.c_str(), spp, frames, W, H);

    VulkanRenderer renderer(W, H);
    if (!renderer.initialize()) {
        std::fprintf(stderr, "FATAL: %s\n", renderer.lastError().c_str());
        return 1;
    }
    Scene scene("Turntable");
    const float S = 5.0f;
    const vec3 white(0.73f), red(0.65f, 0.05f, 0.05f), green(0.12f, 0.45f, 0.15f);
    const vec3 LBB(-S, -S, -S), RBB(S, -S, -S), LTB(-S, S, -S), RTB(S, S, -S);
    const vec3 LBF(-S, -S, S), RBF(S, -S, S), LTF(-S, S, S), RTF(S, S, S);
    if (mode == "env") {
        if (!env_path.empty()) renderer.setEnvironmentMap(env_path);
        else renderer.setEnvironmentPixels(sky(512, 256), 512, 256);
        wall(scene, "Ground", {-20, 0, -20}, {-20, 0, 20}, {20, 0, 20}, {20, 0, -20}, {0, 1, 0}, vec3(0.5f), 0.8f);
    } else {
        wall(scene, "Left", LBB, LTB, LTF, LBF, {1, 0, 0}, red);
        wall(scene, "Right", RBB, RBF, RTF, RTB, {-1, 0, 0}, green);
        wall(scene, "Floor", LBB, LBF, RBF, RBB, {0, 1, 0}, white);
        wall(scene, "Ceiling", LTB, RTB, RTF, LTF, {0, -1, 0}, white);
        if (mode == "mirror") {
            wall(scene, "BackMirror", LBB, RBB, RTB, LTB, {0, 0, 1}, vec3(0.95f), 0.02f, 1.0f);
            wall(scene, "FrontMirror", LBF, RBF, RTF, LTF, {0, 0, -1}, vec3(0.95f), 0.02f, 1.0f);
        } else {
            wall(scene, "Back", LBB, RBB, RTB, LTB, {0, 0, 1}, white);
        }
    }
    std::shared_ptr<Actor> model_actor;
    vec3 model_pos(0.0f);
    auto model = std::make_shared<Model>();
    bool loaded = false;
    if (model_path == "sphere") {
        model = bumpy_sphere(192);
        loaded = true;
    } else {
        loaded = model->loadFromOBJ(model_path);
    }
    if (!loaded) {
        std::fprintf(stderr, "FATAL: cannot load %s\n", model_path.c_str());
        return 1;
    }
    {
        vec3 bmin(1e30f), bmax(-1e30f);
        for (const Vertex& v : model->vertices) {
            bmin = min(bmin, v.position);
            bmax = max(bmax, v.position);
        }
        const float target = (mode == "mirror") ? 5.5f : 4.0f;
        const float scale = target / std::max(bmax.y - bmin.y, 1e-6f);
        const vec3 c = (bmin + bmax) * 0.5f;
        const float floor_y = (mode == "env") ? 0.0f : -S;
        auto actor = scene.createActor("Model");
        model_actor = actor;
        actor->getTransform()->setScale(vec3(scale));
        model_pos = vec3(-c.x * scale, floor_y - bmin.y * scale, -c.z * scale);
        actor->getTransform()->setPosition(model_pos);
        auto mc = actor->addComponent<MeshComponent>();
        mc->setModel(model);
        Material& m = actor->addComponent<MaterialComponent>()->getMaterial();
        m.baseColor = vec3(0.8f, 0.6f, 0.35f);
        m.roughness = 0.35f;
        m.metallic = 1.0f;
    }
    if (mode == "dark") {
        auto fill = scene.createActor("Fill");
        auto l = fill->addComponent<LightComponent>();
        l->setLightType(LightType::Sphere);
        l->setColor({1, 1, 1});
        l->setIntensity(2.0f);
        l->setRadius(0.3f);
        fill->getTransform()->setPosition({4, 4, 4});
    } else {
        auto key = scene.createActor("Key");
        auto l = key->addComponent<LightComponent>();
        l->setLightType(LightType::Sphere);
        l->setColor({1, 0.95f, 0.9f});
        l->setIntensity(mode == "env" ? 8.0f : (mode == "mirror" ? 20.0f : 25.0f));
        l->setRadius(1.0f);
        key->getTransform()->setPosition({3, 4, 3});
        if (mode == "cornell") {
            auto panel = scene.createActor("CeilingPanel");
            auto a = panel->addComponent<LightComponent>();
            a->setLightType(LightType::AreaRect);
            a->setColor({1, 0.98f, 0.92f});
            a->setIntensity(12.0f);
            a->setAreaEdges({3, 0, 0}, {0, 0, 3});
            panel->getTransform()->setPosition({-1.5f, 4.99f, -1.5f});
        }
    }
    renderer.setScene(&scene);
    renderer.setRenderMode(RenderMode::RTOffline);
    if (dmode != DenoiseMode::None) renderer.setDenoiseMode(dmode);
    std::printf("Denoise mode: %s\n", denoiseModeName(renderer.getDenoiseMode()));

    const float radius = (mode == "env") ? 8.0f : (mode == "mirror") ? 4.5f : 4.2f;
    const float height = (mode == "env") ? 1.0f : -3.5f;
    ::mkdir("renders", 0755);
    ::mkdir(out_dir.c_str(), 0755);
    for (int f = 0; f < frames; ++f) {
        const float ang = (float)f / (float)frames * 6.2831853f;
        const float cang = spin ? 0.0f : ang;
        const float cx = radius * std::cos(cang), cz = radius * std::sin(cang);
        auto& cam = renderer.getCamera();
        cam.setPosition({cx, height, cz});
        cam.setRotation(15.0f, degrees(std::atan2(-cz, -cx)));
        cam.setFov(70.0f);
        if (spin && f > 0) {
            auto* tr = model_actor->getTransform();
            tr->setRotation(quat(vec3(0.0f, ang, 0.0f)));
            tr->setPosition(model_pos + vec3(1.5f * std::sin(ang), 0.0f, 0.0f));
            if (rebuild) renderer.setScene(&scene);
            if (!(rebuild ? renderer.updateSceneBuffers() : renderer.updateRTTransforms())) {
                std::fprintf(stderr, "FATAL: frame %d: %s\n", f, renderer.lastError().c_str());
                return 1;
            }
        }
        renderer.resetAccumulation();
        for (int s = 0; s < spp + 3; ++s) renderer.render();
        char name[512];
        std::snprintf(name, sizeof(name), "%s/%s_%04d.png", out_dir.c_str(), mode.c_str(), f);
        const auto px = renderer.getPixelSpan();
        if (px.empty() || !writePNG(name, px.data(), W, H)) {
            std::fprintf(stderr, "FATAL: frame %d: %s\n", f, renderer.lastError().c_str());
            return 1;
        }
        std::printf("Frame %d/%d (%u spp)\n", f + 1, frames, renderer.getSampleCount());
        std::fflush(stdout);
    }
    std::printf("Done: frames in %s\n", out_dir.c_str());
    return 0;
}
