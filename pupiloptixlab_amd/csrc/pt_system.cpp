#include "pt_system.h"

#include <algorithm>
#include <chrono>
#include <cstdio>

namespace Pupil {

std::string Pass::Inspector() noexcept {
    char buf[128];
    std::snprintf(buf, sizeof(buf), "%s: %.3f ms", name.c_str(), m_last_ms);
    return buf;
}

void Pass::Run() noexcept {
    auto t0 = std::chrono::steady_clock::now();
    OnRun();
    m_last_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

bool System::Init(int hip_device) noexcept {
    int rc = pt_create(hip_device, &m_ctx);
    if (rc != PT_OK) {
        m_error = std::string("pt_create: ") + pt_last_error(nullptr);
        m_ctx = nullptr;
        return false;
    }
    return true;
}

void System::AddPass(Pass *pass) noexcept {
    if (pass) m_passes.push_back(pass);
}

bool System::Publish() noexcept {
    pt_scene_get_desc(m_scene, &m_desc);
    bool ok = true;
    for (Pass *p : m_passes) {
        if (auto *ptp = dynamic_cast<pt::PTPass *>(p)) {
            if (!ptp->SetScene(&m_desc)) {
                m_error = ptp->LastError();
                ok = false;
            }
        }
    }
    return ok;
}

bool System::SetScene(const std::string &xml_path) noexcept {
    pt_host_scene *s = nullptr;
    char err[512] = {0};
    if (pt_scene_load_xml(xml_path.c_str(), &s, err, sizeof(err)) != PT_OK) {
        m_error = std::string("scene load failed: ") + err;  // reference: warn and keep the old scene
        return false;
    }
    if (m_scene) pt_scene_destroy(m_scene);
    m_scene = s;
    return Publish();
}

bool System::SetScene(int kind, const pt_gen_params &params) noexcept {
    pt_host_scene *s = nullptr;
    if (pt_scene_generate(kind, &params, &s) != PT_OK) {
        m_error = "scene generation failed";
        return false;
    }
    if (m_scene) pt_scene_destroy(m_scene);
    m_scene = s;
    return Publish();
}

void System::Run(uint32_t frames) noexcept {
    for (uint32_t f = 0; f < frames; ++f)
        for (Pass *p : m_passes) p->Run();
}

void System::Destroy() noexcept {
    m_passes.clear();
    if (m_ctx) pt_destroy(m_ctx);
    m_ctx = nullptr;
    if (m_scene) pt_scene_destroy(m_scene);
    m_scene = nullptr;
}

namespace pt {

PTPass::PTPass(System *system, std::string_view name) noexcept : Pass(name), m_system(system) {
    m_params.spp = 1;
    m_params.rr_depth = 2;
    m_params.accumulate = 1;
    m_params.flags = PT_RENDER_AOV;
}

bool PTPass::SetScene(const pt_scene_desc *scene) noexcept {
    pt_ctx *ctx = m_system ? m_system->Context() : nullptr;
    if (!ctx) {
        m_error = "PTPass::SetScene: system not initialized";
        return false;
    }
    if (pt_set_scene(ctx, scene) != PT_OK || pt_build_bvh(ctx, nullptr) != PT_OK) {
        m_error = pt_last_error(ctx);
        return false;
    }
    m_params.max_depth = scene->max_depth;
    m_dirty = true;
    return true;
}

void PTPass::SetCamera(const float s2c[16], const float c2w[16]) noexcept {
    pt_ctx *ctx = m_system ? m_system->Context() : nullptr;
    if (ctx && pt_set_camera(ctx, s2c, c2w) != PT_OK) m_error = pt_last_error(ctx);
    m_dirty = true;
}

void PTPass::SetMaxDepth(int depth) noexcept {
    depth = std::clamp(depth, 1, 128);
    if ((uint32_t)depth != m_params.max_depth) m_dirty = true;
    m_params.max_depth = (uint32_t)depth;
}

void PTPass::SetAccumulate(bool on) noexcept {
    m_params.accumulate = on ? 1u : 0u;
    m_dirty = true;
}

// PTPass::OnRun (pt_pass.cpp:39-57): restart accumulation when dirty, then
// trace one frame.
void PTPass::OnRun() noexcept {
    pt_ctx *ctx = m_system ? m_system->Context() : nullptr;
    if (!ctx) return;
    if (m_dirty) {
        pt_reset_accumulation(ctx);
        m_sample_cnt = 0;
        m_dirty = false;
    }
    if (pt_render(ctx, &m_params) != PT_OK) {
        m_error = pt_last_error(ctx);
        return;
    }
    m_sample_cnt += m_params.accumulate;
}

std::string PTPass::Inspector() noexcept {
    char buf[160];
    std::snprintf(buf, sizeof(buf), "%s: %.3f ms, sample count %u, max depth %u, accumulate %s", name.c_str(),
                  m_last_ms, m_sample_cnt + 1, m_params.max_depth, m_params.accumulate ? "on" : "off");
    return buf;
}

bool PTPass::ReadBuffer(int which, void *dst, size_t bytes) noexcept {
    pt_ctx *ctx = m_system ? m_system->Context() : nullptr;
    if (!ctx || pt_read_buffer(ctx, which, dst, bytes) != PT_OK) {
        m_error = ctx ? pt_last_error(ctx) : "no context";
        return false;
    }
    return true;
}

}  // namespace pt

DenoisePass::DenoisePass(System *system, const Config &config, std::string_view name) noexcept
    : Pass(name), m_system(system), m_config(config), m_enabled(config.default_enable) {}

// DenoisePass::OnRun (system/denoise_pass.cpp:49-71).
void DenoisePass::OnRun() noexcept {
    if (!m_enabled) return;
    pt_ctx *ctx = m_system ? m_system->Context() : nullptr;
    if (!ctx) return;
    pt_denoise_params p{};
    p.flags = (m_config.use_albedo ? PT_DENOISE_ALBEDO : 0u) | (m_config.use_normal ? PT_DENOISE_NORMAL : 0u);
    if (pt_denoise(ctx, &p) != PT_OK) m_error = pt_last_error(ctx);
}

std::string DenoisePass::Inspector() noexcept {
    char buf[128];
    std::snprintf(buf, sizeof(buf), "%s: %s, %.3f ms, albedo %s, normal %s", name.c_str(), m_enabled ? "on" : "off",
                  m_last_ms, m_config.use_albedo ? "on" : "off", m_config.use_normal ? "on" : "off");
    return buf;
}

}  // namespace Pupil
