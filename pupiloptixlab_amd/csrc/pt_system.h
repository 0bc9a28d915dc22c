// Headless re-creation of the reference's pass / system interface on top of
// the C ABI, so an application written against Pupil::Pass and
// Pupil::System drives the HIP path tracer unchanged.
//
//   Pupil::Pass              system/pass.h:22-43, system/pass.cpp:6-13
//   Pupil::System            system/system.h:22-41, system/system.cpp:33-173
//   Pupil::pt::PTPass        example/path_tracer/pt_pass.h:21-41
//
// Differences: no window, GUI thread or DX12 interop; errors are reported
// through return values and Pass::LastError() instead of logging and
// asserting; System::Run renders a given number of frames and returns.
#pragma once

#include <stdint.h>

#include <memory>
#include <string>
#include <string_view>
#include <vector>

#include "../../include/pupil_hip.h"

namespace Pupil {

enum class EPassTag : uint32_t { None = 0, Pre = 1 << 0, Post = 1 << 1, Asyn = 1 << 2 };

class Pass {
public:
    const std::string name;
    const EPassTag tag;

    Pass(std::string_view name, EPassTag tag = EPassTag::None) noexcept : name(name), tag(tag) {}
    virtual ~Pass() noexcept = default;

    virtual void OnRun() noexcept = 0;
    // Headless inspector: a one-line status instead of an ImGui panel.
    virtual std::string Inspector() noexcept;
    // Timer + OnRun (system/pass.cpp:6-13).
    void Run() noexcept;
    double LastRunMs() const noexcept { return m_last_ms; }

protected:
    double m_last_ms = 0.0;
};

class System {
public:
    bool Init(int hip_device = 0) noexcept;
    void AddPass(Pass *pass) noexcept;  // not owned (system.cpp:129-134)
    // Scene::LoadFromXML + World::LoadScene + SceneLoad event (system.cpp:136-173).
    bool SetScene(const std::string &xml_path) noexcept;
    // Same, for a procedurally generated scene.
    bool SetScene(int scene_kind, const pt_gen_params &params) noexcept;
    // Runs every pass `frames` times on this thread.
    void Run(uint32_t frames) noexcept;
    void Destroy() noexcept;

    pt_ctx *Context() const noexcept { return m_ctx; }
    const pt_scene_desc *Scene() const noexcept { return m_scene ? &m_desc : nullptr; }
    const std::string &LastError() const noexcept { return m_error; }

private:
    bool Publish() noexcept;
    pt_ctx *m_ctx = nullptr;
    pt_host_scene *m_scene = nullptr;
    pt_scene_desc m_desc{};
    std::vector<Pass *> m_passes;
    std::string m_error;
};

namespace pt {

// The path-tracing pass: one frame (one sample per pixel) per OnRun.
class PTPass : public Pass {
public:
    explicit PTPass(System *system, std::string_view name = "Path Tracing") noexcept;
    void OnRun() noexcept override;
    std::string Inspector() noexcept override;

    // SceneLoad handler (pt_pass.cpp:127-240): uploads the world and builds the BVH.
    bool SetScene(const pt_scene_desc *scene) noexcept;
    // EWorldEvent::CameraChange (pt_pass.cpp:242-249).
    void SetCamera(const float sample_to_camera[16], const float camera_to_world[16]) noexcept;

    // Inspector knobs (pt_pass.cpp:256-268): max depth clamped to 1..128.
    void SetMaxDepth(int depth) noexcept;
    void SetAccumulate(bool on) noexcept;
    void SetMisMode(uint32_t mode) noexcept { m_params.mis_mode = mode; m_dirty = true; }

    // Named buffers of the pass ("final result", "pt accum buffer", ...).
    bool ReadBuffer(int which, void *dst, size_t bytes) noexcept;
    const std::string &LastError() const noexcept { return m_error; }
    uint32_t SampleCount() const noexcept { return m_sample_cnt; }

private:
    System *m_system;
    pt_render_params m_params{};
    bool m_dirty = true;
    uint32_t m_sample_cnt = 0;
    std::string m_error;
};

}  // namespace pt

// DenoisePass (system/denoise_pass.h:14-43): filters the path tracer's
// accumulation buffer into "final result" each frame it is enabled, with the
// albedo / normal AOVs as guides when the config asks for them.
class DenoisePass : public Pass {
public:
    struct Config {
        bool default_enable = true;
        bool use_albedo = true;
        bool use_normal = true;
    };
    explicit DenoisePass(System *system, const Config &config, std::string_view name = "Denoise") noexcept;
    void OnRun() noexcept override;
    std::string Inspector() noexcept override;
    void SetEnabled(bool on) noexcept { m_enabled = on; }
    const std::string &LastError() const noexcept { return m_error; }

private:
    System *m_system;
    Config m_config;
    bool m_enabled;
    std::string m_error;
};

}  // namespace Pupil
