// kuafu.hpp — public facade of the MI355X build (kuafu.hpp:31-93 of the reference).
//
// Same class, methods and semantics: construct with a Config, edit scenes,
// call run() once per frame, read frames with downloadLatestFrame().  Every
// GPU action (uploads, BLAS/TLAS builds, tracing, denoising) happens inside
// run() through the C ABI of libkuafu_amd (include/kuafu_amd.h).  The SDL
// viewer is not part of this build: the renderer is offscreen only.
#pragma once
#include <memory>
#include <stdexcept>
#include <vector>

#include "core/camera.hpp"
#include "core/config.hpp"
#include "core/geometry.hpp"
#include "core/light.hpp"
#include "core/scene.hpp"

struct kf_denoiser;

namespace kuafu {

class Window;
class Gui;

/// Frame orchestration (Context::update + Context::render of the reference).
class Context {
public:
    ~Context();
    std::shared_ptr<Config> pConfig;
    std::vector<std::unique_ptr<Scene>> mScenes;
    Scene* mCurrentScene = nullptr;
    void init();
    void render();
    Camera* getCamera() { return mCurrentScene ? mCurrentScene->getCamera() : nullptr; }

private:
    bool mInitialized = false;
    kf_denoiser* mDenoiser = nullptr;
};

class Kuafu {
    Context mContext;
    bool mRunning = true;
    void reset();

public:
    explicit Kuafu(std::shared_ptr<Config> config = nullptr);

    void run();
    [[nodiscard]] bool isRunning() const;
    [[nodiscard]] std::vector<uint8_t> downloadLatestFrame(Camera* cam);

    void setWindow(std::shared_ptr<Window> other);
    void setWindow(int width, int height, const char* title = "App", uint32_t flags = 0);
    [[nodiscard]] std::shared_ptr<Window> getWindow() const { return nullptr; }
    void setGui(std::shared_ptr<Gui> gui);

    inline auto& getConfig() { return *mContext.pConfig; }
    inline Scene* getScene() { return mContext.mCurrentScene; }

    void setScene(Scene* scene);
    Scene* createScene();
    void removeScene(Scene* scene);
};

}  // namespace kuafu
