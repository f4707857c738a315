// Kuafu and Context of the kuafu:: API (kuafu.cpp:27-128, context.cpp:298-684
// of the reference): one run() = camera update, scene upload / BVH build or
// per-frame TLAS update, frame counter, trace, optional denoise.
#include <cstring>
#include <stdexcept>

#include "../../include/kuafu.hpp"
#include "../../include/kuafu_amd.h"

namespace kuafu {

static void kf_ok(int rc) {
    if (rc != 0) throw std::runtime_error(kf_last_error());
}

Context::~Context() {
    mScenes.clear();
    if (mDenoiser) kf_denoiser_destroy(mDenoiser);
}

void Context::init() {
    if (mInitialized) return;
    kf_ok(kf_init(0));
    mInitialized = true;
}

void Context::render() {
    init();
    Scene* scene = mCurrentScene;
    KF_ASSERT(scene, "No scene to render!");
    Camera* cam = scene->getCamera();
    KF_ASSERT(cam, "Trying to render with an invalid camera!");
    // Context::update (context.cpp:298-350), incrementally
    scene->syncDevice();
    if (pConfig->mAccumulateFrames)
        ++global::frameCount;
    else
        global::frameCount = -1;
    // push constants (context.cpp:605-614) and camera UBO (scene.cpp:236-250)
    kf_push_constants pc{};
    glm::vec4 cc = scene->getClearColor();
    float ccv[4] = {cc.x, cc.y, cc.z, cc.w};
    memcpy(pc.clear_color, ccv, sizeof(ccv));
    pc.frame_count = global::frameCount;
    pc.spp = pConfig->mPerPixelSampleRate;
    pc.path_depth = pConfig->mPathDepth;
    pc.use_env_map = scene->mUseEnvironmentMap ? 1u : 0u;
    pc.russian_roulette = pConfig->mRussianRoulette ? 1u : 0u;
    pc.rr_min_bounces = pConfig->mRussianRouletteMinBounces;
    pc.nee = pConfig->mNextEventEstimation ? 1u : 0u;
    pc.nee_min_bounces = pConfig->mNextEventEstimationMinBounces;
    kf_camera_ubo ubo;
    glm::vec3 p = cam->getPosition(), f = cam->getFront();
    glm::vec3 up = glm::vec3(cam->getViewInverseMatrix()[1][0], cam->getViewInverseMatrix()[1][1],
                             cam->getViewInverseMatrix()[1][2]);
    float pv[3] = {p.x, p.y, p.z}, fv[3] = {f.x, f.y, f.z}, uv[3] = {cam->mDirUp.x, cam->mDirUp.y, cam->mDirUp.z};
    (void)up;
    kf_ok(kf_camera_ubo_compute(pv, fv, uv, cam->getWidth(), cam->getHeight(), cam->getFocalX(), cam->getFocalY(),
                                cam->getPrincipalPointX(), cam->getPrincipalPointY(), cam->getSkew(),
                                cam->getAperture(), cam->getFocalLength(), &ubo));
    kf_render_opts opts{};
    opts.seed = pConfig->mSeed;
    kf_renderer* r = cam->renderer();
    kf_ok(kf_render(r, scene->mDevice, &ubo, &pc, &opts, nullptr));
    cam->mRendered = true;
    // denoise stage (context.cpp:650-670): the output overwrites rgba
    if (pConfig->mUseDenoiser) {
        if (!mDenoiser) kf_ok(kf_denoiser_create(nullptr, &mDenoiser));
        kf_ok(kf_denoiser_apply(mDenoiser, r));
    }
}

Kuafu::Kuafu(std::shared_ptr<Config> config) {
    mContext.pConfig = std::make_shared<Config>(config ? *config : Config());
    mContext.mScenes.emplace_back(new Scene(mContext.pConfig));
    mContext.mCurrentScene = mContext.mScenes.back().get();
    mContext.mCurrentScene->initialized = true;
    if (mContext.pConfig->mPresent) {
        KF_INFO("Present mode requested: this build renders offscreen; frames are read with downloadLatestFrame().");
        auto cam = mContext.mCurrentScene->createCamera(mContext.pConfig->mInitialWidth,
                                                        mContext.pConfig->mInitialHeight);
        mContext.mCurrentScene->setCamera(cam);
    } else {
        KF_INFO("Offscreen mode enabled.");
    }
    if (mContext.pConfig->getAssetsPath().empty()) mContext.pConfig->setAssetsPath(Config::sDefaultAssetsPath);
}

bool Kuafu::isRunning() const {
    if (!mRunning) KF_INFO("Shutting down Kuafu.");
    return mRunning;
}

void Kuafu::run() {
    if (!mRunning) return;
    mContext.getCamera()->update();
    mContext.render();
}

std::vector<uint8_t> Kuafu::downloadLatestFrame(Camera* cam) {
    KF_ASSERT(cam, "downloadLatestFrame: invalid camera");
    return cam->downloadLatestFrame();
}

void Kuafu::setWindow(std::shared_ptr<Window>) {
    throw std::runtime_error("Windowed mode is not available in this build (offscreen rendering only).");
}
void Kuafu::setWindow(int, int, const char*, uint32_t) {
    throw std::runtime_error("Windowed mode is not available in this build (offscreen rendering only).");
}
void Kuafu::setGui(std::shared_ptr<Gui>) {
    throw std::runtime_error("The GUI is not available in this build (offscreen rendering only).");
}

void Kuafu::reset() { global::frameCount = -1; }

void Kuafu::setScene(Scene* scene) {
    if (mContext.mCurrentScene == scene && scene->initialized) return;
    KF_INFO("Switching scene, this is heavy...");
    KF_ASSERT(scene, "Trying to set an invalid scene!");
    auto ret = std::find_if(mContext.mScenes.begin(), mContext.mScenes.end(),
                            [scene](auto& s) { return scene == s.get(); });
    KF_ASSERT(ret != mContext.mScenes.end(), "???");
    scene->init();
    mContext.mCurrentScene = scene;
    global::frameCount = -1;
}

Scene* Kuafu::createScene() {
    mContext.mScenes.emplace_back(new Scene(mContext.pConfig));
    return mContext.mScenes.back().get();
}

void Kuafu::removeScene(Scene* scene) {
    KF_ASSERT(scene, "Trying to remove an invalid scene!");
    auto ret = std::find_if(mContext.mScenes.begin(), mContext.mScenes.end(),
                            [scene](auto& s) { return scene == s.get(); });
    KF_ASSERT(ret != mContext.mScenes.end(), "???");
    if (mContext.mCurrentScene == scene) mContext.mCurrentScene = nullptr;
    mContext.mScenes.erase(ret);
}

}  // namespace kuafu
