// Camera of the kuafu:: API (camera.hpp:52-136 of the reference): pose,
// intrinsics, view/projection matrices and downloadLatestFrame().  Each camera
// owns its RGBA32F render targets (rgba, albedo, normal) in HBM.
#pragma once
#include <cstdint>
#include <memory>
#include <vector>

#include "../kuafu_glm.hpp"

struct kf_renderer;

namespace kuafu {

class Camera {
public:
    Camera(const Camera&) = delete;
    Camera(const Camera&&) = delete;
    Camera& operator=(Camera) = delete;
    Camera& operator=(const Camera&) = delete;
    Camera& operator=(const Camera&&) = delete;
    ~Camera();

    void update();
    void resetView();

    auto getPosition() const -> const glm::vec3& { return mPosition; }
    auto getFront() const -> const glm::vec3& { return mDirFront; }
    void setPosition(const glm::vec3& position);
    void setFront(const glm::vec3& front);
    void setUp(const glm::vec3& up);

    [[nodiscard]] inline float getAperture() const { return mAperture; }
    [[nodiscard]] inline float getFocalLength() const { return mFocalLength; }

    void setSize(int width, int height);
    void setPose(glm::mat4 pose);
    void setFullPerspective(float width, float height, float fx, float fy, float cx, float cy, float skew);
    glm::mat4 getPose() const;

    [[nodiscard]] auto getWidth() const { return mWidth; }
    [[nodiscard]] auto getHeight() const { return mHeight; }

    auto getViewMatrix() const -> const glm::mat4& { return mViewMatrix; }
    auto getProjectionMatrix() const -> const glm::mat4& { return mProjMatrix; }
    auto getViewInverseMatrix() const -> glm::mat4 { return glm::inverse(mViewMatrix); }
    auto getProjectionInverseMatrix() const -> glm::mat4 { return glm::inverse(mProjMatrix); }

    void updateViewMatrix();
    void updateProjectionMatrix();
    void processMouse(float xOffset, float yOffset);
    void processKeyboard();

    inline void clearRenderTargets() { mTargetsStale = true; }

    [[nodiscard]] inline float getPrincipalPointX() const { return mCx; }
    [[nodiscard]] inline float getPrincipalPointY() const { return mCy; }
    [[nodiscard]] inline float getFocalX() const { return mFx; }
    [[nodiscard]] inline float getFocalY() const { return mFy; }
    [[nodiscard]] inline float getNear() const { return mNear; }
    [[nodiscard]] inline float getFar() const { return mFar; }
    [[nodiscard]] inline float getSkew() const { return mSkew; }

    /// B8G8R8A8_SRGB bytes of the last rendered frame (width * height * 4).
    std::vector<uint8_t> downloadLatestFrame();
    /// RGBA32F target of the last frame: 0 rgba, 1 albedo, 2 normal.
    std::vector<float> downloadTarget(int target);

    bool mFirst = true;

private:
    friend class Scene;
    friend class Context;
    Camera(int width, int height, const glm::vec3& position = {0.0F, 0.0F, 3.0F});
    kf_renderer* renderer();

    int mWidth;
    int mHeight;
    glm::vec3 mPosition;
    float mFx, mFy, mCx, mCy;
    float mSkew = 0;
    glm::mat4 mViewMatrix = glm::mat4(1.0F);
    glm::mat4 mProjMatrix = glm::mat4(1.0F);
    glm::vec3 mDirUp = {0.0F, 0.0F, 1.0F};
    glm::vec3 mDirRight = {0.0F, -1.0F, 0.0F};
    glm::vec3 mDirFront = {1.0F, 0.0F, 0.0F};
    float mAperture = 0.0F;
    float mFocalLength = 5.0F;
    const float mFar = 100.F;
    const float mNear = 0.1F;
    glm::vec3 mInitialPosition;
    glm::vec3 mPrevPosition;
    bool mTargetsStale = true;
    kf_renderer* mRenderer = nullptr;
    bool mRendered = false;
};

}  // namespace kuafu
