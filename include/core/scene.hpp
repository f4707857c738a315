// Scene of the kuafu:: API (scene.hpp:47-211 of the reference): geometries,
// instances, lights, cameras, clear colour and environment.  Edits only mark
// the scene dirty; uploads and BVH builds happen in the next Kuafu::run().
#pragma once
#include <algorithm>
#include <memory>
#include <stdexcept>
#include <string>
#include <string_view>
#include <vector>

#include "camera.hpp"
#include "config.hpp"
#include "geometry.hpp"
#include "light.hpp"

struct kf_scene;

namespace kuafu {
class Context;
class Kuafu;

class Scene {
public:
    friend Context;
    friend Kuafu;

    Scene() = delete;
    ~Scene();

    auto getGeometries() const -> const std::vector<std::shared_ptr<Geometry>>&;
    auto getGeometryInstances() const -> const std::vector<std::shared_ptr<GeometryInstance>>&;
    auto getGeometryInstance(size_t index) const -> std::shared_ptr<GeometryInstance>;

    void submitGeometryInstance(std::shared_ptr<GeometryInstance> geometryInstance);
    void submitGeometryInstance(const GeometryInstance& geometryInstance);
    void setGeometryInstances(const std::vector<std::shared_ptr<GeometryInstance>>& geometryInstances);
    void removeGeometryInstance(const std::shared_ptr<GeometryInstance>& geometryInstance);
    void removeGeometryInstances(const std::vector<std::shared_ptr<GeometryInstance>>& geometryInstances);
    void clearGeometryInstances();

    void submitGeometry(std::shared_ptr<Geometry> geometry);
    void submitGeometry(const Geometry& geometry);
    void setGeometries(const std::vector<std::shared_ptr<Geometry>>& geometries);
    void removeGeometry(std::shared_ptr<Geometry> geometry);
    void removeGeometry(uint32_t geometryIndex);
    void clearGeometries();

    void setClearColor(const glm::vec4& clearColor);
    inline auto getClearColor() const { return mClearColor; }

    auto findGeometry(std::string_view path) const -> std::shared_ptr<Geometry>;

    /// Environment: the reference loads a KTX cube map; this build takes six
    /// RGBA8 sRGB faces (+X,-X,+Y,-Y,+Z,-Z) or a two-colour gradient.
    void setEnvironmentMap(std::string_view path);
    void setEnvironmentCubemap(const std::vector<std::vector<uint8_t>>& faces, uint32_t size);
    void setEnvironmentGradient(const glm::vec3& bottom, const glm::vec3& top);
    void removeEnvironmentMap();

    Camera* createCamera(int width, int height);
    void removeCamera(Camera* camera);
    void setCamera(Camera* camera);
    Camera* getCamera() const { return mCurrentCamera; }

    inline auto getGeometryInstanceCount() { return mGeometries.size(); }
    inline void markGeometriesChanged() { mUploadGeometries = true; }
    inline void markGeometryInstancesChanged() { mUploadGeometryInstancesToBuffer = true; }

    inline void setDirectionalLight(std::shared_ptr<DirectionalLight> light) { pDirectionalLight = light; }
    inline void removeDirectionalLight() { pDirectionalLight = nullptr; }
    void addPointLight(const std::shared_ptr<PointLight>& light);
    void removePointLight(const std::shared_ptr<PointLight>& light);
    void addActiveLight(const std::shared_ptr<ActiveLight>& light);
    void removeActiveLight(const std::shared_ptr<ActiveLight>& light);

    /// RGBA8 (sRGB) texture from memory; returns its index for NiceMaterial /
    /// ActiveLight texture ids.  The reference only loads textures from files.
    int addTexture(const std::vector<uint8_t>& rgba8, uint32_t width, uint32_t height);

    /// The device scene (acceleration structures and buffers) behind this
    /// Scene; null before the first run().  No counterpart in the reference.
    const kf_scene* getDeviceScene() const { return mDevice; }
    /// Device scenes built from scratch so far (full uploads and BLAS
    /// builds); incremental updates do not count.  No counterpart in the
    /// reference, which rebuilds every BLAS on any change.
    size_t getDeviceBuilds() const { return mDeviceBuilds; }

    void init();

private:
    explicit Scene(std::shared_ptr<Config> pConfig);
    void buildDevice();   // upload + BLAS/TLAS build (Context::update)
    void updateDevice();  // per-frame TLAS rebuild + lights
    void syncDevice();    // the cheapest of the two that brings the device up to date
    bool refitDevice();   // geometries changed in place only: refit their BLAS, re-upload materials

  public:
    // what the device scene was built from (to tell in-place changes apart)
    struct DeviceGeometry {
        const Geometry* geometry;
        size_t vertices, indices;
        uint64_t topology;  // hash of indices and material indices
        bool opaque, hidden;
    };

  private:
    std::vector<DeviceGeometry> mDeviceGeometries;
    size_t mDeviceMaterials = 0, mDeviceTextures = 0, mDeviceBuilds = 0;
    uint32_t mEnvVersion = 0, mDeviceEnvVersion = 0;  // environment maps set so far / in the device scene

    bool initialized = false;
    glm::vec4 mClearColor = glm::vec4(0.F, 0.F, 0.F, 1.F);
    std::vector<std::shared_ptr<Geometry>> mGeometries;
    std::vector<std::shared_ptr<GeometryInstance>> mGeometryInstances;
    std::shared_ptr<DirectionalLight> pDirectionalLight;
    std::vector<std::shared_ptr<PointLight>> pPointLights;
    std::vector<std::shared_ptr<ActiveLight>> pActiveLights;
    std::string mEnvironmentMapTexturePath;
    bool mUseEnvironmentMap = false;
    int mEnvKind = 0;
    std::vector<std::vector<uint8_t>> mEnvFaces;
    uint32_t mEnvSize = 0;
    float mEnvGradient[8] = {0};
    std::vector<std::vector<uint8_t>> mTextures;
    std::vector<std::pair<uint32_t, uint32_t>> mTextureSizes;
    bool mUploadGeometryInstancesToBuffer = false;
    bool mUploadEnvironmentMap = false;
    bool mUploadGeometries = false;
    std::vector<std::unique_ptr<Camera>> mRegisteredCameras;
    Camera* mCurrentCamera = nullptr;
    std::shared_ptr<Config> pConfig = nullptr;
    kf_scene* mDevice = nullptr;
};

}  // namespace kuafu
