// Camera and Scene of the kuafu:: API over the C ABI of libkuafu_amd.
//
// Camera: pose / intrinsics as camera.cpp:8-139 of the reference; the CameraUBO
// is produced by kf_camera_ubo_compute; render targets are a kf_renderer.
// Scene: bookkeeping and limits as scene.cpp:46-173; buildDevice() is
// uploadGeometries + uploadGeometryInstances + createBottomLevelAS + buildTlas
// (context.cpp:323-336), updateDevice() is the per-frame updateTlas +
// uploadLightBuffers (context.cpp:337-343, scene.cpp:252-292).  Unlike the
// reference, which rebuilds every BLAS whenever geometries or instances
// change, syncDevice() rebuilds from scratch only when the set of
// geometries, their triangles, the material count or the textures changed:
// instance changes rebuild the TLAS alone, and geometries re-marked
// uninitialized with the same triangles have their BLAS refit.
#include <cstring>
#include <stdexcept>

#include "../../include/core/scene.hpp"
#include "../../include/kuafu_amd.h"

namespace kuafu {

static void kf_ok(int rc) {
    if (rc != 0) throw std::runtime_error(kf_last_error());
}

// ------------------------------------------------------------------ Camera
Camera::Camera(int width, int height, const glm::vec3& position)
    : mWidth(width), mHeight(height), mPosition(position), mInitialPosition(position), mPrevPosition(position) {
    mCx = mWidth * 0.5f;
    mCy = mHeight * 0.5f;
    mFx = mFy = mWidth * 0.5f;
    updateProjectionMatrix();
    resetView();
}

Camera::~Camera() {
    if (mRenderer) kf_renderer_destroy(mRenderer);
}

void Camera::resetView() {
    mPosition = mInitialPosition;
    mDirUp = {0.0F, 0.0F, 1.0F};
    mDirRight = {0.0F, -1.0F, 0.0F};
    mDirFront = {1.0F, 0.0F, 0.0F};
    updateViewMatrix();
    global::frameCount = -1;
}

glm::mat4 Camera::getPose() const {
    return {{mDirFront[0], mDirFront[1], mDirFront[2], 0},
            {-mDirRight[0], -mDirRight[1], -mDirRight[2], 0},
            {mDirUp[0], mDirUp[1], mDirUp[2], 0},
            {mPosition[0], mPosition[1], mPosition[2], 1}};
}

void Camera::update() {
    if (mPrevPosition != mPosition) {
        global::frameCount = -1;
        mPrevPosition = mPosition;
    }
    processKeyboard();
    updateViewMatrix();
}

void Camera::setPosition(const glm::vec3& position) {
    mPosition = position;
    updateViewMatrix();
}
void Camera::setFront(const glm::vec3& front) {
    mDirFront = front;
    updateViewMatrix();
}
void Camera::setUp(const glm::vec3& up) {
    mDirUp = up;
    updateViewMatrix();
}
void Camera::setSize(int width, int height) {
    mWidth = width;
    mHeight = height;
    mCx = mWidth / 2.f;
    mCy = mHeight / 2.f;
    updateProjectionMatrix();
}
void Camera::updateViewMatrix() { mViewMatrix = glm::lookAt(mPosition, mPosition + mDirFront, mDirUp); }

void Camera::setFullPerspective(float width, float height, float fx, float fy, float cx, float cy, float skew) {
    mWidth = static_cast<int>(width);
    mHeight = static_cast<int>(height);
    mFx = fx;
    mFy = fy;
    mCx = cx;
    mCy = cy;
    mSkew = skew;
    updateProjectionMatrix();
}

void Camera::updateProjectionMatrix() {
    // camera.cpp:111-128
    mProjMatrix = glm::mat4(0.0f);
    mProjMatrix[0][0] = (2.f * mFx) / mWidth;
    mProjMatrix[1][0] = -2 * mSkew / mWidth;
    mProjMatrix[1][1] = -(2.f * mFy) / mHeight;
    mProjMatrix[2][0] = -2.f * mCx / mWidth + 1;
    mProjMatrix[2][1] = -2.f * mCy / mHeight + 1;
    mProjMatrix[2][2] = -mFar / (mFar - mNear);
    mProjMatrix[2][3] = -1.f;
    mProjMatrix[3][2] = -mFar * mNear / (mFar - mNear);
    mProjMatrix[3][3] = 0.f;
    clearRenderTargets();
}

void Camera::setPose(glm::mat4 pose) {
    // camera.cpp:131-139: no view update until the next Camera::update()
    mPosition = {pose[3][0], pose[3][1], pose[3][2]};
    mDirFront = {-pose[2][0], -pose[2][1], -pose[2][2]};
    mDirUp = {pose[1][0], pose[1][1], pose[1][2]};
    global::frameCount = -1;
}

void Camera::processMouse(float xOffset, float yOffset) {
    // viewer interaction: rotate the front vector about up and right
    glm::mat4 urot = glm::rotate(glm::mat4(1.0), xOffset * 0.01f, mDirUp);
    glm::vec4 t = urot * glm::vec4(mDirFront, 0);
    mDirFront = {t[0], t[1], t[2]};
    (void)yOffset;
    global::frameCount = -1;
}

void Camera::processKeyboard() {}  // no window, no keys: the build is offscreen

kf_renderer* Camera::renderer() {
    if (mRenderer && mTargetsStale) {
        kf_renderer_destroy(mRenderer);
        mRenderer = nullptr;
    }
    if (!mRenderer) {
        kf_ok(kf_renderer_create((uint32_t)mWidth, (uint32_t)mHeight, &mRenderer));
        mRendered = false;
    }
    mTargetsStale = false;
    return mRenderer;
}

std::vector<uint8_t> Camera::downloadLatestFrame() {
    if (!mRenderer || !mRendered) throw std::runtime_error("Invalid call to Camera::downloadLatestFrame");
    std::vector<uint8_t> out((size_t)mWidth * mHeight * 4);
    kf_ok(kf_renderer_download_bgra8(mRenderer, out.data()));
    return out;
}

std::vector<float> Camera::downloadTarget(int target) {
    if (!mRenderer || !mRendered) throw std::runtime_error("Invalid call to Camera::downloadTarget");
    std::vector<float> out((size_t)mWidth * mHeight * 4);
    kf_ok(kf_renderer_read(mRenderer, target, out.data()));
    return out;
}

// ------------------------------------------------------------------ Scene
Scene::Scene(std::shared_ptr<Config> config) : pConfig(std::move(config)) { setCamera(createCamera(1, 1)); }

Scene::~Scene() {
    if (mDevice) kf_scene_destroy(mDevice);
}

auto Scene::getGeometries() const -> const std::vector<std::shared_ptr<Geometry>>& { return mGeometries; }
auto Scene::getGeometryInstances() const -> const std::vector<std::shared_ptr<GeometryInstance>>& {
    return mGeometryInstances;
}
auto Scene::getGeometryInstance(size_t index) const -> std::shared_ptr<GeometryInstance> {
    if (index < mGeometryInstances.size()) return mGeometryInstances[index];
    throw std::runtime_error("Geometry Instances out of bound.");
}

void Scene::submitGeometryInstance(std::shared_ptr<GeometryInstance> gi) {
    if (mGeometryInstances.size() > pConfig->mMaxGeometryInstances)
        throw std::runtime_error("Failed to submit geometry instance because instance buffer size has been exceeded.");
    for (size_t i = 0; i < mGeometries.size(); i++)
        if (mGeometries[i] == gi->geometry) gi->geometryIndex = (int)i;
    KF_ASSERT(gi->geometryIndex >= 0, "Geometry not submitted!");
    mGeometryInstances.push_back(gi);
    markGeometryInstancesChanged();
}
void Scene::submitGeometryInstance(const GeometryInstance& gi) {
    submitGeometryInstance(std::make_shared<GeometryInstance>(gi));
}
void Scene::setGeometryInstances(const std::vector<std::shared_ptr<GeometryInstance>>& gis) {
    mGeometryInstances.clear();
    mGeometryInstances.reserve(gis.size());
    for (auto& g : gis) submitGeometryInstance(g);
    markGeometryInstancesChanged();
}
void Scene::removeGeometryInstance(const std::shared_ptr<GeometryInstance>& gi) {
    KF_ASSERT(gi, "Deleting an invalid geometry instance!");
    mGeometryInstances.erase(std::remove(mGeometryInstances.begin(), mGeometryInstances.end(), gi),
                             mGeometryInstances.end());
    markGeometryInstancesChanged();
}
void Scene::removeGeometryInstances(const std::vector<std::shared_ptr<GeometryInstance>>& gis) {
    for (auto& g : gis) removeGeometryInstance(g);
}
void Scene::clearGeometryInstances() {
    mGeometryInstances.clear();
    markGeometryInstancesChanged();
}

void Scene::submitGeometry(std::shared_ptr<Geometry> g) {
    if (mGeometries.size() >= pConfig->mMaxGeometry)
        throw std::runtime_error("Failed to submit geometry because geometries buffer size has been exceeded.");
    mGeometries.push_back(std::move(g));
    markGeometriesChanged();
}
void Scene::submitGeometry(const Geometry& g) { submitGeometry(std::make_shared<Geometry>(g)); }
void Scene::setGeometries(const std::vector<std::shared_ptr<Geometry>>& gs) {
    mGeometries.clear();
    mGeometries.reserve(gs.size());
    for (auto& g : gs) submitGeometry(g);
    markGeometriesChanged();
}
void Scene::removeGeometry(std::shared_ptr<Geometry>) { throw std::runtime_error("Not implemented!"); }
void Scene::removeGeometry(uint32_t) { throw std::runtime_error("Not implemented!"); }
void Scene::clearGeometries() {
    mGeometries.clear();
    mGeometryInstances.clear();
    global::textureIndex = 0;
    markGeometriesChanged();
    markGeometryInstancesChanged();
}

void Scene::setClearColor(const glm::vec4& c) {
    mClearColor = c;
    global::frameCount = -1;
}

auto Scene::findGeometry(std::string_view path) const -> std::shared_ptr<Geometry> {
    for (auto& g : mGeometries)
        if (g->path == path) return g;
    return nullptr;
}

void Scene::setEnvironmentMap(std::string_view path) {
    // KTX cube maps (scene.cpp:301-303, vkCore.hpp:1852-1970) are read here,
    // when set, instead of at the next upload; an empty path selects the
    // clear colour
    mEnvironmentMapTexturePath = path;
    if (path.empty()) {
        mUseEnvironmentMap = false;
        return;
    }
    const std::string p(path);
    const bool ktx = (p.size() >= 4 && p.compare(p.size() - 4, 4, ".ktx") == 0) ||
                     (p.size() >= 5 && p.compare(p.size() - 5, 5, ".ktx2") == 0);
    if (!ktx) throw std::runtime_error("cubemap format not supported: " + p);
    uint32_t size = 0;
    kf_ok(kf_ktx_cubemap_read(p.c_str(), &size, nullptr, 0));
    std::vector<uint8_t> all(6ull * 4 * size * size);
    kf_ok(kf_ktx_cubemap_read(p.c_str(), &size, all.data(), all.size()));
    std::vector<std::vector<uint8_t>> faces(6);
    const size_t face = 4ull * size * size;
    for (int f = 0; f < 6; ++f) faces[f].assign(all.begin() + f * face, all.begin() + (f + 1) * face);
    setEnvironmentCubemap(faces, size);
}
void Scene::setEnvironmentCubemap(const std::vector<std::vector<uint8_t>>& faces, uint32_t size) {
    if (faces.size() != 6) throw std::runtime_error("setEnvironmentCubemap: 6 faces expected");
    for (auto& f : faces)
        if (f.size() != 4ull * size * size) throw std::runtime_error("setEnvironmentCubemap: bad face size");
    mEnvFaces = faces;
    mEnvSize = size;
    mEnvKind = KF_ENV_CUBEMAP;
    ++mEnvVersion;
    mUseEnvironmentMap = true;
    markGeometriesChanged();
}
void Scene::setEnvironmentGradient(const glm::vec3& bottom, const glm::vec3& top) {
    float g[8] = {bottom.x, bottom.y, bottom.z, 0, top.x, top.y, top.z, 0};
    memcpy(mEnvGradient, g, sizeof(g));
    mEnvKind = KF_ENV_GRADIENT;
    ++mEnvVersion;
    mUseEnvironmentMap = true;
    markGeometriesChanged();
}
void Scene::removeEnvironmentMap() { mUseEnvironmentMap = false; }

Camera* Scene::createCamera(int width, int height) {
    mRegisteredCameras.emplace_back(new Camera(width, height, glm::vec3(0.f, 0.f, 0.f)));
    return mRegisteredCameras.back().get();
}
void Scene::removeCamera(Camera* camera) {
    KF_ASSERT(camera, "Trying to remove an invalid camera!");
    auto it = std::find_if(mRegisteredCameras.begin(), mRegisteredCameras.end(),
                           [camera](auto& c) { return camera == c.get(); });
    KF_ASSERT(it != mRegisteredCameras.end(), "Trying to remove an camera which does not belong to the scene!");
    if (mCurrentCamera == camera) {
        KF_INFO("Removing the active camera. This may cause problems.");
        mCurrentCamera = nullptr;
    }
    mRegisteredCameras.erase(it);
}
void Scene::setCamera(Camera* camera) {
    KF_ASSERT(camera, "Trying to set an invalid camera!");
    auto it = std::find_if(mRegisteredCameras.begin(), mRegisteredCameras.end(),
                           [camera](auto& c) { return camera == c.get(); });
    KF_ASSERT(it != mRegisteredCameras.end(), "Trying to set a camera that does not belong to the scene!");
    mCurrentCamera = camera;
}

void Scene::addPointLight(const std::shared_ptr<PointLight>& light) {
    if (pPointLights.size() >= global::maxPointLights)
        KF_WARN("Reached max point light number. The light will not be added!");
    else
        pPointLights.push_back(light);
}
void Scene::removePointLight(const std::shared_ptr<PointLight>& light) {
    KF_ASSERT(light, "Deleting an invalid light!");
    pPointLights.erase(std::remove(pPointLights.begin(), pPointLights.end(), light), pPointLights.end());
}
void Scene::addActiveLight(const std::shared_ptr<ActiveLight>& light) {
    if (pActiveLights.size() >= global::maxActiveLights) {
        KF_WARN("Reached max active light number. The light will not be added!");
    } else {
        pActiveLights.push_back(light);
        markGeometriesChanged();
    }
}
void Scene::removeActiveLight(const std::shared_ptr<ActiveLight>& light) {
    KF_ASSERT(light, "Deleting an invalid light!");
    pActiveLights.erase(std::remove(pActiveLights.begin(), pActiveLights.end(), light), pActiveLights.end());
}

int Scene::addTexture(const std::vector<uint8_t>& rgba8, uint32_t width, uint32_t height) {
    if (rgba8.size() != 4ull * width * height) throw std::runtime_error("addTexture: size mismatch");
    mTextures.push_back(rgba8);
    mTextureSizes.emplace_back(width, height);
    markGeometriesChanged();
    return (int)mTextures.size() - 1;
}

void Scene::init() {
    markGeometriesChanged();
    markGeometryInstancesChanged();
    initialized = true;
}

// Texture reference of a material / light: "mem:<index>" selects a texture
// added with addTexture(); file paths need an image decoder, which this build
// does not ship, so they fall back to the constants like a failed load in
// scene.cpp:335-412.
static int texture_ref(const std::string& path, size_t n_tex, const char* what) {
    if (path.empty()) return -1;
    if (path.rfind("mem:", 0) == 0) {
        int i = std::atoi(path.c_str() + 4);
        if (i >= 0 && (size_t)i < n_tex) return i;
    }
    KF_WARN(std::string("Failed to load ") + what + " texture: " + path + ", constant value will be used!");
    return -1;
}

static void fill_lights(const Scene& s, std::shared_ptr<DirectionalLight> dl,
                        const std::vector<std::shared_ptr<PointLight>>& pls,
                        const std::vector<std::shared_ptr<ActiveLight>>& als, size_t n_tex, kf_dir_light& d,
                        kf_point_lights& p, kf_active_lights& a) {
    (void)s;
    memset(&d, 0, sizeof(d));
    memset(&p, 0, sizeof(p));
    memset(&a, 0, sizeof(a));
    if (dl) {
        glm::vec3 n = glm::normalize(dl->direction);
        float dd[4] = {n.x, n.y, n.z, dl->softness}, rg[4] = {dl->color.x, dl->color.y, dl->color.z, dl->strength};
        memcpy(d.direction, dd, sizeof(dd));
        memcpy(d.rgbs, rg, sizeof(rg));
    }
    for (size_t i = 0; i < pls.size() && i < global::maxPointLights; ++i) {
        KF_ASSERT(pls[i], "Invalid point light!");
        float pr[4] = {pls[i]->position.x, pls[i]->position.y, pls[i]->position.z, pls[i]->radius};
        float rg[4] = {pls[i]->color.x, pls[i]->color.y, pls[i]->color.z, pls[i]->strength};
        memcpy(p.posr[i], pr, sizeof(pr));
        memcpy(p.rgbs[i], rg, sizeof(rg));
    }
    for (size_t i = 0; i < als.size() && i < global::maxActiveLights; ++i) {
        KF_ASSERT(als[i], "Invalid active light!");
        auto& L = *als[i];
        if (!L.texPath.empty()) {
            L.texID = texture_ref(L.texPath, n_tex, "active light");
            if (L.texID < 0) KF_WARN("Active light texture unavailable, degrade to spot light!");
        } else if (L.texID >= (int)n_tex) {
            L.texID = -1;
        }
        if (L.softness > 0) KF_WARN("FIXME: softness of active light is incorrectly implemented!");
        float col[3] = {L.color.x, L.color.y, L.color.z};
        kf_ok(kf_active_light_fill(&a, (int)i, glm::value_ptr(L.viewMat), col, L.fov, L.strength, L.softness,
                                   L.texID));
    }
}

// 64-bit FNV-1a over the triangles and material indices of a geometry.
static uint64_t topology_hash(const Geometry& g) {
    uint64_t h = 1469598103934665603ull;
    auto mix = [&](const std::vector<uint32_t>& v) {
        for (uint32_t x : v)
            for (int k = 0; k < 4; ++k) {
                h ^= (x >> (8 * k)) & 0xffu;
                h *= 1099511628211ull;
            }
    };
    mix(g.indices);
    mix(g.matIndex);
    return h;
}

static Scene::DeviceGeometry device_geometry(const Geometry& g) {
    return {&g, g.vertices.size(), g.indices.size(), topology_hash(g), g.isOpaque, g.hideRender};
}

// The global material list in the NiceMaterialSSBO layout (scene.cpp:354-400).
static std::vector<kf_material> device_materials(size_t n_tex) {
    std::vector<kf_material> mats;
    for (auto& m : global::materials) {
        kf_material s{};
        float d[4] = {m.diffuseColor.x, m.diffuseColor.y, m.diffuseColor.z, 0};
        float e[4] = {m.emission.x, m.emission.y, m.emission.z, m.emissionStrength};
        memcpy(s.diffuse, d, sizeof(d));
        memcpy(s.emission, e, sizeof(e));
        s.alpha = m.alpha;
        s.metallic = m.metallic;
        s.specular = m.specular;
        s.roughness = m.roughness;
        s.ior = m.ior;
        s.transmission = m.transmission;
        s.diffuse_tex = texture_ref(m.diffuseTexPath, n_tex, "diffuse");
        s.metallic_tex = texture_ref(m.metallicTexPath, n_tex, "metallic");
        s.roughness_tex = texture_ref(m.roughnessTexPath, n_tex, "roughness");
        s.transmission_tex = texture_ref(m.transmissionTexPath, n_tex, "transmission");
        mats.push_back(s);
    }
    return mats;
}

static std::vector<kf_instance> device_instances(const std::vector<std::shared_ptr<GeometryInstance>>& instances) {
    std::vector<kf_instance> inst;
    for (auto& gi : instances) {
        kf_instance k{};
        memcpy(k.transform, glm::value_ptr(gi->transform), sizeof(k.transform));
        k.geometry_index = (uint32_t)gi->geometryIndex;
        inst.push_back(k);
    }
    return inst;
}

void Scene::buildDevice() {
    // flatten geometries (reference layouts) and the global material list
    std::vector<uint32_t> voff{0}, toff{0}, flags;
    std::vector<kf_vertex> verts;
    std::vector<uint32_t> idx, mi;
    for (auto& g : mGeometries) {
        size_t nt = g->indices.size() / 3;
        verts.resize(verts.size() + g->vertices.size());
        memcpy(&verts[verts.size() - g->vertices.size()], g->vertices.data(), g->vertices.size() * sizeof(Vertex));
        idx.insert(idx.end(), g->indices.begin(), g->indices.begin() + 3 * nt);
        if (g->matIndex.size() < nt) throw std::runtime_error("Geometry without a material per triangle");
        mi.insert(mi.end(), g->matIndex.begin(), g->matIndex.begin() + nt);
        voff.push_back((uint32_t)verts.size());
        toff.push_back((uint32_t)(idx.size() / 3));
        flags.push_back((g->isOpaque ? KF_GEOM_OPAQUE : 0u) | (g->hideRender ? KF_GEOM_HIDDEN : 0u));
        g->initialized = true;
    }
    size_t n_tex = mTextures.size();
    std::vector<kf_material> mats = device_materials(n_tex);
    std::vector<kf_instance> inst = device_instances(mGeometryInstances);
    std::vector<kf_texture_desc> tex;
    for (size_t i = 0; i < n_tex; ++i) tex.push_back({mTextureSizes[i].first, mTextureSizes[i].second, mTextures[i].data()});
    std::vector<kf_texture_desc> faces;
    for (auto& f : mEnvFaces) faces.push_back({mEnvSize, mEnvSize, f.data()});
    kf_dir_light d;
    kf_point_lights p;
    kf_active_lights a;
    fill_lights(*this, pDirectionalLight, pPointLights, pActiveLights, n_tex, d, p, a);

    kf_scene_desc desc{};
    desc.num_geometries = (uint32_t)mGeometries.size();
    desc.geom_vertex_offset = voff.data();
    desc.geom_tri_offset = toff.data();
    desc.geom_flags = flags.data();
    desc.vertices = verts.data();
    desc.indices = idx.data();
    desc.mat_index = mi.data();
    desc.num_instances = (uint32_t)inst.size();
    desc.instances = inst.data();
    desc.num_materials = (uint32_t)mats.size();
    desc.materials = mats.data();
    desc.num_textures = (uint32_t)tex.size();
    desc.textures = tex.data();
    desc.dir_light = &d;
    desc.point_lights = &p;
    desc.active_lights = &a;
    desc.env_kind = (uint32_t)mEnvKind;
    desc.env_faces = faces.empty() ? nullptr : faces.data();
    memcpy(desc.env_gradient, mEnvGradient, sizeof(mEnvGradient));
    kf_scene* s = nullptr;
    kf_ok(kf_scene_create(&desc, &s));
    if (mDevice) kf_scene_destroy(mDevice);
    mDevice = s;
    mUploadGeometries = false;
    mUploadGeometryInstancesToBuffer = false;
    ++mDeviceBuilds;
    mDeviceGeometries.clear();
    for (auto& g : mGeometries) mDeviceGeometries.push_back(device_geometry(*g));
    mDeviceMaterials = global::materials.size();
    mDeviceTextures = mTextures.size();
    mDeviceEnvVersion = mEnvVersion;
}

bool Scene::refitDevice() {
    if (mGeometries.size() != mDeviceGeometries.size() || global::materials.size() != mDeviceMaterials ||
        mTextures.size() != mDeviceTextures || mEnvVersion != mDeviceEnvVersion)
        return false;
    for (size_t i = 0; i < mGeometries.size(); ++i) {
        const DeviceGeometry now = device_geometry(*mGeometries[i]), was = mDeviceGeometries[i];
        if (now.geometry != was.geometry || now.vertices != was.vertices || now.indices != was.indices ||
            now.topology != was.topology || now.opaque != was.opaque || now.hidden != was.hidden)
            return false;
    }
    for (size_t i = 0; i < mGeometries.size(); ++i) {
        Geometry& g = *mGeometries[i];
        if (g.initialized) continue;
        kf_ok(kf_scene_update_geometry(mDevice, (uint32_t)i, reinterpret_cast<const kf_vertex*>(g.vertices.data()),
                                       (uint32_t)g.vertices.size()));
        g.initialized = true;
    }
    std::vector<kf_material> mats = device_materials(mTextures.size());
    kf_ok(kf_scene_update_materials(mDevice, mats.data(), (uint32_t)mats.size()));
    mUploadGeometries = false;
    return true;
}

void Scene::syncDevice() {
    if (!mDevice || (mUploadGeometries && !refitDevice())) {
        buildDevice();
        return;
    }
    updateDevice();  // the TLAS (instances may have changed) and the lights
    mUploadGeometryInstancesToBuffer = false;
}

void Scene::updateDevice() {
    std::vector<kf_instance> inst = device_instances(mGeometryInstances);
    kf_ok(kf_scene_update_instances(mDevice, inst.data(), (uint32_t)inst.size()));
    kf_dir_light d;
    kf_point_lights p;
    kf_active_lights a;
    fill_lights(*this, pDirectionalLight, pPointLights, pActiveLights, mTextures.size(), d, p, a);
    kf_ok(kf_scene_set_lights(mDevice, &d, &p, &a));
}

}  // namespace kuafu
