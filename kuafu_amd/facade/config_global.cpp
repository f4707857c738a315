// Config setters (config.cpp:31-155 of the reference) and process globals.
#include <algorithm>
#include <cstdio>

#include "../../include/core/config.hpp"
#include "../../include/core/context/global.hpp"
#include "../../include/core/geometry.hpp"

namespace kuafu {

namespace global {
int frameCount = -1;
std::string assetsPath;
uint32_t materialIndex = 0;
uint32_t textureIndex = 0;
std::vector<NiceMaterial> materials;
}  // namespace global

void logInfo(const std::string& m) { std::fprintf(stderr, "[kuafu] [info] %s\n", m.c_str()); }
void logWarn(const std::string& m) { std::fprintf(stderr, "[kuafu] [warning] %s\n", m.c_str()); }

std::string Config::sDefaultAssetsPath;
void Config::setDefaultAssetsPath(std::string p) { sDefaultAssetsPath = std::move(p); }

void Config::setPathDepth(uint32_t recursionDepth) {
    // the reference clamps, warns, then assigns the unclamped value (config.cpp:31-41)
    if (recursionDepth > mMaxPathDepth) KF_WARN("Exceeded maximum path depth. Using highest possible value instead.");
    mPathDepth = recursionDepth;
}
void Config::setNextEventEstimation(bool flag) { mNextEventEstimation = flag; }
void Config::setNextEventEstimationMinBounces(uint32_t minBounces) { mNextEventEstimationMinBounces = minBounces; }
void Config::setRussianRoulette(bool flag) { mRussianRoulette = flag; }
void Config::setRussianRouletteMinBounces(uint32_t minBounces) { mRussianRouletteMinBounces = minBounces; }

void Config::setAssetsPath(int argc, char* argv[]) {
    mAssetsPath.clear();
    for (int i = 0; i < argc; ++i) mAssetsPath += argv[i];
    std::replace(mAssetsPath.begin(), mAssetsPath.end(), '\\', '/');
    mAssetsPath = mAssetsPath.substr(0, mAssetsPath.find_last_of('/') + 1);
    global::assetsPath = mAssetsPath;
}
void Config::setAssetsPath(std::string_view path) {
    if (path.empty()) return;
    mAssetsPath = path;
    std::replace(mAssetsPath.begin(), mAssetsPath.end(), '\\', '/');
    if (mAssetsPath.back() != '/') mAssetsPath += '/';
    global::assetsPath = mAssetsPath;
}
void Config::setAutomaticPipelineRefresh(bool flag) { mAutomaticPipelineRefresh = flag; }

void Config::setGeometryInstanceLimit(uint32_t amount) {
    if (amount == 0) {
        ++amount;
        KF_WARN("Can not use value 0 for the maximum amount of geometry instances. Using 1 instead.");
    }
    mMaxGeometryInstances = amount;
    mMaxGeometryInstancesChanged = true;
}
void Config::setGeometryLimit(size_t amount) {
    if (amount < 16 || amount % 4 != 0) amount = 16;
    mMaxGeometry = ++amount;
    mMaxGeometryChanged = true;
}
void Config::setTextureLimit(size_t amount) {
    if (amount == 0) KF_WARN("Can not use value 0 for the maximum amount of textures. Using 1 instead.");
    mMaxTextures = ++amount;
    mMaxTexturesChanged = true;
}
void Config::setMaterialLimit(size_t amount) {
    if (amount == 0) KF_WARN("Can not use value 0 for the maximum amount of materials. Using 1 instead.");
    mMaxMaterials = ++amount;
}
void Config::setUseDenoiser(bool useDenoiser) { mUseDenoiser = useDenoiser; }
void Config::setPerPixelSampleRate(uint32_t sampleRate) { mPerPixelSampleRate = sampleRate; }
void Config::setAccumulatingFrames(bool flag) { mAccumulateFrames = flag; }
void Config::updateVariance(bool flag) { mUpdateVariance = flag; }

}  // namespace kuafu
