// Process-global state of the kuafu:: API (global.hpp:27-60 of the reference):
// the frame counter of the running mean and the append-only material list.
#pragma once
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace kuafu {
struct NiceMaterial;

namespace global {
extern int frameCount;
extern std::string assetsPath;
extern uint32_t materialIndex;
extern uint32_t textureIndex;
extern std::vector<NiceMaterial> materials;

const size_t maxResources = 2;
const size_t maxPointLights = 32;
const size_t maxActiveLights = 8;
}  // namespace global

void logInfo(const std::string& m);
void logWarn(const std::string& m);
}  // namespace kuafu

#define KF_INFO(msg) ::kuafu::logInfo(msg)
#define KF_WARN(msg) ::kuafu::logWarn(msg)
#define KF_ASSERT(cond, msg)                              \
    do {                                                  \
        if (!(cond)) throw std::runtime_error(msg);       \
    } while (0)
