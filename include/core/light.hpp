// Lights of the kuafu:: API with the reference's public fields and defaults
// (light.hpp:9-39).  Up to 32 point lights and 8 active lights per scene.
#pragma once
#include <string>

#include "../kuafu_glm.hpp"

namespace kuafu {

struct DirectionalLight {
    glm::vec3 direction = {0, 0, -1};
    glm::vec3 color = {1, 1, 1};
    float strength = 0.0;
    float softness = 0.0;
};

struct PointLight {
    glm::vec3 position = {0, 0, 0};
    glm::vec3 color = {1, 1, 1};
    float radius = 1;
    float strength = 0.0;
};

/// Spot light / projector: a camera-like light with a view matrix, a square
/// frustum of `fov` (radians) and an optional pattern texture.
struct ActiveLight {
    glm::mat4 viewMat{};
    glm::vec3 color = {1, 1, 1};
    float fov = 0.0;
    float strength = 0.0;
    float softness = 0.0;
    std::string texPath;
    int texID = -1;
};

}  // namespace kuafu
