// Geometry data of the kuafu:: API (geometry.hpp:11-108 of the reference):
// NiceMaterial, Geometry, GeometryInstance and the mesh generators, with the
// reference's public fields and defaults.
#pragma once
#include <memory>
#include <string>
#include <string_view>
#include <vector>

#include "context/global.hpp"
#include "context/vertex.hpp"

namespace kuafu {

struct NiceMaterial {
    glm::vec3 diffuseColor = glm::vec3(1.0F);
    float alpha = 1.0F;

    std::string diffuseTexPath;
    std::string metallicTexPath;
    std::string roughnessTexPath;
    std::string transmissionTexPath;

    float metallic = 0.0F;
    float specular = 0.5F;
    float roughness = 0.5F;
    float ior = 1.4F;
    float transmission = 0.0F;

    glm::vec3 emission = glm::vec3(1.0F);
    float emissionStrength = 0.0F;

    friend bool operator==(const NiceMaterial& m1, const NiceMaterial& m2);
};

struct Geometry {
    void setMaterial(const NiceMaterial& material);
    void recalculateNormals();

    std::vector<Vertex> vertices;    ///< all vertices of the geometry
    std::vector<uint32_t> indices;   ///< 3 per triangle
    std::vector<uint32_t> matIndex;  ///< global material index per triangle
    std::string path;
    bool initialized = false;
    bool dynamic = false;
    bool isOpaque = true;
    bool hideRender = false;
};

struct GeometryInstance {
    void setTransform(const glm::mat4& transform);

    glm::mat4 transform = glm::mat4(1.0F);  ///< object -> world
    int geometryIndex = -1;
    std::shared_ptr<Geometry> geometry = nullptr;
};

/// Wavefront OBJ import (triangulated faces, v/vt/vn); other formats need assimp.
std::vector<std::shared_ptr<Geometry>> loadScene(std::string_view fname, bool dynamic);
std::shared_ptr<Geometry> loadObj(std::string_view path, bool dynamic = false);

std::shared_ptr<GeometryInstance> instance(const std::shared_ptr<Geometry>& geometry,
                                           const glm::mat4& transform = glm::mat4(1.0F));

std::shared_ptr<Geometry> createYZPlane(bool dynamic = true, NiceMaterial mat = {});
std::shared_ptr<Geometry> createCube(bool dynamic = true, NiceMaterial mat = {});
std::shared_ptr<Geometry> createSphere(bool dynamic = true, NiceMaterial mat = {});
std::shared_ptr<Geometry> createCapsule(float halfHeight = 1., float radius = 1., bool dynamic = true,
                                        NiceMaterial mat = {});

}  // namespace kuafu
