// Vertex of the kuafu:: API (vertex.hpp:28-34 of the reference): 48 bytes,
// unpacked on the device as 3 x vec4 (PathTrace.rchit:52-64).
#pragma once
#include <functional>

#include "../../kuafu_glm.hpp"

namespace kuafu {

struct Vertex {
    glm::vec3 pos;       ///< position
    glm::vec3 normal;    ///< normal
    glm::vec3 color;     ///< colour
    glm::vec2 texCoord;  ///< texture coordinate
    float padding0 = 0;  ///< padding to 48 bytes

    bool operator==(const Vertex& o) const {
        return pos == o.pos && color == o.color && texCoord == o.texCoord && normal == o.normal;
    }
};
static_assert(sizeof(Vertex) == 48, "Vertex must keep the 48-byte layout");

}  // namespace kuafu

namespace std {
template <> struct hash<kuafu::Vertex> {
    size_t operator()(const kuafu::Vertex& v) const {
        size_t a = hash<glm::vec3>()(v.pos), b = hash<glm::vec3>()(v.color), c = hash<glm::vec2>()(v.texCoord),
               d = hash<glm::vec3>()(v.normal);
        return ((((a ^ (b << 1)) >> 1) ^ c) << 1) ^ d;
    }
};
}  // namespace std
