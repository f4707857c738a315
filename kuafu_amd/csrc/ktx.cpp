// ktx.cpp — reader for KTX cube maps of 8-bit RGBA texels (host code).
//
// Replaces the libktx call behind vkCore::Cubemap::init (vkCore.hpp:1852-1970),
// which loads the environment map of Scene::setEnvironmentMap (scene.cpp:301-303)
// and uploads base-level faces as R8G8B8A8_SRGB whatever the file says.  Only
// what that path can use is read: a cube map (6 faces, no array layers, depth
// 1) of square faces, uncompressed 8-bit RGBA, base level.  Both container
// versions of the Khronos specification are accepted:
//   KTX 1: 12-byte identifier, 13 u32 header fields (endianness, glType,
//          glTypeSize, glFormat, glInternalFormat, glBaseInternalFormat,
//          width, height, depth, array elements, faces, mip levels, bytes of
//          key/value data), key/value data, then per level a u32 image size
//          (one face for a non-array cube map) and the faces, each padded to
//          4 bytes;
//   KTX 2: 12-byte identifier, u32 vkFormat, typeSize, width, height, depth,
//          layerCount, faceCount, levelCount, supercompressionScheme, the
//          u32/u64 index (DFD, KVD, SGD), then a level index of u64 (offset,
//          length, uncompressed length) per level; level 0 holds the faces in
//          order.
// Faces come out in the cube-map layer order +X, -X, +Y, -Y, +Z, -Z.
#include <cstdint>
#include <cstring>
#include <fstream>
#include <iterator>
#include <string>
#include <vector>

#include "kf_internal.h"

namespace kf {
namespace {

constexpr uint8_t kId1[12] = {0xAB, 'K', 'T', 'X', ' ', '1', '1', 0xBB, '\r', '\n', 0x1A, '\n'};
constexpr uint8_t kId2[12] = {0xAB, 'K', 'T', 'X', ' ', '2', '0', 0xBB, '\r', '\n', 0x1A, '\n'};

constexpr uint32_t GL_UNSIGNED_BYTE = 0x1401, GL_RGBA = 0x1908, GL_RGBA8 = 0x8058, GL_SRGB8_ALPHA8 = 0x8C43;
constexpr uint32_t VK_FORMAT_R8G8B8A8_UNORM = 37, VK_FORMAT_R8G8B8A8_SRGB = 43;

uint32_t u32(const std::vector<uint8_t>& b, size_t at, bool swap) {
    if (at + 4 > b.size()) throw Error("KTX: file truncated");
    uint32_t v;
    memcpy(&v, &b[at], 4);
    return swap ? __builtin_bswap32(v) : v;
}
uint64_t u64(const std::vector<uint8_t>& b, size_t at) {
    if (at + 8 > b.size()) throw Error("KTX: file truncated");
    uint64_t v;
    memcpy(&v, &b[at], 8);
    return v;
}

struct Cube {
    uint32_t size = 0;
    std::vector<uint8_t> faces;  // 6 * size * size * 4
};

Cube parse(const std::vector<uint8_t>& b) {
    if (b.size() < 12) throw Error("KTX: not a KTX file");
    Cube c;
    if (memcmp(b.data(), kId1, 12) == 0) {
        const uint32_t endian = u32(b, 12, false);
        if (endian != 0x04030201u && endian != 0x01020304u) throw Error("KTX: bad endianness field");
        const bool sw = endian == 0x01020304u;
        const uint32_t gl_type = u32(b, 16, sw), gl_format = u32(b, 24, sw), gl_internal = u32(b, 28, sw);
        const uint32_t w = u32(b, 36, sw), h = u32(b, 40, sw), d = u32(b, 44, sw);
        const uint32_t layers = u32(b, 48, sw), nfaces = u32(b, 52, sw), kvd = u32(b, 60, sw);
        if (gl_type != GL_UNSIGNED_BYTE || gl_format != GL_RGBA ||
            (gl_internal != GL_RGBA8 && gl_internal != GL_SRGB8_ALPHA8))
            throw Error("KTX: only uncompressed 8-bit RGBA cube maps are supported");
        if (nfaces != 6 || layers != 0 || d != 0 || w != h || w == 0)
            throw Error("KTX: expected a cube map of 6 square faces");
        size_t at = 64 + (size_t)kvd;
        const uint32_t image = u32(b, at, sw);  // bytes of one face
        at += 4;
        const size_t face = 4ull * w * h;
        if (image != face) throw Error("KTX: unexpected image size");
        c.size = w;
        c.faces.resize(6 * face);
        for (int f = 0; f < 6; ++f) {
            if (at + face > b.size()) throw Error("KTX: file truncated");
            memcpy(&c.faces[f * face], &b[at], face);
            at += (face + 3) & ~size_t(3);
        }
        return c;
    }
    if (memcmp(b.data(), kId2, 12) == 0) {
        const uint32_t fmt = u32(b, 12, false), w = u32(b, 20, false), h = u32(b, 24, false);
        const uint32_t d = u32(b, 28, false), layers = u32(b, 32, false), nfaces = u32(b, 36, false);
        const uint32_t levels = u32(b, 40, false), scheme = u32(b, 44, false);
        if (fmt != VK_FORMAT_R8G8B8A8_SRGB && fmt != VK_FORMAT_R8G8B8A8_UNORM)
            throw Error("KTX: only uncompressed 8-bit RGBA cube maps are supported");
        if (scheme != 0) throw Error("KTX: supercompressed files are not supported");
        if (nfaces != 6 || layers != 0 || d != 0 || w != h || w == 0)
            throw Error("KTX: expected a cube map of 6 square faces");
        (void)levels;
        // level index after the 48-byte header and the 32-byte index
        const uint64_t off = u64(b, 80), len = u64(b, 88);
        const size_t face = 4ull * w * h;
        if (len < 6 * face || off + 6 * face > b.size()) throw Error("KTX: file truncated");
        c.size = w;
        c.faces.assign(b.begin() + (ptrdiff_t)off, b.begin() + (ptrdiff_t)(off + 6 * face));
        return c;
    }
    throw Error("KTX: not a KTX file");
}

}  // namespace
}  // namespace kf

using namespace kf;

extern "C" int kf_ktx_cubemap_read(const char* path, uint32_t* size, uint8_t* faces, uint64_t max_bytes) {
    try {
        if (!path || !size) throw Error("kf_ktx_cubemap_read: null argument");
        std::ifstream f(path, std::ios::binary);
        if (!f) throw Error(std::string("KTX: cannot open ") + path);
        std::vector<uint8_t> b((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
        const Cube c = parse(b);
        *size = c.size;
        if (faces) {
            if (max_bytes < c.faces.size()) throw Error("kf_ktx_cubemap_read: buffer too small");
            memcpy(faces, c.faces.data(), c.faces.size());
        }
        return 0;
    } catch (const std::exception& e) {
        set_error(e.what());
        return -1;
    }
}
