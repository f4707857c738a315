// Mesh data of the kuafu:: API: materials, instances and the generators of
// geometry.cpp:241-537 in the reference (same surfaces: the YZ plane with
// normal +X, the [-1,1]^3 cube, the 50x50 latitude-longitude sphere about the
// X axis, the capsule along X), written as parametric generators.  Every
// triangle is wound so that cross(v1 - v0, v2 - v0) points out of the surface.
#include <cctype>
#include <cmath>
#include <cstring>
#include <fstream>
#include <iterator>
#include <map>
#include <sstream>
#include <stdexcept>
#include <tuple>

#include "../../include/core/geometry.hpp"

namespace kuafu {

bool operator==(const NiceMaterial& a, const NiceMaterial& b) {
    return a.diffuseColor == b.diffuseColor && a.alpha == b.alpha && a.diffuseTexPath == b.diffuseTexPath &&
           a.metallicTexPath == b.metallicTexPath && a.roughnessTexPath == b.roughnessTexPath &&
           a.transmissionTexPath == b.transmissionTexPath && a.metallic == b.metallic && a.specular == b.specular &&
           a.roughness == b.roughness && a.ior == b.ior && a.emission == b.emission &&
           a.emissionStrength == b.emissionStrength;
}

// Geometry::setMaterial (geometry.cpp:241-249): append to the global list and
// point every triangle at it.
void Geometry::setMaterial(const NiceMaterial& material) {
    isOpaque = material.alpha >= 1;
    global::materials.push_back(material);
    global::materialIndex++;
    matIndex.assign(indices.size() / 3, global::materialIndex - 1);
}

void GeometryInstance::setTransform(const glm::mat4& t) { transform = t; }

std::shared_ptr<GeometryInstance> instance(const std::shared_ptr<Geometry>& geometry, const glm::mat4& transform) {
    if (!geometry) throw std::runtime_error("instance(): null geometry");
    auto r = std::make_shared<GeometryInstance>();
    r->geometry = geometry;
    r->transform = transform;
    return r;
}

// Area-weighted face normals accumulated on the vertices (geometry.cpp:259-290).
void Geometry::recalculateNormals() {
    for (auto& v : vertices) v.normal = glm::vec3(0.0f);
    for (size_t i = 0; i + 2 < indices.size(); i += 3) {
        Vertex& a = vertices[indices[i]];
        Vertex& b = vertices[indices[i + 1]];
        Vertex& c = vertices[indices[i + 2]];
        glm::vec3 n = glm::cross(b.pos - a.pos, c.pos - a.pos);
        float l = glm::length(n);
        if (!(l > 0.0f)) continue;
        n = n / l;
        a.normal += n;
        b.normal += n;
        c.normal += n;
    }
    for (auto& v : vertices)
        if (v.normal != glm::vec3(0.0f)) v.normal = glm::normalize(v.normal);
}

static std::shared_ptr<Geometry> finish(std::shared_ptr<Geometry> g, bool dynamic, NiceMaterial mat) {
    g->path = "";
    g->initialized = false;
    g->dynamic = dynamic;
    g->isOpaque = mat.alpha >= 1.0F;
    g->matIndex.assign(g->indices.size() / 3, global::materialIndex++);
    global::materials.push_back(std::move(mat));
    return g;
}

static Vertex vtx(glm::vec3 p, glm::vec3 n, glm::vec2 uv) {
    Vertex v{};
    v.pos = p;
    v.normal = n;
    v.texCoord = uv;
    return v;
}

std::shared_ptr<Geometry> createYZPlane(bool dynamic, NiceMaterial mat) {
    auto g = std::make_shared<Geometry>();
    const glm::vec3 n(1, 0, 0);
    // corners (y, z) counter-clockwise seen from +X
    g->vertices = {vtx({0, -1, -1}, n, {1, 1}), vtx({0, 1, -1}, n, {0, 1}), vtx({0, 1, 1}, n, {1, 0}),
                   vtx({0, -1, 1}, n, {0, 0})};
    g->indices = {0, 1, 2, 0, 2, 3};
    return finish(g, dynamic, std::move(mat));
}

std::shared_ptr<Geometry> createCube(bool dynamic, NiceMaterial mat) {
    auto g = std::make_shared<Geometry>();
    // one quad per face: normal n, in-plane axes u, v with cross(u, v) = n
    const int faces[6][3][3] = {{{1, 0, 0}, {0, 1, 0}, {0, 0, 1}},  {{-1, 0, 0}, {0, 0, 1}, {0, 1, 0}},
                                {{0, 1, 0}, {0, 0, 1}, {1, 0, 0}},  {{0, -1, 0}, {1, 0, 0}, {0, 0, 1}},
                                {{0, 0, 1}, {1, 0, 0}, {0, 1, 0}},  {{0, 0, -1}, {0, 1, 0}, {1, 0, 0}}};
    for (auto& f : faces) {
        glm::vec3 n(f[0][0], f[0][1], f[0][2]), u(f[1][0], f[1][1], f[1][2]), v(f[2][0], f[2][1], f[2][2]);
        uint32_t b = (uint32_t)g->vertices.size();
        const float c[4][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1}};
        for (auto& k : c)
            g->vertices.push_back(vtx(n + u * k[0] + v * k[1], n, {0.5f * (k[0] + 1), 0.5f * (k[1] + 1)}));
        g->indices.insert(g->indices.end(), {b, b + 1, b + 2, b, b + 2, b + 3});
    }
    return finish(g, dynamic, std::move(mat));
}

std::shared_ptr<Geometry> createSphere(bool dynamic, NiceMaterial mat) {
    auto g = std::make_shared<Geometry>();
    const uint32_t stacks = 50, slices = 50;
    const float radius = 1.f, pi = glm::pi<float>();
    // rings at latitude phi around the X axis, poles at -X and +X
    for (uint32_t i = 1; i < stacks; ++i) {
        float phi = pi / stacks * i - pi / 2;
        for (uint32_t j = 0; j < slices; ++j) {
            float th = pi * 2 / slices * j;
            g->vertices.push_back(vtx({std::sin(phi) * radius, std::cos(th) * std::cos(phi) * radius,
                                       std::sin(th) * std::cos(phi) * radius},
                                      {}, {float(j) / slices, float(i) / stacks}));
        }
    }
    auto ring = [&](uint32_t i, uint32_t j) { return (i - 1) * slices + (j % slices); };
    for (uint32_t i = 1; i + 1 < stacks; ++i)
        for (uint32_t j = 0; j < slices; ++j) {
            uint32_t a = ring(i, j), b = ring(i, j + 1), c = ring(i + 1, j), d = ring(i + 1, j + 1);
            g->indices.insert(g->indices.end(), {a, d, c, a, b, d});
        }
    uint32_t south = (uint32_t)g->vertices.size();
    g->vertices.push_back(vtx({-radius, 0, 0}, {}, {0.5f, 0.0f}));
    uint32_t north = (uint32_t)g->vertices.size();
    g->vertices.push_back(vtx({radius, 0, 0}, {}, {0.5f, 1.0f}));
    for (uint32_t j = 0; j < slices; ++j) {
        g->indices.insert(g->indices.end(), {south, ring(1, j + 1), ring(1, j)});
        g->indices.insert(g->indices.end(), {north, ring(stacks - 1, j), ring(stacks - 1, j + 1)});
    }
    // orient every triangle outward (the sphere is centred at the origin)
    for (size_t t = 0; t < g->indices.size(); t += 3) {
        glm::vec3 a = g->vertices[g->indices[t]].pos, b = g->vertices[g->indices[t + 1]].pos,
                  c = g->vertices[g->indices[t + 2]].pos;
        if (glm::dot(glm::cross(b - a, c - a), a + b + c) < 0) std::swap(g->indices[t + 1], g->indices[t + 2]);
    }
    finish(g, dynamic, std::move(mat));
    g->recalculateNormals();
    return g;
}

std::shared_ptr<Geometry> createCapsule(float halfLength, float radius, bool dynamic, NiceMaterial mat) {
    auto g = std::make_shared<Geometry>();
    const int segments = 32, halfRings = 8, rings = 2 * halfRings;
    const float pi = glm::pi<float>();
    // pole (+X), rings r = 1..rings-1 (the equator ring twice: +X and -X cap), pole (-X)
    uint32_t top = 0;
    g->vertices.push_back(vtx({radius + halfLength, 0, 0}, {1, 0, 0}, {0.5f, 1.0f}));
    std::vector<uint32_t> ringStart;
    for (int r = 1; r <= rings; ++r) {
        int rr = r <= halfRings ? r : r - 1;  // r = halfRings and halfRings+1 both sit on the equator
        float th = pi * rr / rings;
        float x = std::cos(th), yz = std::sin(th);
        float off = r <= halfRings ? halfLength : -halfLength;
        ringStart.push_back((uint32_t)g->vertices.size());
        for (int s = 0; s <= segments; ++s) {
            float ph = pi * 2 * s / segments;
            glm::vec3 n(x, yz * std::cos(ph), yz * std::sin(ph));
            g->vertices.push_back(vtx(n * radius + glm::vec3(off, 0, 0), n,
                                      {float(s) / segments, 1.0f - float(r) / (rings + 1)}));
        }
    }
    uint32_t bottom = (uint32_t)g->vertices.size();
    g->vertices.push_back(vtx({-radius - halfLength, 0, 0}, {-1, 0, 0}, {0.5f, 0.0f}));
    for (int s = 0; s < segments; ++s) g->indices.insert(g->indices.end(), {top, ringStart[0] + s, ringStart[0] + s + 1});
    for (size_t r = 0; r + 1 < ringStart.size(); ++r)
        for (int s = 0; s < segments; ++s) {
            uint32_t a = ringStart[r] + s, b = ringStart[r] + s + 1, c = ringStart[r + 1] + s, d = ringStart[r + 1] + s + 1;
            g->indices.insert(g->indices.end(), {a, c, d, a, d, b});
        }
    uint32_t last = ringStart.back();
    for (int s = 0; s < segments; ++s) g->indices.insert(g->indices.end(), {bottom, last + s + 1, last + s});
    // orient outward: compare with the vertex normals
    for (size_t t = 0; t < g->indices.size(); t += 3) {
        const Vertex &a = g->vertices[g->indices[t]], &b = g->vertices[g->indices[t + 1]], &c = g->vertices[g->indices[t + 2]];
        glm::vec3 n = glm::cross(b.pos - a.pos, c.pos - a.pos);
        if (glm::dot(n, a.normal + b.normal + c.normal) < 0) std::swap(g->indices[t + 1], g->indices[t + 2]);
    }
    return finish(g, dynamic, std::move(mat));
}

// STL, binary (80-byte header, u32 facet count, 50 bytes per facet: normal,
// three vertices, u16 attribute) or ASCII (solid / facet normal / outer loop /
// vertex x3 / endloop / endfacet / endsolid).  As assimp imports it for the
// reference's loadScene (geometry.cpp:46-232): one mesh per solid, three
// vertices of its own per facet carrying the facet normal (the import flags do
// not join identical vertices, and GenNormals keeps normals a file has), no
// UVs, and one default material whose only property is a diffuse colour of
// 0.6 grey; the reference turns it into diffuse 0.6, specular 0.5, IOR 1.4,
// roughness 0 (no shininess and not an OBJ file), opaque.  Colours that some
// binary exporters pack into the attribute bytes are not read.
static std::vector<std::shared_ptr<Geometry>> loadStl(const std::string& path, bool dynamic) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("Failed to load scene: cannot open " + path);
    std::string data((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    std::vector<std::vector<glm::vec3>> solids;  // per solid: normal, v0, v1, v2 per facet
    uint32_t n = 0;
    if (data.size() >= 84) std::memcpy(&n, data.data() + 80, 4);
    if (data.size() >= 84 && data.size() == 84 + 50ull * n) {
        solids.emplace_back();
        for (uint32_t f = 0; f < n; ++f) {
            float v[12];
            std::memcpy(v, data.data() + 84 + 50ull * f, sizeof(v));
            for (int k = 0; k < 4; ++k) solids.back().push_back(glm::vec3(v[3 * k], v[3 * k + 1], v[3 * k + 2]));
        }
    } else {
        std::istringstream ss(data);
        std::string tok;
        if (!(ss >> tok) || tok != "solid") throw std::runtime_error("Failed to load scene: not an STL file, " + path);
        solids.emplace_back();
        glm::vec3 nrm(0.0f);
        int nv = 0;
        while (ss >> tok) {
            if (tok == "solid") {
                solids.emplace_back();
            } else if (tok == "facet") {
                ss >> tok;  // "normal"
                if (tok != "normal" || !(ss >> nrm.x >> nrm.y >> nrm.z))
                    throw std::runtime_error("Failed to load scene: bad facet in " + path);
                solids.back().push_back(nrm);
                nv = 0;
            } else if (tok == "vertex") {
                glm::vec3 p;
                if (solids.back().empty() || !(ss >> p.x >> p.y >> p.z) || nv >= 3)
                    throw std::runtime_error("Failed to load scene: bad vertex in " + path);
                solids.back().push_back(p);
                ++nv;
            } else if (tok == "endfacet") {
                if (nv != 3) throw std::runtime_error("Failed to load scene: facet without 3 vertices in " + path);
            }
        }
        for (auto& sd : solids)  // a facet cut off by the end of the file
            if (sd.size() % 4) throw std::runtime_error("Failed to load scene: truncated facet in " + path);
    }
    NiceMaterial m;
    m.diffuseColor = glm::vec3(0.6f);
    m.specular = 0.5f;
    m.ior = 1.4f;
    m.roughness = 0.0f;
    m.metallic = 0.0f;
    m.transmission = 0.0f;
    m.alpha = 1.0f;
    m.emission = glm::vec3(0.0f);
    m.emissionStrength = 1.0f;
    std::vector<std::shared_ptr<Geometry>> out;
    uint32_t mat = ~0u;
    for (auto& sd : solids) {
        if (sd.size() < 4) continue;  // a solid without facets has no mesh
        auto g = std::make_shared<Geometry>();
        for (size_t f = 0; f + 4 <= sd.size(); f += 4)
            for (int k = 1; k <= 3; ++k) {
                Vertex vx{};
                vx.pos = sd[f + k];
                vx.normal = sd[f];
                g->indices.push_back((uint32_t)g->vertices.size());
                g->vertices.push_back(vx);
            }
        g->initialized = false;
        g->dynamic = dynamic;
        g->isOpaque = true;
        if (mat == ~0u) {  // one material for the file, shared by its solids
            mat = global::materialIndex++;
            global::materials.push_back(m);
        }
        g->matIndex.assign(g->indices.size() / 3, mat);
        g->path = path;
        out.push_back(std::move(g));
    }
    if (out.empty()) throw std::runtime_error("A mesh in the file has no triangles: " + path);
    return out;
}

// Wavefront OBJ: v / vt / vn / f (polygons fanned), UVs flipped vertically and
// missing normals generated, like the reference's assimp flags
// (Triangulate | GenNormals | FlipUVs, geometry.cpp:52-54).  STL: loadStl.
std::vector<std::shared_ptr<Geometry>> loadScene(std::string_view fname, bool dynamic) {
    std::string path(fname);
    std::string ext = path.size() >= 4 ? path.substr(path.size() - 4) : std::string();
    for (auto& ch : ext) ch = (char)std::tolower((unsigned char)ch);
    if (ext == ".stl") return loadStl(path, dynamic);
    if (ext != ".obj")
        throw std::runtime_error("Failed to load scene: only .obj and .stl are supported in this build, " + path);
    std::ifstream in(path);
    if (!in) throw std::runtime_error("Failed to load scene: cannot open " + path);
    std::vector<glm::vec3> P, N;
    std::vector<glm::vec2> T;
    auto g = std::make_shared<Geometry>();
    std::map<std::tuple<int, int, int>, uint32_t> cache;
    bool haveNormals = true;
    std::string line;
    auto fix = [](int i, size_t n) { return i < 0 ? (int)n + i : i - 1; };
    while (std::getline(in, line)) {
        std::istringstream ss(line);
        std::string tag;
        ss >> tag;
        if (tag == "v") {
            glm::vec3 p;
            ss >> p.x >> p.y >> p.z;
            P.push_back(p);
        } else if (tag == "vn") {
            glm::vec3 n;
            ss >> n.x >> n.y >> n.z;
            N.push_back(n);
        } else if (tag == "vt") {
            glm::vec2 t;
            ss >> t.x >> t.y;
            T.push_back(glm::vec2(t.x, 1.0f - t.y));
        } else if (tag == "f") {
            std::vector<uint32_t> poly;
            std::string tok;
            while (ss >> tok) {
                int vi = 0, ti = 0, ni = 0;
                size_t a = tok.find('/');
                vi = std::stoi(tok.substr(0, a));
                if (a != std::string::npos) {
                    size_t b = tok.find('/', a + 1);
                    std::string ts = tok.substr(a + 1, b == std::string::npos ? std::string::npos : b - a - 1);
                    if (!ts.empty()) ti = std::stoi(ts);
                    if (b != std::string::npos) ni = std::stoi(tok.substr(b + 1));
                }
                int v = fix(vi, P.size()), t = ti ? fix(ti, T.size()) : -1, n = ni ? fix(ni, N.size()) : -1;
                if (v < 0 || v >= (int)P.size()) throw std::runtime_error("Failed to load scene: bad face in " + path);
                if (n < 0) haveNormals = false;
                auto key = std::make_tuple(v, t, n);
                auto it = cache.find(key);
                if (it == cache.end()) {
                    Vertex vx{};
                    vx.pos = P[v];
                    if (t >= 0 && t < (int)T.size()) vx.texCoord = T[t];
                    if (n >= 0 && n < (int)N.size()) vx.normal = N[n];
                    it = cache.emplace(key, (uint32_t)g->vertices.size()).first;
                    g->vertices.push_back(vx);
                }
                poly.push_back(it->second);
            }
            for (size_t k = 1; k + 1 < poly.size(); ++k) g->indices.insert(g->indices.end(), {poly[0], poly[k], poly[k + 1]});
        }
    }
    if (g->indices.empty()) throw std::runtime_error("A mesh in the file has no triangles: " + path);
    if (!haveNormals) g->recalculateNormals();
    g->path = path;
    NiceMaterial m;
    m.roughness = 1.0f;  // OBJ without shininess -> roughness 1 (geometry.cpp:118-121)
    finish(g, dynamic, m);
    g->path = path;
    return {g};
}

std::shared_ptr<Geometry> loadObj(std::string_view path, bool dynamic) {
    auto s = loadScene(path, dynamic);
    if (s.size() != 1) throw std::runtime_error("complex scene! use loadScene");
    return s.front();
}

}  // namespace kuafu
