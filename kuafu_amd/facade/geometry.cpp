// Mesh data of the kuafu:: API: materials, instances and the generators of
// geometry.cpp:241-537 in the reference (same surfaces: the YZ plane with
// normal +X, the [-1,1]^3 cube, the 50x50 latitude-longitude sphere about the
// X axis, the capsule along X), written as parametric generators.  Every
// triangle is wound so that cross(v1 - v0, v2 - v0) points out of the surface.
#include <cctype>
#include <filesystem>
#include <cmath>
#include <cstring>
#include <fstream>
#include <iterator>
#include <sstream>
#include <stdexcept>

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

// Index of a material equal to m in the process-global list, appended when
// there is none (the duplicate lookup of geometry.cpp:155-175).
static uint32_t intern_material(const NiceMaterial& m) {
    for (size_t j = 0; j < global::materials.size(); ++j)
        if (global::materials[j] == m) return (uint32_t)j;
    global::materials.push_back(m);
    return global::materialIndex++;
}

// One imported mesh: a Geometry of its own vertices and triangles with one
// material, as loadScene builds it (geometry.cpp:181-229).
static std::shared_ptr<Geometry> imported(std::vector<Vertex> v, std::vector<uint32_t> idx, uint32_t mat,
                                          const std::string& path, bool dynamic) {
    auto g = std::make_shared<Geometry>();
    g->path = path;
    g->dynamic = dynamic;
    g->vertices = std::move(v);
    g->indices = std::move(idx);
    g->initialized = false;
    g->matIndex.assign(g->indices.size() / 3, mat);
    g->isOpaque = global::materials[mat].alpha >= 1.0F;
    return g;
}

// STL, binary (80-byte header, u32 facet count, 50 bytes per facet: normal,
// three vertices, u16 attribute) or ASCII (solid / facet normal / outer loop /
// vertex x3 / endloop / endfacet / endsolid).  As assimp imports it for the
// reference's loadScene (geometry.cpp:46-232): three vertices of its own per
// facet carrying the facet normal (the import flags do not join identical
// vertices, and GenNormals keeps normals a file has), no UVs, and one default
// material whose only property is a diffuse colour of 0.6 grey; the
// reference turns it into diffuse 0.6, specular 0.5, IOR 1.4, roughness 0
// (no shininess and not an OBJ file), opaque.  Every solid uses that one
// material, so aiProcess_PreTransformVertices, which merges the meshes of a
// material, returns one mesh for the file.  Colours that some binary
// exporters pack into the attribute bytes are not read.
static std::vector<std::shared_ptr<Geometry>> loadStl(const std::string& path, bool dynamic) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("Failed to load scene: cannot open " + path);
    std::string data((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    std::vector<glm::vec3> facets;  // normal, v0, v1, v2 per facet, solids one after another
    uint32_t n = 0;
    if (data.size() >= 84) std::memcpy(&n, data.data() + 80, 4);
    if (data.size() >= 84 && data.size() == 84 + 50ull * n) {
        for (uint32_t f = 0; f < n; ++f) {
            float v[12];
            std::memcpy(v, data.data() + 84 + 50ull * f, sizeof(v));
            for (int k = 0; k < 4; ++k) facets.push_back(glm::vec3(v[3 * k], v[3 * k + 1], v[3 * k + 2]));
        }
    } else {
        std::istringstream ss(data);
        std::string tok;
        if (!(ss >> tok) || tok != "solid") throw std::runtime_error("Failed to load scene: not an STL file, " + path);
        glm::vec3 nrm(0.0f);
        int nv = -1;  // vertices of the open facet; -1: none open
        while (ss >> tok) {
            if (tok == "facet") {
                ss >> tok;  // "normal"
                if (tok != "normal" || !(ss >> nrm.x >> nrm.y >> nrm.z))
                    throw std::runtime_error("Failed to load scene: bad facet in " + path);
                facets.push_back(nrm);
                nv = 0;
            } else if (tok == "vertex") {
                glm::vec3 p;
                if (nv < 0 || !(ss >> p.x >> p.y >> p.z) || nv >= 3)
                    throw std::runtime_error("Failed to load scene: bad vertex in " + path);
                facets.push_back(p);
                ++nv;
            } else if (tok == "endfacet") {
                if (nv != 3) throw std::runtime_error("Failed to load scene: facet without 3 vertices in " + path);
                nv = -1;
            }
        }
        if (facets.size() % 4)  // a facet cut off by the end of the file
            throw std::runtime_error("Failed to load scene: truncated facet in " + path);
    }
    if (facets.empty()) throw std::runtime_error("A mesh in the file has no triangles: " + path);
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
    std::vector<Vertex> verts;
    std::vector<uint32_t> idx;
    for (size_t f = 0; f + 4 <= facets.size(); f += 4)
        for (int k = 1; k <= 3; ++k) {
            Vertex vx{};
            vx.pos = facets[f + k];
            vx.normal = facets[f];
            idx.push_back((uint32_t)verts.size());
            verts.push_back(vx);
        }
    return {imported(std::move(verts), std::move(idx), intern_material(m), path, dynamic)};
}

// ------------------------------------------------------------------ OBJ + MTL
// Wavefront OBJ with its MTL library, as assimp (5.0.1, the version the
// reference's CMakeLists asks for) imports them for loadScene with
// Triangulate | GenNormals | FlipUVs | PreTransformVertices
// (geometry.cpp:50-54), then converted by geometry.cpp:78-150:
//   * materials: assimp's list starts with its default material, then the
//     MTL's newmtl entries in file order, then names that usemtl asks for but
//     the library lacks (with default values).  A material's defaults are
//     Kd 0.6 grey, Ke 0, d 1, Ns 0, Ni 1;  Kd, Ke, d (opacity), Tr (alpha =
//     1 - Tr), Ns (shininess), Ni (IOR) and map_Kd are read.  The reference
//     makes diffuseColor = Kd, emission = Ke with strength 1, alpha = opacity,
//     ior = Ni, specular 0.5, metallic 0, transmission 0, roughness from the
//     shininess (Ns <= 5: 1, else 1 - 0.025 sqrt(Ns - 5), at least 0.001; no
//     Ns: 1 for an OBJ file) and the diffuse texture path relative to the
//     OBJ's directory; equal materials are shared through the global list.
//   * meshes: faces before any usemtl use the default material; every face
//     corner is a vertex of its own (the OBJ importer does not share them);
//     polygons are fanned; v of the UVs flipped; a face without normals gets
//     its flat face normal normalize(cross(v1 - v0, v2 - v0)) (GenNormals on
//     unshared vertices).  PreTransformVertices merges the meshes of each
//     material: one Geometry per material that has faces, in material order.
namespace {
struct ObjMaterial {
    std::string name;
    glm::vec3 kd = glm::vec3(0.6f), ke = glm::vec3(0.0f);
    float d = 1.0f, ns = 0.0f, ni = 1.0f;
    std::string map_kd;
};

std::string trim(const std::string& t) {
    size_t a = t.find_first_not_of(" \t\r\n"), b = t.find_last_not_of(" \t\r\n");
    return a == std::string::npos ? std::string() : t.substr(a, b - a + 1);
}

// The file of a map_* statement: its last token (options such as -s u v or
// -clamp on come first).
std::string map_file(std::istringstream& ss) {
    std::string tok, last;
    while (ss >> tok) last = tok;
    return last;
}

void read_mtl(const std::filesystem::path& file, std::vector<ObjMaterial>& mats) {
    std::ifstream in(file);
    if (!in) {
        KF_WARN("OBJ material library " + file.string() + " not found, using default materials");
        return;
    }
    std::string line;
    ObjMaterial* cur = nullptr;
    while (std::getline(in, line)) {
        std::istringstream ss(line);
        std::string tag;
        if (!(ss >> tag) || tag[0] == '#') continue;
        if (tag == "newmtl") {
            std::string rest;
            std::getline(ss, rest);
            mats.push_back(ObjMaterial());
            mats.back().name = trim(rest);
            cur = &mats.back();
        } else if (!cur) {
            continue;
        } else if (tag == "Kd") {
            ss >> cur->kd.x >> cur->kd.y >> cur->kd.z;
        } else if (tag == "Ke") {
            ss >> cur->ke.x >> cur->ke.y >> cur->ke.z;
        } else if (tag == "d") {
            ss >> cur->d;
        } else if (tag == "Tr") {
            float tr = 0.0f;
            if (ss >> tr) cur->d = 1.0f - tr;
        } else if (tag == "Ns") {
            ss >> cur->ns;
        } else if (tag == "Ni") {
            ss >> cur->ni;
        } else if (tag == "map_Kd") {
            cur->map_kd = map_file(ss);
        }
    }
}

NiceMaterial convert(const ObjMaterial& om, const std::filesystem::path& dir) {
    NiceMaterial m;
    m.diffuseColor = om.kd;
    m.emission = om.ke;
    m.emissionStrength = 1.0f;
    m.alpha = om.d;
    m.ior = om.ni;
    m.specular = 0.5f;
    m.metallic = 0.0f;
    m.transmission = 0.0f;
    if (om.ns > 0.0f) {
        m.roughness = om.ns <= 5.0f ? 1.0f : 1.0f - std::sqrt(om.ns - 5.0f) * 0.025f;
        if (m.roughness < 0.001f) m.roughness = 0.001f;
    } else {
        m.roughness = 1.0f;
    }
    if (!om.map_kd.empty()) {
        std::filesystem::path t(om.map_kd);
        m.diffuseTexPath = t.is_absolute() ? t.string() : (dir / t).string();
    }
    return m;
}
}  // namespace

static std::vector<std::shared_ptr<Geometry>> loadObjFile(const std::string& path, bool dynamic) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("Failed to load scene: cannot open " + path);
    const std::filesystem::path dir = std::filesystem::absolute(path).parent_path();
    std::vector<ObjMaterial> mats(1);  // [0]: assimp's default material
    mats[0].name = "DefaultMaterial";
    std::vector<glm::vec3> P, N;
    std::vector<glm::vec2> T;
    struct Corner {
        int v, t, n;
    };
    std::vector<std::vector<Corner>> faces;  // triangles
    std::vector<uint32_t> face_mat;
    uint32_t cur = 0;
    std::string line;
    auto fix = [](int i, size_t n) { return i < 0 ? (int)n + i : i - 1; };
    while (std::getline(in, line)) {
        std::istringstream ss(line);
        std::string tag;
        if (!(ss >> tag) || tag[0] == '#') continue;
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
        } else if (tag == "mtllib") {
            std::string rest;
            std::getline(ss, rest);
            std::filesystem::path lib(trim(rest));
            read_mtl(lib.is_absolute() ? lib : dir / lib, mats);
        } else if (tag == "usemtl") {
            std::string rest;
            std::getline(ss, rest);
            const std::string name = trim(rest);
            uint32_t k = 0;
            while (k < mats.size() && mats[k].name != name) ++k;
            if (k == mats.size()) {
                KF_WARN("OBJ material " + name + " not found for " + path + ", using default values");
                mats.push_back(ObjMaterial());
                mats.back().name = name;
            }
            cur = k;
        } else if (tag == "f") {
            std::vector<Corner> poly;
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
                Corner c{fix(vi, P.size()), ti ? fix(ti, T.size()) : -1, ni ? fix(ni, N.size()) : -1};
                if (c.v < 0 || c.v >= (int)P.size()) throw std::runtime_error("Failed to load scene: bad face in " + path);
                if (c.t >= (int)T.size()) c.t = -1;
                if (c.n >= (int)N.size()) c.n = -1;
                poly.push_back(c);
            }
            for (size_t k = 1; k + 1 < poly.size(); ++k) {
                faces.push_back({poly[0], poly[k], poly[k + 1]});
                face_mat.push_back(cur);
            }
        }
    }
    std::vector<std::shared_ptr<Geometry>> out;
    for (uint32_t m = 0; m < mats.size(); ++m) {
        std::vector<Vertex> verts;
        std::vector<uint32_t> idx;
        for (size_t f = 0; f < faces.size(); ++f) {
            if (face_mat[f] != m) continue;
            const auto& c = faces[f];
            const glm::vec3 flat = glm::normalize(glm::cross(P[c[1].v] - P[c[0].v], P[c[2].v] - P[c[0].v]));
            const bool has_n = c[0].n >= 0 && c[1].n >= 0 && c[2].n >= 0;
            for (int k = 0; k < 3; ++k) {
                Vertex vx{};
                vx.pos = P[c[k].v];
                if (c[k].t >= 0) vx.texCoord = T[c[k].t];
                vx.normal = has_n ? N[c[k].n] : flat;
                idx.push_back((uint32_t)verts.size());
                verts.push_back(vx);
            }
        }
        if (idx.empty()) continue;
        out.push_back(imported(std::move(verts), std::move(idx), intern_material(convert(mats[m], dir)), path,
                               dynamic));
    }
    if (out.empty()) throw std::runtime_error("A mesh in the file has no triangles: " + path);
    return out;
}

// loadScene (geometry.cpp:46-232) for Wavefront OBJ (+ MTL) and STL files.
std::vector<std::shared_ptr<Geometry>> loadScene(std::string_view fname, bool dynamic) {
    std::string path(fname);
    std::string ext = path.size() >= 4 ? path.substr(path.size() - 4) : std::string();
    for (auto& ch : ext) ch = (char)std::tolower((unsigned char)ch);
    if (ext == ".stl") return loadStl(path, dynamic);
    if (ext != ".obj")
        throw std::runtime_error("Failed to load scene: only .obj and .stl are supported in this build, " + path);
    return loadObjFile(path, dynamic);
}

std::shared_ptr<Geometry> loadObj(std::string_view path, bool dynamic) {
    auto s = loadScene(path, dynamic);
    if (s.size() != 1) throw std::runtime_error("complex scene! use loadScene");
    return s.front();
}

}  // namespace kuafu
