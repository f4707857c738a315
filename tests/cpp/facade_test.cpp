// Test program of the kuafu:: C++ facade (include/kuafu.hpp).
//   facade_test cpu          API semantics that need no GPU; prints OK
//   facade_test gpu <dir>    renders through Kuafu::run() and writes
//                            <dir>/scene.json (the scene, read back through
//                            public fields only), rgba.bin, bgra.bin
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>

#include "kuafu.hpp"

using namespace kuafu;

#define CHECK(c)                                                                      \
    do {                                                                              \
        if (!(c)) {                                                                   \
            std::fprintf(stderr, "CHECK failed %s:%d: %s\n", __FILE__, __LINE__, #c); \
            std::exit(1);                                                             \
        }                                                                             \
    } while (0)

template <class F> static bool throws(F f) {
    try {
        f();
    } catch (const std::exception&) {
        return true;
    }
    return false;
}

static bool outward(const Geometry& g, glm::vec3 center) {
    for (size_t t = 0; t + 2 < g.indices.size(); t += 3) {
        glm::vec3 a = g.vertices[g.indices[t]].pos, b = g.vertices[g.indices[t + 1]].pos,
                  c = g.vertices[g.indices[t + 2]].pos;
        glm::vec3 n = glm::cross(b - a, c - a);
        if (glm::dot(n, (a + b + c) * (1.0f / 3.0f) - center) <= 0) return false;
    }
    return true;
}

static int cpu_checks() {
    Config cfg;
    CHECK(cfg.getPathDepth() == 8 && cfg.getPerPixelSampleRate() == 32 && cfg.getRussianRoulette());
    CHECK(cfg.getRussianRouletteMinBounces() == 4 && !cfg.isUsingDenoiser() && cfg.isAccumulatingFrames());
    cfg.setPathDepth(20);  // the reference keeps the unclamped value (config.cpp:31-41)
    CHECK(cfg.getPathDepth() == 20);
    cfg.setPresent(false);
    auto pc = std::make_shared<Config>(cfg);
    Kuafu k(pc);  // no device work until run()
    Scene* s = k.getScene();
    CHECK(s && s->getCamera());
    CHECK(k.getConfig().getPathDepth() == 20);

    // generators: counts, outward winding, one material each
    size_t m0 = global::materials.size();
    auto plane = createYZPlane(true);
    auto cube = createCube(true);
    auto sphere = createSphere(true);
    auto caps = createCapsule(1.0f, 0.5f, true);
    CHECK(plane->indices.size() == 6 && cube->indices.size() == 36);
    CHECK(sphere->indices.size() / 3 == 4900);
    CHECK(global::materials.size() == m0 + 4);
    CHECK(outward(*cube, {0, 0, 0}) && outward(*sphere, {0, 0, 0}));
    glm::vec3 n = glm::cross(plane->vertices[plane->indices[1]].pos - plane->vertices[plane->indices[0]].pos,
                             plane->vertices[plane->indices[2]].pos - plane->vertices[plane->indices[0]].pos);
    CHECK(n.x > 0 && n.y == 0 && n.z == 0);  // YZ plane faces +X like its vertex normals
    for (auto& v : caps->vertices) CHECK(std::fabs(glm::length(v.normal) - 1.0f) < 1e-4f);
    for (auto& v : sphere->vertices) CHECK(std::fabs(glm::length(v.pos) - 1.0f) < 1e-5f);

    // scene bookkeeping and limits (scene.cpp:58-120)
    s->submitGeometry(cube);
    auto gi = instance(cube, glm::translate(glm::mat4(1.0f), {1, 2, 3}));
    s->submitGeometryInstance(gi);
    CHECK(gi->geometryIndex == 0 && s->getGeometryInstances().size() == 1);
    CHECK(throws([&] { s->submitGeometryInstance(instance(sphere)); }));  // geometry not submitted
    CHECK(throws([&] { s->removeGeometry(cube); }));                      // "Not implemented!"
    CHECK(throws([&] { s->getGeometryInstance(5); }));
    Config small;
    small.setGeometryLimit(16);  // -> 17 slots
    small.setPresent(false);
    Kuafu k2(std::make_shared<Config>(small));
    for (int i = 0; i < 17; ++i) k2.getScene()->submitGeometry(cube);
    CHECK(throws([&] { k2.getScene()->submitGeometry(cube); }));
    for (int i = 0; i < 40; ++i) s->addPointLight(std::make_shared<PointLight>());  // extra lights are dropped

    // camera: pose round trip and pinhole projection (camera.cpp:92-139)
    Camera* cam = s->createCamera(64, 48);
    s->setCamera(cam);
    cam->setFullPerspective(64, 48, 50, 52, 30, 20, 0);
    cam->setPosition({1, 2, 3});
    cam->setFront({0, 0, -1});
    cam->setUp({0, 1, 0});
    glm::vec4 pv = cam->getViewMatrix() * glm::vec4(1.5f, 2.5f, 1.0f, 1.0f);
    glm::vec4 clip = cam->getProjectionMatrix() * pv;
    float px = (clip.x / clip.w + 1) / 2 * 64, py = (clip.y / clip.w + 1) / 2 * 48;
    CHECK(std::fabs(px - (30 + 50 * 0.5f / 2.0f)) < 1e-3f && std::fabs(py - (20 - 52 * 0.5f / 2.0f)) < 1e-3f);
    global::frameCount = 5;
    cam->setPose(cam->getPose());
    CHECK(global::frameCount == -1);
    CHECK(throws([&] { cam->downloadLatestFrame(); }));  // nothing rendered yet
    CHECK(throws([&] { k.setWindow(100, 100); }));

    // OBJ import (triangulated, a vertex per face corner, flat normals
    // generated when absent, assimp's default material)
    {
        std::ofstream f("/tmp/kf_facade_test.obj");
        f << "v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nvt 0 0\nvt 1 0\nvt 1 1\nvt 0 1\nf 1/1 2/2 3/3 4/4\n";
    }
    auto obj = loadObj("/tmp/kf_facade_test.obj");
    CHECK(obj->indices.size() == 6 && obj->vertices.size() == 6);
    CHECK(std::fabs(obj->vertices[0].normal.z - 1.0f) < 1e-6f && obj->vertices[2].texCoord.y == 0.0f);
    CHECK(obj->vertices[5].pos == glm::vec3(0, 1, 0) && obj->vertices[3].pos == glm::vec3(0, 0, 0));
    {
        const NiceMaterial& m = global::materials[obj->matIndex[0]];
        CHECK(m.diffuseColor == glm::vec3(0.6f) && m.ior == 1.0f && m.roughness == 1.0f && m.specular == 0.5f &&
              m.emission == glm::vec3(0.0f) && m.emissionStrength == 1.0f && m.alpha == 1.0f && obj->isOpaque);
    }
    CHECK(throws([&] { loadObj("/tmp/does_not_exist.obj"); }));

    // OBJ + MTL: one mesh per material in the library's order, materials read
    // from Kd / Ns / d / Ni / map_Kd, shared between loads of the same file
    {
        std::ofstream m("/tmp/kf_facade_two.mtl");
        m << "# two materials and an unused one\nnewmtl red\nKd 0.8 0.1 0.1\nNs 105\n\n"
             "newmtl glass\nKd 1 1 1\nd 0.5\nNi 1.5\nKe 0.5 0.25 0\nmap_Kd -s 1 1 1 tex/glass.png\n"
             "newmtl unused\nKd 0 0 1\nTr 0.25\n";
        std::ofstream f("/tmp/kf_facade_two.obj");
        f << "mtllib kf_facade_two.mtl\nv 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nvn 0 0 1\n"
             "usemtl glass\nf 1//1 3//1 4//1\nusemtl red\nf 1//1 2//1 3//1\nusemtl glass\nf 2 3 4\n";
    }
    size_t nmat = global::materials.size();
    auto two = loadScene("/tmp/kf_facade_two.obj", false);
    CHECK(two.size() == 2 && global::materials.size() == nmat + 2);
    CHECK(two[0]->indices.size() == 3 && two[1]->indices.size() == 6 && two[1]->vertices.size() == 6);
    {
        const NiceMaterial& red = global::materials[two[0]->matIndex[0]];
        const NiceMaterial& glass = global::materials[two[1]->matIndex[0]];
        CHECK(red.diffuseColor == glm::vec3(0.8f, 0.1f, 0.1f) && std::fabs(red.roughness - 0.75f) < 1e-6f &&
              red.ior == 1.0f && red.alpha == 1.0f && red.diffuseTexPath.empty() && two[0]->isOpaque);
        CHECK(glass.alpha == 0.5f && glass.ior == 1.5f && glass.roughness == 1.0f && !two[1]->isOpaque &&
              glass.emission == glm::vec3(0.5f, 0.25f, 0.0f) && glass.emissionStrength == 1.0f);
        CHECK(glass.diffuseTexPath == "/tmp/tex/glass.png");
    }
    CHECK(two[1]->vertices[0].normal == glm::vec3(0, 0, 1) && two[1]->vertices[5].normal == glm::vec3(0, 0, 1));
    auto again = loadScene("/tmp/kf_facade_two.obj", false);
    CHECK(global::materials.size() == nmat + 2 && again[0]->matIndex[0] == two[0]->matIndex[0] &&
          again[1]->matIndex[0] == two[1]->matIndex[0]);
    CHECK(throws([&] { loadObj("/tmp/kf_facade_two.obj"); }));  // complex scene

    // STL import: ASCII (two solids -> one mesh of their shared material) and
    // binary (80-byte header, facet count, 50-byte facets); three vertices of
    // its own per facet carrying the facet normal; assimp's default material
    {
        std::ofstream f("/tmp/kf_facade_test.stl");
        f << "solid a\n facet normal 0 0 1\n  outer loop\n   vertex 0 0 0\n   vertex 1 0 0\n   vertex 1 1 0\n"
             "  endloop\n endfacet\n facet normal 0 0 1\n  outer loop\n   vertex 0 0 0\n   vertex 1 1 0\n"
             "   vertex 0 1 0\n  endloop\n endfacet\nendsolid a\nsolid b\n facet normal 1 0 0\n  outer loop\n"
             "   vertex 2 0 0\n   vertex 2 1 0\n   vertex 2 0 1\n  endloop\n endfacet\nendsolid b\n";
    }
    nmat = global::materials.size();
    auto stl = loadScene("/tmp/kf_facade_test.stl", false);
    CHECK(stl.size() == 1 && stl[0]->vertices.size() == 9 && stl[0]->indices.size() == 9);
    CHECK(stl[0]->vertices[8].pos.z == 1.0f && stl[0]->vertices[6].normal.x == 1.0f);
    CHECK(stl[0]->vertices[4].pos.x == 1.0f && stl[0]->vertices[4].pos.y == 1.0f && stl[0]->vertices[5].normal.z == 1.0f);
    CHECK(stl[0]->matIndex.size() == 3 && stl[0]->matIndex[2] == stl[0]->matIndex[0] && stl[0]->isOpaque);
    CHECK(loadScene("/tmp/kf_facade_test.stl", false)[0]->matIndex[0] == stl[0]->matIndex[0]);
    CHECK(global::materials.size() <= nmat + 1);
    {
        const NiceMaterial& m = global::materials[stl[0]->matIndex[0]];
        CHECK(m.diffuseColor == glm::vec3(0.6f) && m.roughness == 0.0f && m.specular == 0.5f && m.ior == 1.4f &&
              m.emissionStrength == 1.0f && m.emission == glm::vec3(0.0f));
    }
    {
        std::ofstream f("/tmp/kf_facade_test_bin.STL", std::ios::binary);
        char header[80] = "solid binary files may start with solid too";
        f.write(header, 80);
        const uint32_t n = 1;
        f.write((const char*)&n, 4);
        const float v[12] = {0, 1, 0, 0, 0, 0, 0, 0, 1, 1, 0, 0};  // normal, then three vertices
        const uint16_t attr = 0;
        f.write((const char*)v, sizeof(v));
        f.write((const char*)&attr, 2);
    }
    auto bin = loadObj("/tmp/kf_facade_test_bin.STL");
    CHECK(bin->vertices.size() == 3 && bin->indices.size() == 3 && bin->vertices[0].normal.y == 1.0f);
    CHECK(bin->vertices[1].pos.z == 1.0f && bin->vertices[2].pos.x == 1.0f);
    {
        std::ofstream f("/tmp/kf_facade_test_bad.stl");
        f << "solid c\n facet normal 0 0 1\n  outer loop\n   vertex 0 0 0\n   vertex 1 0 0\n  endloop\n endfacet\n";
    }
    CHECK(throws([&] { loadScene("/tmp/kf_facade_test_bad.stl", false); }));
    {
        std::ofstream f("/tmp/kf_facade_test_cut.stl");
        f << "solid d\n facet normal 0 0 1\n  outer loop\n   vertex 0 0 0\n   vertex 1 0 0\n";
    }
    CHECK(throws([&] { loadScene("/tmp/kf_facade_test_cut.stl", false); }));
    CHECK(throws([&] { loadScene("/tmp/kf_facade_test.ply", false); }));

    // KTX environment map (KTX 1: identifier, 13 u32 header fields, no
    // key/value data, image size, 6 faces of 1x1 RGBA8)
    {
        const uint8_t id[12] = {0xAB, 'K', 'T', 'X', ' ', '1', '1', 0xBB, '\r', '\n', 0x1A, '\n'};
        const uint32_t hdr[13] = {0x04030201u, 0x1401u, 1u, 0x1908u, 0x8C43u, 0x1908u, 1u, 1u, 0u, 0u, 6u, 1u, 0u};
        const uint32_t image = 4;
        std::ofstream f("/tmp/kf_facade_test.ktx", std::ios::binary);
        f.write((const char*)id, 12);
        f.write((const char*)hdr, sizeof(hdr));
        f.write((const char*)&image, 4);
        for (int face = 0; face < 6; ++face) {
            const uint8_t px[4] = {(uint8_t)(40 * face), 0, 255, 255};
            f.write((const char*)px, 4);
        }
    }
    s->setEnvironmentMap("/tmp/kf_facade_test.ktx");
    CHECK(throws([&] { s->setEnvironmentMap("/tmp/kf_facade_test.png"); }));  // scene.cpp:306-308
    CHECK(throws([&] { s->setEnvironmentMap("/tmp/does_not_exist.ktx"); }));
    std::cout << "OK" << std::endl;
    return 0;
}

static void json_vec(std::ostream& o, const float* v, int n) {
    o << "[";
    for (int i = 0; i < n; ++i) o << (i ? "," : "") << v[i];
    o << "]";
}

// The scene, read back through public fields only, as JSON for the Python
// side to rebuild over the C ABI.
static void dump_scene(const std::string& dir, Scene* s, const DirectionalLight& dl, const PointLight& pl,
                       const ActiveLight& al, Camera* cam, int frames, int first_frame = 0) {
    std::ofstream j(dir + "/scene.json");
    j.precision(9);
    j << "{\"geometries\":[";
    for (size_t g = 0; g < s->getGeometries().size(); ++g) {
        auto& G = *s->getGeometries()[g];
        j << (g ? "," : "") << "{\"pos\":[";
        for (size_t v = 0; v < G.vertices.size(); ++v) {
            j << (v ? "," : "");
            json_vec(j, &G.vertices[v].pos.x, 3);
        }
        j << "],\"normal\":[";
        for (size_t v = 0; v < G.vertices.size(); ++v) {
            j << (v ? "," : "");
            json_vec(j, &G.vertices[v].normal.x, 3);
        }
        j << "],\"indices\":[";
        for (size_t i = 0; i < G.indices.size(); ++i) j << (i ? "," : "") << G.indices[i];
        j << "],\"material\":" << G.matIndex[0] << "}";
    }
    j << "],\"instances\":[";
    for (size_t i = 0; i < s->getGeometryInstances().size(); ++i) {
        auto& I = *s->getGeometryInstances()[i];
        j << (i ? "," : "") << "{\"geometry\":" << I.geometryIndex << ",\"transform\":";
        json_vec(j, glm::value_ptr(I.transform), 16);
        j << "}";
    }
    j << "],\"materials\":[";
    for (size_t m = 0; m < global::materials.size(); ++m) {
        auto& M = global::materials[m];
        j << (m ? "," : "") << "{\"diffuse\":";
        json_vec(j, &M.diffuseColor.x, 3);
        j << ",\"metallic\":" << M.metallic << ",\"specular\":" << M.specular << ",\"roughness\":" << M.roughness
          << ",\"ior\":" << M.ior << ",\"transmission\":" << M.transmission << ",\"alpha\":" << M.alpha
          << ",\"emission\":";
        json_vec(j, &M.emission.x, 3);
        j << ",\"emissionStrength\":" << M.emissionStrength << "}";
    }
    j << "],\"dir_light\":{\"direction\":";
    json_vec(j, &dl.direction.x, 3);
    j << ",\"color\":";
    json_vec(j, &dl.color.x, 3);
    j << ",\"strength\":" << dl.strength << ",\"softness\":" << dl.softness << "},\"point_light\":{\"position\":";
    json_vec(j, &pl.position.x, 3);
    j << ",\"color\":";
    json_vec(j, &pl.color.x, 3);
    j << ",\"strength\":" << pl.strength << ",\"radius\":" << pl.radius << "},\"active_light\":{\"view\":";
    json_vec(j, glm::value_ptr(al.viewMat), 16);
    j << ",\"color\":";
    json_vec(j, &al.color.x, 3);
    j << ",\"fov\":" << al.fov << ",\"strength\":" << al.strength << "},\"clear_color\":";
    glm::vec4 cc = s->getClearColor();
    json_vec(j, &cc.x, 4);
    j << ",\"camera\":{\"position\":";
    json_vec(j, &cam->getPosition().x, 3);
    j << ",\"front\":";
    json_vec(j, &cam->getFront().x, 3);
    j << ",\"up\":[0,0,1],\"width\":" << cam->getWidth() << ",\"height\":" << cam->getHeight()
      << ",\"fx\":" << cam->getFocalX() << ",\"fy\":" << cam->getFocalY() << ",\"cx\":" << cam->getPrincipalPointX()
      << ",\"cy\":" << cam->getPrincipalPointY() << "},\"spp\":4,\"path_depth\":4,\"frames\":" << frames
      << ",\"first_frame\":" << first_frame << "}";
    j.close();
}

static int gpu_run(const std::string& dir) {
    Config c;
    c.setPresent(false);
    c.setPerPixelSampleRate(4);
    c.setPathDepth(4);
    c.setGeometryLimit(64);
    auto k = std::make_unique<Kuafu>(std::make_shared<Config>(c));
    Scene* s = k->getScene();
    NiceMaterial floorM;
    floorM.diffuseColor = {0.7f, 0.7f, 0.7f};
    floorM.roughness = 0.8f;
    auto floor = createYZPlane(true, floorM);
    NiceMaterial red;
    red.diffuseColor = {0.8f, 0.2f, 0.2f};
    auto cube = createCube(true, red);
    NiceMaterial metal;
    metal.metallic = 1.0f;
    metal.roughness = 0.3f;
    auto sphere = createSphere(true, metal);
    NiceMaterial glass;
    glass.transmission = 1.0f;
    glass.ior = 1.45f;
    auto caps = createCapsule(0.6f, 0.3f, true, glass);
    s->setGeometries({floor, cube, sphere, caps});
    glm::mat4 ft = glm::translate(glm::mat4(1.0f), {0, 0, -0.5f});
    ft = glm::rotate(ft, glm::radians(90.f), {0., -1., 0.});
    ft = glm::scale(ft, glm::vec3(6.0f));
    s->setGeometryInstances({instance(floor, ft), instance(cube, glm::scale(glm::translate(glm::mat4(1.0f), {0, -1.5f, 0}), glm::vec3(0.4f))),
                             instance(sphere, glm::scale(glm::translate(glm::mat4(1.0f), {0.2f, 0.2f, 0.1f}), glm::vec3(0.5f))),
                             instance(caps, glm::rotate(glm::translate(glm::mat4(1.0f), {0.5f, 1.5f, 0.2f}), 0.7f, {0, 0, 1}))});
    auto dl = std::make_shared<DirectionalLight>();
    dl->direction = {-0.3f, 0.2f, -1.0f};
    dl->strength = 2.0f;
    s->setDirectionalLight(dl);
    auto pl = std::make_shared<PointLight>();
    pl->position = {1, -1, 2};
    pl->strength = 6;
    pl->radius = 0;
    s->addPointLight(pl);
    auto al = std::make_shared<ActiveLight>();
    al->viewMat = glm::lookAt(glm::vec3(3, 0, 3), glm::vec3(0, 0, 0), glm::vec3(0, 0, 1));
    al->fov = 0.6f;
    al->strength = 20.0f;
    s->addActiveLight(al);
    s->setClearColor({0.6f, 0.7f, 0.8f, 0.5f});
    Camera* cam = s->createCamera(96, 64);
    s->setCamera(cam);
    cam->setFullPerspective(96, 64, 80, 80, 48, 32, 0);
    cam->setPosition({-4.0f, 0.0f, 1.5f});
    cam->setFront(glm::normalize(glm::vec3(1.0f, 0.0f, -0.3f)));
    cam->setUp({0, 0, 1});
    k->run();
    k->run();
    CHECK(global::frameCount == 1);
    auto bgra = k->downloadLatestFrame(cam);
    auto rgba = cam->downloadTarget(0);
    CHECK(bgra.size() == 96u * 64u * 4u && rgba.size() == 96u * 64u * 4u);
    {
        std::ofstream f(dir + "/rgba.bin", std::ios::binary);
        f.write((const char*)rgba.data(), rgba.size() * 4);
        std::ofstream g(dir + "/bgra.bin", std::ios::binary);
        g.write((const char*)bgra.data(), bgra.size());
    }
    dump_scene(dir, s, *dl, *pl, *al, cam, 2);
    // the denoise stage through the facade (random weights unless a weight file ships)
    k->getConfig().setUseDenoiser(true);
    k->run();
    auto den = cam->downloadTarget(0);
    for (float v : den) CHECK(std::isfinite(v));
    std::cout << "OK" << std::endl;
    return 0;
}

// Incremental device updates: instance changes rebuild the TLAS only, a
// geometry re-marked uninitialized with the same triangles has its BLAS refit
// and materials are re-uploaded, all on the device scene built first; a new
// geometry makes a full build.  The final frame and scene are dumped for the
// Python side to compare with a scene built from scratch over the C ABI.
static int gpu_update(const std::string& dir) {
    Config c;
    c.setPresent(false);
    c.setPerPixelSampleRate(4);
    c.setPathDepth(4);
    c.setGeometryLimit(64);
    c.setAccumulatingFrames(false);
    auto k = std::make_unique<Kuafu>(std::make_shared<Config>(c));
    Scene* s = k->getScene();
    NiceMaterial floorM;
    floorM.diffuseColor = {0.7f, 0.7f, 0.7f};
    floorM.roughness = 0.8f;
    auto floor = createYZPlane(true, floorM);
    NiceMaterial red;
    red.diffuseColor = {0.8f, 0.2f, 0.2f};
    auto cube = createCube(true, red);
    NiceMaterial metal;
    metal.metallic = 1.0f;
    metal.roughness = 0.3f;
    auto sphere = createSphere(true, metal);
    s->setGeometries({floor, cube, sphere});
    glm::mat4 ft = glm::translate(glm::mat4(1.0f), {0, 0, -0.5f});
    ft = glm::rotate(ft, glm::radians(90.f), {0., -1., 0.});
    ft = glm::scale(ft, glm::vec3(6.0f));
    s->setGeometryInstances({instance(floor, ft), instance(cube, glm::scale(glm::translate(glm::mat4(1.0f), {0, -1.5f, 0}), glm::vec3(0.4f))),
                             instance(sphere, glm::scale(glm::translate(glm::mat4(1.0f), {0.2f, 0.2f, 0.1f}), glm::vec3(0.5f)))});
    auto dl = std::make_shared<DirectionalLight>();
    dl->direction = {-0.3f, 0.2f, -1.0f};
    dl->strength = 2.0f;
    s->setDirectionalLight(dl);
    auto pl = std::make_shared<PointLight>();
    pl->position = {1, -1, 2};
    pl->strength = 6;
    pl->radius = 0;
    s->addPointLight(pl);
    auto al = std::make_shared<ActiveLight>();
    al->viewMat = glm::lookAt(glm::vec3(3, 0, 3), glm::vec3(0, 0, 0), glm::vec3(0, 0, 1));
    al->fov = 0.6f;
    al->strength = 20.0f;
    s->addActiveLight(al);
    s->setClearColor({0.6f, 0.7f, 0.8f, 0.5f});
    Camera* cam = s->createCamera(96, 64);
    s->setCamera(cam);
    cam->setFullPerspective(96, 64, 80, 80, 48, 32, 0);
    cam->setPosition({-4.0f, 0.0f, 1.5f});
    cam->setFront(glm::normalize(glm::vec3(1.0f, 0.0f, -0.3f)));
    cam->setUp({0, 0, 1});
    k->run();
    CHECK(s->getDeviceBuilds() == 1);
    const kf_scene* dev = s->getDeviceScene();
    // instances: one moved, one added -> TLAS only
    s->getGeometryInstance(1)->setTransform(glm::scale(glm::translate(glm::mat4(1.0f), {0.3f, -1.2f, 0.1f}), glm::vec3(0.5f)));
    s->submitGeometryInstance(instance(cube, glm::scale(glm::translate(glm::mat4(1.0f), {0.4f, 1.2f, 0.0f}), glm::vec3(0.3f))));
    k->run();
    CHECK(s->getDeviceBuilds() == 1 && s->getDeviceScene() == dev);
    // the sphere deformed in place and a material edited -> refit + material upload
    for (auto& v : sphere->vertices) {
        v.pos.x *= 1.4f;
        v.pos.z += 0.2f * v.pos.y * v.pos.y;
    }
    sphere->initialized = false;
    global::materials[cube->matIndex[0]].diffuseColor = {0.1f, 0.6f, 0.9f};
    s->markGeometriesChanged();
    k->run();
    CHECK(s->getDeviceBuilds() == 1 && s->getDeviceScene() == dev && sphere->initialized);
    auto rgba = cam->downloadTarget(0);
    {
        std::ofstream f(dir + "/rgba.bin", std::ios::binary);
        f.write((const char*)rgba.data(), rgba.size() * 4);
    }
    dump_scene(dir, s, *dl, *pl, *al, cam, 1, global::frameCount);  // -1: frames are not accumulated
    // a new geometry -> full build
    NiceMaterial green;
    green.diffuseColor = {0.2f, 0.8f, 0.2f};
    auto ball = createSphere(true, green);
    s->submitGeometry(ball);
    s->submitGeometryInstance(instance(ball, glm::scale(glm::translate(glm::mat4(1.0f), {-0.5f, 0.5f, 0.0f}), glm::vec3(0.3f))));
    k->run();
    CHECK(s->getDeviceBuilds() == 2);
    std::cout << "OK" << std::endl;
    return 0;
}

int main(int argc, char** argv) {
    try {
        if (argc >= 2 && std::string(argv[1]) == "cpu") return cpu_checks();
        if (argc >= 3 && std::string(argv[1]) == "gpu") return gpu_run(argv[2]);
        if (argc >= 3 && std::string(argv[1]) == "gpu_update") return gpu_update(argv[2]);
        std::fprintf(stderr, "usage: facade_test cpu | gpu <dir> | gpu_update <dir>\n");
        return 2;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "exception: %s\n", e.what());
        return 1;
    }
}
