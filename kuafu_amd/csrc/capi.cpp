// capi.cpp — host side of libkuafu_amd: device/stream, scene upload, host
// math of the facade (camera and light UBO contents), render targets.
//
// Scene upload replaces Scene::uploadGeometries / uploadGeometryInstances /
// uploadLightBuffers (scene.cpp:252-467); the data keeps the reference's
// std430 layouts (include/kuafu_amd.h).  Every device index is validated
// here, on the host, before any kernel can read it.
#include "kf_internal.h"

#include <cstdlib>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>

namespace kf {

namespace {
thread_local std::string g_error;
int g_device = -1;
hipStream_t g_stream = nullptr;
std::mutex g_mu;
}  // namespace

void set_error(const std::string& m) { g_error = m; }

uint32_t traversal_refill() {
    // Idle lanes of a wave that make the persistent traversal kernels fetch
    // new rays (1..64); KUAFU_AMD_REFILL overrides it for tuning.
    static const uint32_t v = [] {
        const char* e = getenv("KUAFU_AMD_REFILL");
        long x = e ? strtol(e, nullptr, 10) : 0;
        return (x >= 1 && x <= 64) ? (uint32_t)x : 16u;
    }();
    return v;
}

hipStream_t stream() {
    if (!g_stream) {
        std::lock_guard<std::mutex> lk(g_mu);
        if (!g_stream) {
            if (g_device < 0) {
                g_device = 0;
                KF_CHECK(hipSetDevice(0));
            }
            KF_CHECK(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
        }
    }
    return g_stream;
}

template <class T> static T* upload(const T* h, size_t n) {
    T* d = nullptr;
    if (!n) return nullptr;
    KF_CHECK(hipMalloc((void**)&d, n * sizeof(T)));
    KF_CHECK(hipMemcpyAsync(d, h, n * sizeof(T), hipMemcpyHostToDevice, stream()));
    return d;
}
// Frees on destruction paths, which cannot throw: a failure is reported by
// kf_last_error, and the error hipFree leaves is consumed here so that it does
// not surface in an unrelated later call.
template <class T> static void release(T*& p) {
    if (p) {
        const hipError_t e = hipFree((void*)p);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            set_error(std::string("hipFree: ") + hipGetErrorString(e));
        }
    }
    p = nullptr;
}

SceneView Scene::view() const {
    SceneView v;
    memset(&v, 0, sizeof(v));
    v.blas = d_blas4;
    v.tri = d_tri;
    v.tlas = d_tlas4;
    v.tlas_root = n_inst ? tlas_root4 : kEmptyRef;
    v.env_root = n_env > 1 ? d_env_root : nullptr;
    v.n_env = n_env;
    v.n_blas_nodes = n_blas4;
    v.n_tri = n_tri;
    v.n_tlas_nodes = n_tlas4;
    v.inst = d_inst;
    v.n_inst = n_inst;
    v.any_alpha = 0;
    for (uint32_t f : h_gflags)
        if (!(f & KF_GEOM_OPAQUE)) v.any_alpha = 1;
    v.verts = d_verts;
    v.idx = d_idx;
    v.matidx = d_matidx;
    v.voff = d_voff;
    v.toff = d_toff;
    v.mats = d_mats;
    v.n_mats = n_mat;
    v.tex_data = d_tex;
    v.tex_info = d_tex_info;
    v.n_tex = n_tex;
    v.env_kind = env_kind;
    v.env_info = d_env_info;
    for (int i = 0; i < 8; ++i) v.env_grad[i] = env_grad[i];
    v.srgb_lut = d_srgb;
    return v;
}

Scene::~Scene() {
    if (g_stream) (void)hipStreamSynchronize(g_stream);
    release(d_verts); release(d_idx); release(d_matidx); release(d_voff); release(d_toff); release(d_tri_geom);
    release(d_mats); release(d_blas); release(d_tri); release(d_blas_root); release(d_tlas); release(d_inst);
    release(d_blas4); release(d_blas_root4); release(d_trav_root4); release(d_mesh); release(d_tlas4); release(d_gflags); release(d_tlas_aux);
    release(d_tlas_err);
    release(d_env_off); release(d_env_node_off); release(d_inst_env); release(d_env_root); release(d_env_cnt);
    release(d_node_off); release(d_blas_cnt);
    if (h_vstage) (void)hipHostFree(h_vstage);
    if (vstage_ev) (void)hipEventDestroy(vstage_ev);
    if (h_stage) (void)hipHostFree(h_stage);
    if (stage_ev) (void)hipEventDestroy(stage_ev);
    d_env_info = nullptr;  // points into d_tex_info, not an allocation of its own
    release(d_inst_src); release(d_ilo); release(d_ihi); release(d_tex); release(d_tex_info); release(d_srgb);
}

// Environments of a batched scene: instance ranges, BVH2 / BVH4 node ranges
// of their TLAS trees (n_e - 1 nodes for n_e >= 2 instances), the environment
// of each instance and the traversal root item of each tree.
static void setup_envs(Scene& s, uint32_t n_env, const uint32_t* off, uint32_t n_inst) {
    if (!off) throw Error("kf_scene_create: num_envs > 1 without env_instance_offset");
    if (off[0] != 0 || off[n_env] != n_inst) throw Error("kf_scene_create: environment offsets must run from 0 to num_instances");
    for (uint32_t e = 0; e < n_env; ++e)
        if (off[e + 1] < off[e]) throw Error("kf_scene_create: environment offsets must be non-decreasing");
    s.n_env = n_env;
    s.h_env_off.assign(off, off + n_env + 1);
    s.h_env_node_off.assign(n_env + 1, 0);
    std::vector<uint32_t> inst_env(n_inst);
    for (uint32_t e = 0; e < n_env; ++e) {
        const uint32_t n_e = off[e + 1] - off[e];
        s.h_env_node_off[e + 1] = s.h_env_node_off[e] + (n_e > 1 ? n_e - 1 : 0);
        for (uint32_t i = off[e]; i < off[e + 1]; ++i) inst_env[i] = e;
    }
    std::vector<int> roots;
    env_roots(s.h_env_off, s.h_env_node_off, roots);
    std::vector<uint32_t> items(n_env);
    for (uint32_t e = 0; e < n_env; ++e) {
        const int r = roots[e];
        items[e] = r == kEmptyRef ? kEmptyItem : (r >= 0 ? (uint32_t)r : ((2u << 30) | (uint32_t)~r));
    }
    s.d_env_off = upload(s.h_env_off.data(), s.h_env_off.size());
    s.d_env_node_off = upload(s.h_env_node_off.data(), s.h_env_node_off.size());
    s.d_inst_env = upload(inst_env.data(), inst_env.size());
    s.d_env_root = upload(items.data(), items.size());
    KF_CHECK(hipMalloc((void**)&s.d_env_cnt, 2 * sizeof(uint32_t) * n_env));
    KF_CHECK(hipMemsetAsync(s.d_env_cnt, 0, 2 * sizeof(uint32_t) * n_env, stream()));
    KF_CHECK(hipStreamSynchronize(stream()));  // the host vectors above are the copies' sources
}

}  // namespace kf

using namespace kf;

extern "C" {

int kf_init(int device) {
    try {
        std::lock_guard<std::mutex> lk(g_mu);
        int n = 0;
        KF_CHECK(hipGetDeviceCount(&n));
        if (device < 0 || device >= n) throw Error("kf_init: no such device");
        if (g_device != device) {
            KF_CHECK(hipSetDevice(device));
            g_device = device;
            if (g_stream) {
                (void)hipStreamDestroy(g_stream);
                g_stream = nullptr;
            }
        }
        if (!g_stream) KF_CHECK(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
        trace_noalloc_setup();
        return 0;
    } catch (const std::exception& e) {
        set_error(e.what());
        return -1;
    }
}

const char* kf_last_error(void) { return g_error.c_str(); }
const char* kf_version(void) { return "kuafu_amd 0.1 (gfx950)"; }

// ---------------------------------------------------------------- host math
// lookAt (right-handed) and the intrinsics projection of camera.cpp:92-128;
// inverses by cofactors in double, rounded to float.
static void inverse4(const double m[16], double out[16]) {
    double inv[16];
    inv[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] + m[9] * m[7] * m[14] + m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
    inv[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] - m[8] * m[7] * m[14] - m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
    inv[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] + m[8] * m[7] * m[13] + m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
    inv[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] - m[8] * m[6] * m[13] - m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
    inv[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] + m[9] * m[2] * m[15] - m[9] * m[3] * m[14] - m[13] * m[2] * m[11] + m[13] * m[3] * m[10];
    inv[5] = m[0] * m[10] * m[15] - m[0] * m[11] * m[14] - m[8] * m[2] * m[15] + m[8] * m[3] * m[14] + m[12] * m[2] * m[11] - m[12] * m[3] * m[10];
    inv[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15] - m[8] * m[3] * m[13] - m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
    inv[13] = m[0] * m[9] * m[14] - m[0] * m[10] * m[13] - m[8] * m[1] * m[14] + m[8] * m[2] * m[13] + m[12] * m[1] * m[10] - m[12] * m[2] * m[9];
    inv[2] = m[1] * m[6] * m[15] - m[1] * m[7] * m[14] - m[5] * m[2] * m[15] + m[5] * m[3] * m[14] + m[13] * m[2] * m[7] - m[13] * m[3] * m[6];
    inv[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15] - m[4] * m[3] * m[14] - m[12] * m[2] * m[7] + m[12] * m[3] * m[6];
    inv[10] = m[0] * m[5] * m[15] - m[0] * m[7] * m[13] - m[4] * m[1] * m[15] + m[4] * m[3] * m[13] + m[12] * m[1] * m[7] - m[12] * m[3] * m[5];
    inv[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14] - m[4] * m[2] * m[13] - m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
    inv[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11] - m[5] * m[3] * m[10] - m[9] * m[2] * m[7] + m[9] * m[3] * m[6];
    inv[7] = m[0] * m[6] * m[11] - m[0] * m[7] * m[10] - m[4] * m[2] * m[11] + m[4] * m[3] * m[10] + m[8] * m[2] * m[7] - m[8] * m[3] * m[6];
    inv[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11] - m[4] * m[3] * m[9] - m[8] * m[1] * m[7] + m[8] * m[3] * m[5];
    inv[15] = m[0] * m[5] * m[10] - m[0] * m[6] * m[9] - m[4] * m[1] * m[10] + m[4] * m[2] * m[9] + m[8] * m[1] * m[6] - m[8] * m[2] * m[5];
    double det = m[0] * inv[0] + m[1] * inv[4] + m[2] * inv[8] + m[3] * inv[12];
    for (int i = 0; i < 16; ++i) out[i] = inv[i] / det;
}

int kf_camera_ubo_compute(const float position[3], const float front[3], const float up[3], int width, int height,
                          float fx, float fy, float cx, float cy, float skew, float aperture, float focal_length,
                          kf_camera_ubo* out) {
    if (!position || !front || !up || !out || width <= 0 || height <= 0) {
        set_error("kf_camera_ubo_compute: bad argument");
        return -1;
    }
    memset(out, 0, sizeof(*out));
    // view = lookAt(pos, pos + front, up)  (camera.cpp:92)
    double e[3] = {position[0], position[1], position[2]};
    double c[3] = {(double)(position[0] + front[0]), (double)(position[1] + front[1]), (double)(position[2] + front[2])};
    double f[3] = {c[0] - e[0], c[1] - e[1], c[2] - e[2]};
    double fl = std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    for (double& v : f) v /= fl;
    double u0[3] = {up[0], up[1], up[2]};
    double s[3] = {f[1] * u0[2] - f[2] * u0[1], f[2] * u0[0] - f[0] * u0[2], f[0] * u0[1] - f[1] * u0[0]};
    double sl = std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    for (double& v : s) v /= sl;
    double u[3] = {s[1] * f[2] - s[2] * f[1], s[2] * f[0] - s[0] * f[2], s[0] * f[1] - s[1] * f[0]};
    double view[16] = {0};
    view[0 * 4 + 0] = s[0]; view[1 * 4 + 0] = s[1]; view[2 * 4 + 0] = s[2];
    view[0 * 4 + 1] = u[0]; view[1 * 4 + 1] = u[1]; view[2 * 4 + 1] = u[2];
    view[0 * 4 + 2] = -f[0]; view[1 * 4 + 2] = -f[1]; view[2 * 4 + 2] = -f[2];
    view[3 * 4 + 0] = -(s[0] * e[0] + s[1] * e[1] + s[2] * e[2]);
    view[3 * 4 + 1] = -(u[0] * e[0] + u[1] * e[1] + u[2] * e[2]);
    view[3 * 4 + 2] = (f[0] * e[0] + f[1] * e[1] + f[2] * e[2]);
    view[3 * 4 + 3] = 1.0;
    // projection (camera.cpp:111-128), far 100, near 0.1
    const double Far = 100.0, Near = 0.1;
    double proj[16] = {0};
    proj[0 * 4 + 0] = (2.0 * fx) / width;
    proj[1 * 4 + 0] = -2.0 * skew / width;
    proj[1 * 4 + 1] = -(2.0 * fy) / height;
    proj[2 * 4 + 0] = -2.0 * cx / width + 1;
    proj[2 * 4 + 1] = -2.0 * cy / height + 1;
    proj[2 * 4 + 2] = -Far / (Far - Near);
    proj[2 * 4 + 3] = -1.0;
    proj[3 * 4 + 2] = -Far * Near / (Far - Near);
    proj[3 * 4 + 3] = 0.0;
    double vinv[16], pinv[16];
    inverse4(view, vinv);
    inverse4(proj, pinv);
    for (int i = 0; i < 16; ++i) {
        out->view[i] = (float)view[i];
        out->proj[i] = (float)proj[i];
        out->view_inverse[i] = (float)vinv[i];
        out->proj_inverse[i] = (float)pinv[i];
    }
    out->position[0] = position[0]; out->position[1] = position[1]; out->position[2] = position[2];
    out->position[3] = aperture;
    out->front[0] = front[0]; out->front[1] = front[1]; out->front[2] = front[2];
    out->front[3] = focal_length;
    for (int i = 0; i < 4; ++i) out->pad1[i] = out->pad2[i] = 1.0f;
    return 0;
}

int kf_world_to_object(const float t[16], float out[12]) {
    if (!t || !out) {
        set_error("kf_world_to_object: null argument");
        return -1;
    }
    // m[r][c] = t[c*4 + r]; 3x3 inverse by cofactors in double; translation
    // -M^-1 t; rounded to float (gl_WorldToObjectEXT).
    double m[3][3], tr[3];
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c) m[r][c] = t[c * 4 + r];
        tr[r] = t[3 * 4 + r];
    }
    double c00 = m[1][1] * m[2][2] - m[1][2] * m[2][1];
    double c01 = m[1][2] * m[2][0] - m[1][0] * m[2][2];
    double c02 = m[1][0] * m[2][1] - m[1][1] * m[2][0];
    double det = m[0][0] * c00 + m[0][1] * c01 + m[0][2] * c02;
    double inv[3][3];
    inv[0][0] = c00 / det;
    inv[0][1] = (m[0][2] * m[2][1] - m[0][1] * m[2][2]) / det;
    inv[0][2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) / det;
    inv[1][0] = c01 / det;
    inv[1][1] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) / det;
    inv[1][2] = (m[0][2] * m[1][0] - m[0][0] * m[1][2]) / det;
    inv[2][0] = c02 / det;
    inv[2][1] = (m[0][1] * m[2][0] - m[0][0] * m[2][1]) / det;
    inv[2][2] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) / det;
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c) out[r * 4 + c] = (float)inv[r][c];
        out[r * 4 + 3] = (float)(-(inv[r][0] * tr[0] + inv[r][1] * tr[1] + inv[r][2] * tr[2]));
    }
    return 0;
}

int kf_active_light_fill(kf_active_lights* ubo, int slot, const float view[16], const float color[3], float fov,
                         float strength, float softness, int tex_id) {
    if (!ubo || !view || !color || slot < 0 || slot >= KF_MAX_ACTIVE_LIGHTS) {
        set_error("kf_active_light_fill: bad argument");
        return -1;
    }
    // scene.cpp:266-281: viewMatInv = inverse(viewMat);
    // projMat = perspective(fov, 1, 0.01, 1000); front = -viewMatInv[2];
    // position = viewMatInv[3]; sftp = (softness, fov, texID, 0).
    double v[16], vi[16];
    for (int i = 0; i < 16; ++i) v[i] = view[i];
    inverse4(v, vi);
    memcpy(ubo->view[slot], view, 16 * sizeof(float));
    double th = std::tan((double)fov / 2.0);
    const double zn = 0.01, zf = 1000.0;
    float* P = ubo->proj[slot];
    for (int i = 0; i < 16; ++i) P[i] = 0.f;
    P[0 * 4 + 0] = (float)(1.0 / (1.0 * th));
    P[1 * 4 + 1] = (float)(1.0 / th);
    P[2 * 4 + 2] = (float)(-(zf + zn) / (zf - zn));
    P[2 * 4 + 3] = -1.0f;
    P[3 * 4 + 2] = (float)(-(2.0 * zf * zn) / (zf - zn));
    ubo->front[slot][0] = (float)-vi[2 * 4 + 0];
    ubo->front[slot][1] = (float)-vi[2 * 4 + 1];
    ubo->front[slot][2] = (float)-vi[2 * 4 + 2];
    ubo->front[slot][3] = 1.0f;
    ubo->rgbs[slot][0] = color[0]; ubo->rgbs[slot][1] = color[1]; ubo->rgbs[slot][2] = color[2];
    ubo->rgbs[slot][3] = strength;
    ubo->position[slot][0] = (float)vi[3 * 4 + 0];
    ubo->position[slot][1] = (float)vi[3 * 4 + 1];
    ubo->position[slot][2] = (float)vi[3 * 4 + 2];
    ubo->position[slot][3] = 0.0f;
    ubo->sftp[slot][0] = softness;
    ubo->sftp[slot][1] = fov;
    ubo->sftp[slot][2] = (float)tex_id;
    ubo->sftp[slot][3] = 0.0f;
    return 0;
}

// ---------------------------------------------------------------- scene
static void upload_textures(Scene& s, const kf_texture_desc* tex, uint32_t n, uint8_t** d_data, uint4** d_info) {
    std::vector<uint4> info(n);
    size_t total = 0;
    for (uint32_t i = 0; i < n; ++i) {
        if (tex[i].width && tex[i].height && !tex[i].rgba8) throw Error("texture without pixels");
        info[i] = make_uint4((uint32_t)total, tex[i].width, tex[i].height, 0);
        total += 4 * (size_t)tex[i].width * tex[i].height;
        if (total > 0xffffffffull) throw Error("textures exceed 4 GiB");
    }
    std::vector<uint8_t> data(std::max<size_t>(total, 4));
    for (uint32_t i = 0; i < n; ++i)
        if (tex[i].width && tex[i].height)
            memcpy(&data[info[i].x], tex[i].rgba8, 4 * (size_t)tex[i].width * tex[i].height);
    *d_data = upload(data.data(), data.size());
    *d_info = upload(info.data(), std::max<uint32_t>(n, 1) == n ? n : 0);
    if (!n) {
        uint4 z = make_uint4(0, 0, 0, 0);
        *d_info = upload(&z, 1);
    }
    (void)s;
}

int kf_scene_create(const kf_scene_desc* d, kf_scene** out) {
    kf_scene* s = nullptr;
    try {
        if (!d || !out) throw Error("kf_scene_create: null argument");
        s = new kf_scene();
        s->n_geom = d->num_geometries;
        if (s->n_geom && (!d->geom_vertex_offset || !d->geom_tri_offset || !d->geom_flags))
            throw Error("kf_scene_create: missing geometry offsets");
        s->h_voff.assign(d->geom_vertex_offset, d->geom_vertex_offset + (s->n_geom ? s->n_geom + 1 : 0));
        s->h_toff.assign(d->geom_tri_offset, d->geom_tri_offset + (s->n_geom ? s->n_geom + 1 : 0));
        if (!s->n_geom) { s->h_voff = {0}; s->h_toff = {0}; }
        s->h_gflags.assign(d->geom_flags, d->geom_flags + s->n_geom);
        s->n_vert = s->h_voff.back();
        s->n_tri = s->h_toff.back();
        s->n_mat = d->num_materials;
        if (s->h_voff[0] != 0 || s->h_toff[0] != 0) throw Error("kf_scene_create: offsets must start at 0");
        // traversal items carry a 30-bit record index
        if (s->n_tri >= (1u << 29) || d->num_instances >= (1u << 29))
            throw Error("kf_scene_create: at most 2^29 triangles and 2^29 instances");
        for (uint32_t g = 0; g < s->n_geom; ++g)
            if (s->h_voff[g + 1] < s->h_voff[g] || s->h_toff[g + 1] < s->h_toff[g])
                throw Error("kf_scene_create: offsets must be non-decreasing");
        // validate indices and material indices (every device read stays in bounds)
        for (uint32_t g = 0; g < s->n_geom; ++g) {
            uint32_t nv = s->h_voff[g + 1] - s->h_voff[g];
            for (size_t t = s->h_toff[g]; t < s->h_toff[g + 1]; ++t) {
                for (int k = 0; k < 3; ++k)
                    if (d->indices[3 * t + k] >= nv) throw Error("kf_scene_create: vertex index out of range");
                if (d->mat_index[t] >= s->n_mat) throw Error("kf_scene_create: material index out of range");
            }
        }
        for (uint32_t i = 0; i < d->num_instances; ++i)
            if (d->instances[i].geometry_index >= s->n_geom)
                throw Error("Geometry index is out of bounds.");  // rt.cpp:118-121
        if (d->num_envs > 1) setup_envs(*s, d->num_envs, d->env_instance_offset, d->num_instances);
        hipStream_t st = stream();
        // vertices as 3 x float4 (the GLSL unpack of rchit:52-64)
        s->d_verts = upload((const float4*)d->vertices, 3 * (size_t)s->n_vert);
        s->d_idx = upload(d->indices, 3 * (size_t)s->n_tri);
        s->d_matidx = upload(d->mat_index, (size_t)s->n_tri);
        s->d_voff = upload(s->h_voff.data(), s->h_voff.size());
        s->d_toff = upload(s->h_toff.data(), s->h_toff.size());
        s->d_mats = upload(d->materials, (size_t)s->n_mat);
        // textures (RGBA8 sRGB, vkCore.hpp:585) and the decode table
        // cube-map faces are appended to the texture pool; env_info points
        // at their records inside d_tex_info
        s->n_tex = d->num_textures;
        std::vector<kf_texture_desc> all;
        if (d->num_textures) all.assign(d->textures, d->textures + d->num_textures);
        if (d->env_kind == KF_ENV_CUBEMAP) {
            if (!d->env_faces) throw Error("kf_scene_create: cubemap without faces");
            for (int f = 0; f < 6; ++f) all.push_back(d->env_faces[f]);
        }
        upload_textures(*s, all.data(), (uint32_t)all.size(), &s->d_tex, &s->d_tex_info);
        if (d->env_kind == KF_ENV_CUBEMAP) s->d_env_info = s->d_tex_info + d->num_textures;
        std::vector<float> lut(256);
        for (int i = 0; i < 256; ++i) {
            double c = i / 255.0;
            lut[i] = (float)(c <= 0.04045 ? c / 12.92 : std::pow((c + 0.055) / 1.055, 2.4));
        }
        s->d_srgb = upload(lut.data(), 256);
        s->env_kind = d->env_kind;
        for (int i = 0; i < 8; ++i) s->env_grad[i] = d->env_gradient[i];
        if (d->dir_light) s->dl = *d->dir_light;
        if (d->point_lights) s->pl = *d->point_lights;
        if (d->active_lights) s->al = *d->active_lights;
        KF_CHECK(hipStreamSynchronize(st));
        hipEvent_t e0, e1;
        KF_CHECK(hipEventCreate(&e0));
        KF_CHECK(hipEventCreate(&e1));
        KF_CHECK(hipEventRecord(e0, st));
        share_geometry(*s, d->vertices, d->indices, d->separate_geometry != 0);
        build_blas(*s);
        build_tlas(*s, d->instances, d->num_instances);
        KF_CHECK(hipEventRecord(e1, st));
        KF_CHECK(hipEventSynchronize(e1));
        KF_CHECK(hipEventElapsedTime(&s->build_ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        *out = s;
        return 0;
    } catch (const std::exception& e) {
        delete s;
        set_error(e.what());
        return -1;
    }
}

void kf_scene_destroy(kf_scene* s) { delete s; }

int kf_scene_update_instances(kf_scene* s, const kf_instance* inst, uint32_t n) {
    try {
        if (!s || (n && !inst)) throw Error("kf_scene_update_instances: null argument");
        for (uint32_t i = 0; i < n; ++i)
            if (inst[i].geometry_index >= s->n_geom) throw Error("Geometry index is out of bounds.");
        build_tlas(*s, inst, n);
        return 0;
    } catch (const std::exception& e) {
        set_error(e.what());
        return -1;
    }
}

int kf_scene_update_geometry(kf_scene* s, uint32_t g, const kf_vertex* vertices, uint32_t n) {
    try {
        if (!s || (n && !vertices)) throw Error("kf_scene_update_geometry: null argument");
        if (g >= s->n_geom) throw Error("kf_scene_update_geometry: no such geometry");
        if (n != s->h_voff[g + 1] - s->h_voff[g])
            throw Error("kf_scene_update_geometry: the vertex count must stay the same (rebuild the scene otherwise)");
        hipStream_t st = stream();
        const size_t bytes = (size_t)n * sizeof(kf_vertex);
        if (bytes) {
            // host vertices -> pinned staging -> device, no wait for the GPU
            // unless the previous update's copy is still pending
            if (!s->vstage_ev) KF_CHECK(hipEventCreateWithFlags(&s->vstage_ev, hipEventDisableTiming));
            if (s->vstage_pending) KF_CHECK(hipEventSynchronize(s->vstage_ev));
            s->vstage_pending = false;
            if (bytes > s->vstage_cap) {
                if (s->h_vstage) KF_CHECK(hipHostFree(s->h_vstage));
                s->h_vstage = nullptr;
                KF_CHECK(hipHostMalloc((void**)&s->h_vstage, bytes, hipHostMallocDefault));
                s->vstage_cap = bytes;
            }
            memcpy(s->h_vstage, vertices, bytes);
            KF_CHECK(hipMemcpyAsync(s->d_verts + 3 * (size_t)s->h_voff[g], s->h_vstage, bytes, hipMemcpyHostToDevice, st));
            KF_CHECK(hipEventRecord(s->vstage_ev, st));
            s->vstage_pending = true;
        }
        refit_blas(*s, g);
        unshare_geometry(*s, g);
        // instance boxes follow the refit BLAS
        std::vector<kf_instance> inst = s->h_inst;
        build_tlas(*s, inst.data(), (uint32_t)inst.size());
        return 0;
    } catch (const std::exception& e) {
        set_error(e.what());
        return -1;
    }
}

int kf_scene_update_materials(kf_scene* s, const kf_material* m, uint32_t n) {
    try {
        if (!s || (n && !m)) throw Error("kf_scene_update_materials: null argument");
        if (n != s->n_mat) throw Error("kf_scene_update_materials: the material count must stay the same");
        if (n) {
            hipStream_t st = stream();
            KF_CHECK(hipMemcpyAsync(s->d_mats, m, n * sizeof(kf_material), hipMemcpyHostToDevice, st));
            KF_CHECK(hipStreamSynchronize(st));
        }
        return 0;
    } catch (const std::exception& e) {
        set_error(e.what());
        return -1;
    }
}

int kf_scene_set_lights(kf_scene* s, const kf_dir_light* d, const kf_point_lights* p, const kf_active_lights* a) {
    if (!s) {
        set_error("kf_scene_set_lights: null scene");
        return -1;
    }
    memset(&s->dl, 0, sizeof(s->dl));
    memset(&s->pl, 0, sizeof(s->pl));
    memset(&s->al, 0, sizeof(s->al));
    if (d) s->dl = *d;
    if (p) s->pl = *p;
    if (a) s->al = *a;
    return 0;
}

static uint32_t subtree_depth(const std::vector<Node>& nodes, int ref) {
    if (ref < 0) return 0;
    uint32_t best = 0;
    std::vector<std::pair<int, uint32_t>> st{{ref, 1}};
    while (!st.empty()) {
        auto [r, dd] = st.back();
        st.pop_back();
        best = std::max(best, dd);
        const Node& n = nodes[r];
        if (n.d.x >= 0) st.push_back({n.d.x, dd + 1});
        if (n.d.y >= 0) st.push_back({n.d.y, dd + 1});
    }
    return best;
}

// BVH4 node count of the TLAS (kept on the device by the asynchronous
// rebuild); also reports a malformed tree.
// BVH4 nodes of all BLAS: geometry g's sit at the start of its range
// [h_node_off[g], h_node_off[g] + n_g - 1); the counts stay on the device.
static uint32_t blas4_count(kf_scene* s) {
    if (!s->d_blas_cnt || !s->n_geom) return 0;
    std::vector<uint32_t> c(2 * s->n_geom);
    KF_CHECK(hipStreamSynchronize(stream()));
    KF_CHECK(hipMemcpy(c.data(), s->d_blas_cnt, c.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
    uint32_t tot = 0;
    for (uint32_t g = 0; g < s->n_geom; ++g)
        if (s->h_toff[g + 1] - s->h_toff[g] >= 2) tot += c[2 * g] - s->h_node_off[g];
    return tot;
}

// With environments the count is the sum over their trees, whose nodes sit
// in per-environment ranges of the array (not packed).
static uint32_t tlas4_count(kf_scene* s) {
    if (!s->d_tlas_aux || s->n_inst < 2) return 0;
    int err = 0;
    KF_CHECK(hipStreamSynchronize(stream()));
    KF_CHECK(hipMemcpy(&err, s->d_tlas_err, sizeof(err), hipMemcpyDeviceToHost));
    if (err) throw Error("TLAS build: malformed tree");
    if (s->n_env > 1) {
        std::vector<uint32_t> c(2 * s->n_env);
        KF_CHECK(hipMemcpy(c.data(), s->d_env_cnt, c.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
        uint32_t tot = 0;
        for (uint32_t e = 0; e < s->n_env; ++e)
            if (s->h_env_off[e + 1] - s->h_env_off[e] >= 2) tot += c[2 * e] - s->h_env_node_off[e];
        return tot;
    }
    uint32_t c = 0;
    KF_CHECK(hipMemcpy(&c, s->d_tlas_aux + 3, sizeof(c), hipMemcpyDeviceToHost));
    return std::min(c, s->tlas4_cap);
}

int kf_scene_bvh_info(kf_scene* s, kf_bvh_info* out) {
    try {
        if (!s || !out) throw Error("kf_scene_bvh_info: null argument");
        out->blas_nodes = s->n_blas_nodes;
        out->tlas_nodes = s->n_env > 1 ? s->h_env_node_off.back() : (s->n_inst > 1 ? s->n_inst - 1 : 0);
        out->triangles = s->n_tri;
        out->instances = s->n_inst;
        out->build_ms = s->build_ms;
        std::vector<Node> nodes(s->n_blas_nodes);
        if (s->n_blas_nodes)
            KF_CHECK(hipMemcpy(nodes.data(), s->d_blas, nodes.size() * sizeof(Node), hipMemcpyDeviceToHost));
        uint32_t md = 0;
        for (int r : s->h_blas_root)
            if (r != kEmptyRef) md = std::max(md, subtree_depth(nodes, r));
        out->max_blas_depth = md;
        std::vector<Node> tn(out->tlas_nodes);
        if (!tn.empty()) KF_CHECK(hipMemcpy(tn.data(), s->d_tlas, tn.size() * sizeof(Node), hipMemcpyDeviceToHost));
        out->max_tlas_depth = 0;
        if (s->n_env > 1) {
            std::vector<int> roots;
            env_roots(s->h_env_off, s->h_env_node_off, roots);
            for (int r : roots)
                if (r != kEmptyRef) out->max_tlas_depth = std::max(out->max_tlas_depth, subtree_depth(tn, r));
        } else if (s->n_inst > 1) {
            out->max_tlas_depth = subtree_depth(tn, s->tlas_root);
        }
        out->blas4_nodes = blas4_count(s);
        out->tlas4_nodes = tlas4_count(s);
        return 0;
    } catch (const std::exception& e) {
        set_error(e.what());
        return -1;
    }
}

int kf_scene_read_bvh4(kf_scene* s, int which, float* nodes, uint32_t max_nodes, uint32_t* n_nodes, int32_t* roots,
                       uint32_t max_roots) {
    try {
        if (!s || (which != 0 && which != 1)) throw Error("kf_scene_read_bvh4: bad argument");
        // a batched scene's TLAS array is returned whole (per-environment ranges)
        const uint32_t nn = which == 0 ? s->n_blas4 : (s->n_env > 1 ? (tlas4_count(s), s->n_tlas4) : tlas4_count(s));
        const Node4* src = which == 0 ? s->d_blas4 : s->d_tlas4;
        if (n_nodes) *n_nodes = nn;
        if (nodes && nn) {
            if (max_nodes < nn) throw Error("kf_scene_read_bvh4: node buffer too small");
            KF_CHECK(hipMemcpy(nodes, src, nn * sizeof(Node4), hipMemcpyDeviceToHost));
        }
        if (roots) {
            if (which == 0) {
                if (max_roots < s->n_geom) throw Error("kf_scene_read_bvh4: root buffer too small");
                for (uint32_t g = 0; g < s->n_geom; ++g) roots[g] = s->h_blas_root4[g];
            } else if (s->n_env > 1) {
                if (max_roots < s->n_env) throw Error("kf_scene_read_bvh4: root buffer too small");
                std::vector<int> r;
                env_roots(s->h_env_off, s->h_env_node_off, r);
                for (uint32_t e = 0; e < s->n_env; ++e) roots[e] = r[e];
            } else {
                if (max_roots < 1) throw Error("kf_scene_read_bvh4: root buffer too small");
                roots[0] = s->n_inst ? s->tlas_root4 : kEmptyRef;
            }
        }
        return 0;
    } catch (const std::exception& e) {
        set_error(e.what());
        return -1;
    }
}

int kf_scene_read_blas(kf_scene* s, uint32_t g, float* nodes, uint32_t max_nodes, uint32_t* n_nodes,
                       int32_t* root_ref, uint32_t* tri_order, uint32_t max_tris) {
    try {
        if (!s || g >= s->n_geom) throw Error("kf_scene_read_blas: bad geometry");
        uint32_t nt = s->h_toff[g + 1] - s->h_toff[g];
        uint32_t nn = nt > 1 ? nt - 1 : 0;
        if (n_nodes) *n_nodes = nn;
        // refs are returned local to the geometry: internal nodes from 0, leaves ~(slot - first slot)
        int node0 = (int)s->h_node_off[g], slot0 = (int)s->h_toff[g];
        auto local = [&](int r) { return r == kEmptyRef ? r : (r >= 0 ? r - node0 : ~(~r - slot0)); };
        if (root_ref) *root_ref = local(s->h_blas_root[g]);
        if (nodes) {
            if (max_nodes < nn) throw Error("kf_scene_read_blas: node buffer too small");
            std::vector<Node> h(nn);
            if (nn) KF_CHECK(hipMemcpy(h.data(), s->d_blas + node0, nn * sizeof(Node), hipMemcpyDeviceToHost));
            for (auto& n : h) {
                n.d.x = local(n.d.x);
                n.d.y = local(n.d.y);
            }
            if (nn) memcpy(nodes, h.data(), nn * sizeof(Node));
        }
        if (tri_order) {
            if (max_tris < nt) throw Error("kf_scene_read_blas: order buffer too small");
            std::vector<float4> tri(3 * (size_t)nt);
            if (nt)
                KF_CHECK(hipMemcpy(tri.data(), s->d_tri + 3 * (size_t)s->h_toff[g], tri.size() * sizeof(float4),
                                   hipMemcpyDeviceToHost));
            for (uint32_t i = 0; i < nt; ++i) {
                uint32_t p;
                memcpy(&p, &tri[3 * (size_t)i].w, 4);
                tri_order[i] = p;
            }
        }
        return 0;
    } catch (const std::exception& e) {
        set_error(e.what());
        return -1;
    }
}

// ---------------------------------------------------------------- renderer
int kf_renderer_create(uint32_t w, uint32_t h, kf_renderer** out) {
    kf_renderer* r = nullptr;
    try {
        if (!out || !w || !h) throw Error("kf_renderer_create: bad argument");
        r = new kf_renderer();
        r->W = w;
        r->H = h;
        size_t n = (size_t)w * h;
        KF_CHECK(hipMalloc((void**)&r->d_rgba, n * sizeof(float4)));
        KF_CHECK(hipMalloc((void**)&r->d_albedo, n * sizeof(float4)));
        KF_CHECK(hipMalloc((void**)&r->d_normal, n * sizeof(float4)));
        KF_CHECK(hipMemsetAsync(r->d_rgba, 0, n * sizeof(float4), stream()));
        KF_CHECK(hipMemsetAsync(r->d_albedo, 0, n * sizeof(float4), stream()));
        KF_CHECK(hipMemsetAsync(r->d_normal, 0, n * sizeof(float4), stream()));
        KF_CHECK(hipStreamSynchronize(stream()));
        *out = r;
        return 0;
    } catch (const std::exception& e) {
        delete r;
        set_error(e.what());
        return -1;
    }
}

void kf_renderer_destroy(kf_renderer* r) {
    if (r && g_stream) (void)hipStreamSynchronize(g_stream);
    delete r;
}

float* kf_renderer_device_target(kf_renderer* r, int target) {
    if (!r) return nullptr;
    return (float*)(target == 0 ? r->d_rgba : target == 1 ? r->d_albedo : target == 2 ? r->d_normal : nullptr);
}

int kf_renderer_read(kf_renderer* r, int target, float* out) {
    try {
        float* p = kf_renderer_device_target(r, target);
        if (!p || !out) throw Error("kf_renderer_read: bad argument");
        KF_CHECK(hipMemcpyAsync(out, p, (size_t)r->W * r->H * sizeof(float4), hipMemcpyDeviceToHost, stream()));
        KF_CHECK(hipStreamSynchronize(stream()));
        return 0;
    } catch (const std::exception& e) {
        set_error(e.what());
        return -1;
    }
}

int kf_renderer_write(kf_renderer* r, int target, const float* in) {
    try {
        float* p = kf_renderer_device_target(r, target);
        if (!p || !in) throw Error("kf_renderer_write: bad argument");
        KF_CHECK(hipMemcpyAsync(p, in, (size_t)r->W * r->H * sizeof(float4), hipMemcpyHostToDevice, stream()));
        KF_CHECK(hipStreamSynchronize(stream()));
        return 0;
    } catch (const std::exception& e) {
        set_error(e.what());
        return -1;
    }
}

int kf_renderer_read_samples(kf_renderer* r, float* out, uint32_t max_floats) {
    try {
        if (!r || !out) throw Error("kf_renderer_read_samples: bad argument");
        size_t n = (size_t)r->samples_spp * r->W * r->H * 3;
        if (!r->samples || n == 0) throw Error("kf_renderer_read_samples: render with keep_samples first");
        if (max_floats < n) throw Error("kf_renderer_read_samples: buffer too small");
        KF_CHECK(hipMemcpyAsync(out, r->samples, n * sizeof(float), hipMemcpyDeviceToHost, stream()));
        KF_CHECK(hipStreamSynchronize(stream()));
        return 0;
    } catch (const std::exception& e) {
        set_error(e.what());
        return -1;
    }
}

}  // extern "C"
