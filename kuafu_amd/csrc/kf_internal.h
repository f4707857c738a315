// kf_internal.h — device data layout and host-side objects of libkuafu_amd.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>

#include "../../include/kuafu_amd.h"

#define KF_CHECK(expr)                                                                    \
    do {                                                                                  \
        hipError_t _e = (expr);                                                           \
        if (_e != hipSuccess) throw kf::Error(std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

namespace kf {

struct Error : std::exception {
    std::string msg;
    explicit Error(std::string m) : msg(std::move(m)) {}
    const char* what() const noexcept override { return msg.c_str(); }
};

void set_error(const std::string& m);
hipStream_t stream();
uint32_t traversal_refill();
void trace_noalloc_setup();  // device address of the no-alloc trace scratch, per device

// tea (Random.glsl:7-21) on the host, for the frame seed.
inline uint32_t tea_host(uint32_t v0, uint32_t v1) {
    uint32_t s0 = 0;
    for (int n = 0; n < 16; ++n) {
        s0 += 0x9e3779b9u;
        v0 += ((v1 << 4) + 0xa341316cu) ^ (v1 + s0) ^ ((v1 >> 5) + 0xc8013ea4u);
        v1 += ((v0 << 4) + 0xad90777du) ^ (v0 + s0) ^ ((v0 >> 5) + 0x7e95761eu);
    }
    return v0;
}

// ------------------------------------------------------------------ BVH layout
// BVH2 node, 64 B, both child boxes stored in the parent so one 64-B fetch
// tests two children:
//   a = (L.lo.x, L.hi.x, L.lo.y, L.hi.y)
//   b = (R.lo.x, R.hi.x, R.lo.y, R.hi.y)
//   c = (L.lo.z, L.hi.z, R.lo.z, R.hi.z)
//   d = (left_ref, right_ref, -, -)
// A child ref >= 0 is an internal node index (global over all BLAS / TLAS
// arrays); < 0 is a leaf: ~leaf where leaf is a triangle slot (BLAS) or an
// instance index (TLAS).
struct alignas(16) Node {
    float4 a, b, c;
    int4 d;
};

// BVH4 node, 128 B, collapsed from the BVH2 above: child c's box is
// (lox[c], hix[c]) x (loy[c], hiy[c]) x (loz[c], hiz[c]); ref[c] is the
// child as a traversal item: bits 31:30 the kind (0 TLAS node, 1 BLAS node,
// 2 instance, 3 triangle slot), bits 29:0 its index; kEmptyItem marks an
// unused slot (its box is all zero).  The traversal reads these; the BVH2
// stays for the builder and its tests.
constexpr uint32_t kEmptyItem = 0xffffffffu;
struct alignas(16) Node4 {
    float4 lox, hix, loy, hiy, loz, hiz;
    int4 ref;
    int4 pad;
};

// Instance record read by the traversal, 192 B.  Lines 0-6 repeat the root
// node of the instance's BLAS in the Node4 layout (child boxes in object
// space, child items), so the traversal step that reaches an instance also
// tests its BLAS root's children, with the object-space ray it has just
// computed; a BLAS of one triangle gets a node with that triangle as its
// only child, an invalid instance a node without children.  Lines 7-9 are
// the world->object rows, line 10 (root_ref, geometry, flags, 0) with flags
// bit0 = opaque, bit1 = valid; line 11 (vertex offset, triangle offset) of
// the arrays the geometry's shading reads and the geometry's own triangle
// offset (its material indices), 0.
struct alignas(16) Inst {
    float4 lox, hix, loy, hiy, loz, hiz;
    int4 ref;
    float4 w2o[3];
    int4 info;
    int4 shade;
};
static_assert(sizeof(Inst) == 192, "instance record layout");

enum : int { kInstOpaque = 1, kInstValid = 2 };
constexpr int kEmptyRef = 0x7fffffff;  // root ref of a BLAS without triangles

// Everything the traversal and shading kernels read, by value.
struct SceneView {
    const Node4* blas;        // BVH4 of all geometries
    const float4* tri;        // 3 float4 per triangle slot, by axis: (v0[a], v1[a], v2[a], -); tri[3s].w = local prim id bits
    const Node4* tlas;        // BVH4 over the instances (one tree per environment)
    int tlas_root;            // BVH4 root ref of environment 0's TLAS
    const uint32_t* env_root; // n_env > 1: root item of each environment's TLAS (kEmptyItem: no instance)
    uint32_t n_env;
    uint32_t n_blas_nodes, n_tri, n_tlas_nodes;  // bounds every traversal index is checked against
    const Inst* inst;
    uint32_t n_inst;
    uint32_t any_alpha;       // some geometry is not opaque: path rays run the alpha test
    // shading data (reference SSBO layouts)
    const float4* verts;      // kf_vertex as 3 x float4
    const uint32_t* idx;      // 3 per triangle (local)
    const uint32_t* matidx;   // per triangle
    const uint32_t* voff;     // per geometry
    const uint32_t* toff;     // per geometry
    const kf_material* mats;
    uint32_t n_mats;
    const uint8_t* tex_data;  // all RGBA8 textures
    const uint4* tex_info;    // (byte offset, width, height, 0)
    uint32_t n_tex;
    uint32_t env_kind;
    const uint4* env_info;    // 6 faces, same encoding as tex_info
    float env_grad[8];
    const float* srgb_lut;    // 256 floats: sRGB byte -> linear
};

// ------------------------------------------------------------------ scene
struct Scene {
    uint32_t n_geom = 0, n_tri = 0, n_vert = 0, n_inst = 0, n_mat = 0;
    std::vector<uint32_t> h_voff, h_toff, h_gflags, h_node_off;
    std::vector<int> h_blas_root;
    // device buffers
    float4* d_verts = nullptr;
    uint32_t *d_idx = nullptr, *d_matidx = nullptr, *d_voff = nullptr, *d_toff = nullptr;
    uint32_t* d_tri_geom = nullptr;
    kf_material* d_mats = nullptr;
    Node* d_blas = nullptr;
    float4* d_tri = nullptr;
    int* d_blas_root = nullptr;
    uint32_t n_blas_nodes = 0;
    Node* d_tlas = nullptr;
    // BVH4 copies read by the traversal
    Node4* d_blas4 = nullptr;                // geometry g's BVH4 nodes in [h_node_off[g], ... + n_g - 1)
    int* d_blas_root4 = nullptr;
    uint32_t* d_node_off = nullptr;          // h_node_off on the device
    uint32_t* d_blas_cnt = nullptr;          // 2 collapse counters per geometry: [2g] = next free BVH4 slot
    std::vector<kf_instance> h_inst;         // instances of the last TLAS build (for rebuilds after a refit)
    uint8_t* h_vstage = nullptr;             // pinned staging of geometry vertex updates
    size_t vstage_cap = 0;
    hipEvent_t vstage_ev = nullptr;
    bool vstage_pending = false;
    std::vector<int> h_blas_root4;           // each geometry's own BVH4 root
    // Geometry whose BVH the traversal uses for geometry g: g itself, or the
    // first geometry with the same triangles (the same vertex positions at
    // every corner, bit for bit), whose BVH is then a BVH of g's triangles
    // with the same local primitive ids.  Shading still reads g's own data.
    std::vector<uint32_t> h_canon;
    int* d_trav_root4 = nullptr;             // BVH4 root of h_canon[g], referenced by instance records
    // Geometry whose vertex and index arrays shading reads for geometry g:
    // g itself, or the first geometry with the same vertex records and
    // indices, bit for bit (materials always come from g's own matIndex).
    std::vector<uint32_t> h_mesh;
    uint2* d_mesh = nullptr;                 // (voff, toff) of h_mesh[g]
    uint32_t n_blas4 = 0;
    Node4* d_tlas4 = nullptr;
    uint32_t n_tlas4 = 0, tlas4_cap = 0;    // n_tlas4: bound for traversal index checks (the capacity)
    int tlas_root4 = kEmptyRef;
    // TLAS updates run without host round trips: host inputs go through a
    // pinned staging buffer, reused once the previous update's copies are done
    uint8_t* h_stage = nullptr;
    size_t stage_cap = 0;
    hipEvent_t stage_ev = nullptr;
    bool stage_pending = false;
    uint32_t* d_gflags = nullptr;            // per-geometry flags (static)
    uint32_t* d_tlas_aux = nullptr;          // [0..1] segment offsets, [2] node offset, [3..4] collapse counters
    int* d_tlas_err = nullptr;
    Inst* d_inst = nullptr;
    kf_instance* d_inst_src = nullptr;
    float4 *d_ilo = nullptr, *d_ihi = nullptr;  // instance world boxes (the TLAS leaf boxes)
    int tlas_root = kEmptyRef;
    // Batched environments: instances [h_env_off[e], h_env_off[e+1]) form
    // environment e, with its BVH2 nodes (and its BVH4 nodes, root first)
    // from h_env_node_off[e].  One environment: the arrays stay empty.
    uint32_t n_env = 1;
    std::vector<uint32_t> h_env_off, h_env_node_off;
    uint32_t *d_env_off = nullptr, *d_env_node_off = nullptr, *d_inst_env = nullptr, *d_env_root = nullptr;
    uint32_t* d_env_cnt = nullptr;           // 2 collapse counters per environment
    uint8_t* d_tex = nullptr;
    uint4* d_tex_info = nullptr;
    uint32_t n_tex = 0;
    uint4* d_env_info = nullptr;
    uint8_t* d_env = nullptr;
    uint32_t env_kind = 0;
    float env_grad[8] = {0};
    float* d_srgb = nullptr;
    kf_dir_light dl{};
    kf_point_lights pl{};
    kf_active_lights al{};
    float build_ms = 0.f;
    std::vector<hipEvent_t> events;

    SceneView view() const;
    ~Scene();
};

// ------------------------------------------------------------------ builders (bvh.hip)
// Batched LBVH over n prims in n_seg segments (prims of segment s occupy
// [seg_off[s], seg_off[s+1])).  Writes nodes (internal nodes of segment s at
// node_off[s] ...), the sorted prim order, and each segment's root ref.
// leaf_is_slot: leaf refs are ~slot (BLAS) instead of ~prim (TLAS).
void lbvh_build(uint32_t n, const float4* d_lo, const float4* d_hi, const uint32_t* d_seg, uint32_t n_seg,
                const uint32_t* d_seg_off, const uint32_t* d_node_off, Node* d_nodes, uint32_t* d_order,
                int* d_root, bool leaf_is_slot, hipStream_t st);
// Collapse a BVH2 into a BVH4 (level by level on the GPU): each BVH4 node
// takes the two children of a BVH2 node and then repeatedly opens its inner
// child of largest surface area until it has four.  roots2[i] (BVH2 root
// refs, host) -> roots4[i] (BVH4 root refs); leaf and empty roots pass
// through.  Returns the number of BVH4 nodes written to out (capacity cap).
uint32_t collapse_bvh4(const Node* d_b2, uint32_t n_b2, const std::vector<int>& roots2, std::vector<int>& roots4,
                       Node4* d_out, uint32_t cap, uint32_t inner_kind, hipStream_t st);
// Root refs of the TLAS trees of environments with instance offsets env_off
// and BVH2 node offsets node_off (known by construction, no device read).
void env_roots(const std::vector<uint32_t>& env_off, const std::vector<uint32_t>& node_off, std::vector<int>& roots);
void build_blas(Scene& s);
void upload_trav_roots(Scene& s);
void upload_mesh_sources(Scene& s);
// Refit geometry g's BLAS (BVH2 + BVH4 + triangle slots) after its
// vertices changed in place; the TLAS must be rebuilt afterwards.
void refit_blas(Scene& s, uint32_t g);
void build_tlas(Scene& s, const kf_instance* h_inst, uint32_t n);
// Geometries with the same triangles (exact positions of every corner in
// order) as an earlier one traverse its BVH (Scene::h_canon), and those with
// the same vertex records and indices shade from its arrays (Scene::h_mesh);
// separate: each keeps its own.
void share_geometry(Scene& s, const kf_vertex* verts, const uint32_t* idx, bool separate);
// Geometry g's vertices changed: it traverses its own BVH and reads its own
// arrays again, and groups it led move to their next member; d_trav_root4
// and d_mesh follow (the TLAS must be rebuilt afterwards).
void unshare_geometry(Scene& s, uint32_t g);

// ------------------------------------------------------------------ renderer
struct Renderer {
    uint32_t W = 0, H = 0;
    float4 *d_rgba = nullptr, *d_albedo = nullptr, *d_normal = nullptr;
    uint8_t* d_bgra = nullptr;
    // wavefront state, sized for `cap` paths and `slots` NEE slots per path
    size_t cap = 0;
    uint32_t slots = 0;
    float4 *ray_o = nullptr, *ray_d = nullptr, *thr = nullptr, *col = nullptr, *rw = nullptr, *em = nullptr;
    float4* hit = nullptr;
    uint32_t* hit_prim = nullptr;
    uint32_t* seed = nullptr;
    uint32_t *queue_a = nullptr, *queue_b = nullptr;
    uint32_t* queue_e = nullptr;               // paths k_extend traces at the next bounce
    float4 *sh_o = nullptr, *sh_d = nullptr;   // shadow queue (o.xyz, slot bits), (d.xyz, tmax)
    uint32_t* sh_root = nullptr;               // shadow queue: TLAS root item (scenes with environments)
    kf_camera_ubo* d_cams = nullptr;           // one camera per environment
    // Camera-ray TLAS: the scene's TLAS nodes followed by kTileNodes nodes
    // per 8x8 pixel tile, and each tile's root item (k_tile_roots)
    Node4* d_ftlas = nullptr;
    size_t ftlas_cap = 0;
    uint32_t* d_tile_root = nullptr;
    size_t tile_cap = 0;
    uint32_t cams_cap = 0;
    float4 *nee_a = nullptr, *nee_b = nullptr; // per slot (c_a.xyz, theta), (c_b.xyz, mode|group bits)
    uint8_t* vis = nullptr;
    float* samples = nullptr;                  // spp * W * H * 3
    size_t samples_cap = 0;
    uint32_t samples_spp = 0;
    unsigned long long* counters = nullptr;    // device counters
    void* d_lights = nullptr;                  // DevLights of the frame
    uint2* d_jump = nullptr;                   // LCG jump-ahead per sample
    uint32_t jump_cap = 0;
    float4* d_acc = nullptr;                   // per-pixel running sample sum
    float4* d_out = nullptr;                   // per-path final radiance of a batch
    size_t out_cap = 0;
    std::vector<hipEvent_t> ev;
    hipStream_t st2 = nullptr;                 // any-hit traversal, beside the main stream
    std::vector<hipEvent_t> sync_ev;           // per bounce: shading done, shadow rays done
    void* d_fv = nullptr;                      // per-frame seed and count read by the kernels
    unsigned long long* h_totals = nullptr;    // pinned host copy of the frame's global counters
    // Frame graph (kf_render_opts.use_graph): every batch of a frame captured
    // once and replayed while everything baked into its kernel arguments
    // (graph_key) is unchanged.
    hipGraphExec_t graph = nullptr;
    std::vector<uint8_t> graph_key;
    uint32_t graph_ext = 0, graph_sh = 0;      // traversal launches in the graph
    uint64_t graph_captures = 0;               // frame graphs captured by this renderer so far
    // Schedule picked by kf_renderer_autotune (0: built-in default).
    uint32_t tune_spb = 0, tune_refill = 0, tune_grab = 0, tune_tail = 0, tune_grab0 = 0, tune_grabs = 0;
    uint32_t tune_pairs = 1;  // 1: k_tail2 may carry the tail bounces of few paths, 0: k_tail only
    uint32_t n_env_last = 1;                   // environments of the last frame (the denoiser's batch)
    ~Renderer();
    void ensure(size_t paths, uint32_t nslots, uint32_t spp);
};

}  // namespace kf

struct kf_scene : kf::Scene {};
struct kf_renderer : kf::Renderer {};
