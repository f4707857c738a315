// denoise.hip — the denoise stage: a BF16 U-Net on MFMA (gfx950).
//
// Replaces DenoiserOptix (denoiser.cpp:24-344): the reference hands rgba,
// albedo and normal to the closed OptiX LDR denoiser and writes the result
// back into rgba (context.cpp:650-670).  Here the same three targets feed a
// U-Net trained on this tracer's own images:
//   5 levels, 32/64/96/128/160 channels, two 3x3 convs + ReLU per level,
//   2x2 average pooling, nearest 2x upsampling, skip concatenation,
//   final 3x3 conv to RGB, residual on log1p(noisy RGB).
// Tensors are NHWC BF16 with FP32 accumulation.  Convolutions are implicit
// GEMMs on v_mfma_f32_16x16x32_bf16:
//   k_conv2  the plain, pooled-output and upsampled-input convolutions
//            (forward, and the unmasked data gradient): 8 waves, 32 x 16
//            tiles, double-buffered LDS filled by loads straight into LDS
//   k_conv   the first layer (packs the float4 render targets itself), the
//            FP32 output layer and the masked data gradients (the same
//            kernel on flipped, transposed weights, ReLU mask of the forward
//            input fused); bias + ReLU fused in both
//   k_wgrad  weight gradient, pixels as the K dimension, split-K with FP32
//            atomics
// Concatenations cost no copies: producers write into channel slices of the
// concatenated tensor (row stride "ld").
#include "kf_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <type_traits>
#include <vector>

namespace kf {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int TB = 256;

__host__ __device__ inline uint16_t f2bf(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;  // NaN
    u += 0x7fffu + ((u >> 16) & 1u);                       // round to nearest even
    return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t b) {
    uint32_t u = (uint32_t)b << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

__device__ inline float sane(float v) { return (v == v) ? fminf(fmaxf(v, -1e4f), 1e4f) : 0.0f; }

// The 9 network input channels of one pixel: log1p(rgb), albedo * pi, normal.
__device__ inline float log_radiance(float v) { return log1pf(fmaxf(sane(v), 0.f)); }
__device__ inline void pack9(float4 c, float4 a, float4 n, float v[9]) {
    v[0] = log_radiance(c.x);
    v[1] = log_radiance(c.y);
    v[2] = log_radiance(c.z);
    v[3] = sane(a.x) * 3.141592f;
    v[4] = sane(a.y) * 3.141592f;
    v[5] = sane(a.z) * 3.141592f;
    v[6] = sane(n.x);
    v[7] = sane(n.y);
    v[8] = sane(n.z);
}

// float4 render targets of N images of h x w stacked vertically, read as the
// network input padded to the convolution's H x W by repeating edge pixels.
struct PackSrc {
    const float4* rgba = nullptr;
    const float4* alb = nullptr;
    const float4* nrm = nullptr;
    int h = 0, w = 0;
};

// ---------------------------------------------------------------- convolution
// y[p, co] = sum_{tap, ci} x[p + off(tap), ci] * w[tap, co, ci]   (zero padding)
// Work item: a tile of 32 x 8 pixels of one image x BN output channels (BN =
// 16, 32 or 64, chosen from the channels actually stored); step: an item and
// one 32-channel chunk of its input.  Per step the block holds in LDS the
// 34 x 10 input halo of its tile and the weights of all 9 taps, so one
// barrier pair covers 9 taps x 4 x BN/16 MFMAs per wave; wave w owns a
// 16 x 4 block of pixels (four 16-pixel MFMA tiles, one per row).
// The grid is persistent: block b takes items b, b + gridDim.x, ... and,
// while its waves run the MFMAs of one step, holds the global loads of the
// next step in registers (written to LDS after the next barrier), so memory
// latency hides behind matrix work instead of behind other blocks.  Weights
// stay in LDS while consecutive steps need the same ones (one input chunk
// and the same output channels).
// Cin must be a multiple of 32; ldx, ldy multiples of 8.
constexpr int CBK = 32, CLD = CBK + 8;
// LDS rows of the convolution tiles: 32 halves (64 bytes) per pixel or
// weight row, 16-byte group q of row r stored at group q ^ ((r / 2) % 4), so
// that the 16-byte reads of an MFMA operand (16 consecutive rows, one group)
// hit distinct banks (measured: no bank conflicts, against ~half the LDS
// cycles with rows padded to 80 bytes).
constexpr int CCLD = CBK;
__device__ inline int cofs(int row, int q) { return row * CCLD + ((q ^ ((row >> 1) & 3)) << 3); }

// Raw buffer resource over `bytes` bytes at p (stride 0, 32-bit data format):
// loads at offsets >= bytes return zeros and touch no memory.
constexpr uint32_t kOOB = 0x80000000u;  // an offset past every resource built here (< 2^31 bytes)
__device__ inline __amdgpu_buffer_rsrc_t buffer_rsrc(const void* p, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}
constexpr int TX = 32, TY = 8, HXX = TX + 2, HYY = TY + 2, NHALO = HXX * HYY;

template <int BN>
struct ConvShape {
    static constexpr int kHaloU4 = NHALO * 4;               // 16-byte groups of the halo
    static constexpr int kWU4 = 9 * BN * 4;                 // ... of one chunk of weights
    static constexpr int kHR = (kHaloU4 + TB - 1) / TB;     // per thread
    static constexpr int kWR = (kWU4 + TB - 1) / TB;
    static constexpr int kLds = NHALO * CCLD + 9 * BN * CCLD + 2 * BN;  // halves (+ BN FP32 biases)
};

// FUSE bits: kFuseUp upsampled input channels, kFusePool pooled output,
// kFusePack the 32-channel network input computed from the render targets
// (k_conv2_pack), kOutF32 FP32 output (Yf) instead of BF16 (Y), kMask output
// masked by M > 0.
constexpr int kFuseUp = 1, kFusePool = 2, kFusePack = 4, kOutF32 = 8, kMask = 16;

// Waves per SIMD the register budget is set for: 3 (168 registers) where that
// needs no spills, else 2 (64 output channels per item, masked data
// gradient).
constexpr int conv_waves(int bn, int fuse) { return (bn > 32 || (fuse & kMask)) ? 2 : 3; }

// Epilogue of a work item, straight from the accumulators (bias from Bsm,
// ReLU, optional mask, BF16 or FP32 store, optional 2x2 average pool).
template <int BN, int FUSE>
__device__ __attribute__((always_inline)) inline void conv_store(const f32x4 (&acc)[4][BN / 16], const float* Bsm,
                                                                int n, int y0, int x0, int co0, int wv, int lane,
                                                                int H, int W, uint16_t* __restrict__ Y, int ldy,
                                                                float* __restrict__ Yf, int ldyf, int cstore,
                                                                const uint16_t* __restrict__ M, int ldm, int relu,
                                                                uint16_t* __restrict__ P) {
    const int fr = lane & 15;
    const struct {
        int n, y0, x0, co0;
    } cur = {n, y0, x0, co0};
    // Epilogue straight from the accumulators: lane holds channels
    // co0 + nb * 16 + 4 (lane / 16) .. + 3 of pixel (lane % 16) of row mt.
    const int xo = cur.x0 + (wv & 1) * 16 + fr;
    // per-tile base (uniform, 64-bit) + per-lane 32-bit offsets
    const size_t pbase = (size_t)(cur.n * H + cur.y0) * W + cur.x0;
    const int xl = (wv & 1) * 16 + fr;
#pragma unroll
    for (int nb = 0; nb < BN / 16; ++nb) {
        const int ch0 = cur.co0 + nb * 16 + (lane >> 4) * 4;
        const float4 b4 = *(const float4*)&Bsm[nb * 16 + (lane >> 4) * 4];
        const float bv[4] = {b4.x, b4.y, b4.z, b4.w};
        uint32_t pool[4][2] = {};
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            const int yo = cur.y0 + 4 * (wv >> 1) + mt;
            if (yo >= H || xo >= W || ch0 >= cstore) continue;
            const int loc = (4 * (wv >> 1) + mt) * W + xl;  // pixel - pbase
            float v[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                v[e] = acc[mt][nb][e] + bv[e];
                if (relu) v[e] = fmaxf(v[e], 0.0f);
                if constexpr ((FUSE & kMask) != 0)
                    if (ch0 + e < cstore && !(bf2f(M[pbase * ldm + (unsigned)(loc * ldm + ch0 + e)]) > 0.0f))
                        v[e] = 0.0f;
            }
            if constexpr ((FUSE & kOutF32) != 0) {
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    if (ch0 + e < cstore) Yf[pbase * ldyf + (unsigned)(loc * ldyf + ch0 + e)] = v[e];
            } else {
                // v_cvt_pk_bf16_f32: round to nearest even, as f2bf
                typedef float f2 __attribute__((ext_vector_type(2)));
                typedef __bf16 b2 __attribute__((ext_vector_type(2)));
                const uint32_t lo = __builtin_bit_cast(uint32_t, __builtin_convertvector(f2{v[0], v[1]}, b2));
                const uint32_t hi = __builtin_bit_cast(uint32_t, __builtin_convertvector(f2{v[2], v[3]}, b2));
                uint16_t* yp = Y + pbase * ldy + (unsigned)(loc * ldy + ch0);
                if (ch0 + 4 <= cstore) {
                    *(uint2*)yp = make_uint2(lo, hi);
                    if constexpr ((FUSE & kFusePool) != 0) {
                        // 2x2 average pool of the rounded outputs: rows mt, mt + 1
                        // here, columns xo, xo + 1 in this lane and the next
                        pool[mt][0] = lo;
                        pool[mt][1] = hi;
                    }
                } else {
                    const uint32_t h4[2] = {lo, hi};
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        if (ch0 + e < cstore) yp[e] = (uint16_t)(h4[e >> 1] >> (16 * (e & 1)));
                }
            }
        }
        if constexpr ((FUSE & kFusePool) != 0) {
            // pooled pixel (yo / 2, xo / 2) from rows 2 mp, 2 mp + 1 of the lanes
            // of columns xo (even) and xo + 1 (the next lane); summed in the
            // order (2y, 2x), (2y, 2x + 1), (2y + 1, 2x), (2y + 1, 2x + 1)
#pragma unroll
            for (int mp = 0; mp < 2; ++mp) {
                const int yo = cur.y0 + 4 * (wv >> 1) + 2 * mp;
                uint32_t q4[4][2];
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    q4[0][h] = pool[2 * mp][h];
                    q4[1][h] = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)pool[2 * mp][h], 0xB1, 0xF, 0xF, false);
                    q4[2][h] = pool[2 * mp + 1][h];
                    q4[3][h] =
                        (uint32_t)__builtin_amdgcn_update_dpp(0, (int)pool[2 * mp + 1][h], 0xB1, 0xF, 0xF, false);
                }
                if ((fr & 1) || yo >= H || xo >= W || ch0 + 4 > cstore) continue;
                float m[4];
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    float sum = 0.0f;
#pragma unroll
                    for (int k = 0; k < 4; ++k) sum += bf2f((uint16_t)(q4[k][e >> 1] >> (16 * (e & 1))));
                    m[e] = 0.25f * sum;
                }
                typedef float f2 __attribute__((ext_vector_type(2)));
                typedef __bf16 b2 __attribute__((ext_vector_type(2)));
                const uint32_t lo = __builtin_bit_cast(uint32_t, __builtin_convertvector(f2{m[0], m[1]}, b2));
                const uint32_t hi = __builtin_bit_cast(uint32_t, __builtin_convertvector(f2{m[2], m[3]}, b2));
                const size_t qbase = (size_t)(cur.n * (H >> 1) + (cur.y0 >> 1)) * (W >> 1) + (cur.x0 >> 1);
                const int qloc = (2 * (wv >> 1) + mp) * (W >> 1) + (xl >> 1);
                *(uint2*)(P + qbase * cstore + (unsigned)(qloc * cstore + ch0)) = make_uint2(lo, hi);
            }
        }
    }
}

template <int BN, int FUSE>
__global__ __launch_bounds__(TB) __attribute__((amdgpu_waves_per_eu(conv_waves(BN, FUSE), conv_waves(BN, FUSE)))) void k_conv(const uint16_t* __restrict__ X, int ldx, int Cin,
                                             const uint16_t* __restrict__ Wt, const float* __restrict__ bias,
                                             int Cout, int N, int H, int W, uint16_t* __restrict__ Y, int ldy,
                                             float* __restrict__ Yf, int ldyf, int cstore,
                                             const uint16_t* __restrict__ M, int ldm, int relu, int n_items,
                                             const uint16_t* __restrict__ X2, int ldx2, int csplit,
                                             uint16_t* __restrict__ P) {
    using SH = ConvShape<BN>;
    __shared__ __attribute__((aligned(16))) uint16_t smem[SH::kLds];
    uint16_t* Hs = smem;
    uint16_t* Ws = smem + NHALO * CCLD;
    float* Bsm = (float*)(smem + NHALO * CCLD + 9 * BN * CCLD);  // biases of the LDS weights' channels
    const int tiles_x = (W + TX - 1) / TX, tiles_y = (H + TY - 1) / TY;
    const int n_tiles = N * tiles_x * tiles_y;
    const int nchunk = Cin / CBK;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int fr = lane & 15;
    const uint4 zero = make_uint4(0, 0, 0, 0);
    struct Item {
        int n, y0, x0, co0;
    };
    auto decode = [&](int i) {  // i is uniform: the fields live in scalar registers
        Item it;
        const int tile = i % n_tiles;
        it.co0 = __builtin_amdgcn_readfirstlane((i / n_tiles) * BN);
        it.x0 = __builtin_amdgcn_readfirstlane((tile % tiles_x) * TX);
        const int r = tile / tiles_x;
        it.y0 = __builtin_amdgcn_readfirstlane((r % tiles_y) * TY);
        it.n = __builtin_amdgcn_readfirstlane(r / tiles_y);
        return it;
    };
    uint4 hreg[SH::kHR], wreg[SH::kWR];
    auto load_halo = [&](const Item& it, int ci0, bool valid) {
        // Halo slot r of this thread: pixel hp = (t + r TB) / 4 of the (TX + 2) x
        // (TY + 2) halo, 16-byte channel group q, read through a buffer
        // resource over the item's image: offsets past its end (pixels off the
        // image, unused slots, or every slot when there is no next step) read
        // zeros without a branch.
        const int y0 = it.y0 - 1, x0 = it.x0 - 1;
        const bool up = (FUSE & kFuseUp) != 0 && ci0 >= csplit;  // the upsampled half-resolution tensor
        const int sh = up ? 1 : 0, Wl = W >> sh, ld = up ? ldx2 : ldx, c0 = up ? ci0 - csplit : ci0;
        const uint32_t img = (uint32_t)(H >> sh) * (uint32_t)Wl * (uint32_t)ld;  // elements of one image
        const uint16_t* base = (up ? X2 : X) + (size_t)it.n * img;
        const __amdgpu_buffer_rsrc_t rs = buffer_rsrc(base, valid ? 2 * img : 0);
#pragma unroll
        for (int r = 0; r < SH::kHR; ++r) {
            const int i = t + r * TB;
            const int hp = i >> 2, q = i & 3;
            const int hy = hp / HXX, hx = hp - hy * HXX;
            const int yy = y0 + hy, xx = x0 + hx;
            const bool ok = i < SH::kHaloU4 && (unsigned)yy < (unsigned)H && (unsigned)xx < (unsigned)W;
            const uint32_t off = ok ? 2u * (uint32_t)(((yy >> sh) * Wl + (xx >> sh)) * ld + c0 + q * 8) : kOOB;
            hreg[r] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 0));
        }
    };
    float breg = 0.0f;
    auto load_w = [&](int co0, int ci0) {
        const __amdgpu_buffer_rsrc_t rb = buffer_rsrc(bias, bias ? 4u * Cout : 0u);
        const bool okb = t < BN && co0 + t < Cout;
        breg = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rb, okb ? 4u * (co0 + t) : kOOB, 0, 0));
        const __amdgpu_buffer_rsrc_t rw = buffer_rsrc(Wt, 18u * Cout * Cin);
#pragma unroll
        for (int r = 0; r < SH::kWR; ++r) {
            const int i = t + r * TB;
            const int q = i & 3, rr = (i >> 2) % BN, tap = (i >> 2) / BN;
            const int co = co0 + rr;
            const bool ok = i < SH::kWU4 && co < Cout;
            const uint32_t off = ok ? 2u * (uint32_t)((tap * Cout + co) * Cin + ci0 + q * 8) : kOOB;
            wreg[r] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(rw, off, 0, 0));
        }
    };
    int i = blockIdx.x, c = 0;
    if (i >= n_items) return;
    Item cur = decode(i);
    load_halo(cur, 0, true);
    load_w(cur.co0, 0);
    bool need_w = true;
    f32x4 acc[4][BN / 16];
    while (true) {
        __syncthreads();  // the previous step's LDS reads (and epilogue) are done
        {
#pragma unroll
            for (int r = 0; r < SH::kHR; ++r) {
                const int k = t + r * TB;
                if (k < SH::kHaloU4) *(uint4*)&Hs[cofs(k >> 2, k & 3)] = hreg[r];
            }
        }
        if (need_w) {
            if (t < BN) Bsm[t] = breg;
#pragma unroll
            for (int r = 0; r < SH::kWR; ++r) {
                const int k = t + r * TB;
                if (k < SH::kWU4) *(uint4*)&Ws[cofs(k >> 2, k & 3)] = wreg[r];
            }
        }
        __syncthreads();
        // the next step; its loads go out between the MFMAs below
        int ni = i, nc = c + 1;
        if (nc == nchunk) {
            nc = 0;
            ni = i + gridDim.x;
        }
        const bool has_next = ni < n_items;
        const Item nxt = has_next ? decode(ni) : cur;
        const bool next_w = has_next && (nxt.co0 != cur.co0 || nc != c);
        if (c == 0) {
#pragma unroll
            for (int mt = 0; mt < 4; ++mt)
#pragma unroll
                for (int nb = 0; nb < BN / 16; ++nb) acc[mt][nb] = f32x4{0.f, 0.f, 0.f, 0.f};
        }
        // Wave w owns pixel rows 4 (w / 2) .. + 3 of the 16 columns (w % 2) * 16 ..;
        // per horizontal offset dx it reads the 6 halo rows it spans once and
        // uses each for up to 3 vertical taps.  The weights are the MFMA's A
        // operand and the pixels its B operand, so each lane ends up holding 4
        // consecutive output channels of one pixel.
#pragma unroll
        for (int dx = 0; dx < 3; ++dx) {
            bf16x8 a[6];
#pragma unroll
            for (int j = 0; j < 6; ++j)
                a[j] = __builtin_bit_cast(
                    bf16x8, *(const uint4*)&Hs[cofs((4 * (wv >> 1) + j) * HXX + (wv & 1) * 16 + fr + dx, lane >> 4)]);
#pragma unroll
            for (int dy = 0; dy < 3; ++dy) {
                const int tap = dy * 3 + dx;
#pragma unroll
                for (int nb = 0; nb < BN / 16; ++nb) {
                    const bf16x8 b =
                        __builtin_bit_cast(bf16x8, *(const uint4*)&Ws[cofs(tap * BN + nb * 16 + fr, lane >> 4)]);
#pragma unroll
                    for (int mt = 0; mt < 4; ++mt)
                        acc[mt][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b, a[mt + dy], acc[mt][nb], 0, 0, 0);
                }
            }
            if (dx == 0) load_halo(nxt, nc * CBK, has_next);
            if (dx == 1 && next_w) load_w(nxt.co0, nc * CBK);
        }
        if (c == nchunk - 1)
            conv_store<BN, FUSE>(acc, Bsm, cur.n, cur.y0, cur.x0, cur.co0, wv, lane, H, W, Y, ldy, Yf, ldyf, cstore, M, ldm,
                                 relu, P);
        if (!has_next) break;
        i = ni;
        c = nc;
        cur = nxt;
        need_w = next_w;
    }
}

// ---------------------------------------------------------- convolution, 8 waves
// The forward convolutions without packed input, mask or FP32 output (plain,
// upsampled input, pooled output), for 32 or 64 output channels per item:
// blocks of 8 waves on 32 x 16 pixel tiles (wave w owns the same 16 x 4
// pixel block as in k_conv, rows 4 (w / 2) ..), with two LDS buffers, each
// holding one step's 34 x 18 halo and weights.  Both go from memory straight
// into LDS (buffer loads with the LDS bit: lane l's 16 bytes land at the
// instruction's LDS base + 16 l, and an offset past the resource writes
// zeros), so the loads of step s + 1 run into one buffer while the waves
// run the MFMAs of step s on the other, without staging registers.  The
// MFMAs, their order and the epilogue are those of k_conv, so the outputs are
// the same bits.
constexpr int TB2 = 512, TY2 = 16, HYY2 = TY2 + 2, NHALO2 = HXX * HYY2;
constexpr int kHaloUnits2 = NHALO2 * 4;                      // 16-byte units of the halo
constexpr int kHaloLoads2 = (kHaloUnits2 + 63) / 64;         // wave loads that cover it
typedef __attribute__((address_space(3))) void lds_void;

template <int BN>
struct Conv2Shape {
    static constexpr int kWUnits = 9 * BN * 4;               // 16-byte units of one chunk's weights
    static constexpr int kWLoads = kWUnits / 64;
    static constexpr int kHalo = kHaloLoads2 * 64 * 8;       // halves; the last load's spill-over included
    static constexpr int kBuf = kHalo + kWUnits * 8 + 2 * 64;  // + 64 FP32 biases
    static_assert(kWUnits % 64 == 0, "whole wave loads of weights");
};

template <int BN, int FUSE>
__global__ __launch_bounds__(TB2) void k_conv2(const uint16_t* __restrict__ X, int ldx, int Cin,
                                              const uint16_t* __restrict__ Wt, const float* __restrict__ bias,
                                              int Cout, int N, int H, int W, uint16_t* __restrict__ Y, int ldy,
                                              float* __restrict__ Yf, int ldyf, int cstore, int relu, int n_items,
                                              const uint16_t* __restrict__ X2, int ldx2, int csplit,
                                              uint16_t* __restrict__ P) {
    static_assert(((FUSE & ~(kFuseUp | kFusePool)) == 0 && (BN == 32 || BN == 64)) || (FUSE == kOutF32 && BN == 16),
                  "k_conv2 variants");
    using SH = Conv2Shape<BN>;
    // two separate arrays, so that the compiler tells the buffers apart and
    // the MFMA reads of one need not wait for the loads into the other
    __shared__ __attribute__((aligned(16))) uint16_t buf0[SH::kBuf];
    __shared__ __attribute__((aligned(16))) uint16_t buf1[SH::kBuf];
    const int tiles_x = (W + TX - 1) / TX, tiles_y = (H + TY2 - 1) / TY2;
    const int n_tiles = N * tiles_x * tiles_y;
    const int nchunk = Cin / CBK;
    const int t = threadIdx.x, lane = t & 63, wv = __builtin_amdgcn_readfirstlane(t >> 6);
    const int fr = lane & 15;
    struct Item {
        int n, y0, x0, co0;
    };
    auto decode = [&](int i) {  // i is uniform: the fields live in scalar registers
        Item it;
        const int tile = i % n_tiles;
        it.co0 = __builtin_amdgcn_readfirstlane((i / n_tiles) * BN);
        it.x0 = __builtin_amdgcn_readfirstlane((tile % tiles_x) * TX);
        const int r = tile / tiles_x;
        it.y0 = __builtin_amdgcn_readfirstlane((r % tiles_y) * TY2);
        it.n = __builtin_amdgcn_readfirstlane(r / tiles_y);
        return it;
    };
    // The loads of one step into buffer B: halo unit u = load j, lane l is
    // LDS row u / 4 (a halo pixel), 16-byte slot u % 4, which holds channel
    // group (u % 4) ^ ((row / 2) % 4) of that pixel (the swizzle of cofs).
    auto issue = [&](uint16_t* B, const Item& it, int ci0, bool with_w) __attribute__((always_inline)) {
        const bool up = (FUSE & kFuseUp) != 0 && ci0 >= csplit;  // the upsampled half-resolution tensor
        const int sh = up ? 1 : 0, Wl = W >> sh, ld = up ? ldx2 : ldx, c0 = up ? ci0 - csplit : ci0;
        const uint32_t img = (uint32_t)(H >> sh) * (uint32_t)Wl * (uint32_t)ld;
        const __amdgpu_buffer_rsrc_t rs = buffer_rsrc((up ? X2 : X) + (size_t)it.n * img, 2 * img);
        for (int j = wv; j < kHaloLoads2; j += TB2 / 64) {
            const int u = j * 64 + lane, r = u >> 2;
            const int q = (u & 3) ^ ((r >> 1) & 3);
            const int hy = r / HXX, hx = r - hy * HXX;
            const int yy = it.y0 - 1 + hy, xx = it.x0 - 1 + hx;
            const bool ok = u < kHaloUnits2 && (unsigned)yy < (unsigned)H && (unsigned)xx < (unsigned)W;
            const uint32_t off = ok ? 2u * (uint32_t)(((yy >> sh) * Wl + (xx >> sh)) * ld + c0 + q * 8) : kOOB;
            __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, (lds_void*)(B + j * 512), 16, off, 0, 0, 0);
        }
        if (with_w) {
            const __amdgpu_buffer_rsrc_t rw = buffer_rsrc(Wt, 18u * Cout * Cin);
            uint16_t* Ws = B + SH::kHalo;
            for (int j = wv; j < SH::kWLoads; j += TB2 / 64) {
                const int u = j * 64 + lane, r = u >> 2;
                const int q = (u & 3) ^ ((r >> 1) & 3);
                const int tap = r / BN, co = it.co0 + (r - tap * BN);
                const uint32_t off = co < Cout ? 2u * (uint32_t)((tap * Cout + co) * Cin + ci0 + q * 8) : kOOB;
                __builtin_amdgcn_raw_ptr_buffer_load_lds(rw, (lds_void*)(Ws + j * 512), 16, off, 0, 0, 0);
            }
            if (wv == TB2 / 64 - 1) {  // the biases of the item's channels, 4 bytes per lane
                const __amdgpu_buffer_rsrc_t rb = buffer_rsrc(bias, bias ? 4u * Cout : 0u);
                const uint32_t off = it.co0 + lane < Cout ? 4u * (uint32_t)(it.co0 + lane) : kOOB;
                __builtin_amdgcn_raw_ptr_buffer_load_lds(rb, (lds_void*)(Ws + SH::kWUnits * 8), 4, off, 0, 0, 0);
            }
        }
    };
    int i = blockIdx.x, c = 0;
    if (i >= n_items) return;
    Item cur = decode(i);
    // weights (output block, chunk) held by each buffer: a step whose weights
    // a buffer already holds loads only the halo
    int wkey0 = cur.co0 * 1024, wkey1 = -1;
    issue(buf0, cur, 0, true);
    f32x4 acc[4][BN / 16];
    // One step: MFMAs of (cur, c) from buffer A while the next step's loads
    // go into buffer B.  Returns false after the block's last step.
    auto step = [&](uint16_t* A, uint16_t* B, int& wkeyB) __attribute__((always_inline)) -> bool {
        int ni = i, nc = c + 1;
        if (nc == nchunk) {
            nc = 0;
            ni = i + gridDim.x;
        }
        const bool has_next = ni < n_items;
        const Item nxt = has_next ? decode(ni) : cur;
        __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this wave's loads into A have landed
        __syncthreads();                     // every wave's; and B is no longer read
        if (has_next) {
            const int k = nxt.co0 * 1024 + nc;
            issue(B, nxt, nc * CBK, k != wkeyB);
            wkeyB = k;
        }
        const uint16_t* Hs = A;
        const uint16_t* Ws = A + SH::kHalo;
        if (c == 0) {
#pragma unroll
            for (int mt = 0; mt < 4; ++mt)
#pragma unroll
                for (int nb = 0; nb < BN / 16; ++nb) acc[mt][nb] = f32x4{0.f, 0.f, 0.f, 0.f};
        }
        // per horizontal offset dx: the 6 halo rows and the 3 x BN / 16 weight
        // fragments of its taps are read up front, so that the reads are in
        // flight together and the MFMAs (in k_conv's order) wait on them once
#pragma unroll
        for (int dx = 0; dx < 3; ++dx) {
            bf16x8 a[6], bw[3][BN / 16];
#pragma unroll
            for (int j = 0; j < 6; ++j)
                a[j] = __builtin_bit_cast(
                    bf16x8, *(const uint4*)&Hs[cofs((4 * (wv >> 1) + j) * HXX + (wv & 1) * 16 + fr + dx, lane >> 4)]);
#pragma unroll
            for (int dy = 0; dy < 3; ++dy)
#pragma unroll
                for (int nb = 0; nb < BN / 16; ++nb)
                    bw[dy][nb] = __builtin_bit_cast(
                        bf16x8, *(const uint4*)&Ws[cofs((dy * 3 + dx) * BN + nb * 16 + fr, lane >> 4)]);
            __builtin_amdgcn_sched_barrier(0);  // keep the reads ahead of the MFMAs
#pragma unroll
            for (int dy = 0; dy < 3; ++dy)
#pragma unroll
                for (int nb = 0; nb < BN / 16; ++nb)
#pragma unroll
                    for (int mt = 0; mt < 4; ++mt)
                        acc[mt][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bw[dy][nb], a[mt + dy], acc[mt][nb], 0, 0, 0);
        }
        if (c == nchunk - 1)
            conv_store<BN, FUSE>(acc, (const float*)(Ws + SH::kWUnits * 8), cur.n, cur.y0, cur.x0, cur.co0, wv, lane, H,
                                 W, Y, ldy, Yf, ldyf, cstore, nullptr, 0, relu, P);
        if (!has_next) return false;
        i = ni;
        c = nc;
        cur = nxt;
        return true;
    };
    while (step(buf0, buf1, wkey1) && step(buf1, buf0, wkey0)) {
    }
}

// ------------------------------------------- first layer, 8 waves, packed input
// The network's first convolution (9 input channels packed from the float4
// render targets, at most 32 outputs, one weight set) on k_conv2's 32 x 16
// tiles.  The three targets' halo pixels go from memory straight into LDS
// (raw, two buffers: the next tile's loads run while this one is packed and
// convolved); the threads pack them into the BF16 halo (channels 0-8; the
// zero channel groups 2 and 3 are written once), then the MFMAs and the
// epilogue are k_conv's: each output pixel gets the same MFMA sequence, so
// the outputs do not depend on the tiling.
// Rows of the first layer's tiles.  16: one 8-wave block per CU (1.07 ms on
// 128 x 512x512); 8 rows fit two 4-wave blocks per CU and were slower (1.31 ms).
constexpr int kPackRows = 16;
template <int TYP>  // tile rows: 4 of them per pair of waves
struct PackShape {
    static constexpr int kThreads = TYP * 32, kWaves = kThreads / 64;
    static constexpr int kHalo = HXX * (TYP + 2);           // halo pixels
    static constexpr int kRawLoads = (kHalo + 63) / 64;     // wave loads per target over the halo
    static constexpr int kRawUnits = kRawLoads * 64;        // float4 per target and buffer (the last load's spill-over included)
};

template <int TYP>
__global__ __launch_bounds__(PackShape<TYP>::kThreads) void k_conv2_pack(const uint16_t* __restrict__ Wt, const float* __restrict__ bias,
                                                   int Cout, int N, int H, int W, uint16_t* __restrict__ Y, int ldy,
                                                   int cstore, int relu, int n_items, PackSrc pk) {
    constexpr int BN = 32;
    using SH = Conv2Shape<BN>;
    using PS = PackShape<TYP>;
    constexpr int kRawUnits = PS::kRawUnits, kRawLoads = PS::kRawLoads, NH = PS::kHalo, NT = PS::kThreads;
    __shared__ __attribute__((aligned(16))) float4 raw0[3 * kRawUnits];
    __shared__ __attribute__((aligned(16))) float4 raw1[3 * kRawUnits];
    __shared__ __attribute__((aligned(16))) uint16_t Hs[NH * CCLD];
    __shared__ __attribute__((aligned(16))) uint16_t Ws[SH::kWUnits * 8 + 2 * 64];  // + 64 FP32 biases (one per lane)
    const int tiles_x = (W + TX - 1) / TX, tiles_y = (H + TYP - 1) / TYP;
    const int n_tiles = N * tiles_x * tiles_y;
    const int t = threadIdx.x, lane = t & 63, wv = __builtin_amdgcn_readfirstlane(t >> 6);
    const int fr = lane & 15;
    struct Item {
        int n, y0, x0;
    };
    auto decode = [&](int i) {  // i is uniform: the fields live in scalar registers
        Item it;
        const int tile = i % n_tiles;
        it.x0 = __builtin_amdgcn_readfirstlane((tile % tiles_x) * TX);
        const int r = tile / tiles_x;
        it.y0 = __builtin_amdgcn_readfirstlane((r % tiles_y) * TYP);
        it.n = __builtin_amdgcn_readfirstlane(r / tiles_y);
        return it;
    };
    // halo pixel u of the item (row u / (TX + 2), column u % (TX + 2)) into
    // unit u of each target's part of buffer B; pixels off the H x W input
    // read zeros, those past the targets' h x w repeat the edge pixel
    auto issue = [&](float4* B, const Item& it) __attribute__((always_inline)) {
        const uint32_t img = (uint32_t)pk.h * (uint32_t)pk.w;
        const float4* src[3] = {pk.rgba, pk.alb, pk.nrm};
        for (int j = wv; j < kRawLoads; j += PS::kWaves) {
            const int u = j * 64 + lane;
            const int hy = u / HXX, hx = u - hy * HXX;
            const int yy = it.y0 - 1 + hy, xx = it.x0 - 1 + hx;
            const bool ok = u < NH && (unsigned)yy < (unsigned)H && (unsigned)xx < (unsigned)W;
            const uint32_t off = ok ? 16u * (uint32_t)(min(yy, pk.h - 1) * pk.w + min(xx, pk.w - 1)) : kOOB;
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const __amdgpu_buffer_rsrc_t rs = buffer_rsrc(src[a] + (size_t)it.n * img, 16u * img);
                __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, (lds_void*)(B + a * kRawUnits + j * 64), 16, off, 0, 0, 0);
            }
        }
    };
    int i = blockIdx.x;
    if (i >= n_items) return;
    Item cur = decode(i);
    issue(raw0, cur);
    {  // the one weight set and its biases (k_conv2's layout), and the zero channel groups
        const __amdgpu_buffer_rsrc_t rw = buffer_rsrc(Wt, 18u * Cout * CBK);
        for (int j = wv; j < SH::kWLoads; j += PS::kWaves) {
            const int u = j * 64 + lane, r = u >> 2;
            const int q = (u & 3) ^ ((r >> 1) & 3);
            const int tap = r / BN, co = r - tap * BN;
            const uint32_t off = co < Cout ? 2u * (uint32_t)((tap * Cout + co) * CBK + q * 8) : kOOB;
            __builtin_amdgcn_raw_ptr_buffer_load_lds(rw, (lds_void*)(Ws + j * 512), 16, off, 0, 0, 0);
        }
        if (wv == PS::kWaves - 1) {
            const __amdgpu_buffer_rsrc_t rb = buffer_rsrc(bias, bias ? 4u * Cout : 0u);
            const uint32_t off = lane < Cout ? 4u * (uint32_t)lane : kOOB;
            __builtin_amdgcn_raw_ptr_buffer_load_lds(rb, (lds_void*)(Ws + SH::kWUnits * 8), 4, off, 0, 0, 0);
        }
        for (int r = t; r < NH; r += NT) {
            *(uint4*)&Hs[cofs(r, 2)] = make_uint4(0u, 0u, 0u, 0u);
            *(uint4*)&Hs[cofs(r, 3)] = make_uint4(0u, 0u, 0u, 0u);
        }
    }
    const float* Bsm = (const float*)(Ws + SH::kWUnits * 8);
    // One step: the tile whose raw halo is in A, while the next tile's loads
    // go into B (issued first, so that they run beside the pack, the MFMAs
    // and the stores); the outputs of the previous tile are stored before
    // this tile's MFMAs, so that the wait at the next step's start finds them
    // long under way.  Returns false after the block's last tile.
    f32x4 acc[4][BN / 16];
    Item prev = cur;
    bool have_prev = false;
    auto step = [&](const float4* A, float4* B) __attribute__((always_inline)) -> bool {
        const int ni = i + gridDim.x;
        const bool has_next = ni < n_items;
        const Item nxt = has_next ? decode(ni) : cur;
        __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this wave's loads into A (and the weights) have landed
        __syncthreads();                     // every wave's; B and Hs are no longer read
        if (has_next) issue(B, nxt);
        for (int r = t; r < NH; r += NT) {
            float v[9];
            pack9(A[r], A[kRawUnits + r], A[2 * kRawUnits + r], v);
            // v_cvt_pk_bf16_f32 rounds to nearest even as f2bf (no NaN after sane)
            typedef float f2 __attribute__((ext_vector_type(2)));
            typedef __bf16 b2 __attribute__((ext_vector_type(2)));
            uint32_t o[5];
#pragma unroll
            for (int j = 0; j < 5; ++j)
                o[j] = __builtin_bit_cast(uint32_t, __builtin_convertvector(f2{v[2 * j], j < 4 ? v[2 * j + 1] : 0.f}, b2));
            *(uint4*)&Hs[cofs(r, 0)] = make_uint4(o[0], o[1], o[2], o[3]);
            *(uint4*)&Hs[cofs(r, 1)] = make_uint4(o[4], 0u, 0u, 0u);
        }
        // Hs is packed: every wave's LDS writes are done.  A barrier without
        // __syncthreads' fence, which would wait for the loads into B too.
        asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
        if (have_prev)
            conv_store<BN, 0>(acc, Bsm, prev.n, prev.y0, prev.x0, 0, wv, lane, H, W, Y, ldy, nullptr, 0, cstore,
                              nullptr, 0, relu, nullptr);
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
#pragma unroll
            for (int nb = 0; nb < BN / 16; ++nb) acc[mt][nb] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int dx = 0; dx < 3; ++dx) {
            bf16x8 a[6], bw[3][BN / 16];
#pragma unroll
            for (int j = 0; j < 6; ++j)
                a[j] = __builtin_bit_cast(
                    bf16x8, *(const uint4*)&Hs[cofs((4 * (wv >> 1) + j) * HXX + (wv & 1) * 16 + fr + dx, lane >> 4)]);
#pragma unroll
            for (int dy = 0; dy < 3; ++dy)
#pragma unroll
                for (int nb = 0; nb < BN / 16; ++nb)
                    bw[dy][nb] = __builtin_bit_cast(
                        bf16x8, *(const uint4*)&Ws[cofs((dy * 3 + dx) * BN + nb * 16 + fr, lane >> 4)]);
            __builtin_amdgcn_sched_barrier(0);  // keep the reads ahead of the MFMAs
#pragma unroll
            for (int dy = 0; dy < 3; ++dy)
#pragma unroll
                for (int nb = 0; nb < BN / 16; ++nb)
#pragma unroll
                    for (int mt = 0; mt < 4; ++mt)
                        acc[mt][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bw[dy][nb], a[mt + dy], acc[mt][nb], 0, 0, 0);
        }
        prev = cur;
        have_prev = true;
        if (!has_next) return false;
        i = ni;
        cur = nxt;
        return true;
    };
    while (step(raw0, raw1) && step(raw1, raw0)) {
    }
    conv_store<BN, 0>(acc, Bsm, prev.n, prev.y0, prev.x0, 0, wv, lane, H, W, Y, ldy, nullptr, 0, cstore, nullptr, 0,
                      relu, nullptr);
}

// dW[tap, co, ci] += sum_p dY[p, co] * X[p + off(tap), ci] over one pixel chunk.
// Tile 64 co x 64 ci, K step 32 pixels; LDS tiles are stored channel-major
// (pixels contiguous) so each lane's 8 K values are one 16-byte read.
__global__ __launch_bounds__(TB) void k_wgrad(const uint16_t* __restrict__ X, int ldx, int Cin,
                                              const uint16_t* __restrict__ dY, int ldy, int Cout, int N, int H,
                                              int W, int chunk, float* __restrict__ dW) {
    __shared__ __attribute__((aligned(16))) uint16_t As[64 * CLD];  // [co][px]
    __shared__ __attribute__((aligned(16))) uint16_t Bs[64 * CLD];  // [ci][px]
    const int P = N * H * W;
    const int tap = blockIdx.y;
    const int n_ci_t = (Cin + 63) / 64;
    const int co0 = (blockIdx.z / n_ci_t) * 64, ci0 = (blockIdx.z % n_ci_t) * 64;
    const int pbeg = blockIdx.x * chunk, pend = min(P, pbeg + chunk);
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int dy = tap / 3 - 1, dx = tap % 3 - 1;
    f32x4 acc[4];
    for (int i = 0; i < 4; ++i) acc[i] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int lr = t >> 3, lc = (t & 7) * 8;  // pixel row 0..31, channel chunk
    const int fr = lane & 15, fk = (lane >> 4) * 8;
    for (int pk = pbeg; pk < pend; pk += 32) {
        const int p = pk + lr;
        uint16_t av[8], bv[8];
        bool pv = p < pend;
        uint4 a4 = make_uint4(0, 0, 0, 0), b4 = make_uint4(0, 0, 0, 0);
        if (pv && co0 + lc < Cout) a4 = *(const uint4*)(dY + (size_t)p * ldy + co0 + lc);
        if (pv && ci0 + lc < Cin) {
            int n_ = p / (H * W), rem = p - n_ * H * W, y = rem / W, x = rem - y * W;
            int yy = y + dy, xx = x + dx;
            if (yy >= 0 && yy < H && xx >= 0 && xx < W)
                b4 = *(const uint4*)(X + (size_t)((n_ * H + yy) * W + xx) * ldx + ci0 + lc);
        }
        memcpy(av, &a4, 16);
        memcpy(bv, &b4, 16);
        __syncthreads();
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            As[(lc + j) * CLD + lr] = av[j];
            Bs[(lc + j) * CLD + lr] = bv[j];
        }
        __syncthreads();
        bf16x8 a = __builtin_bit_cast(bf16x8, *(const uint4*)&As[(wv * 16 + fr) * CLD + fk]);
#pragma unroll
        for (int nb = 0; nb < 4; ++nb) {
            bf16x8 b = __builtin_bit_cast(bf16x8, *(const uint4*)&Bs[(nb * 16 + fr) * CLD + fk]);
            acc[nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[nb], 0, 0, 0);
        }
    }
#pragma unroll
    for (int nb = 0; nb < 4; ++nb)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int co = co0 + wv * 16 + (lane >> 4) * 4 + i;
            const int ci = ci0 + nb * 16 + (lane & 15);
            if (co < Cout && ci < Cin) atomicAdd(&dW[((size_t)tap * Cout + co) * Cin + ci], acc[nb][i]);
        }
}

__global__ void k_bias_grad(const uint16_t* __restrict__ dY, int ldy, int Cout, int P, int chunk, float* db) {
    int pbeg = blockIdx.x * chunk, pend = min(P, pbeg + chunk);
    for (int c = threadIdx.x; c < Cout; c += blockDim.x) {
        float s = 0.0f;
        for (int p = pbeg; p < pend; ++p) s += bf2f(dY[(size_t)p * ldy + c]);
        atomicAdd(&db[c], s);
    }
}

// ---------------------------------------------------------------- glue kernels
// dS[hi] += (S[hi] > 0) * dP[lo] / 4   (gradient of the pooled skip tensor)
__global__ void k_pool_bwd(const uint16_t* __restrict__ dPl, int C, const uint16_t* __restrict__ S, int lds,
                           uint16_t* __restrict__ dS, int ldd, int N, int H, int W) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    int C8 = C / 8;
    if (i >= (size_t)N * H * W * C8) return;
    int c8 = (int)(i % C8);
    size_t p = i / C8;
    int x = (int)(p % W), y = (int)((p / W) % H), n = (int)(p / ((size_t)W * H));
    size_t lo = ((size_t)(n * (H / 2) + y / 2) * (W / 2) + x / 2) * C + c8 * 8;
    uint4 gv = *(const uint4*)(dPl + lo), sv = *(const uint4*)(S + p * lds + c8 * 8),
          dv = *(const uint4*)(dS + p * ldd + c8 * 8);
    uint16_t g[8], s[8], d[8];
    memcpy(g, &gv, 16);
    memcpy(s, &sv, 16);
    memcpy(d, &dv, 16);
    for (int j = 0; j < 8; ++j)
        if (bf2f(s[j]) > 0.0f) d[j] = f2bf(bf2f(d[j]) + 0.25f * bf2f(g[j]));
    memcpy(&dv, d, 16);
    *(uint4*)(dS + p * ldd + c8 * 8) = dv;
}

// nearest 2x upsampling into a channel slice: Y[n, y, x, c] = X[n, y/2, x/2, c]
__global__ void k_up(const uint16_t* __restrict__ X, int C, int N, int H, int W, uint16_t* __restrict__ Y, int ldy) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    int C8 = C / 8, Hh = 2 * H, Wh = 2 * W;
    if (i >= (size_t)N * Hh * Wh * C8) return;
    int c8 = (int)(i % C8);
    size_t p = i / C8;
    int x = (int)(p % Wh), y = (int)((p / Wh) % Hh), n = (int)(p / ((size_t)Wh * Hh));
    uint4 v = *(const uint4*)(X + ((size_t)(n * H + y / 2) * W + x / 2) * C + c8 * 8);
    *(uint4*)(Y + p * ldy + c8 * 8) = v;
}

// dX[lo] = (X[lo] > 0) * sum of the 4 children of dY (upsampling backward + ReLU mask)
__global__ void k_up_bwd(const uint16_t* __restrict__ dY, int ldy, int C, const uint16_t* __restrict__ Xf,
                         int N, int H, int W, uint16_t* __restrict__ dX) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    int C8 = C / 8;
    if (i >= (size_t)N * H * W * C8) return;
    int c8 = (int)(i % C8);
    size_t p = i / C8;
    int x = (int)(p % W), y = (int)((p / W) % H), n = (int)(p / ((size_t)W * H));
    float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int k = 0; k < 4; ++k) {
        size_t hi = ((size_t)(n * 2 * H + 2 * y + (k >> 1)) * (2 * W) + 2 * x + (k & 1)) * ldy + c8 * 8;
        uint4 v = *(const uint4*)(dY + hi);
        uint16_t e[8];
        memcpy(e, &v, 16);
        for (int j = 0; j < 8; ++j) s[j] += bf2f(e[j]);
    }
    uint4 xv = *(const uint4*)(Xf + p * C + c8 * 8);
    uint16_t xe[8], o[8];
    memcpy(xe, &xv, 16);
    for (int j = 0; j < 8; ++j) o[j] = bf2f(xe[j]) > 0.0f ? f2bf(s[j]) : (uint16_t)0;
    uint4 ov;
    memcpy(&ov, o, 16);
    *(uint4*)(dX + p * C + c8 * 8) = ov;
}

// Network input from packed records (n, h, w, 9): rgb radiance, albedo (as
// stored in the albedo target, baseColor / pi), normal.
__global__ void k_pack_records(const float* __restrict__ in, int P, uint16_t* __restrict__ X, float* __restrict__ lin) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    const float* r = in + (size_t)i * 9;
    float v[9];
    pack9(make_float4(r[0], r[1], r[2], 0.f), make_float4(r[3], r[4], r[5], 0.f), make_float4(r[6], r[7], r[8], 0.f),
          v);
    uint16_t o[32];
    for (int k = 0; k < 32; ++k) o[k] = k < 9 ? f2bf(v[k]) : (uint16_t)0;
    for (int k = 0; k < 4; ++k) *(uint4*)(X + (size_t)i * 32 + 8 * k) = *(const uint4*)&o[8 * k];
    for (int k = 0; k < 3; ++k) lin[(size_t)i * 4 + k] = v[k];
    lin[(size_t)i * 4 + 3] = 0.f;
}

constexpr float kLogClamp = 10.0f;

// denoised linear RGB = expm1(min(y + log1p(noisy), 10)), cropped to H x W;
// out may be rgba itself (each pixel is read before it is written, by the
// same thread)
__global__ void k_unpack(const float* __restrict__ y, const float4* rgba, int N, int H, int W, int Hp, int Wp,
                         float4* out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)N * H * W) return;
    const int img = (int)(i / ((size_t)H * W)), j = (int)(i % ((size_t)H * W));
    int x = j % W, yy = j / W;
    size_t s = ((size_t)img * Hp + yy) * Wp + x;
    const float4 c = rgba[i];
    const float lin[3] = {log_radiance(c.x), log_radiance(c.y), log_radiance(c.z)};
    float r[3];
    for (int k = 0; k < 3; ++k) r[k] = expm1f(fminf(y[s * 4 + k] + lin[k], kLogClamp));
    out[i] = make_float4(r[0], r[1], r[2], 1.0f);
}

__global__ void k_unpack_records(const float* __restrict__ y, const float* __restrict__ lin, int P, float* __restrict__ out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    for (int k = 0; k < 3; ++k) out[(size_t)i * 3 + k] = expm1f(fminf(y[(size_t)i * 4 + k] + lin[(size_t)i * 4 + k], kLogClamp));
}

// L = mean |yhat - log1p(t)| + mean (o - t)^2 / (t^2 + 0.01), o = expm1(yhat),
// yhat = y + log1p(noisy).  Writes dL/dy (BF16, channels 0-2 of a 32-wide row).
__global__ void k_loss(const float* __restrict__ y, const float* __restrict__ lin, const float* __restrict__ target,
                       int P, uint16_t* __restrict__ dy, float* __restrict__ loss) {
    __shared__ float red[TB];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.0f;
    if (i < P) {
        uint16_t g[32];
        for (int k = 0; k < 32; ++k) g[k] = 0;
        const float inv = 1.0f / (3.0f * (float)P);
        for (int k = 0; k < 3; ++k) {
            float yh = y[(size_t)i * 4 + k] + lin[(size_t)i * 4 + k];
            float t = fmaxf(sane(target[(size_t)i * 3 + k]), 0.0f);
            float lt = log1pf(t);
            bool clamped = yh > kLogClamp;
            float yc = fminf(yh, kLogClamp);
            float o = expm1f(yc);
            float den = t * t + 0.01f;
            l += fabsf(yh - lt) + (o - t) * (o - t) / den;
            float d = (yh > lt ? 1.0f : (yh < lt ? -1.0f : 0.0f)) + (clamped ? 0.0f : 2.0f * (o - t) / den * (o + 1.0f));
            g[k] = f2bf(d * inv);
        }
        for (int k = 0; k < 4; ++k) *(uint4*)(dy + (size_t)i * 32 + 8 * k) = *(const uint4*)&g[8 * k];
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int s = TB / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss, red[0] / (3.0f * (float)P));
}

__global__ void k_zero_f32(float* p, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

// Adam on the FP32 master weights.
__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       size_t n, float lr, float b1, float b2, float eps, float c1, float c2) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float gi = g[i];
    if (!(gi == gi)) gi = 0.0f;
    float mi = b1 * m[i] + (1.0f - b1) * gi;
    float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// BF16 compute copies of one conv: Wb[tap][co][ci] and the data-gradient
// weights Wd[tap][ci][co] = W[8 - tap][co][ci].
__global__ void k_weights_bf16(const float* __restrict__ w32, int Cin, int Cout, uint16_t* __restrict__ Wb,
                               uint16_t* __restrict__ Wd) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int n = 9 * Cin * Cout;
    if (i >= n) return;
    int ci = i % Cin, co = (i / Cin) % Cout, tap = i / (Cin * Cout);
    uint16_t b = f2bf(w32[i]);
    Wb[i] = b;
    Wd[((size_t)(8 - tap) * Cin + ci) * Cout + co] = b;
}

inline unsigned blocks(size_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

template <class T> T* dalloc(size_t n) {
    T* p = nullptr;
    KF_CHECK(hipMalloc((void**)&p, std::max<size_t>(n, 1) * sizeof(T)));
    return p;
}

// ---------------------------------------------------------------- launchers
// One 3x3 convolution: output tiles of 16 channels when at most 16 are
// stored (the RGB output layer), 64 when 64 divides them, else 32.
// Persistent grid of k_conv<BN>: the blocks that are resident at once (from
// its LDS and registers), or fewer when there are fewer items.
template <int BN, int FUSE>
unsigned conv_grid(unsigned n_items) {
    static unsigned resident = 0;
    if (!resident) {
        int dev = 0, cus = 256, nb = 0;
        KF_CHECK(hipGetDevice(&dev));
        KF_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        KF_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_conv<BN, FUSE>, TB, 0));
        resident = (unsigned)std::max(nb, 1) * (unsigned)cus;
    }
    return std::min(n_items, resident);
}

template <int BN, int FUSE>
unsigned conv2_grid(unsigned n_items) {
    static unsigned resident = 0;
    if (!resident) {
        int dev = 0, cus = 256, nb = 0;
        KF_CHECK(hipGetDevice(&dev));
        KF_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        KF_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_conv2<BN, FUSE>, TB2, 0));
        resident = (unsigned)std::max(nb, 1) * (unsigned)cus;
    }
    return std::min(n_items, resident);
}

template <int TYP>
unsigned conv2_pack_grid(unsigned n_items) {
    static unsigned resident = 0;
    if (!resident) {
        int dev = 0, cus = 256, nb = 0;
        KF_CHECK(hipGetDevice(&dev));
        KF_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        KF_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_conv2_pack<TYP>, PackShape<TYP>::kThreads, 0));
        resident = (unsigned)std::max(nb, 1) * (unsigned)cus;
    }
    return std::min(n_items, resident);
}

// Optional fusions: input channels csplit.. come from X2 (ldx2), an H/2 x W/2
// tensor upsampled by nearest neighbour; P receives the 2x2 average pool of
// the bf16 outputs (H/2 x W/2 x cstore).
struct ConvFuse {
    const uint16_t* X2 = nullptr;
    int ldx2 = 0, csplit = 1 << 30;
    uint16_t* P = nullptr;
    PackSrc pk;  // when pk.rgba is set, the input is packed from the render targets
};

void conv_launch(const uint16_t* X, int ldx, int Cin, const uint16_t* Wt, const float* bias, int Cout, int N, int H,
                 int W, uint16_t* Y, int ldy, float* Yf, int ldyf, int cstore, const uint16_t* M, int ldm, int relu,
                 hipStream_t st, const ConvFuse& fz = ConvFuse()) {
    if (Cin % CBK != 0) throw Error("conv: input channels must be a multiple of 32");
    if (fz.X2 && (H % 2 || W % 2 || fz.csplit % CBK || fz.ldx2 % 8 || Cin - fz.csplit > fz.ldx2))
        throw Error("conv: bad upsampled input");
    if (fz.P && (H % 2 || W % 2 || !Y || Yf || cstore != Cout || cstore % 16))
        throw Error("conv: bad pooled output");
    const unsigned tiles = (unsigned)N * blocks(H, TY) * blocks(W, TX);
    if ((uint64_t)tiles * blocks(cstore, 16) >= (1u << 31)) throw Error("conv: too many tiles");
    // the kernel indexes pixels, weights and rows of a tile with 32-bit integers
    const uint64_t ld_max = std::max<uint64_t>(std::max(ldy, ldyf), std::max(ldm, ldx));
    if ((uint64_t)N * H * W >= (1u << 31) || 18ull * Cout * Cin >= (1u << 31) || (uint64_t)TY2 * W * ld_max >= (1u << 31) ||
        2ull * H * W * ldx >= (1u << 31) || (fz.X2 && 2ull * (H / 2) * (W / 2) * fz.ldx2 >= (1u << 31)))
        throw Error("conv: tensor too large for 32-bit indexing (2 GiB per image)");
    const int fuse = (fz.X2 ? kFuseUp : 0) | (fz.P ? kFusePool : 0) | (fz.pk.rgba ? kFusePack : 0);
    if (fuse & (fuse - 1)) throw Error("conv: one fusion at a time");
    if (fuse && (M || Yf || cstore <= 16)) throw Error("conv: fusions write plain BF16 outputs of > 16 channels");
    if (fz.pk.rgba && (Cin != CBK || Cout > 32 || fz.pk.h > H || fz.pk.w > W || fz.pk.h < 1 || fz.pk.w < 1))
        throw Error("conv: bad packed input");
    const int flags = fuse | (Yf ? kOutF32 : 0) | (M ? kMask : 0);
    auto go = [&](auto bn, auto fl) {
        constexpr int BN = decltype(bn)::value, F = decltype(fl)::value;
        const unsigned n = tiles * blocks(cstore, BN);
        k_conv<BN, F><<<conv_grid<BN, F>(n), TB, 0, st>>>(X, ldx, Cin, Wt, bias, Cout, N, H, W, Y, ldy, Yf, ldyf,
                                                         cstore, M, ldm, relu, (int)n, fz.X2, fz.ldx2, fz.csplit,
                                                         fz.P);
    };
    using I16 = std::integral_constant<int, 16>;
    using I32 = std::integral_constant<int, 32>;
    using I64 = std::integral_constant<int, 64>;
    auto outs = [&](auto bn) {  // the output kinds of an unfused convolution
        switch (flags) {
            case 0: go(bn, std::integral_constant<int, 0>()); break;
            case kOutF32: go(bn, std::integral_constant<int, kOutF32>()); break;
            case kMask: go(bn, std::integral_constant<int, kMask>()); break;
            default: go(bn, std::integral_constant<int, kOutF32 | kMask>()); break;
        }
    };
    auto go2 = [&](auto bn, auto fl) {
        constexpr int BN = decltype(bn)::value, F = decltype(fl)::value;
        const unsigned n = (unsigned)N * blocks(H, TY2) * blocks(W, TX) * blocks(cstore, BN);
        k_conv2<BN, F><<<conv2_grid<BN, F>(n), TB2, 0, st>>>(X, ldx, Cin, Wt, bias, Cout, N, H, W, Y, ldy, Yf, ldyf,
                                                            cstore, relu, (int)n, fz.X2, fz.ldx2, fz.csplit, fz.P);
    };
    // k_conv2 takes the plain, upsampled-input and pooled-output layers of
    // more than 16 output channels, 64 per item where they divide the output
    // (twice the MFMAs per staged halo), and the FP32 output layer (16 per
    // item); k_conv2_pack the packed first layer; k_conv the masked data
    // gradients and the BF16 layers of at most 16 outputs.  (Keeping the
    // 32-channel full-resolution layers on k_conv, which had looked faster
    // for two of them under the profiler, made the whole denoise of
    // 128 x 512x512 slower: 14.16-14.18 against 13.96-14.07 ms.)
    if (!Yf && !M && !fz.pk.rgba && cstore > 16) {
        const bool wide = cstore % 64 == 0;
        auto pick = [&](auto fl) {
            if (wide) go2(I64(), fl);
            else go2(I32(), fl);
        };
        if (fuse == kFuseUp) pick(std::integral_constant<int, kFuseUp>());
        else if (fuse == kFusePool) pick(std::integral_constant<int, kFusePool>());
        else pick(std::integral_constant<int, 0>());
    } else if (cstore <= 16 && flags == kOutF32) {
        go2(I16(), std::integral_constant<int, kOutF32>());
    } else if (cstore <= 16) {
        outs(I16());
    } else if (fuse == kFusePack) {  // 32 inputs, at most 32 outputs: one weight set
        constexpr int TYP = kPackRows;
        const unsigned n = (unsigned)N * blocks(H, TYP) * blocks(W, TX);
        k_conv2_pack<TYP><<<conv2_pack_grid<TYP>(n), PackShape<TYP>::kThreads, 0, st>>>(Wt, bias, Cout, N, H, W, Y, ldy,
                                                                                      cstore, relu, (int)n, fz.pk);
    } else {
        outs(I32());  // masked data gradient
    }
    KF_CHECK(hipGetLastError());
}

void conv_fwd(const uint16_t* X, int ldx, int Cin, const uint16_t* Wb, const float* bias, int Cout, int N, int H,
              int W, uint16_t* Y, int ldy, bool relu, hipStream_t st, float* Yf = nullptr, int ldyf = 0,
              int cstore = -1, const ConvFuse& fz = ConvFuse()) {
    conv_launch(X, ldx, Cin, Wb, bias, Cout, N, H, W, Y, ldy, Yf, ldyf, cstore < 0 ? Cout : std::min(cstore, Cout),
                nullptr, 0, relu ? 1 : 0, st, fz);
}
// dX = conv(dY, Wd) masked by (Xf > 0) when Xf is given
void conv_dgrad(const uint16_t* dY, int ldy, int Cout, const uint16_t* Wd, int Cin, int N, int H, int W, uint16_t* dX,
                int ldx, const uint16_t* Xf, int ldxf, hipStream_t st) {
    conv_launch(dY, ldy, Cout, Wd, nullptr, Cin, N, H, W, dX, ldx, nullptr, 0, Cin, Xf, ldxf, 0, st);
}
void conv_wgrad(const uint16_t* X, int ldx, int Cin, const uint16_t* dY, int ldy, int Cout, int N, int H, int W,
                float* dW, float* db, hipStream_t st) {
    const int P = N * H * W;
    int chunk = 2048;
    while ((size_t)blocks(P, chunk) * 9 * blocks(Cout, 64) * blocks(Cin, 64) < 2048 && chunk > 64) chunk /= 2;
    dim3 grid(blocks(P, chunk), 9, blocks(Cout, 64) * blocks(Cin, 64));
    k_wgrad<<<grid, TB, 0, st>>>(X, ldx, Cin, dY, ldy, Cout, N, H, W, chunk, dW);
    int bchunk = 1024;
    k_bias_grad<<<blocks(P, bchunk), TB, 0, st>>>(dY, ldy, Cout, P, bchunk, db);
}

// ---------------------------------------------------------------- network
constexpr int kLevels = 5;
constexpr int kCh[kLevels] = {32, 64, 96, 128, 160};
constexpr int kInCh = 32;  // 9 used

struct ConvDef {
    int cin, cout;
};

// conv order: e0a e0b e1a e1b e2a e2b e3a e3b ba bb d3a d3b d2a d2b d1a d1b d0a d0b out
std::vector<ConvDef> unet_convs() {
    std::vector<ConvDef> c;
    int prev = kInCh;
    for (int l = 0; l < kLevels; ++l) {
        c.push_back({prev, kCh[l]});
        c.push_back({kCh[l], kCh[l]});
        prev = kCh[l];
    }
    for (int l = kLevels - 2; l >= 0; --l) {
        c.push_back({kCh[l] + kCh[l + 1], kCh[l]});
        c.push_back({kCh[l], kCh[l]});
    }
    c.push_back({kCh[0], 32});  // output conv, channels 0-2 used
    return c;
}

struct Acts {
    int N = 0, H = 0, W = 0;
    uint16_t* x0 = nullptr;
    float* lin = nullptr;
    uint16_t *t[kLevels] = {}, *cat[kLevels - 1] = {}, *pl[kLevels - 1] = {}, *bt = nullptr;
    uint16_t *da[kLevels - 1] = {}, *u[kLevels - 1] = {};
    float* y = nullptr;
    // gradients
    uint16_t* dy = nullptr;
    uint16_t *du[kLevels - 1] = {}, *dda[kLevels - 1] = {}, *dcat[kLevels - 1] = {}, *dt[kLevels] = {};
    uint16_t *dpl[kLevels - 1] = {}, *dbt = nullptr;
    float* target = nullptr;
    float* inrec = nullptr;
    float* outrec = nullptr;
    bool grads = false;
    std::vector<void*> all;

    template <class T> T* A(size_t n) {
        T* p = dalloc<T>(n);
        all.push_back(p);
        return p;
    }
    void release() {
        for (void* p : all) (void)hipFree(p);
        all.clear();
        N = H = W = 0;
    }
    void ensure(int n, int h, int w, bool need_grads) {
        if (n == N && h == H && w == W && (grads || !need_grads)) return;
        release();
        N = n, H = h, W = w;
        grads = need_grads;
        auto P = [&](int l) { return (size_t)n * (h >> l) * (w >> l); };
        x0 = A<uint16_t>(P(0) * kInCh);
        lin = A<float>(P(0) * 4);
        y = A<float>(P(0) * 4);
        for (int l = 0; l < kLevels; ++l) t[l] = A<uint16_t>(P(l) * kCh[l]);
        for (int l = 0; l < kLevels - 1; ++l) {
            cat[l] = A<uint16_t>(P(l) * (kCh[l] + kCh[l + 1]));
            pl[l] = A<uint16_t>(P(l + 1) * kCh[l]);
            da[l] = A<uint16_t>(P(l) * kCh[l]);
            u[l] = A<uint16_t>(P(l) * kCh[l]);
        }
        bt = A<uint16_t>(P(kLevels - 1) * kCh[kLevels - 1]);
        if (need_grads) {
            dy = A<uint16_t>(P(0) * 32);
            for (int l = 0; l < kLevels; ++l) dt[l] = A<uint16_t>(P(l) * kCh[l]);
            for (int l = 0; l < kLevels - 1; ++l) {
                du[l] = A<uint16_t>(P(l) * kCh[l]);
                dda[l] = A<uint16_t>(P(l) * kCh[l]);
                dcat[l] = A<uint16_t>(P(l) * (kCh[l] + kCh[l + 1]));
                dpl[l] = A<uint16_t>(P(l + 1) * kCh[l]);
            }
            dbt = A<uint16_t>(P(kLevels - 1) * kCh[kLevels - 1]);
            target = A<float>(P(0) * 3);
        }
    }
    ~Acts() { release(); }
};

}  // namespace

struct Unet {
    std::vector<ConvDef> defs;
    std::vector<size_t> woff, boff;  // offsets into the flat parameter vector
    size_t nparams = 0;
    float *w32 = nullptr, *g32 = nullptr, *m32 = nullptr, *v32 = nullptr;
    std::vector<uint16_t*> wb, wd;
    float* d_loss = nullptr;
    int step = 0;
    Acts act;
    float* rec_in = nullptr;
    size_t rec_in_cap = 0;
    float* rec_tg = nullptr;
    size_t rec_tg_cap = 0;
    float* rec_out = nullptr;
    size_t rec_out_cap = 0;

    Unet() {
        defs = unet_convs();
        for (auto& d : defs) {
            woff.push_back(nparams);
            nparams += (size_t)9 * d.cin * d.cout;
            boff.push_back(nparams);
            nparams += d.cout;
        }
        w32 = dalloc<float>(nparams);
        g32 = dalloc<float>(nparams);
        m32 = dalloc<float>(nparams);
        v32 = dalloc<float>(nparams);
        for (auto& d : defs) {
            wb.push_back(dalloc<uint16_t>((size_t)9 * d.cin * d.cout));
            wd.push_back(dalloc<uint16_t>((size_t)9 * d.cin * d.cout));
        }
        d_loss = dalloc<float>(1);
    }
    ~Unet() {
        act.release();
        for (void* p : {(void*)w32, (void*)g32, (void*)m32, (void*)v32, (void*)d_loss, (void*)rec_in, (void*)rec_tg,
                        (void*)rec_out})
            if (p) (void)hipFree(p);
        for (auto p : wb) (void)hipFree(p);
        for (auto p : wd) (void)hipFree(p);
    }

    void sync_bf16(hipStream_t st) {
        for (size_t k = 0; k < defs.size(); ++k) {
            int n = 9 * defs[k].cin * defs[k].cout;
            k_weights_bf16<<<blocks(n, TB), TB, 0, st>>>(w32 + woff[k], defs[k].cin, defs[k].cout, wb[k], wd[k]);
        }
    }

    void reset_optimizer(hipStream_t st) {
        k_zero_f32<<<blocks(nparams, TB), TB, 0, st>>>(m32, nparams);
        k_zero_f32<<<blocks(nparams, TB), TB, 0, st>>>(v32, nparams);
        step = 0;
    }

    void set_params_host(const std::vector<float>& h) {
        hipStream_t st = stream();
        KF_CHECK(hipMemcpyAsync(w32, h.data(), nparams * sizeof(float), hipMemcpyHostToDevice, st));
        sync_bf16(st);
        reset_optimizer(st);
        KF_CHECK(hipStreamSynchronize(st));
    }

    void init_random(uint32_t seed) {
        std::vector<float> h(nparams, 0.0f);
        uint64_t s = 0x9E3779B97F4A7C15ull ^ seed;
        auto next = [&]() {  // splitmix64
            s += 0x9E3779B97F4A7C15ull;
            uint64_t z = s;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            return z ^ (z >> 31);
        };
        auto unif = [&]() { return ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); };
        for (size_t k = 0; k < defs.size(); ++k) {
            bool last = k + 1 == defs.size();
            int cin_eff = k == 0 ? 9 : defs[k].cin;
            double std = last ? 1e-3 : std::sqrt(2.0 / (9.0 * cin_eff));
            for (int tap = 0; tap < 9; ++tap)
                for (int co = 0; co < defs[k].cout; ++co)
                    for (int ci = 0; ci < defs[k].cin; ++ci) {
                        double u1 = unif(), u2 = unif();
                        double g = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
                        bool used = (k != 0 || ci < 9) && (!last || co < 3);
                        h[woff[k] + ((size_t)tap * defs[k].cout + co) * defs[k].cin + ci] = used ? (float)(g * std) : 0.f;
                    }
        }
        set_params_host(h);
    }

    // Pixel stride of the skip tensor in act.cat[l]: the concatenation's when
    // the backward pass needs it whole, else the skip channels alone (dense
    // rows, no partly used cache lines).
    int skip_stride(int l) const { return act.grads ? kCh[l] + kCh[l + 1] : kCh[l]; }

    // forward pass on act.x0 (or the render targets pk) -> act.y
    void forward(hipStream_t st, const PackSrc* pk = nullptr) {
        const int N = act.N;
        auto H = [&](int l) { return act.H >> l; };
        auto W = [&](int l) { return act.W >> l; };
        int k = 0;
        const uint16_t* in = act.x0;
        int ldin = kInCh, cin = kInCh;
        for (int l = 0; l < kLevels; ++l) {
            ConvFuse fin;
            if (l == 0 && pk) fin.pk = *pk;
            conv_fwd(in, ldin, cin, wb[k], w32 + boff[k], kCh[l], N, H(l), W(l), act.t[l], kCh[l], true, st, nullptr,
                     0, -1, fin);
            ++k;
            if (l < kLevels - 1) {
                ConvFuse fz;
                fz.P = act.pl[l];  // the 2x2 pool comes out of this convolution's epilogue
                conv_fwd(act.t[l], kCh[l], kCh[l], wb[k], w32 + boff[k], kCh[l], N, H(l), W(l), act.cat[l],
                         skip_stride(l), true, st, nullptr, 0, -1, fz);
                ++k;
                in = act.pl[l];
                ldin = cin = kCh[l];
            } else {
                conv_fwd(act.t[l], kCh[l], kCh[l], wb[k], w32 + boff[k], kCh[l], N, H(l), W(l), act.bt, kCh[l], true,
                         st);
                ++k;
            }
        }
        const uint16_t* low = act.bt;
        for (int l = kLevels - 2; l >= 0; --l) {
            int ldc = kCh[l] + kCh[l + 1];
            if (act.grads) {  // the backward pass reads the whole concatenation
                size_t n8 = (size_t)N * H(l) * W(l) * kCh[l + 1] / 8;
                k_up<<<blocks(n8, TB), TB, 0, st>>>(low, kCh[l + 1], N, H(l + 1), W(l + 1), act.cat[l] + kCh[l], ldc);
            }
            // skip channels from act.cat[l], the upsampled ones straight from low
            ConvFuse fz;
            fz.X2 = low;
            fz.ldx2 = kCh[l + 1];
            fz.csplit = kCh[l];
            conv_fwd(act.cat[l], skip_stride(l), ldc, wb[k], w32 + boff[k], kCh[l], N, H(l), W(l), act.da[l], kCh[l], true, st,
                     nullptr, 0, -1, fz);
            ++k;
            conv_fwd(act.da[l], kCh[l], kCh[l], wb[k], w32 + boff[k], kCh[l], N, H(l), W(l), act.u[l], kCh[l], true,
                     st);
            ++k;
            low = act.u[l];
        }
        conv_fwd(act.u[0], kCh[0], kCh[0], wb[k], w32 + boff[k], 32, N, H(0), W(0), nullptr, 0, false, st, act.y, 4, 3);
    }

    // backward from act.dy; accumulates into g32 (zeroed here)
    void backward(hipStream_t st) {
        const int N = act.N;
        auto H = [&](int l) { return act.H >> l; };
        auto W = [&](int l) { return act.W >> l; };
        k_zero_f32<<<blocks(nparams, TB), TB, 0, st>>>(g32, nparams);
        int k = (int)defs.size() - 1;
        // output conv
        conv_wgrad(act.u[0], kCh[0], kCh[0], act.dy, 32, 32, N, H(0), W(0), g32 + woff[k], g32 + boff[k], st);
        conv_dgrad(act.dy, 32, 32, wd[k], kCh[0], N, H(0), W(0), act.du[0], kCh[0], act.u[0], kCh[0], st);
        --k;
        for (int l = 0; l < kLevels - 1; ++l) {
            int ldc = kCh[l] + kCh[l + 1];
            // d_lb: input da[l], grad du[l]
            conv_wgrad(act.da[l], kCh[l], kCh[l], act.du[l], kCh[l], kCh[l], N, H(l), W(l), g32 + woff[k],
                       g32 + boff[k], st);
            conv_dgrad(act.du[l], kCh[l], kCh[l], wd[k], kCh[l], N, H(l), W(l), act.dda[l], kCh[l], act.da[l], kCh[l],
                       st);
            --k;
            // d_la: input cat[l], grad dda[l]
            conv_wgrad(act.cat[l], ldc, ldc, act.dda[l], kCh[l], kCh[l], N, H(l), W(l), g32 + woff[k], g32 + boff[k],
                       st);
            conv_dgrad(act.dda[l], kCh[l], kCh[l], wd[k], ldc, N, H(l), W(l), act.dcat[l], ldc, act.cat[l], ldc, st);
            --k;
            // upsampled half of the concat -> gradient of the level below's output
            const uint16_t* lowf = (l + 1 < kLevels - 1) ? act.u[l + 1] : act.bt;
            uint16_t* dlow = (l + 1 < kLevels - 1) ? act.du[l + 1] : act.dbt;
            size_t n8 = (size_t)N * H(l + 1) * W(l + 1) * kCh[l + 1] / 8;
            k_up_bwd<<<blocks(n8, TB), TB, 0, st>>>(act.dcat[l] + kCh[l], ldc, kCh[l + 1], lowf, N, H(l + 1), W(l + 1),
                                                   dlow);
        }
        // bottleneck bb, ba
        int L = kLevels - 1;
        conv_wgrad(act.t[L], kCh[L], kCh[L], act.dbt, kCh[L], kCh[L], N, H(L), W(L), g32 + woff[k], g32 + boff[k], st);
        conv_dgrad(act.dbt, kCh[L], kCh[L], wd[k], kCh[L], N, H(L), W(L), act.dt[L], kCh[L], act.t[L], kCh[L], st);
        --k;
        conv_wgrad(act.pl[L - 1], kCh[L - 1], kCh[L - 1], act.dt[L], kCh[L], kCh[L], N, H(L), W(L), g32 + woff[k],
                   g32 + boff[k], st);
        conv_dgrad(act.dt[L], kCh[L], kCh[L], wd[k], kCh[L - 1], N, H(L), W(L), act.dpl[L - 1], kCh[L - 1], nullptr, 0,
                   st);
        --k;
        for (int l = kLevels - 2; l >= 0; --l) {
            int ldc = kCh[l] + kCh[l + 1];
            // pooling backward into the skip half of dcat (masked by the skip values)
            size_t n8 = (size_t)N * H(l) * W(l) * kCh[l] / 8;
            k_pool_bwd<<<blocks(n8, TB), TB, 0, st>>>(act.dpl[l], kCh[l], act.cat[l], ldc, act.dcat[l], ldc, N, H(l),
                                                     W(l));
            // e_lb: input t[l], grad dcat[l][:, 0:C]
            conv_wgrad(act.t[l], kCh[l], kCh[l], act.dcat[l], ldc, kCh[l], N, H(l), W(l), g32 + woff[k], g32 + boff[k],
                       st);
            conv_dgrad(act.dcat[l], ldc, kCh[l], wd[k], kCh[l], N, H(l), W(l), act.dt[l], kCh[l], act.t[l], kCh[l], st);
            --k;
            // e_la: input pl[l-1] (or x0), grad dt[l]
            const uint16_t* in = l > 0 ? act.pl[l - 1] : act.x0;
            int cin = l > 0 ? kCh[l - 1] : kInCh;
            conv_wgrad(in, cin, cin, act.dt[l], kCh[l], kCh[l], N, H(l), W(l), g32 + woff[k], g32 + boff[k], st);
            if (l > 0)
                conv_dgrad(act.dt[l], kCh[l], kCh[l], wd[k], cin, N, H(l), W(l), act.dpl[l - 1], cin, nullptr, 0, st);
            --k;
        }
    }

    float loss_and_grad(hipStream_t st) {
        k_zero_f32<<<1, 32, 0, st>>>(d_loss, 1);
        int P = act.N * act.H * act.W;
        k_loss<<<blocks(P, TB), TB, 0, st>>>(act.y, act.lin, act.target, P, act.dy, d_loss);
        backward(st);
        float l = 0;
        KF_CHECK(hipMemcpyAsync(&l, d_loss, sizeof(float), hipMemcpyDeviceToHost, st));
        KF_CHECK(hipStreamSynchronize(st));
        return l;
    }

    void upload_records(const float* in, const float* tg, int n, int h, int w, bool grads) {
        hipStream_t st = stream();
        act.ensure(n, h, w, grads);
        size_t P = (size_t)n * h * w;
        if (P * 9 > rec_in_cap) {
            if (rec_in) KF_CHECK(hipFree(rec_in));
            rec_in = dalloc<float>(P * 9);
            rec_in_cap = P * 9;
        }
        KF_CHECK(hipMemcpyAsync(rec_in, in, P * 9 * sizeof(float), hipMemcpyHostToDevice, st));
        k_pack_records<<<blocks(P, TB), TB, 0, st>>>(rec_in, (int)P, act.x0, act.lin);
        if (tg) KF_CHECK(hipMemcpyAsync(act.target, tg, P * 3 * sizeof(float), hipMemcpyHostToDevice, st));
    }
};

}  // namespace kf

struct kf_denoiser : kf::Unet {};

using namespace kf;

namespace {

constexpr char kMagic[8] = {'K', 'F', 'U', 'N', 'E', 'T', '0', '1'};

std::string default_weights_path() {
    const char* env = getenv("KUAFU_AMD_WEIGHTS");
    if (env && *env) return env;
    // the weight file ships next to the library: <pkg>/weights/unet.kfw
    std::string so;
    {
        std::ifstream maps("/proc/self/maps");
        std::string line;
        while (std::getline(maps, line)) {
            size_t pos = line.find("libkuafu_amd.so");
            if (pos != std::string::npos) {
                size_t s = line.find('/');
                so = line.substr(s, pos + strlen("libkuafu_amd.so") - s);
                break;
            }
        }
    }
    if (so.empty()) return "";
    return so.substr(0, so.rfind('/')) + "/weights/unet.kfw";
}

bool read_weights(const std::string& path, const Unet& u, std::vector<float>& out) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    char magic[8];
    f.read(magic, 8);
    if (!f || memcmp(magic, kMagic, 8) != 0) throw Error("bad denoiser weight file: " + path);
    uint32_t nconv = 0;
    f.read((char*)&nconv, 4);
    if (nconv != u.defs.size()) throw Error("denoiser weight file has a different network: " + path);
    out.assign(u.nparams, 0.0f);
    for (size_t k = 0; k < u.defs.size(); ++k) {
        uint32_t ci = 0, co = 0;
        f.read((char*)&ci, 4);
        f.read((char*)&co, 4);
        if ((int)ci != u.defs[k].cin || (int)co != u.defs[k].cout) throw Error("denoiser weight shape mismatch");
        size_t nw = (size_t)9 * ci * co;
        std::vector<uint16_t> wb(nw);
        f.read((char*)wb.data(), nw * 2);
        for (size_t i = 0; i < nw; ++i) out[u.woff[k] + i] = bf2f(wb[i]);
        f.read((char*)&out[u.boff[k]], co * sizeof(float));
    }
    if (!f) throw Error("truncated denoiser weight file: " + path);
    return true;
}

}  // namespace

extern "C" {

int kf_denoiser_create(const char* weights_path, kf_denoiser** out) {
    kf_denoiser* d = nullptr;
    try {
        if (!out) throw Error("kf_denoiser_create: null argument");
        d = new kf_denoiser();
        std::vector<float> h;
        std::string path = weights_path ? weights_path : default_weights_path();
        if (!path.empty() && read_weights(path, *d, h)) {
            d->set_params_host(h);
        } else {
            if (weights_path) throw Error("cannot open denoiser weights: " + path);
            d->init_random(0);
        }
        *out = d;
        return 0;
    } catch (const std::exception& e) {
        delete d;
        set_error(e.what());
        return -1;
    }
}

void kf_denoiser_destroy(kf_denoiser* d) {
    if (d) (void)hipStreamSynchronize(stream());
    delete d;
}

int kf_denoiser_init_random(kf_denoiser* d, uint32_t seed) {
    try {
        if (!d) throw Error("null denoiser");
        d->init_random(seed);
        return 0;
    } catch (const std::exception& e) {
        set_error(e.what());
        return -1;
    }
}

int kf_denoiser_num_params(kf_denoiser* d, uint64_t* n) {
    if (!d || !n) {
        set_error("null argument");
        return -1;
    }
    *n = d->nparams;
    return 0;
}

int kf_denoiser_get_params(kf_denoiser* d, float* out) {
    try {
        if (!d || !out) throw Error("null argument");
        KF_CHECK(hipMemcpyAsync(out, d->w32, d->nparams * sizeof(float), hipMemcpyDeviceToHost, stream()));
        KF_CHECK(hipStreamSynchronize(stream()));
        return 0;
    } catch (const std::exception& e) {
        set_error(e.what());
        return -1;
    }
}

int kf_denoiser_set_params(kf_denoiser* d, const float* in) {
    try {
        if (!d || !in) throw Error("null argument");
        d->set_params_host(std::vector<float>(in, in + d->nparams));
        return 0;
    } catch (const std::exception& e) {
        set_error(e.what());
        return -1;
    }
}

int kf_denoiser_save(kf_denoiser* d, const char* path) {
    try {
        if (!d || !path) throw Error("null argument");
        std::vector<float> h(d->nparams);
        KF_CHECK(hipMemcpy(h.data(), d->w32, d->nparams * sizeof(float), hipMemcpyDeviceToHost));
        std::ofstream f(path, std::ios::binary);
        if (!f) throw Error(std::string("cannot write ") + path);
        f.write(kMagic, 8);
        uint32_t nconv = (uint32_t)d->defs.size();
        f.write((const char*)&nconv, 4);
        for (size_t k = 0; k < d->defs.size(); ++k) {
            uint32_t ci = d->defs[k].cin, co = d->defs[k].cout;
            f.write((const char*)&ci, 4);
            f.write((const char*)&co, 4);
            size_t nw = (size_t)9 * ci * co;
            std::vector<uint16_t> wb(nw);
            for (size_t i = 0; i < nw; ++i) wb[i] = f2bf(h[d->woff[k] + i]);
            f.write((const char*)wb.data(), nw * 2);
            f.write((const char*)&h[d->boff[k]], co * sizeof(float));
        }
        return 0;
    } catch (const std::exception& e) {
        set_error(e.what());
        return -1;
    }
}

int kf_denoiser_load(kf_denoiser* d, const char* path) {
    try {
        if (!d || !path) throw Error("null argument");
        std::vector<float> h;
        if (!read_weights(path, *d, h)) throw Error(std::string("cannot open ") + path);
        d->set_params_host(h);
        return 0;
    } catch (const std::exception& e) {
        set_error(e.what());
        return -1;
    }
}

int kf_denoiser_infer_device(kf_denoiser* d, const float* d_rgba, const float* d_albedo, const float* d_normal,
                             uint32_t width, uint32_t height, float* d_out, float* kernel_ms) {
    return kf_denoiser_infer_device_batch(d, d_rgba, d_albedo, d_normal, width, height, 1, d_out, kernel_ms);
}

int kf_denoiser_infer_device_batch(kf_denoiser* d, const float* d_rgba, const float* d_albedo,
                                   const float* d_normal, uint32_t width, uint32_t height, uint32_t n_images,
                                   float* d_out, float* kernel_ms) {
    try {
        if (!d || !d_rgba || !d_albedo || !d_normal || !d_out || !width || !height || !n_images)
            throw Error("bad argument");
        hipStream_t st = stream();
        const int N = (int)n_images;
        int Hp = (int)((height + 15) / 16 * 16), Wp = (int)((width + 15) / 16 * 16);
        if ((uint64_t)N * Hp * Wp >= (1ull << 31)) throw Error("kf_denoiser_infer_device_batch: batch too large");
        d->act.ensure(N, Hp, Wp, false);
        hipEvent_t e0 = nullptr, e1 = nullptr;  // timing events only when the caller asks for the time
        if (kernel_ms) {
            KF_CHECK(hipEventCreate(&e0));
            KF_CHECK(hipEventCreate(&e1));
            KF_CHECK(hipEventRecord(e0, st));
        }
        // the first convolution reads the render targets itself
        PackSrc pk;
        pk.rgba = (const float4*)d_rgba;
        pk.alb = (const float4*)d_albedo;
        pk.nrm = (const float4*)d_normal;
        pk.h = (int)height;
        pk.w = (int)width;
        d->forward(st, &pk);
        k_unpack<<<blocks((size_t)N * width * height, TB), TB, 0, st>>>(d->act.y, (const float4*)d_rgba, N,
                                                                        (int)height, (int)width, Hp, Wp,
                                                                        (float4*)d_out);
        KF_CHECK(hipGetLastError());
        if (kernel_ms) {
            KF_CHECK(hipEventRecord(e1, st));
            KF_CHECK(hipEventSynchronize(e1));
            float ms = 0;
            KF_CHECK(hipEventElapsedTime(&ms, e0, e1));
            *kernel_ms = ms;
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
        } else {
            KF_CHECK(hipStreamSynchronize(st));
        }
        return 0;
    } catch (const std::exception& e) {
        set_error(e.what());
        return -1;
    }
}

int kf_denoiser_apply(kf_denoiser* d, kf_renderer* r) {
    if (!d || !r) {
        set_error("kf_denoiser_apply: null argument");
        return -1;
    }
    // the output overwrites rgba (denoiser.cpp:225-251, context.cpp:669-670);
    // a frame of environments is denoised as a batch of their images
    const uint32_t n = std::max<uint32_t>(r->n_env_last, 1u);
    return kf_denoiser_infer_device_batch(d, (const float*)r->d_rgba, (const float*)r->d_albedo,
                                          (const float*)r->d_normal, r->W, r->H / n, n, (float*)r->d_rgba, nullptr);
}

int kf_denoiser_forward(kf_denoiser* d, const float* inputs, uint32_t n, uint32_t h, uint32_t w, float* out) {
    try {
        if (!d || !inputs || !out || h % 16 || w % 16 || !n) throw Error("kf_denoiser_forward: bad argument");
        hipStream_t st = stream();
        d->upload_records(inputs, nullptr, (int)n, (int)h, (int)w, d->act.grads);
        d->forward(st);
        size_t P = (size_t)n * h * w;
        if (P * 3 > d->rec_out_cap) {
            if (d->rec_out) KF_CHECK(hipFree(d->rec_out));
            d->rec_out = dalloc<float>(P * 3);
            d->rec_out_cap = P * 3;
        }
        k_unpack_records<<<blocks(P, TB), TB, 0, st>>>(d->act.y, d->act.lin, (int)P, d->rec_out);
        KF_CHECK(hipMemcpyAsync(out, d->rec_out, P * 3 * sizeof(float), hipMemcpyDeviceToHost, st));
        KF_CHECK(hipStreamSynchronize(st));
        return 0;
    } catch (const std::exception& e) {
        set_error(e.what());
        return -1;
    }
}

int kf_denoiser_gradients(kf_denoiser* d, const float* inputs, const float* targets, uint32_t n, uint32_t h,
                          uint32_t w, float* loss, float* grads) {
    try {
        if (!d || !inputs || !targets || h % 16 || w % 16 || !n) throw Error("kf_denoiser_gradients: bad argument");
        hipStream_t st = stream();
        d->upload_records(inputs, targets, (int)n, (int)h, (int)w, true);
        d->forward(st);
        float l = d->loss_and_grad(st);
        if (loss) *loss = l;
        if (grads) {
            KF_CHECK(hipMemcpyAsync(grads, d->g32, d->nparams * sizeof(float), hipMemcpyDeviceToHost, st));
            KF_CHECK(hipStreamSynchronize(st));
        }
        return 0;
    } catch (const std::exception& e) {
        set_error(e.what());
        return -1;
    }
}

int kf_denoiser_train_step(kf_denoiser* d, const float* inputs, const float* targets, uint32_t n, uint32_t h,
                           uint32_t w, float lr, float* loss) {
    try {
        if (!d || !inputs || !targets || h % 16 || w % 16 || !n) throw Error("kf_denoiser_train_step: bad argument");
        hipStream_t st = stream();
        d->upload_records(inputs, targets, (int)n, (int)h, (int)w, true);
        d->forward(st);
        float l = d->loss_and_grad(st);
        d->step++;
        const float b1 = 0.9f, b2 = 0.999f;
        float c1 = 1.0f - std::pow(b1, (float)d->step), c2 = 1.0f - std::pow(b2, (float)d->step);
        k_adam<<<blocks(d->nparams, TB), TB, 0, st>>>(d->w32, d->g32, d->m32, d->v32, d->nparams, lr, b1, b2, 1e-8f,
                                                      c1, c2);
        d->sync_bf16(st);
        KF_CHECK(hipGetLastError());
        if (loss) *loss = l;
        return 0;
    } catch (const std::exception& e) {
        set_error(e.what());
        return -1;
    }
}

int kf_conv3x3_forward(const uint16_t* x, uint32_t cin, const uint16_t* w, const float* bias, uint32_t cout,
                       uint32_t n, uint32_t h, uint32_t wd, int relu, uint16_t* y) {
    try {
        if (!x || !w || !y || cin % 32 || !cout || !n || !h || !wd) throw Error("kf_conv3x3_forward: bad argument");
        hipStream_t st = stream();
        size_t P = (size_t)n * h * wd;
        uint32_t ldy = (cout + 7) / 8 * 8;
        uint16_t* dx = dalloc<uint16_t>(P * cin);
        uint16_t* dw = dalloc<uint16_t>((size_t)9 * cin * cout);
        float* db = dalloc<float>(cout);
        uint16_t* dy = dalloc<uint16_t>(P * ldy);
        KF_CHECK(hipMemcpyAsync(dx, x, P * cin * 2, hipMemcpyHostToDevice, st));
        KF_CHECK(hipMemcpyAsync(dw, w, (size_t)9 * cin * cout * 2, hipMemcpyHostToDevice, st));
        if (bias) KF_CHECK(hipMemcpyAsync(db, bias, cout * 4, hipMemcpyHostToDevice, st));
        conv_fwd(dx, cin, cin, dw, bias ? db : nullptr, cout, n, h, wd, dy, ldy, relu != 0, st);
        KF_CHECK(hipGetLastError());
        std::vector<uint16_t> tmp(P * ldy);
        KF_CHECK(hipMemcpyAsync(tmp.data(), dy, P * ldy * 2, hipMemcpyDeviceToHost, st));
        KF_CHECK(hipStreamSynchronize(st));
        for (size_t p = 0; p < P; ++p) memcpy(y + p * cout, &tmp[p * ldy], cout * 2);
        for (void* q : {(void*)dx, (void*)dw, (void*)db, (void*)dy}) KF_CHECK(hipFree(q));
        return 0;
    } catch (const std::exception& e) {
        set_error(e.what());
        return -1;
    }
}

int kf_conv3x3_backward(const uint16_t* x, uint32_t cin, const uint16_t* w, uint32_t cout, const uint16_t* dy,
                        uint32_t n, uint32_t h, uint32_t wd, int mask_by_x, uint16_t* dx, float* dw, float* db) {
    try {
        if (!x || !w || !dy || cin % 32 || cout % 32 || !n || !h || !wd) throw Error("kf_conv3x3_backward: bad argument");
        hipStream_t st = stream();
        size_t P = (size_t)n * h * wd, NW = (size_t)9 * cin * cout;
        uint16_t* d_x = dalloc<uint16_t>(P * cin);
        uint16_t* d_dy = dalloc<uint16_t>(P * cout);
        uint16_t* d_dx = dalloc<uint16_t>(P * cin);
        float* d_w32 = dalloc<float>(NW);
        uint16_t* d_wb = dalloc<uint16_t>(NW);
        uint16_t* d_wd = dalloc<uint16_t>(NW);
        float* d_dw = dalloc<float>(NW);
        float* d_db = dalloc<float>(cout);
        std::vector<float> w32(NW);
        for (size_t i = 0; i < NW; ++i) w32[i] = bf2f(w[i]);
        KF_CHECK(hipMemcpyAsync(d_x, x, P * cin * 2, hipMemcpyHostToDevice, st));
        KF_CHECK(hipMemcpyAsync(d_dy, dy, P * cout * 2, hipMemcpyHostToDevice, st));
        KF_CHECK(hipMemcpyAsync(d_w32, w32.data(), NW * 4, hipMemcpyHostToDevice, st));
        k_weights_bf16<<<blocks(NW, TB), TB, 0, st>>>(d_w32, (int)cin, (int)cout, d_wb, d_wd);
        k_zero_f32<<<blocks(NW, TB), TB, 0, st>>>(d_dw, NW);
        k_zero_f32<<<blocks(cout, TB), TB, 0, st>>>(d_db, cout);
        conv_dgrad(d_dy, cout, cout, d_wd, cin, n, h, wd, d_dx, cin, mask_by_x ? d_x : nullptr, cin, st);
        conv_wgrad(d_x, cin, cin, d_dy, cout, cout, n, h, wd, d_dw, d_db, st);
        KF_CHECK(hipGetLastError());
        if (dx) KF_CHECK(hipMemcpyAsync(dx, d_dx, P * cin * 2, hipMemcpyDeviceToHost, st));
        if (dw) KF_CHECK(hipMemcpyAsync(dw, d_dw, NW * 4, hipMemcpyDeviceToHost, st));
        if (db) KF_CHECK(hipMemcpyAsync(db, d_db, cout * 4, hipMemcpyDeviceToHost, st));
        KF_CHECK(hipStreamSynchronize(st));
        for (void* q : {(void*)d_x, (void*)d_dy, (void*)d_dx, (void*)d_w32, (void*)d_wb, (void*)d_wd, (void*)d_dw,
                        (void*)d_db})
            KF_CHECK(hipFree(q));
        return 0;
    } catch (const std::exception& e) {
        set_error(e.what());
        return -1;
    }
}

}  // extern "C"
