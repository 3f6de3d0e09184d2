// nn.hip -- kernels of the 2-D U-Net that proposes the eEM starting image
// (gfx950): 3x3 convolution forward / backward-data / backward-weight on BF16
// MFMA with FP32 accumulation, 2x2 average pooling, nearest 2x upsampling,
// ReLU masks, the 1x1 output head with its loss, Adam, and the resampling of
// an FBP image onto the eEM grid.
//
// Activations are NHWC bf16 with an explicit channel stride `ld` and channel
// offset, so a skip concatenation is two writers into one buffer.  Master
// weights, gradients and optimizer state are FP32; the convolutions read a
// bf16 copy of their weights, layout [Cout][9][Cin] (tap = 3*(dy+1) + (dx+1)).
//
// MFMA v_mfma_f32_32x32x16_bf16, D[32x32] += A[32x16] B[16x32]:
//   lane l supplies A[l % 32][8(l/32) .. 8(l/32)+7] and B[8(l/32) .. +7][l % 32],
//   and holds D[8(i/4) + 4(l/32) + (i % 4)][l % 32] in accumulator i (0..15).
// v_mfma_f32_16x16x32_bf16: lane l supplies A[l % 16][8(l/16) .. +7] and
//   B[8(l/16) .. +7][l % 16], holds D[4(l/16) + i][l % 16], i < 4.
// raft_nn_mfma_probe runs both on host-given operands so these layouts are
// checked before anything relies on them.
#include "raft_internal.hpp"
#include "../../include/raft_nn.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>

namespace raft {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
// native 16-byte vector for values held in registers across a loop (the
// struct-with-union uint4 defeats scalar replacement there and lands in scratch)
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ u32x4 as_u32x4(uint4 v) { return u32x4{v.x, v.y, v.z, v.w}; }
typedef unsigned short u16;

// FP32 -> bf16, round to nearest even: one v_cvt_pk_bf16_f32 per two values
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ unsigned bf16_pack2(float lo, float hi)
{
    const bf16x2 v = {(__bf16)lo, (__bf16)hi};
    return __builtin_bit_cast(unsigned, v);
}
__device__ __forceinline__ u16 bf16_bits(float f) { return __builtin_bit_cast(u16, (__bf16)f); }

// value of lane l | 1 in lane l (for even l): a DPP quad permutation [1, 1, 3, 3]
__device__ __forceinline__ float odd_neighbour(float v)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xF5, 0xF, 0xF, false));
}
__device__ __forceinline__ float bf16_float(u16 b) { return __uint_as_float((unsigned)b << 16); }

__device__ __forceinline__ bf16x8 as_bf16x8(uint4 v)
{
    bf16x8 r;
    __builtin_memcpy(&r, &v, 16);
    return r;
}
__device__ __forceinline__ bf16x8 lds8(const u16* p) { return as_bf16x8(*reinterpret_cast<const uint4*>(p)); }

// ---------------------------------------------------------------------------
// MFMA probe: raw per-lane results of both shapes on row-major operands
//   A16 [16][32], B16 [32][16] -> out16[lane][4]; A32 [32][16], B32 [16][32] -> out32[lane][16]
// ---------------------------------------------------------------------------
__global__ void k_mfma_probe(const u16* A16, const u16* B16, float* out16, const u16* A32, const u16* B32,
                             float* out32)
{
    const int l = threadIdx.x;
    u16 ta[8], tb[8];
    for (int j = 0; j < 8; ++j) {
        ta[j] = A16[(l % 16) * 32 + 8 * (l / 16) + j];
        tb[j] = B16[(8 * (l / 16) + j) * 16 + (l % 16)];
    }
    bf16x8 a, b;
    __builtin_memcpy(&a, ta, 16);
    __builtin_memcpy(&b, tb, 16);
    f32x4 c = {0.f, 0.f, 0.f, 0.f};
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    for (int i = 0; i < 4; ++i) out16[l * 4 + i] = c[i];

    for (int j = 0; j < 8; ++j) {
        ta[j] = A32[(l % 32) * 16 + 8 * (l / 32) + j];
        tb[j] = B32[(8 * (l / 32) + j) * 32 + (l % 32)];
    }
    __builtin_memcpy(&a, ta, 16);
    __builtin_memcpy(&b, tb, 16);
    f32x16 d;
    for (int i = 0; i < 16; ++i) d[i] = 0.f;
    d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, d, 0, 0, 0);
    for (int i = 0; i < 16; ++i) out32[l * 16 + i] = d[i];
}

// 8 bf16 channels (16 bytes) <-> 8 floats.  The elementwise kernels move 8
// channels per thread; every channel count, stride and offset they see is a
// multiple of 8.
__device__ __forceinline__ void unpack8(uint4 v, float f[8])
{
    const unsigned e[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        f[2 * j] = __uint_as_float(e[j] << 16);
        f[2 * j + 1] = __uint_as_float(e[j] & 0xffff0000u);
    }
}
__device__ __forceinline__ uint4 pack8(const float f[8])
{
    return make_uint4(bf16_pack2(f[0], f[1]), bf16_pack2(f[2], f[3]), bf16_pack2(f[4], f[5]), bf16_pack2(f[6], f[7]));
}
__device__ __forceinline__ uint4 ld8(const u16* p) { return *reinterpret_cast<const uint4*>(p); }
__device__ __forceinline__ void st8(u16* p, uint4 v) { *reinterpret_cast<uint4*>(p) = v; }


// ---------------------------------------------------------------------------
// 3x3 convolution, forward (and backward-data with flipped weights):
//   y[n][h][w][cy0 + co] = act(b[co] + sum_{dy,dx,ci} x[n][h+dy][w+dx][cx0 + ci] W[co][tap][ci])
// zero outside the image.  Block = 4 waves; tile = TH = 4*WR rows x 32 columns
// of one image x CO_T = 32*WC output channels; wave w owns rows WR*w .. +WR-1.
// Per 32-channel input chunk the block stages the (TH+2) x 34 halo of x and
// the chunk's weights in LDS (80-byte pixel / weight-row stride: 32 channels
// plus 16 bytes of padding), then runs 9 taps x 2 k-halves of 32x32x16 MFMAs
// with A = weights (rows = output channels) and B = pixels (columns).
// ---------------------------------------------------------------------------
constexpr int CONV_PS = 40; // LDS stride (bf16 elements) of a pixel / weight row

struct ConvArgs {
    const u16* x;
    const u16* w; // [Cout][9][Cin]
    const float* b;
    u16* y;
    int N, H, W, Cin, Cout, ldx, cx0, ldy, cy0, relu;
    // fused input sources: input channels [0, split) come from x -- at H x W,
    // or (pool) as the 2x2 average of x at 2H x 2W; channels [split, Cin) come
    // from x2 at H/2 x W/2, read with nearest 2x upsampling (a skip
    // concatenation [x | up2(x2)] that is never written out)
    const u16* x2;
    int ldx2, cx2, split, pool;
    // output head folded into the epilogue (persistent form, 32 output channels;
    // y is then not written): pred[p] = res[p] + hb + sum_c hw[c] bf16(act(.)[c]),
    // clamped from below to floor_value when clamp -- what k_head computes from
    // the stored bf16 activations, in the same order
    const float* hw = nullptr;
    const float* hb = nullptr;
    const float* res = nullptr;
    float* pred = nullptr;
    float floor_value = 0.0f;
    int clamp = 0;
    // first layer folded into the halo load (persistent form, Cin = 32 = its
    // Cout): the halo's 32 channels are act0(b0 + sum_tap bf16(img0 * scale0)
    // w0[co][tap]) computed from the FP32 image, zero outside the image
    const float* img0 = nullptr;
    float scale0 = 1.0f;
    const u16* w0 = nullptr;
    int ldw0 = 0;
    const float* b0 = nullptr;
    int relu0 = 0;
    // also write the 2x2 average of the output (persistent form; H, W even):
    // ypool[(n * H/2 + y) * W/2 + x][cp0 + co] at channel stride ldp
    u16* ypool = nullptr;
    int ldp = 0, cp0 = 0;
};

// 16-byte unit (8 input channels c .. c+7) of the convolution input at pixel
// (gy, gx) of image n, inside the image.  c and split are multiples of 8 and
// split a multiple of 32, so the branch is uniform over a 32-channel chunk.
// The pooled and upsampled values equal k_pool2 / k_up2 outputs bit for bit.
// 2x2 average of x (at 2H x 2W) for input channels c .. c+7 (c < split) at
// pixel (gy, gx), as k_pool2 computes it; every load is consumed here
__device__ __forceinline__ uint4 conv_pooled_src(const ConvArgs& a, int n, int gy, int gx, int c)
{
    const int W2 = 2 * a.W;
    const size_t base = ((size_t)n * 2 * a.H + 2 * gy) * W2 + 2 * gx;
    const size_t src[4] = {base, base + 1, base + W2, base + W2 + 1};
    float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f}, f[8];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        unpack8(ld8(a.x + src[k] * a.ldx + a.cx0 + c), f);
#pragma unroll
        for (int j = 0; j < 8; ++j) acc[j] += f[j];
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) acc[j] *= 0.25f;
    return pack8(acc);
}

__device__ __forceinline__ uint4 conv_src(const ConvArgs& a, int n, int gy, int gx, int c)
{
    if (c >= a.split) {
        const int Hl = a.H >> 1, Wl = a.W >> 1;
        const size_t p = ((size_t)n * Hl + (gy >> 1)) * Wl + (gx >> 1);
        return *reinterpret_cast<const uint4*>(a.x2 + p * a.ldx2 + a.cx2 + (c - a.split));
    }
    if (a.pool) return conv_pooled_src(a, n, gy, gx, c);
    const size_t p = ((size_t)n * a.H + gy) * a.W + gx;
    return *reinterpret_cast<const uint4*>(a.x + p * a.ldx + a.cx0 + c);
}

// x / d by a 24-bit multiply-shift, exact for the ranges used below (checked
// exhaustively: /34 below 4096, /36 below 3311, /9 below 8192)
__device__ __forceinline__ int div34(int x) { return (int)(__umul24((unsigned)x, 1928u) >> 16); }
__device__ __forceinline__ int div36(int x) { return (int)(__umul24((unsigned)x, 1821u) >> 16); }
__device__ __forceinline__ int div9(int x) { return (int)(__umul24((unsigned)x, 7282u) >> 16); }

template <int WR, int WC>
constexpr int conv_lds_bytes()
{
    return ((4 * WR + 2) * 34 + 9 * 32 * WC) * CONV_PS * 2;
}

template <int WR, int WC>
__global__ __launch_bounds__(256) void k_conv3x3(ConvArgs a)
{
    constexpr int TH = 4 * WR, HW = 34, HH = TH + 2, CO_T = 32 * WC;
    extern __shared__ uint4 conv_lds[];
    u16* s_in = reinterpret_cast<u16*>(conv_lds);
    u16* s_w = s_in + HH * HW * CONV_PS;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int l32 = lane & 31, hi = lane >> 5;
    const int tiles_y = (a.H + TH - 1) / TH;
    const int tx0 = blockIdx.x * 32;
    const int ty0 = (blockIdx.y % tiles_y) * TH;
    const int n = blockIdx.y / tiles_y;
    const int co0 = blockIdx.z * CO_T;

    f32x16 acc[WR][WC];
#pragma unroll
    for (int r = 0; r < WR; ++r)
#pragma unroll
        for (int c = 0; c < WC; ++c)
#pragma unroll
            for (int i = 0; i < 16; ++i) acc[r][c][i] = 0.f;

    const size_t img = (size_t)n * a.H * a.W;
    for (int c0 = 0; c0 < a.Cin; c0 += 32) {
        __syncthreads();
        for (int u = tid; u < HH * HW * 4; u += 256) {
            const int pix = u >> 2, q = u & 3;
            const int gy = ty0 + pix / HW - 1, gx = tx0 + pix % HW - 1;
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if (gy >= 0 && gy < a.H && gx >= 0 && gx < a.W) v = conv_src(a, n, gy, gx, c0 + 8 * q);
            *reinterpret_cast<uint4*>(s_in + pix * CONV_PS + 8 * q) = v;
        }
        for (int u = tid; u < 9 * CO_T * 4; u += 256) {
            const int q = u & 3, tap = (u >> 2) % 9, co = (u >> 2) / 9;
            const uint4 v =
                *reinterpret_cast<const uint4*>(a.w + ((size_t)(co0 + co) * 9 + tap) * a.Cin + c0 + 8 * q);
            *reinterpret_cast<uint4*>(s_w + (tap * CO_T + co) * CONV_PS + 8 * q) = v;
        }
        __syncthreads();
#pragma unroll 1
        for (int tap = 0; tap < 9; ++tap) {
            const int dy = tap / 3, dx = tap % 3; // halo offsets
#pragma unroll
            for (int kk = 0; kk < 32; kk += 16) {
                bf16x8 bp[WR], aw[WC];
#pragma unroll
                for (int r = 0; r < WR; ++r)
                    bp[r] = lds8(s_in + ((wave * WR + r + dy) * HW + l32 + dx) * CONV_PS + kk + 8 * hi);
#pragma unroll
                for (int c = 0; c < WC; ++c) aw[c] = lds8(s_w + (tap * CO_T + 32 * c + l32) * CONV_PS + kk + 8 * hi);
#pragma unroll
                for (int r = 0; r < WR; ++r)
#pragma unroll
                    for (int c = 0; c < WC; ++c)
                        acc[r][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(aw[c], bp[r], acc[r][c], 0, 0, 0);
            }
        }
    }
    // lane holds pixel column tx0 + l32 and channels co0 + 32c + 8g + 4hi + (0..3)
    if (tx0 + l32 >= a.W) return;
#pragma unroll
    for (int r = 0; r < WR; ++r) {
        if (ty0 + wave * WR + r >= a.H) break;
        const size_t p = img + (size_t)(ty0 + wave * WR + r) * a.W + tx0 + l32;
        u16* yp = a.y + p * a.ldy + a.cy0;
#pragma unroll
        for (int c = 0; c < WC; ++c)
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                const int co = co0 + 32 * c + 8 * g + 4 * hi;
                float o[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    float v = acc[r][c][4 * g + i] + (a.b ? a.b[co + i] : 0.0f);
                    if (a.relu) v = fmaxf(v, 0.0f);
                    o[i] = v;
                }
                uint2 pk;
                pk.x = bf16_pack2(o[0], o[1]);
                pk.y = bf16_pack2(o[2], o[3]);
                *reinterpret_cast<uint2*>(yp + co) = pk;
            }
    }
}

// ---------------------------------------------------------------------------
// 3x3 convolution, persistent form (same arithmetic and layouts as k_conv3x3).
// A block walks `tiles_per_block` consecutive tiles of one output-channel
// tile and, for each, the Cin/32 input chunks.  While one (tile, chunk) item is
// computed from LDS, the next item's halo (and its weights when the chunk
// changes; with a single chunk they stay resident) is loaded into registers
// and written to LDS after a barrier, so global latency overlaps the MFMAs.
// The tap loop runs dx outer: the WR + 2 pixel fragments of one halo column
// serve all three dy taps (3 WR MFMAs per WR + 2 LDS fragment reads).
// ---------------------------------------------------------------------------
template <int WR, int WC>
struct ConvP {
    static constexpr int TH = 4 * WR, HW = 34, HH = TH + 2, CO_T = 32 * WC;
    static constexpr int HALO_UNITS = HH * HW * 4; // 16-byte units
    static constexpr int W_UNITS = 9 * CO_T * 4;
    static constexpr int HU = (HALO_UNITS + 255) / 256;
    static constexpr int WU = (W_UNITS + 255) / 256;
    // halo, weights, then the tile's CO_T biases (FP32)
    static constexpr int LDS = (HH * HW + 9 * CO_T) * CONV_PS * 2 + CO_T * 4;
    // first-layer mode: the FP32 image patch under the halo's 3x3 taps
    static constexpr int PW = HW + 2, PH = HH + 2, PATCH = PH * PW, PU = (PATCH + 255) / 256;
    static constexpr int HALO_GROUPS = (HH * HW + 31) / 32;
    static constexpr int LDS_FIRST = LDS + PATCH * 4;
};

struct ConvPArgs {
    ConvArgs c;
    int tiles_x, tiles_y, total_tiles, tiles_per_block;
};

// the 16x64 tile (WR = 4, WC = 2) holds one block per CU in LDS (95 KB) anyway,
// so it may use a whole SIMD's registers (accumulators in AGPRs, no spills)
template <int WR, int WC, bool UNROLL_DX, bool HEAD = false, bool FIRST = false, bool POOL = false>
__global__ __launch_bounds__(256, (WR * WC >= 8) ? 1 : (WR * WC <= 2 ? 3 : 2)) void k_conv3x3p(ConvPArgs pa)
{
    using K = ConvP<WR, WC>;
    const ConvArgs& a = pa.c;
    extern __shared__ uint4 convp_lds[];
    u16* s_in = reinterpret_cast<u16*>(convp_lds);
    u16* s_w = s_in + K::HH * K::HW * CONV_PS;
    float* s_bias = reinterpret_cast<float*>(s_w + 9 * K::CO_T * CONV_PS);
    float* s_patch = s_bias + K::CO_T; // first-layer mode
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int l32 = lane & 31, hi = lane >> 5;
    const int co0 = blockIdx.y * K::CO_T;
    const int nch = a.Cin / 32;
    const int t0 = blockIdx.x * pa.tiles_per_block;
    const int t1 = min(t0 + pa.tiles_per_block, pa.total_tiles);
    const int items = (t1 - t0) * nch;

    // the tile's biases live in LDS (read by the epilogue only), not in registers;
    // the first barrier of the item loop makes them visible
    for (int i = tid; i < K::CO_T; i += 256) s_bias[i] = a.b ? a.b[co0 + i] : 0.0f;
    // bias of accumulator element e (channel 8 (e / 4) + 4 hi + e % 4) of output-channel block c
    auto bias_of = [&](int c, int e) __attribute__((always_inline)) {
        return s_bias[32 * c + 8 * (e >> 2) + 4 * hi + (e & 3)];
    };

    // head: the 1x1 weights and bias, read once (uniform, kept in scalar registers)
    float hwv[HEAD ? 33 : 1];
    if constexpr (HEAD) {
#pragma unroll
        for (int j = 0; j < 32; ++j)
            hwv[j] = __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, a.hw[j])));
        hwv[32] = __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, a.hb[0])));
    }

    // first-layer mode: lane (l32, hi) supplies the weights of first-layer
    // channel l32 for taps 8hi .. 8hi+7 (zero past tap 8) and receives
    // channels 8g + 4hi + i of a halo pixel
    bf16x8 a0;
    float b0v[16];
    if constexpr (FIRST) {
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int tap = 8 * hi + j;
            a0[j] = tap < 9 ? __builtin_bit_cast(__bf16, a.w0[((size_t)l32 * 9 + tap) * a.ldw0]) : (__bf16)0.0f;
        }
#pragma unroll
        for (int i = 0; i < 16; ++i) b0v[i] = a.b0 ? a.b0[8 * (i / 4) + 4 * hi + (i % 4)] : 0.0f;
    }

    // Item cursors (uniform): chunk, tile column / row, image.  Items run in
    // order, so a cursor advances by increments; only the first tile is divided out.
    struct Cursor {
        int ch, tx, ty, n;
    };
    auto cursor_at = [&](int t) __attribute__((always_inline)) {
        Cursor c;
        c.ch = 0;
        c.tx = t % pa.tiles_x;
        c.ty = (t / pa.tiles_x) % pa.tiles_y;
        c.n = t / (pa.tiles_x * pa.tiles_y);
        return c;
    };
    auto advance = [&](Cursor& c) __attribute__((always_inline)) {
        if (++c.ch == nch) {
            c.ch = 0;
            if (++c.tx == pa.tiles_x) {
                c.tx = 0;
                if (++c.ty == pa.tiles_y) {
                    c.ty = 0;
                    ++c.n;
                }
            }
        }
    };
    Cursor lc = cursor_at(t0); // next item to load
    Cursor cc = lc;            // item being computed

    u32x4 hreg[FIRST ? 1 : K::HU], wreg[K::WU];
    float preg[FIRST ? K::PU : 1];
    unsigned vmask = 0; // bit k: unit k of the prefetched item lies inside the image
    // Every unit's load is issued unconditionally from a clamped in-image address
    // and out-of-image units are zeroed when stored to LDS, so the loads of an item
    // are all in flight together (with the loads under per-lane branches, the
    // compiler waited for each one before reusing registers on the other side of
    // the branch).  The source of a 32-channel chunk (x, the upsampled x2, or the
    // 2x2 average of x) is uniform over the item; the 2x2 average keeps per-lane
    // branches and consumes its four loads at once.
    // weight unit u: row x = u / 4 = co * 9 + tap of [Cout][9][Cin], 8 channels at 8 (u % 4)
    auto load_item = [&](const Cursor& c, bool weights) __attribute__((always_inline)) {
        const int ch = c.ch, tx0 = c.tx * 32, ty0 = c.ty * K::TH, n = c.n;
        vmask = 0;
        if constexpr (FIRST) {
            // image patch rows ty0-2 .. ty0+TH+1, columns tx0-2 .. tx0+33
            const float* ib = a.img0 + (size_t)n * a.H * a.W;
#pragma unroll
            for (int k = 0; k < K::PU; ++k) {
                const int u = min(tid + 256 * k, K::PATCH - 1);
                const int py = div36(u);
                const int gy = ty0 - 2 + py, gx = tx0 - 2 + (u - py * K::PW);
                const bool valid = tid + 256 * k < K::PATCH && (unsigned)gy < (unsigned)a.H && (unsigned)gx < (unsigned)a.W;
                const int cy = min(max(gy, 0), a.H - 1), cx = min(max(gx, 0), a.W - 1);
                preg[k] = ib[__umul24((unsigned)cy, (unsigned)a.W) + (unsigned)cx];
                vmask |= (unsigned)valid << k;
            }
        } else {
            const u16* xb = a.x + (size_t)n * a.H * a.W * a.ldx + a.cx0;
            const int c0 = 32 * ch;
            const bool up = __builtin_amdgcn_readfirstlane(c0 >= a.split ? 1 : 0) != 0;
            if (!up && a.pool) {
#pragma unroll
                for (int k = 0; k < K::HU; ++k) {
                    const int u = tid + 256 * k;
                    uint4 v = make_uint4(0u, 0u, 0u, 0u);
                    if (u < K::HALO_UNITS) {
                        const int pix = u >> 2, q = u & 3;
                        const int hy = div34(pix);
                        const int gy = ty0 + hy - 1, gx = tx0 + (pix - hy * K::HW) - 1;
                        if ((unsigned)gy < (unsigned)a.H && (unsigned)gx < (unsigned)a.W)
                            v = conv_pooled_src(a, n, gy, gx, c0 + 8 * q);
                    }
                    hreg[k] = as_u32x4(v);
                }
                vmask = ~0u;
            } else {
                // x at H x W, or x2 at H/2 x W/2 read with nearest 2x upsampling
                const int sh = up ? 1 : 0;
                const unsigned ws = (unsigned)(a.W >> sh), ld = (unsigned)(up ? a.ldx2 : a.ldx);
                const u16* base = up ? a.x2 + (size_t)n * (a.H >> 1) * (a.W >> 1) * a.ldx2 + a.cx2 + (c0 - a.split)
                                     : xb + c0;
#pragma unroll
                for (int k = 0; k < K::HU; ++k) {
                    const int u = tid + 256 * k;
                    const int pix = u >> 2, q = u & 3;
                    const int hy = div34(pix);
                    const int gy = ty0 + hy - 1, gx = tx0 + (pix - hy * K::HW) - 1;
                    const bool valid = (unsigned)gy < (unsigned)a.H && (unsigned)gx < (unsigned)a.W;
                    const int cy = min(max(gy, 0), a.H - 1), cx = min(max(gx, 0), a.W - 1);
                    const unsigned sp = __umul24((unsigned)(cy >> sh), ws) + (unsigned)(cx >> sh);
                    hreg[k] = *reinterpret_cast<const u32x4*>(base + __umul24(sp, ld) + 8 * q);
                    vmask |= (unsigned)valid << k;
                }
            }
        }
        if (weights) {
            const u16* wb = a.w + 32 * ch;
#pragma unroll
            for (int k = 0; k < K::WU; ++k) {
                const int u = tid + 256 * k;
                if (u < K::W_UNITS)
                    wreg[k] = *reinterpret_cast<const u32x4*>(
                        wb + __umul24((unsigned)(co0 * 9 + (u >> 2)), (unsigned)a.Cin) + 8 * (u & 3));
            }
        }
    };
    // Every prefetch register is waited for here, also in lanes whose unit lies
    // past the end: a load left pending on one path makes the compiler wait for
    // all outstanding memory operations (the tile's output stores included) at
    // the next write of that register.
    auto store_item = [&](bool weights) __attribute__((always_inline)) {
        if constexpr (FIRST) {
#pragma unroll
            for (int k = 0; k < K::PU; ++k) asm volatile("" ::"v"(preg[k]));
        } else {
#pragma unroll
            for (int k = 0; k < K::HU; ++k) asm volatile("" ::"v"(hreg[k]));
        }
#pragma unroll
        for (int k = 0; k < K::WU; ++k) asm volatile("" ::"v"(wreg[k]));
        if constexpr (FIRST) {
#pragma unroll
            for (int k = 0; k < K::PU; ++k) {
                const int u = tid + 256 * k;
                // bf16-rounded scaled image, zero outside
                if (u < K::PATCH)
                    s_patch[u] = (vmask >> k) & 1u ? bf16_float(bf16_bits(preg[k] * a.scale0)) : 0.0f;
            }
        } else {
#pragma unroll
            for (int k = 0; k < K::HU; ++k) {
                const int u = tid + 256 * k;
                const u32x4 zero = {0u, 0u, 0u, 0u};
                if (u < K::HALO_UNITS)
                    *reinterpret_cast<u32x4*>(s_in + (u >> 2) * CONV_PS + 8 * (u & 3)) = (vmask >> k) & 1u ? hreg[k] : zero;
            }
        }
        if (weights) {
#pragma unroll
            for (int k = 0; k < K::WU; ++k) {
                const int u = tid + 256 * k;
                if (u < K::W_UNITS) {
                    const int x = u >> 2, co = div9(x), tap = x - 9 * co;
                    *reinterpret_cast<u32x4*>(s_w + (tap * K::CO_T + co) * CONV_PS + 8 * (u & 3)) = wreg[k];
                }
            }
        }
    };

    f32x16 acc[WR][WC];
#pragma unroll
    for (int r = 0; r < WR; ++r)
#pragma unroll
        for (int c = 0; c < WC; ++c)
#pragma unroll
            for (int i = 0; i < 16; ++i) acc[r][c][i] = 0.f;

    float resv[WR]; // head: the tile's residual values
    // bias, activation and stores of a finished tile (then its accumulators are cleared)
    auto finish_tile = [&](const Cursor& t) __attribute__((always_inline)) {
        const int tx0 = t.tx * 32, ty0 = t.ty * K::TH;
        const size_t img = (size_t)t.n * a.H * a.W;
        if constexpr (HEAD && WC == 1) {
            // lane (l32, hi) holds channels 8g + 4hi + i of pixel column l32;
            // the hi = 0 lane gathers its partner's 16 and sums all 32 in order
#pragma unroll
            for (int r = 0; r < WR; ++r) {
                if (ty0 + wave * WR + r >= a.H) break; // wave-uniform
                float hv[16], pv[16];
#pragma unroll
                for (int i = 0; i < 16; ++i) {
                    float v = acc[r][0][i] + bias_of(0, i);
                    if (a.relu) v = fmaxf(v, 0.0f);
                    hv[i] = bf16_float(bf16_bits(v));
                }
#pragma unroll
                for (int i = 0; i < 16; ++i) pv[i] = __shfl_xor(hv[i], 32);
                if (hi == 0 && tx0 + l32 < a.W) {
                    float z = hwv[32];
#pragma unroll
                    for (int g = 0; g < 4; ++g) {
#pragma unroll
                        for (int i = 0; i < 4; ++i) z = fmaf(hwv[8 * g + i], hv[4 * g + i], z);
#pragma unroll
                        for (int i = 0; i < 4; ++i) z = fmaf(hwv[8 * g + 4 + i], pv[4 * g + i], z);
                    }
                    const size_t p = img + (size_t)(ty0 + wave * WR + r) * a.W + tx0 + l32;
                    float pr = resv[r] + z;
                    if (a.clamp) pr = fmaxf(pr, a.floor_value);
                    a.pred[p] = pr;
                }
            }
        } else if (tx0 + l32 < a.W) {
#pragma unroll
            for (int r = 0; r < WR; ++r) {
                if (ty0 + wave * WR + r < a.H) {
                    const size_t p = img + (size_t)(ty0 + wave * WR + r) * a.W + tx0 + l32;
                    u16* yp = a.y + p * a.ldy + a.cy0;
                    if constexpr (WR * WC > 2) {
                        // 16-byte stores when the output rows allow them (uniform; the
                        // 32x32-tile forms keep 8-byte stores and their register budget
                        // for three blocks per CU)
                        const bool st16 =
                            ((reinterpret_cast<size_t>(a.y) & 15) | (size_t)(a.ldy & 7) | (size_t)(a.cy0 & 7)) == 0;
#pragma unroll
                        for (int c = 0; c < WC; ++c) {
                            uint2 pk[4]; // channels 8 g + 4 hi + (0..3), packed
#pragma unroll
                            for (int g = 0; g < 4; ++g) {
                                float o[4];
#pragma unroll
                                for (int i = 0; i < 4; ++i) {
                                    float v = acc[r][c][4 * g + i] + bias_of(c, 4 * g + i);
                                    if (a.relu) v = fmaxf(v, 0.0f);
                                    o[i] = v;
                                }
                                pk[g].x = bf16_pack2(o[0], o[1]);
                                pk[g].y = bf16_pack2(o[2], o[3]);
                            }
                            if (st16) {
                                // the lane pair (l32, l32 + 32) trades halves so that lane hi
                                // holds channels 8 (2k + hi) .. +7 of group pair k: one
                                // 16-byte store per pair instead of two 8-byte ones
#pragma unroll
                                for (int k = 0; k < 2; ++k) {
                                    const uint2 keep = hi ? pk[2 * k + 1] : pk[2 * k];
                                    const uint2 give = hi ? pk[2 * k] : pk[2 * k + 1];
                                    const unsigned gx = (unsigned)__shfl_xor((int)give.x, 32);
                                    const unsigned gy = (unsigned)__shfl_xor((int)give.y, 32);
                                    const uint4 v =
                                        hi ? make_uint4(gx, gy, keep.x, keep.y) : make_uint4(keep.x, keep.y, gx, gy);
                                    *reinterpret_cast<uint4*>(yp + co0 + 32 * c + 8 * (2 * k + hi)) = v;
                                }
                            } else {
#pragma unroll
                                for (int g = 0; g < 4; ++g)
                                    *reinterpret_cast<uint2*>(yp + co0 + 32 * c + 8 * g + 4 * hi) = pk[g];
                            }
                        }
                    } else {
                        // 32x32 tiles: each 8-byte group stored as soon as it is packed
                        // (fewer live registers)
#pragma unroll
                        for (int c = 0; c < WC; ++c)
#pragma unroll
                            for (int g = 0; g < 4; ++g) {
                                float o[4];
#pragma unroll
                                for (int i = 0; i < 4; ++i) {
                                    float v = acc[r][c][4 * g + i] + bias_of(c, 4 * g + i);
                                    if (a.relu) v = fmaxf(v, 0.0f);
                                    o[i] = v;
                                }
                                uint2 pk;
                                pk.x = bf16_pack2(o[0], o[1]);
                                pk.y = bf16_pack2(o[2], o[3]);
                                *reinterpret_cast<uint2*>(yp + co0 + 32 * c + 8 * g + 4 * hi) = pk;
                            }
                    }
                }
            }
        }
        if constexpr (POOL && !HEAD) {
            // 2x2 average of the stored (bf16) outputs of row pairs (2k, 2k+1)
            // and column pairs (l32, l32 + 1), summed in k_pool2's order
            // (+0, x00, x01, x10, x11), written by the even-column lanes
#pragma unroll
            for (int r = 1; r < WR; r += 2) {
                const int row = ty0 + wave * WR + r - 1;
                if (row >= a.H) break; // wave-uniform; H even, so row + 1 < H
                const bool store = (l32 & 1) == 0 && tx0 + l32 < a.W;
                const size_t pp = ((size_t)t.n * (a.H >> 1) + (row >> 1)) * (a.W >> 1) + ((tx0 + l32) >> 1);
                u16* qp = a.ypool + pp * a.ldp + a.cp0;
#pragma unroll
                for (int c = 0; c < WC; ++c)
#pragma unroll
                    for (int g = 0; g < 4; ++g) {
                        float q[4];
#pragma unroll
                        for (int i = 0; i < 4; ++i) {
                            float v0 = acc[r - 1][c][4 * g + i] + bias_of(c, 4 * g + i);
                            float v1 = acc[r][c][4 * g + i] + bias_of(c, 4 * g + i);
                            if (a.relu) {
                                v0 = fmaxf(v0, 0.0f);
                                v1 = fmaxf(v1, 0.0f);
                            }
                            v0 = bf16_float(bf16_bits(v0));
                            v1 = bf16_float(bf16_bits(v1));
                            const float n0 = odd_neighbour(v0), n1 = odd_neighbour(v1);
                            float sum = 0.0f;
                            sum += v0;
                            sum += n0;
                            sum += v1;
                            sum += n1;
                            q[i] = sum * 0.25f;
                        }
                        if (store) {
                            uint2 pk;
                            pk.x = bf16_pack2(q[0], q[1]);
                            pk.y = bf16_pack2(q[2], q[3]);
                            *reinterpret_cast<uint2*>(qp + co0 + 32 * c + 8 * g + 4 * hi) = pk;
                        }
                    }
            }
        }
#pragma unroll
        for (int r = 0; r < WR; ++r)
#pragma unroll
            for (int c = 0; c < WC; ++c)
#pragma unroll
                for (int i = 0; i < 16; ++i) acc[r][c][i] = 0.f;
    };

    if (items > 0) {
        load_item(lc, true);
        advance(lc);
    }
    // A tile's outputs are written at the start of the next item, before that
    // item's prefetch loads: global stores share the loads' completion counter,
    // so stores issued after the loads would be waited for at the next LDS
    // store; issued first, they complete while the MFMAs run.
    Cursor pc = cc; // tile whose outputs are pending
    bool pending = false;
    for (int it = 0; it < items; ++it) {
        const int ch = cc.ch;
        __syncthreads();
        store_item(it == 0 || nch > 1);
        __syncthreads();
        if (pending) {
            finish_tile(pc);
            pending = false;
        }
        if (it + 1 < items) {
            load_item(lc, nch > 1);
            advance(lc);
        }
        if constexpr (FIRST) {
            // the halo's first-layer channels, one MFMA per 32 halo pixels
            const int tx0 = cc.tx * 32, ty0 = cc.ty * K::TH;
            for (int grp = wave; grp < K::HALO_GROUPS; grp += 4) {
                const int hp = 32 * grp + l32;
                const bool valid = hp < K::HH * K::HW;
                const int hy = valid ? div34(hp) : 0, hx = valid ? hp - hy * K::HW : 0;
                bf16x8 bv;
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const int tap = 8 * hi + j;
                    bv[j] = tap < 9 ? (__bf16)s_patch[(hy + tap / 3) * K::PW + hx + tap % 3] : (__bf16)0.0f;
                }
                f32x16 d;
#pragma unroll
                for (int i = 0; i < 16; ++i) d[i] = 0.0f;
                d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, bv, d, 0, 0, 0);
                const int gy = ty0 - 1 + hy, gx = tx0 - 1 + hx;
                const bool inside = valid && gy >= 0 && gy < a.H && gx >= 0 && gx < a.W;
                if (valid) {
#pragma unroll
                    for (int g = 0; g < 4; ++g) {
                        float o[4];
#pragma unroll
                        for (int i = 0; i < 4; ++i) {
                            float v = d[4 * g + i] + b0v[4 * g + i];
                            if (a.relu0) v = fmaxf(v, 0.0f);
                            o[i] = inside ? v : 0.0f;
                        }
                        uint2 pk;
                        pk.x = bf16_pack2(o[0], o[1]);
                        pk.y = bf16_pack2(o[2], o[3]);
                        *reinterpret_cast<uint2*>(s_in + hp * CONV_PS + 8 * g + 4 * hi) = pk;
                    }
                }
            }
            __syncthreads();
        }
        // head: the tile's residual values are fetched before its MFMAs
        if constexpr (HEAD && WC == 1) {
            if (ch == nch - 1) {
                const int tx0 = cc.tx * 32, ty0 = cc.ty * K::TH;
                const size_t img = (size_t)cc.n * a.H * a.W;
#pragma unroll
                for (int r = 0; r < WR; ++r) {
                    const int row = ty0 + wave * WR + r;
                    resv[r] = (a.res && row < a.H && tx0 + l32 < a.W) ? a.res[img + (size_t)row * a.W + tx0 + l32]
                                                                       : 0.0f;
                }
            }
        }
#pragma unroll(UNROLL_DX ? 3 : 1)
        for (int dx = 0; dx < 3; ++dx) {
#pragma unroll
            for (int kk = 0; kk < 32; kk += 16) {
                bf16x8 bp[WR + 2];
#pragma unroll
                for (int rr = 0; rr < WR + 2; ++rr)
                    bp[rr] = lds8(s_in + ((wave * WR + rr) * K::HW + l32 + dx) * CONV_PS + kk + 8 * hi);
#pragma unroll
                for (int dy = 0; dy < 3; ++dy) {
                    bf16x8 aw[WC];
#pragma unroll
                    for (int c = 0; c < WC; ++c)
                        aw[c] = lds8(s_w + ((3 * dy + dx) * K::CO_T + 32 * c + l32) * CONV_PS + kk + 8 * hi);
#pragma unroll
                    for (int r = 0; r < WR; ++r)
#pragma unroll
                        for (int c = 0; c < WC; ++c)
                            acc[r][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(aw[c], bp[r + dy], acc[r][c], 0, 0, 0);
                }
            }
        }
        if (ch == nch - 1) {
            pc = cc;
            pending = true;
        }
        advance(cc);
    }
    if (pending) finish_tile(pc);
}

// wt[ci][8 - tap][co] = w[co][tap][ci]: the backward-data weights
__global__ void k_conv_weight_flip(const u16* w, u16* wt, int Cout, int Cin)
{
    const int total = Cout * 9 * Cin;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < total; e += gridDim.x * blockDim.x) {
        const int co = e / (9 * Cin), tap = (e / Cin) % 9, ci = e % Cin;
        wt[((size_t)ci * 9 + (8 - tap)) * Cout + co] = w[e];
    }
}

// ---------------------------------------------------------------------------
// 3x3 convolution, backward-weight:
//   dW[co][tap][ci] = sum_{n,h,w} g[n][h][w][co] x[n][h+dy][w+dx][ci],  db[co] = sum g[..][co]
// GEMM with M = 32 output channels, N = 32 input channels, K = pixels.  A
// block owns one (co, ci) tile pair and a contiguous range of 4-row x 32-column
// pixel tiles; per tile it stages g transposed ([co][pixel]) and three
// column-shifted transposed copies of the x halo ([dx][ci][row][col]) so that
// each lane's 8 consecutive K values are one aligned 16-byte LDS read.  Wave w
// reduces tile row w for all 9 taps; the 4 waves are summed through LDS and
// the block writes its FP32 partial [9][32][32] (+ 32 bias sums) to
// part[blockIdx.x]; k_wgrad_reduce adds the partials in a fixed order.
// ---------------------------------------------------------------------------
constexpr int WG_TH = 4;
constexpr int WG_ROW = 32 + 8;                  // LDS row stride (elements) of a transposed row
constexpr int WG_G = 32 * WG_TH * WG_ROW;       // g: [32 co][4 rows][WG_ROW]
constexpr int WG_X = 32 * (WG_TH + 2) * WG_ROW; // one shifted copy: [32 ci][6 rows][WG_ROW]

struct WgradArgs {
    const u16* x;
    const u16* g;
    float* part; // [gridDim.x][Cout/32][Cin/32][9*32*32 + 32]
    int N, H, W, Cin, Cout, ldx, cx0, ldg, cg0;
    int tiles_per_block, total_tiles;
};

__global__ __launch_bounds__(256) void k_conv3x3_wgrad(WgradArgs a)
{
    extern __shared__ uint4 wg_lds[];
    u16* s_g = reinterpret_cast<u16*>(wg_lds);
    u16* s_x = s_g + WG_G; // three copies
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int l32 = lane & 31, hi = lane >> 5;
    const int cot = blockIdx.y, cit = blockIdx.z;
    const int tiles_x = (a.W + 31) / 32, tiles_y = (a.H + WG_TH - 1) / WG_TH;

    f32x16 acc[9];
#pragma unroll
    for (int t = 0; t < 9; ++t)
#pragma unroll
        for (int i = 0; i < 16; ++i) acc[t][i] = 0.f;
    float bsum = 0.0f;

    const int t_begin = blockIdx.x * a.tiles_per_block;
    const int t_end = min(t_begin + a.tiles_per_block, a.total_tiles);
    for (int t = t_begin; t < t_end; ++t) {
        const int tx0 = (t % tiles_x) * 32;
        const int ty0 = ((t / tiles_x) % tiles_y) * WG_TH;
        const size_t img = (size_t)(t / (tiles_x * tiles_y)) * a.H * a.W;
        __syncthreads();
        // g: 4 rows x 16 column pairs x 4 channel octets = 256 units, one per thread
        {
            const int q = tid & 3, pr = (tid >> 2) & 15, row = tid >> 6;
            const int gx = tx0 + 2 * pr;
            const size_t p = img + (size_t)(ty0 + row) * a.W + gx;
            const int off = a.cg0 + 32 * cot + 8 * q;
            uint4 v0 = make_uint4(0u, 0u, 0u, 0u), v1 = v0;
            if (ty0 + row < a.H) {
                if (gx < a.W) v0 = *reinterpret_cast<const uint4*>(a.g + p * a.ldg + off);
                if (gx + 1 < a.W) v1 = *reinterpret_cast<const uint4*>(a.g + (p + 1) * a.ldg + off);
            }
            const unsigned e0[4] = {v0.x, v0.y, v0.z, v0.w}, e1[4] = {v1.x, v1.y, v1.z, v1.w};
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const unsigned lo = (e0[j >> 1] >> (16 * (j & 1))) & 0xffffu;
                const unsigned hv = (e1[j >> 1] >> (16 * (j & 1))) & 0xffffu;
                *reinterpret_cast<unsigned*>(s_g + ((8 * q + j) * WG_TH + row) * WG_ROW + 2 * pr) = lo | (hv << 16);
            }
        }
        // x halo: 3 shifts x 6 rows x 16 column pairs x 4 octets = 1152 units
        for (int u = tid; u < 3 * (WG_TH + 2) * 16 * 4; u += 256) {
            const int q = u & 3, pr = (u >> 2) & 15, hr = (u >> 6) % (WG_TH + 2), dx = (u >> 6) / (WG_TH + 2);
            const int gy = ty0 + hr - 1;
            const int gx0 = tx0 + 2 * pr + dx - 1;
            uint4 v0 = make_uint4(0u, 0u, 0u, 0u), v1 = v0;
            if (gy >= 0 && gy < a.H) {
                const size_t rowp = img + (size_t)gy * a.W;
                const int off = a.cx0 + 32 * cit + 8 * q;
                if (gx0 >= 0 && gx0 < a.W) v0 = *reinterpret_cast<const uint4*>(a.x + (rowp + gx0) * a.ldx + off);
                if (gx0 + 1 >= 0 && gx0 + 1 < a.W)
                    v1 = *reinterpret_cast<const uint4*>(a.x + (rowp + gx0 + 1) * a.ldx + off);
            }
            const unsigned e0[4] = {v0.x, v0.y, v0.z, v0.w}, e1[4] = {v1.x, v1.y, v1.z, v1.w};
            u16* dst = s_x + dx * WG_X;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const unsigned lo = (e0[j >> 1] >> (16 * (j & 1))) & 0xffffu;
                const unsigned hv = (e1[j >> 1] >> (16 * (j & 1))) & 0xffffu;
                *reinterpret_cast<unsigned*>(dst + ((8 * q + j) * (WG_TH + 2) + hr) * WG_ROW + 2 * pr) =
                    lo | (hv << 16);
            }
        }
        __syncthreads();
        // wave = tile row; K = its 32 pixels in two halves of 16
#pragma unroll
        for (int tap = 0; tap < 9; ++tap) {
            const int dy = tap / 3, dx = tap % 3;
#pragma unroll
            for (int kk = 0; kk < 32; kk += 16) {
                const bf16x8 A = lds8(s_g + (l32 * WG_TH + wave) * WG_ROW + kk + 8 * hi);
                const bf16x8 B = lds8(s_x + dx * WG_X + (l32 * (WG_TH + 2) + wave + dy) * WG_ROW + kk + 8 * hi);
                acc[tap] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A, B, acc[tap], 0, 0, 0);
            }
        }
        if (cit == 0 && tid < 32) {
            for (int row = 0; row < WG_TH; ++row)
                for (int c = 0; c < 32; ++c) bsum += bf16_float(s_g[(tid * WG_TH + row) * WG_ROW + c]);
        }
    }
    // sum the 4 waves' accumulators through LDS (reusing the staging area)
    __syncthreads();
    float* red = reinterpret_cast<float*>(wg_lds); // [9][32][32]
    for (int w = 0; w < 4; ++w) {
        if (wave == w) {
#pragma unroll
            for (int tap = 0; tap < 9; ++tap)
#pragma unroll
                for (int i = 0; i < 16; ++i) {
                    const int co = 8 * (i / 4) + 4 * hi + (i % 4);
                    float* d = red + (tap * 32 + co) * 32 + l32;
                    *d = (w == 0 ? 0.0f : *d) + acc[tap][i];
                }
        }
        __syncthreads();
    }
    float* out = a.part + ((size_t)(blockIdx.x * gridDim.y + cot) * gridDim.z + cit) * (9 * 32 * 32 + 32);
    for (int e = tid; e < 9 * 32 * 32; e += 256) out[e] = red[e];
    if (tid < 32) out[9 * 32 * 32 + tid] = bsum;
}

// dW[co][tap][ci] (+)= sum over partials in order; db likewise
__global__ void k_wgrad_reduce(const float* part, int nparts, int ncot, int ncit, float* dw, float* db, int Cin,
                               int accumulate)
{
    const int per = 9 * 32 * 32 + 32;
    const int total = ncot * ncit * per;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < total; e += gridDim.x * blockDim.x) {
        const int tile = e / per, k = e % per;
        const int cot = tile / ncit, cit = tile % ncit;
        float s = 0.0f;
        for (int b = 0; b < nparts; ++b) s += part[(size_t)b * total + e];
        if (k < 9 * 32 * 32) {
            const int tap = k / 1024, co = (k / 32) % 32, ci = k % 32;
            float* d = dw + ((size_t)(32 * cot + co) * 9 + tap) * Cin + 32 * cit + ci;
            *d = accumulate ? *d + s : s;
        } else if (cit == 0 && db) {
            float* d = db + 32 * cot + (k - 9 * 32 * 32);
            *d = accumulate ? *d + s : s;
        }
    }
}

// ---------------------------------------------------------------------------
// elementwise / layout kernels (NHWC bf16, channel stride + offset everywhere)
// ---------------------------------------------------------------------------
struct View {
    u16* p;
    int ld, c0;
};
struct CView {
    const u16* p;
    int ld, c0;
};

// out = g * (act > 0)
__global__ void k_relu_mask(CView g, CView act, View out, long long P, int C)
{
    const int C8 = C / 8;
    const long long total = P * C8;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const long long p = e / C8;
        const int c = 8 * (int)(e % C8);
        const uint4 gv = ld8(g.p + p * g.ld + g.c0 + c), av = ld8(act.p + p * act.ld + act.c0 + c);
        const unsigned ge[4] = {gv.x, gv.y, gv.z, gv.w}, ae[4] = {av.x, av.y, av.z, av.w};
        unsigned o[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const unsigned lo = __uint_as_float(ae[j] << 16) > 0.0f ? (ge[j] & 0xffffu) : 0u;
            const unsigned hv = __uint_as_float(ae[j] & 0xffff0000u) > 0.0f ? (ge[j] & 0xffff0000u) : 0u;
            o[j] = lo | hv;
        }
        st8(out.p + p * out.ld + out.c0 + c, make_uint4(o[0], o[1], o[2], o[3]));
    }
}

// 2x2 average pool: y (Ho x Wo) from x (2Ho x 2Wo)
__global__ void k_pool2(CView x, View y, int N, int Ho, int Wo, int C)
{
    const int C8 = C / 8;
    const long long total = (long long)N * Ho * Wo * C8;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const int c = 8 * (int)(e % C8);
        const long long q = e / C8;
        const int w = (int)(q % Wo), h = (int)((q / Wo) % Ho);
        const long long n = q / ((long long)Wo * Ho);
        const long long base = (n * 2 * Ho + 2 * h) * 2 * Wo + 2 * w;
        const long long src[4] = {base, base + 1, base + 2 * Wo, base + 2 * Wo + 1};
        float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f}, f[8];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            unpack8(ld8(x.p + src[k] * x.ld + x.c0 + c), f);
#pragma unroll
            for (int j = 0; j < 8; ++j) acc[j] += f[j];
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) acc[j] *= 0.25f;
        st8(y.p + q * y.ld + y.c0 + c, pack8(acc));
    }
}

// Gradient reaching a skip tensor s (post-ReLU) from both of its readers:
//   out = (gcat + pool_backward(gpool)) * (s > 0),  pool_backward(g)[2h+a][2w+b] = g[h][w] / 4
__global__ void k_skip_grad(CView gcat, CView gpool, CView act, View out, int N, int H, int W, int C)
{
    const int C8 = C / 8;
    const long long total = (long long)N * H * W * C8;
    const int Ho = H / 2, Wo = W / 2;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const int c = 8 * (int)(e % C8);
        const long long p = e / C8;
        const int w = (int)(p % W), h = (int)((p / W) % H);
        const long long n = p / ((long long)W * H);
        const long long pp = (n * Ho + h / 2) * Wo + w / 2;
        float a8[8], g8[8], q8[8];
        unpack8(ld8(act.p + p * act.ld + act.c0 + c), a8);
        unpack8(ld8(gcat.p + p * gcat.ld + gcat.c0 + c), g8);
        unpack8(ld8(gpool.p + pp * gpool.ld + gpool.c0 + c), q8);
#pragma unroll
        for (int j = 0; j < 8; ++j) g8[j] = a8[j] > 0.0f ? g8[j] + 0.25f * q8[j] : 0.0f;
        st8(out.p + p * out.ld + out.c0 + c, pack8(g8));
    }
}

// nearest 2x upsampling (Hl x Wl -> 2Hl x 2Wl) and its backward (sum of the 4)
__global__ void k_up2(CView x, View y, int N, int Hl, int Wl, int C, int backward)
{
    const int C8 = C / 8;
    const long long total = (long long)N * Hl * Wl * C8;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const int c = 8 * (int)(e % C8);
        const long long q = e / C8;
        const int w = (int)(q % Wl), h = (int)((q / Wl) % Hl);
        const long long n = q / ((long long)Wl * Hl);
        const long long base = (n * 2 * Hl + 2 * h) * 2 * Wl + 2 * w;
        const long long hp[4] = {base, base + 1, base + 2 * Wl, base + 2 * Wl + 1};
        if (!backward) {
            const uint4 v = ld8(x.p + q * x.ld + x.c0 + c);
#pragma unroll
            for (int k = 0; k < 4; ++k) st8(y.p + hp[k] * y.ld + y.c0 + c, v);
        } else {
            float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f}, f[8];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                unpack8(ld8(x.p + hp[k] * x.ld + x.c0 + c), f);
#pragma unroll
                for (int j = 0; j < 8; ++j) acc[j] += f[j];
            }
            st8(y.p + q * y.ld + y.c0 + c, pack8(acc));
        }
    }
}

// Network input: x[p][0] = bf16(img[p] * scale), channels 1..31 zero (ld = 32)
__global__ void k_input(const float* img, u16* x, long long P, float scale)
{
    for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < P; p += (long long)gridDim.x * blockDim.x) {
        uint4 v = make_uint4((unsigned)bf16_bits(img[p] * scale), 0u, 0u, 0u);
        uint4* d = reinterpret_cast<uint4*>(x + p * 32);
        d[0] = v;
        v.x = 0u;
        d[1] = v;
        d[2] = v;
        d[3] = v;
    }
}

// Output head (32 -> 1, 1x1) with residual and MSE loss.
//   z = b + sum_c w[c] h[p][c];  pred = res[p] + z
// With a target: loss += sum (pred - t)^2 (FP64), gz = 2 (pred - t) / P,
// gh[p][c] = gz w[c] (bf16), and per-block partials of dw (= sum gz h) and db.
// Without a target the prediction may be clamped from below (positivity of an
// eEM starting image).
struct HeadArgs {
    const u16* h;
    int ldh, ch0;
    const float* w;   // [32]
    const float* b;   // [1]
    const float* res; // [P] or null
    float* pred;      // [P]
    const float* target;
    u16* gh;
    int ldgh, cgh0;
    float* part; // [gridDim.x][33]
    double* loss;
    long long P;
    float floor_value;
    int clamp;
};

__global__ __launch_bounds__(256) void k_head(HeadArgs a)
{
    __shared__ float s_red[4][33];
    __shared__ double s_loss[4];
    float w[32];
#pragma unroll
    for (int c = 0; c < 32; ++c) w[c] = a.w[c];
    const float b = a.b[0];
    float dw[32];
#pragma unroll
    for (int c = 0; c < 32; ++c) dw[c] = 0.0f;
    float db = 0.0f;
    double loss = 0.0;
    const float invP = 1.0f / (float)a.P;
    for (long long p = blockIdx.x * 256LL + threadIdx.x; p < a.P; p += (long long)gridDim.x * 256) {
        const uint4* hp = reinterpret_cast<const uint4*>(a.h + p * a.ldh + a.ch0);
        float hv[32];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint4 v = hp[k];
            const unsigned e[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                hv[8 * k + 2 * j] = __uint_as_float(e[j] << 16);
                hv[8 * k + 2 * j + 1] = __uint_as_float(e[j] & 0xffff0000u);
            }
        }
        float z = b;
#pragma unroll
        for (int c = 0; c < 32; ++c) z = fmaf(w[c], hv[c], z);
        float pr = (a.res ? a.res[p] : 0.0f) + z;
        if (!a.target) {
            if (a.clamp) pr = fmaxf(pr, a.floor_value);
            a.pred[p] = pr;
            continue;
        }
        a.pred[p] = pr;
        const float d = pr - a.target[p];
        loss += (double)d * (double)d;
        const float gz = 2.0f * d * invP;
        db += gz;
        unsigned o[16];
#pragma unroll
        for (int c = 0; c < 32; c += 2) {
            dw[c] = fmaf(gz, hv[c], dw[c]);
            dw[c + 1] = fmaf(gz, hv[c + 1], dw[c + 1]);
            o[c / 2] = (unsigned)bf16_bits(gz * w[c]) | ((unsigned)bf16_bits(gz * w[c + 1]) << 16);
        }
        uint4* gp = reinterpret_cast<uint4*>(a.gh + p * a.ldgh + a.cgh0);
#pragma unroll
        for (int k = 0; k < 4; ++k) gp[k] = make_uint4(o[4 * k], o[4 * k + 1], o[4 * k + 2], o[4 * k + 3]);
    }
    if (!a.target) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c < 32; ++c) {
        float v = dw[c];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
        if (lane == 0) s_red[wave][c] = v;
    }
    for (int o = 32; o > 0; o >>= 1) {
        db += __shfl_xor(db, o);
        loss += __shfl_xor(loss, o);
    }
    if (lane == 0) {
        s_red[wave][32] = db;
        s_loss[wave] = loss;
    }
    __syncthreads();
    if (threadIdx.x < 33)
        a.part[blockIdx.x * 33 + threadIdx.x] =
            s_red[0][threadIdx.x] + s_red[1][threadIdx.x] + s_red[2][threadIdx.x] + s_red[3][threadIdx.x];
    if (threadIdx.x == 0) a.loss[blockIdx.x + 1] = s_loss[0] + s_loss[1] + s_loss[2] + s_loss[3];
}

// dw, db from the head partials (block k < 33 reduces column k) and loss[0]
// = sum of the per-block losses (block 33); fixed order, so deterministic.
__global__ __launch_bounds__(256) void k_head_reduce(const float* part, const int nparts, float* dw, float* db,
                                                     double* loss, int accumulate)
{
    __shared__ double sd[256];
    const int k = blockIdx.x;
    double s = 0.0;
    for (int b = threadIdx.x; b < nparts; b += 256) s += k < 33 ? (double)part[b * 33 + k] : loss[b + 1];
    sd[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) sd[threadIdx.x] += sd[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    if (k == 33) {
        loss[0] = sd[0];
        return;
    }
    float* d = k < 32 ? dw + k : db;
    *d = accumulate ? *d + (float)sd[0] : (float)sd[0];
}

// Adam (bias-corrected) on FP32 master weights; refreshes the bf16 copy
__global__ void k_adam(float* p, const float* g, float* m, float* v, u16* pb, long long n, float lr, float b1,
                       float b2, float eps, float c1, float c2, float gscale)
{
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
        const float gr = g[e] * gscale;
        const float mm = b1 * m[e] + (1.0f - b1) * gr;
        const float vv = b2 * v[e] + (1.0f - b2) * gr * gr;
        m[e] = mm;
        v[e] = vv;
        const float np = p[e] - lr * (mm / c1) / (sqrtf(vv / c2) + eps);
        p[e] = np;
        if (pb) pb[e] = bf16_bits(np);
    }
}

__global__ void k_to_bf16(const float* in, u16* out, long long n)
{
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x)
        out[e] = bf16_bits(in[e]);
}

// Bilinear resampling with edge clamping: out[s][i][j] = img[s](Y(i), X(j)) with
// X(j) = ax + bx j, Y(i) = ay + by i in input pixel units.
__global__ void k_resample(const float* img, float* out, int S, int Hi, int Wi, int Ho, int Wo, double ax, double bx,
                           double ay, double by)
{
    const long long total = (long long)S * Ho * Wo;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const int j = (int)(e % Wo), i = (int)((e / Wo) % Ho);
        const long long s = e / ((long long)Wo * Ho);
        double X = fma((double)j, bx, ax), Y = fma((double)i, by, ay);
        X = fmin(fmax(X, 0.0), (double)(Wi - 1));
        Y = fmin(fmax(Y, 0.0), (double)(Hi - 1));
        const int x0 = (int)X, y0 = (int)Y;
        const int x1 = min(x0 + 1, Wi - 1), y1 = min(y0 + 1, Hi - 1);
        const float fx = (float)(X - x0), fy = (float)(Y - y0);
        const float* b = img + s * (long long)Hi * Wi;
        const float v00 = b[(long long)y0 * Wi + x0], v01 = b[(long long)y0 * Wi + x1];
        const float v10 = b[(long long)y1 * Wi + x0], v11 = b[(long long)y1 * Wi + x1];
        const float top = fmaf(fx, v01 - v00, v00), bot = fmaf(fx, v11 - v10, v10);
        out[e] = fmaf(fy, bot - top, top);
    }
}

int grid_for(long long n, int block = 256)
{
    const long long b = (n + block - 1) / block;
    return (int)(b > 8192 ? 8192 : (b < 1 ? 1 : b));
}

// ---------------------------------------------------------------------------
// First 3x3 convolution straight from the FP32 image (one input channel):
//   y[p][cy0 + co] = act(b[co] + sum_tap bf16(img[p + off(tap)] * scale) W[co][tap][0])
// with W read from a [Cout][9][ldw] bf16 tensor (the layer's weights, whose
// other input channels multiply the zero-padded channels of the generic
// path).  One thread per pixel, 8 output channels per 16-byte store; products
// of two bf16 values are exact in FP32 and are summed over taps in order.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_conv_first(const float* img, float scale, const u16* w, int ldw,
                                                   const float* b, u16* y, int ldy, int cy0, int N, int H, int W,
                                                   int Cout, int relu)
{
    __shared__ float s_w[9 * 64];
    __shared__ float s_b[64];
    for (int e = threadIdx.x; e < 9 * Cout; e += blockDim.x) {
        const int co = e / 9, tap = e % 9;
        s_w[tap * Cout + co] = bf16_float(w[((size_t)co * 9 + tap) * ldw]);
    }
    for (int e = threadIdx.x; e < Cout; e += blockDim.x) s_b[e] = b ? b[e] : 0.0f;
    __syncthreads();
    const long long P = (long long)N * H * W;
    for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < P; p += (long long)gridDim.x * blockDim.x) {
        const int x = (int)(p % W), yy = (int)((p / W) % H);
        float in[9];
#pragma unroll
        for (int tap = 0; tap < 9; ++tap) {
            const int dy = tap / 3 - 1, dx = tap % 3 - 1;
            const int gy = yy + dy, gx = x + dx;
            in[tap] = (gy >= 0 && gy < H && gx >= 0 && gx < W) ? bf16_float(bf16_bits(img[p + dy * W + dx] * scale))
                                                                : 0.0f;
        }
        u16* yp = y + p * ldy + cy0;
        for (int c8 = 0; c8 < Cout; c8 += 8) {
            float o[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                float acc = 0.0f;
#pragma unroll
                for (int tap = 0; tap < 9; ++tap) acc += in[tap] * s_w[tap * Cout + c8 + j];
                float v = acc + s_b[c8 + j];
                o[j] = relu ? fmaxf(v, 0.0f) : v;
            }
            st8(yp + c8, pack8(o));
        }
    }
}


hipStream_t S(void* s) { return static_cast<hipStream_t>(s); }

int launched() { return hipGetLastError() == hipSuccess ? RAFT_OK : RAFT_EHIP; }

template <int WR, int WC>
int launch_conv(const ConvArgs& a, hipStream_t st)
{
    constexpr int lds = conv_lds_bytes<WR, WC>();
    static std::atomic<bool> attr[64]; // per device: the LDS limit is raised once
    int attr_dev = 0;
    if (hipGetDevice(&attr_dev) != hipSuccess || attr_dev < 0 || attr_dev >= 64) return RAFT_EHIP;
    if (!attr[attr_dev].load(std::memory_order_relaxed)) {
        if (hipFuncSetAttribute(reinterpret_cast<const void*>(&k_conv3x3<WR, WC>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, lds) != hipSuccess)
            return RAFT_EHIP;
        attr[attr_dev].store(true, std::memory_order_relaxed);
    }
    dim3 grid((a.W + 31) / 32, ((a.H + 4 * WR - 1) / (4 * WR)) * a.N, a.Cout / (32 * WC));
    if ((long long)((a.H + 4 * WR - 1) / (4 * WR)) * a.N > 65535) return RAFT_EINVAL;
    hipLaunchKernelGGL((k_conv3x3<WR, WC>), grid, dim3(256), lds, st, a);
    return launched();
}

template <int WR, int WC, bool UNROLL_DX = true, bool HEAD = false, bool FIRST = false, bool POOL = false>
int launch_convp(const ConvArgs& a, hipStream_t st)
{
    using K = ConvP<WR, WC>;
    constexpr int lds = FIRST ? K::LDS_FIRST : K::LDS;
    static std::atomic<bool> attr[64]; // per device: the LDS limit is raised once
    int attr_dev = 0;
    if (hipGetDevice(&attr_dev) != hipSuccess || attr_dev < 0 || attr_dev >= 64) return RAFT_EHIP;
    if (!attr[attr_dev].load(std::memory_order_relaxed)) {
        if (hipFuncSetAttribute(reinterpret_cast<const void*>(&k_conv3x3p<WR, WC, UNROLL_DX, HEAD, FIRST, POOL>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, lds) != hipSuccess)
            return RAFT_EHIP;
        attr[attr_dev].store(true, std::memory_order_relaxed);
    }
    // the kernel's 24-bit multiplies and 32-bit per-image offsets need these
    const long long hw = (long long)a.H * a.W;
    if (hw >= (1LL << 24) || hw * a.ldx >= (1LL << 31) || (long long)a.Cout * 9 * a.Cin >= (1LL << 31) ||
        (a.x2 && (long long)(a.H >> 1) * (a.W >> 1) * a.ldx2 >= (1LL << 31)))
        return RAFT_EINVAL;
    ConvPArgs pa;
    pa.c = a;
    pa.tiles_x = (a.W + 31) / 32;
    pa.tiles_y = (a.H + K::TH - 1) / K::TH;
    const long long total = (long long)a.N * pa.tiles_y * pa.tiles_x;
    if (total > (1LL << 30)) return RAFT_EINVAL;
    pa.total_tiles = (int)total;
    const int ny = a.Cout / K::CO_T;
    // one resident round: as many blocks as the device holds at once (CUs x
    // blocks per CU for this kernel's registers and LDS), tiles split evenly
    static std::atomic<int> slots[64]; // per device, 0 = not yet queried
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return RAFT_EHIP;
    int nslots = slots[dev].load(std::memory_order_relaxed);
    if (nslots == 0) {
        int occ = 0, cus = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, reinterpret_cast<const void*>(
                                                                   &k_conv3x3p<WR, WC, UNROLL_DX, HEAD, FIRST, POOL>),
                                                         256, lds) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            return RAFT_EHIP;
        nslots = std::max(1, occ) * std::max(1, cus);
        slots[dev].store(nslots, std::memory_order_relaxed);
    }
    const long long per = (total * ny + nslots - 1) / nslots;
    pa.tiles_per_block = (int)std::max(1LL, per);
    const long long nbx = (total + pa.tiles_per_block - 1) / pa.tiles_per_block;
    hipLaunchKernelGGL((k_conv3x3p<WR, WC, UNROLL_DX, HEAD, FIRST, POOL>), dim3((unsigned)nbx, ny), dim3(256), lds, st,
                       pa);
    return launched();
}

struct WgradPlan {
    int nblk, per, tiles, pairs;
};

bool wgrad_plan(int N, int H, int W, int Cin, int Cout, WgradPlan& p)
{
    if (N < 1 || W < 1 || H < 1 || Cin % 32 || Cout % 32 || Cin < 32 || Cout < 32) return false;
    p.tiles = N * ((H + WG_TH - 1) / WG_TH) * ((W + 31) / 32);
    p.pairs = (Cout / 32) * (Cin / 32);
    int nblk = (512 + p.pairs - 1) / p.pairs;
    if (nblk > p.tiles) nblk = p.tiles;
    p.per = (p.tiles + nblk - 1) / nblk;
    p.nblk = (p.tiles + p.per - 1) / p.per;
    return true;
}

} // namespace
} // namespace raft

using namespace raft;

extern "C" {

int raft_nn_mfma_probe(const unsigned short* A16, const unsigned short* B16, float* out16, const unsigned short* A32,
                       const unsigned short* B32, float* out32, void* stream)
{
    hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, S(stream), A16, B16, out16, A32, B32, out32);
    return launched();
}

static int conv_dispatch(const ConvArgs& a, int variant, hipStream_t st)
{
    // measured on gfx950 (batch 32, 512^2 .. 64^2 layers): the persistent 8x64
    // form for 64+ output channels, the persistent 8x32 form (dx loop not
    // unrolled) for 32
    if (variant == 0) variant = (a.Cout % 64 == 0) ? 7 : 13;
    switch (variant) {
    case 6:
        return launch_convp<4, 1>(a, st);
    case 7:
        if (a.Cout % 64) return RAFT_EINVAL;
        return launch_convp<2, 2>(a, st);
    case 8:
        if (a.Cout % 64) return RAFT_EINVAL;
        return launch_convp<4, 2>(a, st);
    case 9:
        return launch_convp<2, 1>(a, st);
    case 10:
        return launch_convp<4, 1, false>(a, st);
    case 11:
        if (a.Cout % 64) return RAFT_EINVAL;
        return launch_convp<2, 2, false>(a, st);
    case 12:
        if (a.Cout % 64) return RAFT_EINVAL;
        return launch_convp<4, 2, false>(a, st);
    case 13:
        return launch_convp<2, 1, false>(a, st);
    case 1:
        if (a.Cout % 64) return RAFT_EINVAL;
        return launch_conv<2, 2>(a, st);
    case 2:
        return launch_conv<4, 1>(a, st);
    case 3:
        return launch_conv<1, 1>(a, st);
    case 4:
        if (a.Cout % 64) return RAFT_EINVAL;
        return launch_conv<4, 2>(a, st);
    case 5:
        return launch_conv<2, 1>(a, st);
    default:
        return RAFT_EINVAL;
    }
}

// persistent forms that also write the 2x2-pooled output (ConvArgs::ypool)
static int conv_dispatch_pool(const ConvArgs& a, int variant, hipStream_t st)
{
    if (variant == 0) variant = (a.Cout % 64 == 0) ? 7 : 13;
    if ((variant == 7 || variant == 8 || variant == 11 || variant == 12) && a.Cout % 64) return RAFT_EINVAL;
    switch (variant) {
    case 6:
        return launch_convp<4, 1, true, false, false, true>(a, st);
    case 7:
        return launch_convp<2, 2, true, false, false, true>(a, st);
    case 8:
        return launch_convp<4, 2, true, false, false, true>(a, st);
    case 9:
        return launch_convp<2, 1, true, false, false, true>(a, st);
    case 10:
        return launch_convp<4, 1, false, false, false, true>(a, st);
    case 11:
        return launch_convp<2, 2, false, false, false, true>(a, st);
    case 12:
        return launch_convp<4, 2, false, false, false, true>(a, st);
    case 13:
        return launch_convp<2, 1, false, false, false, true>(a, st);
    default:
        return RAFT_EINVAL;
    }
}

static bool pool_out_ok(unsigned short* yp, int ldp, int cp0, int H, int W, int Cout)
{
    return !yp || (H % 2 == 0 && W % 2 == 0 && ldp % 4 == 0 && cp0 % 4 == 0 && cp0 + Cout <= ldp);
}

int raft_nn_conv3x3(const unsigned short* x, int ldx, int cx0, const unsigned short* w, const float* b,
                    unsigned short* y, int ldy, int cy0, int N, int H, int W, int Cin, int Cout, int relu, int variant,
                    void* stream)
{
    if (N < 1 || H < 1 || W < 1 || Cin < 32 || Cout < 32 || Cin % 32 || Cout % 32 || ldx % 8 || cx0 % 8 ||
        ldy % 4 || cy0 % 4 || cx0 + Cin > ldx || cy0 + Cout > ldy)
        return RAFT_EINVAL;
    ConvArgs a{x, w, b, y, N, H, W, Cin, Cout, ldx, cx0, ldy, cy0, relu, nullptr, 0, 0, Cin, 0};
    return conv_dispatch(a, variant, S(stream));
}

int raft_nn_conv3x3_ex(const unsigned short* x, int ldx, int cx0, int pool, const unsigned short* x2, int ldx2,
                       int cx2, int split, const unsigned short* w, const float* b, unsigned short* y, int ldy, int cy0,
                       int N, int H, int W, int Cin, int Cout, int relu, int variant, unsigned short* yp, int ldp,
                       int cp0, void* stream)
{
    if (!pool_out_ok(yp, ldp, cp0, H, W, Cout)) return RAFT_EINVAL;
    if (N < 1 || H < 1 || W < 1 || Cin < 32 || Cout < 32 || Cin % 32 || Cout % 32 || ldx % 8 || cx0 % 8 ||
        ldy % 4 || cy0 % 4 || cy0 + Cout > ldy || split < 32 || split > Cin || split % 32 || cx0 + split > ldx)
        return RAFT_EINVAL;
    if (split < Cin && (!x2 || H % 2 || W % 2 || ldx2 % 8 || cx2 % 8 || cx2 + (Cin - split) > ldx2))
        return RAFT_EINVAL;
    ConvArgs a{x, w, b, y, N, H, W, Cin, Cout, ldx, cx0, ldy, cy0, relu, split < Cin ? x2 : nullptr, ldx2, cx2, split,
               pool ? 1 : 0};
    if (!yp) return conv_dispatch(a, variant, S(stream));
    a.ypool = yp;
    a.ldp = ldp;
    a.cp0 = cp0;
    return conv_dispatch_pool(a, variant, S(stream));
}

int raft_nn_conv3x3_head(const unsigned short* x, int ldx, int cx0, const unsigned short* w, const float* b, int N,
                         int H, int W, int Cin, int relu, int variant, const float* hw, const float* hb,
                         const float* res, float* pred, float floor_value, int clamp, void* stream)
{
    if (N < 1 || H < 1 || W < 1 || Cin < 32 || Cin % 32 || ldx % 8 || cx0 % 8 || cx0 + Cin > ldx || !hw || !hb ||
        !pred)
        return RAFT_EINVAL;
    ConvArgs a{x, w, b, nullptr, N, H, W, Cin, 32, ldx, cx0, 0, 0, relu, nullptr, 0, 0, Cin, 0};
    a.hw = hw;
    a.hb = hb;
    a.res = res;
    a.pred = pred;
    a.floor_value = floor_value;
    a.clamp = clamp ? 1 : 0;
    // the persistent forms with one 32-channel output tile
    switch (variant) {
    case 6:
        return launch_convp<4, 1, true, true>(a, S(stream));
    case 9:
        return launch_convp<2, 1, true, true>(a, S(stream));
    case 10:
        return launch_convp<4, 1, false, true>(a, S(stream));
    case 0:
    case 13:
        return launch_convp<2, 1, false, true>(a, S(stream));
    default:
        return RAFT_EINVAL;
    }
}

int raft_nn_conv_first2(const float* img, float scale, const unsigned short* w0, int ldw0, const float* b0, int relu0,
                        const unsigned short* w, const float* b, unsigned short* y, int ldy, int cy0, int N, int H,
                        int W, int relu, int variant, unsigned short* yp, int ldp, int cp0, void* stream)
{
    if (N < 1 || H < 1 || W < 1 || ldw0 < 1 || ldy % 4 || cy0 % 4 || cy0 + 32 > ldy || !img || !w0 || !w ||
        !pool_out_ok(yp, ldp, cp0, H, W, 32))
        return RAFT_EINVAL;
    ConvArgs a{nullptr, w, b, y, N, H, W, 32, 32, 32, 0, ldy, cy0, relu, nullptr, 0, 0, 32, 0};
    a.img0 = img;
    a.scale0 = scale;
    a.w0 = w0;
    a.ldw0 = ldw0;
    a.b0 = b0;
    a.relu0 = relu0 ? 1 : 0;
    if (yp) {
        a.ypool = yp;
        a.ldp = ldp;
        a.cp0 = cp0;
        switch (variant) {
        case 6:
            return launch_convp<4, 1, true, false, true, true>(a, S(stream));
        case 9:
            return launch_convp<2, 1, true, false, true, true>(a, S(stream));
        case 10:
            return launch_convp<4, 1, false, false, true, true>(a, S(stream));
        case 0:
        case 13:
            return launch_convp<2, 1, false, false, true, true>(a, S(stream));
        default:
            return RAFT_EINVAL;
        }
    }
    switch (variant) {
    case 6:
        return launch_convp<4, 1, true, false, true>(a, S(stream));
    case 9:
        return launch_convp<2, 1, true, false, true>(a, S(stream));
    case 10:
        return launch_convp<4, 1, false, false, true>(a, S(stream));
    case 0:
    case 13:
        return launch_convp<2, 1, false, false, true>(a, S(stream));
    default:
        return RAFT_EINVAL;
    }
}

int raft_nn_conv_first(const float* img, float scale, const unsigned short* w, int ldw, const float* b,
                       unsigned short* y, int ldy, int cy0, int N, int H, int W, int Cout, int relu, void* stream)
{
    if (N < 1 || H < 1 || W < 1 || Cout < 8 || Cout > 64 || Cout % 8 || ldw < 1 || ldy % 8 || cy0 % 8 ||
        cy0 + Cout > ldy)
        return RAFT_EINVAL;
    hipLaunchKernelGGL(k_conv_first, dim3(grid_for((long long)N * H * W)), dim3(256), 0, S(stream), img, scale, w, ldw,
                       b, y, ldy, cy0, N, H, W, Cout, relu);
    return launched();
}

int raft_nn_conv_weight_flip(const unsigned short* w, unsigned short* wt, int Cout, int Cin, void* stream)
{
    hipLaunchKernelGGL(k_conv_weight_flip, dim3(grid_for((long long)Cout * 9 * Cin)), dim3(256), 0, S(stream), w, wt,
                       Cout, Cin);
    return launched();
}

long long raft_nn_wgrad_workspace(int N, int H, int W, int Cin, int Cout)
{
    WgradPlan p;
    if (!wgrad_plan(N, H, W, Cin, Cout, p)) return -1;
    return (long long)p.nblk * p.pairs * (9 * 32 * 32 + 32);
}

int raft_nn_conv3x3_wgrad(const unsigned short* x, int ldx, int cx0, const unsigned short* g, int ldg, int cg0,
                          float* dw, float* db, float* work, long long work_floats, int N, int H, int W, int Cin,
                          int Cout, int accumulate, void* stream)
{
    WgradPlan p;
    if (!wgrad_plan(N, H, W, Cin, Cout, p) || ldx % 8 || cx0 % 8 || ldg % 8 || cg0 % 8 || cx0 + Cin > ldx ||
        cg0 + Cout > ldg)
        return RAFT_EINVAL;
    if (work_floats < (long long)p.nblk * p.pairs * (9 * 32 * 32 + 32)) return RAFT_EINVAL;
    WgradArgs a{x, g, work, N, H, W, Cin, Cout, ldx, cx0, ldg, cg0, p.per, p.tiles};
    constexpr int lds = (WG_G + 3 * WG_X) * 2;
    static_assert(lds >= 9 * 32 * 32 * 4, "the wave reduction reuses the staging LDS");
    static std::atomic<bool> attr[64]; // per device: the LDS limit is raised once
    int attr_dev = 0;
    if (hipGetDevice(&attr_dev) != hipSuccess || attr_dev < 0 || attr_dev >= 64) return RAFT_EHIP;
    if (!attr[attr_dev].load(std::memory_order_relaxed)) {
        if (hipFuncSetAttribute(reinterpret_cast<const void*>(&k_conv3x3_wgrad),
                                hipFuncAttributeMaxDynamicSharedMemorySize, lds) != hipSuccess)
            return RAFT_EHIP;
        attr[attr_dev].store(true, std::memory_order_relaxed);
    }
    hipLaunchKernelGGL(k_conv3x3_wgrad, dim3(p.nblk, Cout / 32, Cin / 32), dim3(256), lds, S(stream), a);
    if (hipGetLastError() != hipSuccess) return RAFT_EHIP;
    const long long outn = (long long)p.pairs * (9 * 32 * 32 + 32);
    hipLaunchKernelGGL(k_wgrad_reduce, dim3(grid_for(outn)), dim3(256), 0, S(stream), work, p.nblk, Cout / 32,
                       Cin / 32, dw, db, Cin, accumulate);
    return launched();
}

static bool al8(int a, int b, int c, int d) { return a % 8 == 0 && b % 8 == 0 && c % 8 == 0 && d % 8 == 0; }

int raft_nn_relu_mask(const unsigned short* g, int ldg, int cg0, const unsigned short* act, int lda, int ca0,
                      unsigned short* out, int ldo, int co0, long long P, int C, void* stream)
{
    if (C < 8 || !al8(C, ldg, cg0, lda) || !al8(ca0, ldo, co0, 0)) return RAFT_EINVAL;
    hipLaunchKernelGGL(k_relu_mask, dim3(grid_for(P * C / 8)), dim3(256), 0, S(stream), CView{g, ldg, cg0},
                       CView{act, lda, ca0}, View{out, ldo, co0}, P, C);
    return launched();
}

int raft_nn_pool2(const unsigned short* x, int ldx, int cx0, unsigned short* y, int ldy, int cy0, int N, int Ho,
                  int Wo, int C, void* stream)
{
    if (C < 8 || !al8(C, ldx, cx0, ldy) || cy0 % 8) return RAFT_EINVAL;
    hipLaunchKernelGGL(k_pool2, dim3(grid_for((long long)N * Ho * Wo * C / 8)), dim3(256), 0, S(stream),
                       CView{x, ldx, cx0}, View{y, ldy, cy0}, N, Ho, Wo, C);
    return launched();
}

int raft_nn_skip_grad(const unsigned short* gcat, int ldc, int cc0, const unsigned short* gpool, int ldp, int cp0,
                      const unsigned short* act, int lda, int ca0, unsigned short* out, int ldo, int co0, int N, int H,
                      int W, int C, void* stream)
{
    if (H % 2 || W % 2 || C < 8 || !al8(C, ldc, cc0, ldp) || !al8(cp0, lda, ca0, ldo) || co0 % 8) return RAFT_EINVAL;
    hipLaunchKernelGGL(k_skip_grad, dim3(grid_for((long long)N * H * W * C / 8)), dim3(256), 0, S(stream),
                       CView{gcat, ldc, cc0}, CView{gpool, ldp, cp0}, CView{act, lda, ca0}, View{out, ldo, co0}, N, H,
                       W, C);
    return launched();
}

int raft_nn_up2(const unsigned short* x, int ldx, int cx0, unsigned short* y, int ldy, int cy0, int N, int Hl, int Wl,
                int C, int backward, void* stream)
{
    if (C < 8 || !al8(C, ldx, cx0, ldy) || cy0 % 8) return RAFT_EINVAL;
    hipLaunchKernelGGL(k_up2, dim3(grid_for((long long)N * Hl * Wl * C / 8)), dim3(256), 0, S(stream), CView{x, ldx, cx0},
                       View{y, ldy, cy0}, N, Hl, Wl, C, backward);
    return launched();
}

int raft_nn_input(const float* img, unsigned short* x, long long P, float scale, void* stream)
{
    hipLaunchKernelGGL(k_input, dim3(grid_for(P)), dim3(256), 0, S(stream), img, x, P, scale);
    return launched();
}

// Workspaces of raft_nn_head with a target: floats for the dw/db partials and
// doubles for the loss (slot 0 receives the total).
long long raft_nn_head_workspace(long long P) { return (long long)grid_for(P) * 33; }
long long raft_nn_head_loss_slots(long long P) { return (long long)grid_for(P) + 1; }

int raft_nn_head(const unsigned short* h, int ldh, int ch0, const float* w, const float* b, const float* res,
                 float* pred, const float* target, unsigned short* gh, int ldgh, int cgh0, float* dw, float* db,
                 double* loss, float* work, long long P, float floor_value, int clamp, int accumulate, void* stream)
{
    if (P < 1 || ldh % 8 || ch0 % 8 || ch0 + 32 > ldh) return RAFT_EINVAL;
    if (target && (ldgh % 8 || cgh0 % 8 || cgh0 + 32 > ldgh || !gh || !dw || !db || !loss || !work))
        return RAFT_EINVAL;
    const int nb = grid_for(P);
    HeadArgs a{h, ldh, ch0, w, b, res, pred, target, gh, ldgh, cgh0, work, loss, P, floor_value, clamp};
    hipLaunchKernelGGL(k_head, dim3(nb), dim3(256), 0, S(stream), a);
    if (hipGetLastError() != hipSuccess) return RAFT_EHIP;
    if (!target) return RAFT_OK;
    hipLaunchKernelGGL(k_head_reduce, dim3(34), dim3(256), 0, S(stream), work, nb, dw, db, loss, accumulate);
    return launched();
}

int raft_nn_adam(float* p, const float* g, float* m, float* v, unsigned short* pb, long long n, float lr, float b1,
                 float b2, float eps, int t, float gscale, void* stream)
{
    if (t < 1 || n < 0) return RAFT_EINVAL;
    const float c1 = (float)(1.0 - std::pow((double)b1, (double)t));
    const float c2 = (float)(1.0 - std::pow((double)b2, (double)t));
    hipLaunchKernelGGL(k_adam, dim3(grid_for(n)), dim3(256), 0, S(stream), p, g, m, v, pb, n, lr, b1, b2, eps, c1, c2,
                       gscale);
    return launched();
}

int raft_nn_to_bf16(const float* in, unsigned short* out, long long n, void* stream)
{
    hipLaunchKernelGGL(k_to_bf16, dim3(grid_for(n)), dim3(256), 0, S(stream), in, out, n);
    return launched();
}

int raft_nn_resample(const float* img, float* out, int S_, int Hi, int Wi, int Ho, int Wo, double ax, double bx,
                     double ay, double by, void* stream)
{
    if (S_ < 1 || Hi < 1 || Wi < 1 || Ho < 1 || Wo < 1) return RAFT_EINVAL;
    hipLaunchKernelGGL(k_resample, dim3(grid_for((long long)S_ * Ho * Wo)), dim3(256), 0, S(stream), img, out, S_, Hi,
                       Wi, Ho, Wo, ax, bx, ay, by);
    return launched();
}

} // extern "C"
