// backproject.hip -- pixel-driven backprojection for FBP-RT and the eEM update
// (gfx950, wave64), plus the layout kernels around it.
//
// Slice interleaving.  Every slice of a call shares the geometry, so the
// kernel computes each bin index once and gathers the same bin of a group of
// slices with one 16-byte LDS read: staged rows are stored slice-interleaved,
// [group][angle][bin][SB] with SB = 4 (FP32 rows) or 8 (FP16 rows).
//
// One workgroup = 256 threads = one 64 x 32 pixel tile of one slice group;
// each thread owns 4 x 2 pixels at (tx + 16p, ty + 16q), so a wave covers a
// 16 x 4 pixel patch for every (p, q) and its 64 bin lookups of one angle fall
// into at most 16|cos| + 4|sin| + 2 consecutive 16-byte LDS entries.  Angles
// are processed in chunks of BP_CHUNK: the first threads compute the bin
// window the tile touches for one angle each (from the tile corners), then
// all threads copy the windows (<= BP_SLOT entries per angle, coalesced
// 16-byte loads) into LDS, then every thread sweeps the chunk:
//     u = fma(X_p, cA, fma(Y_q, cB, c0));  T = (int)u;  acc[0..SB) += win[T]
// in IEEE double for the bin coordinate (bit-identical to the CPU oracle's
// expression) and FP32 for the values.  The FP32-row kernels also come in a
// "pair" form (PAIR = 2): one workgroup takes two 4-slice groups, stages both
// groups' windows side by side in LDS and reads both entries of a bin (two
// 16-byte LDS reads, the second at an immediate offset), so every FP64 bin
// index and LDS address serves 8 slices.  Each slice's sum is formed in the
// same order in both forms, so they agree bit for bit.  The staged rows already carry
// dtheta_a * angular scale and zeros outside the valid bin range, so the
// inner loop has no range test and no multiply.  Per-chunk partial sums are
// added to the running total (two-level summation over angles).
//
// Reference behaviour restated: BackProjection_SS, fbp.cu:8-76 (+ crop of
// remove_paddR2R, padding.cu:245-268); kernel_backprojectionWithUpdate,
// em_RT_kernel.cu:249-299; kernel_backprojectionOfOnes, em_RT_kernel.cu:303-350.
#include "raft_internal.hpp"

#include <algorithm>
#include <cstring>

namespace raft {

namespace {

struct BpKArgs {
    const uint4* q;     // [G][A][W] entries of SB values
    float* out;         // MODE 0: [S][n][n]; MODE 1: interleaved [G][n][n][4]
    const float* x;     // MODE 1: interleaved current image [G][n][n][4] (may alias out)
    const float* norm;  // MODE 1: backones [n][n] (shared by all slices), or per slice interleaved if norm_il
    float* xt;          // MODE 1: interleaved transposed image [G][n][n][4] (or null)
    const double* cA;
    const double* cB;
    double xc, c0;
    int n, A, tlo, W, S;
    int tiles_x;
    int norm_il;
};

// accumulate one 16-byte staged entry into SB running sums
template <typename QT> struct Acc;
template <> struct Acc<float> {
    static constexpr int SB = 4;
    __device__ static __forceinline__ void add(float* acc, const uint4& e)
    {
        acc[0] += __uint_as_float(e.x);
        acc[1] += __uint_as_float(e.y);
        acc[2] += __uint_as_float(e.z);
        acc[3] += __uint_as_float(e.w);
    }
};
// FP16 entries: acc += (float)half in one v_fma_mix_f32 (f16 source selected from
// the low or high half of the dword, times 1.0, plus the f32 accumulator)
__device__ __forceinline__ float mix_lo(unsigned v, float acc)
{
    float r;
    asm("v_fma_mix_f32 %0, %1, 1.0, %2 op_sel_hi:[1,0,0]" : "=v"(r) : "v"(v), "v"(acc));
    return r;
}
__device__ __forceinline__ float mix_hi(unsigned v, float acc)
{
    float r;
    asm("v_fma_mix_f32 %0, %1, 1.0, %2 op_sel:[1,0,0] op_sel_hi:[1,0,0]" : "=v"(r) : "v"(v), "v"(acc));
    return r;
}
template <> struct Acc<__half> {
    static constexpr int SB = 8;
    __device__ static __forceinline__ void add(float* acc, const uint4& e)
    {
        acc[0] = mix_lo(e.x, acc[0]);
        acc[1] = mix_hi(e.x, acc[1]);
        acc[2] = mix_lo(e.y, acc[2]);
        acc[3] = mix_hi(e.y, acc[3]);
        acc[4] = mix_lo(e.z, acc[4]);
        acc[5] = mix_hi(e.z, acc[5]);
        acc[6] = mix_lo(e.w, acc[6]);
        acc[7] = mix_hi(e.w, acc[7]);
    }
};

// LDS of one workgroup: the staged windows of CHUNK angles and their bin ranges
template <int PAIR> constexpr int bp_slot() { return PAIR == 2 ? BP_SLOT_PAIR : BP_SLOT; }
template <int CHUNK, int PAIR> struct BpSmem {
    uint4 win[PAIR * CHUNK * bp_slot<PAIR>()]; // PAIR halves of CHUNK * slot entries
    int tl[CHUNK], w[CHUNK], boff[CHUNK];
};

// One 64 x 32 pixel tile of slice groups PAIR g .. PAIR g + PAIR - 1 (a
// missing second group of a pair reads zeros and stores nothing).  Starts
// with a barrier (the LDS may still be read by the previous tile of a
// persistent workgroup) and touches LDS only through sm.
template <typename QT, int MODE, int CHUNK, int PAIR, int SPF>
__device__ __forceinline__ void bp_tile(const BpKArgs& a, int tile, int gp, BpSmem<CHUNK, PAIR>& sm)
{
    static_assert(PAIR == 1 || (PAIR == 2 && Acc<QT>::SB == 4), "pairs are formed of FP32-row groups");
    constexpr int SB1 = Acc<QT>::SB;   // slices per 16-byte entry
    constexpr int SB = SB1 * PAIR;     // slices per pixel of this workgroup
    constexpr int SLOT = bp_slot<PAIR>();  // LDS entries per staged angle window
    constexpr int HALF = CHUNK * SLOT * (int)sizeof(uint4); // byte offset of the second half
    const int g = gp * PAIR;           // first slice group
    const bool second = PAIR == 2 && (g + 1) * SB1 < a.S; // the pair's second group exists
    constexpr int TX = BP_TILE_X, TY = BP_TILE_Y;
    uint4* win = sm.win;
    int* s_tl = sm.tl;
    int* s_w = sm.w;
    int* s_boff = sm.boff;

    const int tid = threadIdx.x;
    const int tx = tid & 15, ty = tid >> 4;
    const int i0 = (tile % a.tiles_x) * TX;
    const int j0 = (tile / a.tiles_x) * TY;

    double X[4], Y[2];
#pragma unroll
    for (int p = 0; p < 4; ++p) X[p] = (double)(i0 + tx + 16 * p) - a.xc;
#pragma unroll
    for (int q = 0; q < 2; ++q) Y[q] = (double)(j0 + ty + 16 * q) - a.xc;

    const uint4* qg = a.q + (size_t)g * a.A * a.W;
    const int thi_g = a.tlo + a.W - 1;

    float tot[2][4][SB];
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
        for (int p = 0; p < 4; ++p)
#pragma unroll
            for (int s = 0; s < SB; ++s) tot[q][p][s] = 0.0f;

    for (int a0 = 0; a0 < a.A; a0 += CHUNK) {
        const int nA = __builtin_amdgcn_readfirstlane(min(CHUNK, a.A - a0)); // uniform: scalar loop counter
        __syncthreads();
        // ---- per-angle bin windows of this tile (one thread per angle) ----
        if (tid < nA) {
            const int ang = a0 + tid;
            const double cA = a.cA[ang], cB = a.cB[ang];
            const double Xmin = (double)i0 - a.xc, Xmax = (double)(i0 + TX - 1) - a.xc;
            const double Ymin = (double)j0 - a.xc, Ymax = (double)(j0 + TY - 1) - a.xc;
            const double u00 = fma(Xmin, cA, fma(Ymin, cB, a.c0));
            const double u10 = fma(Xmax, cA, fma(Ymin, cB, a.c0));
            const double u01 = fma(Xmin, cA, fma(Ymax, cB, a.c0));
            const double u11 = fma(Xmax, cA, fma(Ymax, cB, a.c0));
            const double umin = fmin(fmin(u00, u10), fmin(u01, u11));
            const double umax = fmax(fmax(u00, u10), fmax(u01, u11));
            // any pixel's trunc(u) lies in [floor(umin)-1, floor(umax)+2]
            const int tl = max((int)floor(umin) - 1, a.tlo);
            const int th = min((int)floor(umax) + 2, thi_g);
            s_tl[tid] = tl;
            s_w[tid] = th - tl + 1;
            s_boff[tid] = (tid * SLOT - tl) * (int)sizeof(uint4); // wraps for large tl: used modulo 2^32

        }
        __syncthreads();
        // ---- stage the windows: 16-byte coalesced loads ----
        for (int e = tid; e < nA * SLOT; e += 256) {
            const int al = e / SLOT;
            const int w = e - al * SLOT;
            uint4 v = make_uint4(0u, 0u, 0u, 0u), v2 = make_uint4(0u, 0u, 0u, 0u);
            const size_t src = (size_t)(a0 + al) * a.W + (s_tl[al] - a.tlo) + w;
            if (w < s_w[al]) {
                v = qg[src];
                if (second) v2 = qg[src + (size_t)a.A * a.W];
            }
            win[e] = v;
            if (PAIR == 2) win[CHUNK * SLOT + e] = v2;
        }
        __syncthreads();
        // ---- sweep ----
        // FP32 rows: partial sums over groups of BP_GROUP angles added to the
        // totals (two-level summation, for the 1e-5 bar); the grouping does not
        // depend on CHUNK (a multiple of BP_GROUP), so every CHUNK variant
        // gives bit-identical sums.  FP16 rows sum into the totals directly
        // (their error is the FP16 storage; 64 fewer registers, occupancy 3).
        constexpr bool TWO_LEVEL = SB1 == 4;
        static_assert(CHUNK % BP_GROUP == 0, "the angle chunk must hold whole summation groups");
        const char* wbase = reinterpret_cast<const char*>(win);
#pragma unroll 1
        for (int ag = 0; ag < nA; ag += BP_GROUP) {
            const int ng = __builtin_amdgcn_readfirstlane(min(BP_GROUP, nA - ag));
            float part[2][4][TWO_LEVEL ? SB : 1];
#pragma unroll
            for (int q = 0; q < 2; ++q)
#pragma unroll
                for (int p = 0; p < 4; ++p)
#pragma unroll
                    for (int s = 0; s < (TWO_LEVEL ? SB : 1); ++s) part[q][p][s] = 0.0f;
            // One angle of the sweep: 8 FP64 bin indices, 8 (pair: 16) gathers.
            // The angles' FP64 coefficients are the same for every lane, so they
            // come through scalar loads from the tables in memory (no LDS
            // instruction: the gathers are what bounds the sweep); the window
            // offsets stay in LDS (they depend on the tile).
            auto sweep_angle = [&](int al, double cA, double cB) __attribute__((always_inline)) {
                // byte offset of bin 0 of this angle's window: entry (al, T - tl)
                const unsigned boff = (unsigned)s_boff[al];
                int T[2][4];
#pragma unroll
                for (int q = 0; q < 2; ++q) {
                    const double rt = fma(Y[q], cB, a.c0);
#pragma unroll
                    for (int p = 0; p < 4; ++p) T[q][p] = (int)fma(X[p], cA, rt);
                }
                uint4 v[2][4][PAIR];
#pragma unroll
                for (int q = 0; q < 2; ++q)
#pragma unroll
                    for (int p = 0; p < 4; ++p) {
                        const char* e = wbase + (((unsigned)T[q][p] << 4) + boff);
#pragma unroll
                        for (int h = 0; h < PAIR; ++h) v[q][p][h] = *reinterpret_cast<const uint4*>(e + h * HALF);
                    }
#pragma unroll
                for (int q = 0; q < 2; ++q)
#pragma unroll
                    for (int p = 0; p < 4; ++p)
#pragma unroll
                        for (int h = 0; h < PAIR; ++h)
                            Acc<QT>::add((TWO_LEVEL ? part[q][p] : tot[q][p]) + h * SB1, v[q][p][h]);
            };
            const int aend = ag + ng;
            if constexpr (SPF == 0) {
                // coefficients loaded where they are used, 4 angles per trip
#pragma unroll 4
                for (int al = ag; al < aend; ++al) sweep_angle(al, a.cA[a0 + al], a.cB[a0 + al]);
            } else {
                // scalar-prefetch form: the next trip's 4 coefficients are loaded
                // while this trip's angles are swept (indices clamped into the
                // chunk, so every load is in bounds).  Fewer registers (128 at
                // 16 angles per chunk: four workgroups per CU).
                double nxA[4], nxB[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int i = a0 + min(ag + u, nA - 1);
                    nxA[u] = a.cA[i];
                    nxB[u] = a.cB[i];
                }
#pragma unroll 1
                for (int al4 = ag; al4 < aend; al4 += 4) {
                    double cA4[4], cB4[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        cA4[u] = nxA[u];
                        cB4[u] = nxB[u];
                        const int i = a0 + min(al4 + 4 + u, nA - 1);
                        nxA[u] = a.cA[i];
                        nxB[u] = a.cB[i];
                    }
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        if (al4 + u >= aend) break;
                        sweep_angle(al4 + u, cA4[u], cB4[u]);
                    }
                }
            }
            if constexpr (TWO_LEVEL) {
#pragma unroll
                for (int q = 0; q < 2; ++q)
#pragma unroll
                    for (int p = 0; p < 4; ++p)
#pragma unroll
                        for (int s = 0; s < SB; ++s) tot[q][p][s] += part[q][p][s];
            }
        }
    }

    // ---- epilogue ----
    const size_t nn = (size_t)a.n * a.n;
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const int j = j0 + ty + 16 * q;
        if (j >= a.n) continue;
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            const int i = i0 + tx + 16 * p;
            if (i >= a.n) continue;
            if (MODE == 1) {
                // eEM update on the interleaved image (4 slices per group here)
#pragma unroll
                for (int h = 0; h < PAIR; ++h) {
                    const int gh = g + h;
                    if (h > 0 && !second) break;
                    const size_t pix = ((size_t)gh * nn + (size_t)j * a.n + i) * 4;
                    float4 nv;
                    if (a.norm_il) {
                        nv = *reinterpret_cast<const float4*>(a.norm + pix);
                    } else {
                        const float n1 = a.norm[(size_t)j * a.n + i];
                        nv = make_float4(n1, n1, n1, n1);
                    }
                    // slices at or past S (the padding of the last group, or slices
                    // whose iterations are done in a scheduled run) keep their values
                    float4 xv = *reinterpret_cast<const float4*>(a.x + pix);
                    const int s0 = gh * 4;
                    const float* t4 = tot[q][p] + 4 * h;
                    xv.x = s0 + 0 < a.S ? xv.x * t4[0] / nv.x : xv.x;
                    xv.y = s0 + 1 < a.S ? xv.y * t4[1] / nv.y : xv.y;
                    xv.z = s0 + 2 < a.S ? xv.z * t4[2] / nv.z : xv.z;
                    xv.w = s0 + 3 < a.S ? xv.w * t4[3] / nv.w : xv.w;
                    *reinterpret_cast<float4*>(a.out + pix) = xv;
                    if (a.xt) *reinterpret_cast<float4*>(a.xt + ((size_t)gh * nn + (size_t)i * a.n + j) * 4) = xv;
                }
            } else {
#pragma unroll
                for (int s = 0; s < SB; ++s) {
                    const int sl = g * SB1 + s;
                    if (sl < a.S) a.out[(size_t)sl * nn + (size_t)j * a.n + i] = tot[q][p][s];
                }
            }
        }
    }
}

// the pair form is held to 256 VGPRs: two waves per SIMD, two workgroups per CU
template <typename QT, int MODE, int CHUNK, int PAIR, int SPF>
__global__ __launch_bounds__(256, PAIR) void k_backproject(BpKArgs a)
{
    __shared__ BpSmem<CHUNK, PAIR> sm;
    bp_tile<QT, MODE, CHUNK, PAIR, SPF>(a, blockIdx.x, blockIdx.y, sm);
}

// Persistent work-queue form: a resident grid takes items from a device
// counter; item i of launch segment b (items [first[b], first[b + 1])) is
// tile (i - first[b]) % tiles of slice group (or pair of groups)
// (i - first[b]) / tiles of the volume described by descs[b].  The host orders the segments by cost per
// item (angles), largest first.  The last workgroup to leave clears the
// counters for the next launch.
template <typename QT, int MODE, int CHUNK, int PAIR, int SPF>
__global__ __launch_bounds__(256, PAIR) void k_backproject_queue(const BpKArgs* __restrict__ descs, BpQueue qu)
{
    __shared__ BpSmem<CHUNK, PAIR> sm;
    __shared__ int s_item;
    for (;;) {
        __syncthreads(); // every thread has read the previous item
        if (threadIdx.x == 0)
            s_item = __hip_atomic_fetch_add(qu.counter, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        const int item = __builtin_amdgcn_readfirstlane(s_item);
        if (item >= qu.first[qu.nseg]) break;
        int b = 0;
        while (b + 1 < qu.nseg && item >= qu.first[b + 1]) ++b;
        b = __builtin_amdgcn_readfirstlane(b);
        const int local = item - qu.first[b];
        const int tiles = qu.tiles[b];
        bp_tile<QT, MODE, CHUNK, PAIR, SPF>(descs[b], local % tiles, local / tiles, sm);
    }
    if (threadIdx.x == 0) {
        __threadfence();
        const int d = __hip_atomic_fetch_add(qu.counter + 1, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (d == (int)gridDim.x - 1) {
            __hip_atomic_store(qu.counter, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(qu.counter + 1, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// rows for B(1), interleaved by 4: q[a][w][0..4) = wa[a] where 0 <= tlo + w < nbins
__global__ void k_fill_rows(float4* q, const float* wa, int A, int W, int tlo, int nbins)
{
    const size_t total = (size_t)A * W;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const int w = (int)(e % W);
        const int ang = (int)(e / W);
        const int T = tlo + w;
        const float v = (T >= 0 && T < nbins) ? wa[ang] : 0.0f;
        q[e] = make_float4(v, v, v, v);
    }
}

// rows from planar data [S][A][nbins]: q[g][a][w][s] = data[4g+s][a][tlo+w] * wa[a]
__global__ void k_rows_from_planar(float4* q, const float* data, const float* wa, int S, int A, int W, int tlo,
                                   int nbins)
{
    const int G = (S + 3) / 4;
    const size_t total = (size_t)G * A * W;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const int w = (int)(e % W);
        const int ang = (int)((e / W) % A);
        const int g = (int)(e / ((size_t)W * A));
        const int T = tlo + w;
        float v[4];
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const int sl = 4 * g + s;
            v[s] = (T >= 0 && T < nbins && sl < S) ? data[((size_t)sl * A + ang) * nbins + T] * wa[ang] : 0.0f;
        }
        q[e] = make_float4(v[0], v[1], v[2], v[3]);
    }
}

// planar [S][rows][cols] <-> interleaved [G][rows][cols][4] (optionally transposed:
// interleaved [G][cols][rows][4]); missing slices of the last group read as 0.
// 32 x 32 tiles through LDS.
template <int TO_IL, int TRANSPOSE>
__global__ __launch_bounds__(256) void k_layout(const float* in, float* out, int S, int rows, int cols)
{
    __shared__ float4 t[32][33];
    const int g = blockIdx.z;
    const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5; // 32 x 8
    const size_t plane = (size_t)rows * cols;
    if (TO_IL) {
#pragma unroll
        for (int k = 0; k < 32; k += 8) {
            const int r = r0 + ty + k, c = c0 + tx;
            float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
            if (r < rows && c < cols) {
                const size_t o = (size_t)r * cols + c;
                const int s = g * 4;
                v.x = (s + 0 < S) ? in[(size_t)(s + 0) * plane + o] : 0.f;
                v.y = (s + 1 < S) ? in[(size_t)(s + 1) * plane + o] : 0.f;
                v.z = (s + 2 < S) ? in[(size_t)(s + 2) * plane + o] : 0.f;
                v.w = (s + 3 < S) ? in[(size_t)(s + 3) * plane + o] : 0.f;
            }
            t[ty + k][tx] = v;
        }
        __syncthreads();
        float4* dst = reinterpret_cast<float4*>(out) + (size_t)g * plane;
#pragma unroll
        for (int k = 0; k < 32; k += 8) {
            if (TRANSPOSE) {
                const int c = c0 + ty + k, r = r0 + tx;
                if (r < rows && c < cols) dst[(size_t)c * rows + r] = t[tx][ty + k];
            } else {
                const int r = r0 + ty + k, c = c0 + tx;
                if (r < rows && c < cols) dst[(size_t)r * cols + c] = t[ty + k][tx];
            }
        }
    } else {
        // interleaved [G][rows][cols][4] -> planar [S][rows][cols] (no transpose)
        const float4* src = reinterpret_cast<const float4*>(in) + (size_t)g * plane;
#pragma unroll
        for (int k = 0; k < 32; k += 8) {
            const int r = r0 + ty + k, c = c0 + tx;
            if (r < rows && c < cols) {
                const float4 v = src[(size_t)r * cols + c];
                const size_t o = (size_t)r * cols + c;
                const int s = g * 4;
                if (s + 0 < S) out[(size_t)(s + 0) * plane + o] = v.x;
                if (s + 1 < S) out[(size_t)(s + 1) * plane + o] = v.y;
                if (s + 2 < S) out[(size_t)(s + 2) * plane + o] = v.z;
                if (s + 3 < S) out[(size_t)(s + 3) * plane + o] = v.w;
            }
        }
    }
}

// LDS read-rate probe: every lane reads ITERS conflict-free 16-byte entries
// (a wave reads 64 consecutive entries), two integer adds per read keep the
// loads live; every index is masked to the 4096-entry array
constexpr int LDSP_ITERS = 4096;
constexpr int LDSP_TILE = 4096;
__global__ __launch_bounds__(256) void k_lds_probe(unsigned* out)
{
    __shared__ uint4 t[LDSP_TILE];
    for (int i = threadIdx.x; i < LDSP_TILE; i += 256) t[i] = make_uint4(i, i + 1, i + 2, i + 3);
    __syncthreads();
    unsigned acc = 0, acc2 = 0;
    unsigned idx = threadIdx.x;
#pragma unroll 8
    for (int it = 0; it < LDSP_ITERS; ++it) {
        idx = (idx + 64) & (LDSP_TILE - 1);
        const uint4 v = t[idx];
        acc = acc + v.x + v.y;
        acc2 = acc2 + v.z + v.w;
    }
    out[blockIdx.x * 256 + threadIdx.x] = acc ^ acc2;
}

} // namespace

double lds_read_probe(int* cus_out, int* clock_khz)
{
    int dev = 0, cus = 0, clk = 0;
    RAFT_HIP(hipGetDevice(&dev));
    RAFT_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    RAFT_HIP(hipDeviceGetAttribute(&clk, hipDeviceAttributeClockRate, dev));
    const int blocks = cus * 8;
    DevBuf out((size_t)blocks * 256 * sizeof(unsigned));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    RAFT_HIP(hipEventCreate(&e0));
    RAFT_HIP(hipEventCreate(&e1));
    float best = 0.0f;
    for (int rep = 0; rep < 3; ++rep) {
        RAFT_HIP(hipEventRecord(e0, nullptr));
        hipLaunchKernelGGL(k_lds_probe, dim3(blocks), dim3(256), 0, nullptr, out.as<unsigned>());
        RAFT_HIP(hipGetLastError());
        RAFT_HIP(hipEventRecord(e1, nullptr));
        RAFT_HIP(hipEventSynchronize(e1));
        float ms = 0.0f;
        RAFT_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (rep > 0 && (best == 0.0f || ms < best)) best = ms;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (cus_out) *cus_out = cus;
    if (clock_khz) *clock_khz = clk;
    const double bytes = (double)blocks * 256 * LDSP_ITERS * 16;
    return bytes / ((double)best * 1e-3);
}

// the staged rows a launch over b.S slices reads must lie in the buffer behind q
static void bp_check_rows(const BpArgs& b, int store)
{
    const int sb = (store == RAFT_STORE_F16 && b.mode == 0) ? 8 : 4;
    const size_t need = (size_t)((b.S + sb - 1) / sb) * (size_t)b.A * (size_t)b.W * 16;
    if (b.q_bytes && need > b.q_bytes)
        fail(RAFT_EINVAL, "internal: backprojection would read " + std::to_string(need) + " bytes of staged rows from a " +
                              std::to_string(b.q_bytes) + "-byte buffer");
}

static BpKArgs bp_kargs(const BpArgs& b)
{
    BpKArgs k;
    k.q = static_cast<const uint4*>(b.q);
    k.out = b.out; k.x = b.x; k.norm = b.norm; k.xt = b.xt; k.norm_il = b.norm_il;
    k.cA = b.cA; k.cB = b.cB; k.xc = b.xc; k.c0 = b.c0;
    k.n = b.n; k.A = b.A; k.tlo = b.tlo; k.W = b.W; k.S = b.S;
    k.tiles_x = (b.n + BP_TILE_X - 1) / BP_TILE_X;
    return k;
}

int bp_tiles(int n) { return ((n + BP_TILE_X - 1) / BP_TILE_X) * ((n + BP_TILE_Y - 1) / BP_TILE_Y); }

size_t bp_desc_size() { return sizeof(BpKArgs); }

void bp_desc_fill(const BpArgs& b, int store, void* dst)
{
    bp_check_rows(b, store);
    const BpKArgs k = bp_kargs(b);
    std::memcpy(dst, &k, sizeof k);
}

// kernel variant: (FP32 rows, store) x (FBP sums, eEM update) x angles per LDS
// round x (one group or a pair of FP32-row groups per workgroup)
template <int CHUNK, int PAIR, int SPF = 0> static const void* bp_kernel(int store, int mode, bool queue)
{
    if (mode == 1)
        return queue ? reinterpret_cast<const void*>(&k_backproject_queue<float, 1, CHUNK, PAIR, SPF>)
                     : reinterpret_cast<const void*>(&k_backproject<float, 1, CHUNK, PAIR, SPF>);
    if (store == RAFT_STORE_F16) {
        if (PAIR != 1) fail(RAFT_EINVAL, "internal: the pair form takes FP32 rows");
        return queue ? reinterpret_cast<const void*>(&k_backproject_queue<__half, 0, CHUNK, 1, SPF>)
                     : reinterpret_cast<const void*>(&k_backproject<__half, 0, CHUNK, 1, SPF>);
    }
    return queue ? reinterpret_cast<const void*>(&k_backproject_queue<float, 0, CHUNK, PAIR, SPF>)
                 : reinterpret_cast<const void*>(&k_backproject<float, 0, CHUNK, PAIR, SPF>);
}

static const void* bp_kernel_any(int store, int mode, int variant, bool queue)
{
    if (!bp_variant_supported(variant, mode == 1 ? RAFT_STORE_F32 : store))
        fail(RAFT_EINVAL, "internal: unsupported backprojection variant " + std::to_string(variant));
    switch (variant == 0 ? BP_CHUNK : variant) {
    case 16: return bp_kernel<16, 1>(store, mode, queue);
    case 32: return bp_kernel<32, 1>(store, mode, queue);
    case 48: return bp_kernel<48, 1>(store, mode, queue);
    case BP_SPF | 16: return bp_kernel<16, 1, 1>(store, mode, queue);
    case BP_SPF | 32: return bp_kernel<32, 1, 1>(store, mode, queue);
    case BP_PAIR | 16: return bp_kernel<16, 2>(store, mode, queue);
    default: return bp_kernel<32, 2>(store, mode, queue); // BP_PAIR | 32
    }
}

bool bp_variant_supported(int variant, int store)
{
    switch (variant) {
    case 0:
    case 16:
    case 32:
    case 48: return true;
    case BP_SPF | 16:
    case BP_SPF | 32: return true;
    case BP_PAIR | 16:
    case BP_PAIR | 32: return store == RAFT_STORE_F32;
    default: return false;
    }
}

int bp_slices_per_item(int store, int mode, int variant)
{
    const int sb = (store == RAFT_STORE_F16 && mode == 0) ? BP_SLICES_F16 : BP_SLICES_F32;
    return (variant & BP_PAIR) ? 2 * sb : sb;
}

void launch_backproject(const BpArgs& b, int store, hipStream_t st)
{
    if (b.S <= 0 || b.n <= 0) return;
    bp_check_rows(b, store);
    BpKArgs k = bp_kargs(b);
    const int per = bp_slices_per_item(store, b.mode, b.chunk);
    const int groups = (b.S + per - 1) / per;
    if (groups > 65535) fail(RAFT_EUNSUPPORTED, "too many slice groups for one backprojection launch");
    const void* fn = bp_kernel_any(store, b.mode, b.chunk, false);
    void* args[] = {&k};
    RAFT_HIP(hipLaunchKernel(fn, dim3(bp_tiles(b.n), groups), dim3(256), args, 0, st));
}

int bp_queue_grid(int store, int mode, int chunk)
{
    const void* fn = bp_kernel_any(store, mode, chunk, true);
    int dev = 0, cus = 0, per = 0;
    RAFT_HIP(hipGetDevice(&dev));
    RAFT_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    RAFT_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fn, 256, 0));
    return std::max(1, cus * std::max(1, per));
}

void launch_backproject_queue(const void* d_descs, const BpQueue& q, int store, int mode, int chunk, int grid,
                              hipStream_t st)
{
    if (q.nseg <= 0 || q.first[q.nseg] <= 0) return;
    if (q.nseg > QUEUE_MAXSEG) fail(RAFT_EINVAL, "internal: too many queue segments");
    const void* fn = bp_kernel_any(store, mode, chunk, true);
    BpQueue qq = q;
    const BpKArgs* descs = static_cast<const BpKArgs*>(d_descs);
    void* args[] = {&descs, &qq};
    RAFT_HIP(hipLaunchKernel(fn, dim3(std::max(1, std::min(grid, q.first[q.nseg]))), dim3(256), args, 0, st));
}

void launch_rows_from_planar(float* q, const float* data, const float* wa, int S, int A, int W, int tlo, int nbins,
                             hipStream_t st)
{
    const size_t total = (size_t)((S + 3) / 4) * A * W;
    if (!total) return;
    int blocks = (int)std::min<size_t>((total + 255) / 256, 65536);
    hipLaunchKernelGGL(k_rows_from_planar, dim3(blocks), dim3(256), 0, st, reinterpret_cast<float4*>(q), data, wa, S,
                       A, W, tlo, nbins);
    RAFT_HIP(hipGetLastError());
}

void launch_fill_rows(float* q, const float* wa, int A, int W, int tlo, int nbins, hipStream_t st)
{
    const size_t total = (size_t)A * W;
    if (!total) return;
    int blocks = (int)std::min<size_t>((total + 255) / 256, 65536);
    hipLaunchKernelGGL(k_fill_rows, dim3(blocks), dim3(256), 0, st, reinterpret_cast<float4*>(q), wa, A, W, tlo,
                       nbins);
    RAFT_HIP(hipGetLastError());
}

void launch_to_interleaved(const float* in, float* out, int S, int rows, int cols, bool transpose, hipStream_t st)
{
    if (S <= 0 || rows <= 0 || cols <= 0) return;
    const int G = (S + 3) / 4;
    dim3 grid((cols + 31) / 32, (rows + 31) / 32, G);
    if (transpose) hipLaunchKernelGGL((k_layout<1, 1>), grid, dim3(256), 0, st, in, out, S, rows, cols);
    else hipLaunchKernelGGL((k_layout<1, 0>), grid, dim3(256), 0, st, in, out, S, rows, cols);
    RAFT_HIP(hipGetLastError());
}

void launch_from_interleaved(const float* in, float* out, int S, int rows, int cols, hipStream_t st)
{
    if (S <= 0 || rows <= 0 || cols <= 0) return;
    const int G = (S + 3) / 4;
    dim3 grid((cols + 31) / 32, (rows + 31) / 32, G);
    hipLaunchKernelGGL((k_layout<0, 0>), grid, dim3(256), 0, st, in, out, S, rows, cols);
    RAFT_HIP(hipGetLastError());
}

} // namespace raft
