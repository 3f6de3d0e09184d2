// backproject.hip -- pixel-driven backprojection for FBP-RT and the eEM update
// (gfx950, wave64).
//
// One workgroup = 256 threads = one 64 x 64 pixel tile of one slice; each
// thread owns 4 x 4 pixels at (tx + 16p, ty + 16q), so a wave covers a 16 x 4
// pixel patch for every (p, q) and its 64 bin lookups of one angle fall into
// at most 16|cos| + 4|sin| + 2 consecutive LDS words (bank-conflict free).
// Angles are processed in chunks of BP_CHUNK: each wave stages the bin
// window that the tile touches for its share of the chunk's angles into LDS
// (<= BP_SLOT words per angle, coalesced global loads), then every thread
// sweeps the chunk:
//     u = fma(X_p, cA, fma(Y_q, cB, c0));  T = (int)u;  acc += win[T]
// in IEEE double for the bin coordinate (bit-identical to the CPU oracle's
// expression), FP32 for the values.  The staged rows already carry
// dtheta_a * angular scale and zeros outside the valid bin range, so the
// inner loop has no range test and no multiply.  Per-chunk partial sums are
// added to the running total (two-level summation over angles).
//
// Reference behaviour restated: BackProjection_SS, fbp.cu:8-76 (+ crop of
// remove_paddR2R, padding.cu:245-268); kernel_backprojectionWithUpdate,
// em_RT_kernel.cu:249-299; kernel_backprojectionOfOnes, em_RT_kernel.cu:303-350.
#include "raft_internal.hpp"

namespace raft {

namespace {

template <typename QT> __device__ __forceinline__ float to_f32(QT v);
template <> __device__ __forceinline__ float to_f32<float>(float v) { return v; }
template <> __device__ __forceinline__ float to_f32<__half>(__half v) { return __half2float(v); }

template <typename QT> __device__ __forceinline__ QT zero_q();
template <> __device__ __forceinline__ float zero_q<float>() { return 0.0f; }
template <> __device__ __forceinline__ __half zero_q<__half>() { return __float2half(0.0f); }

struct BpKArgs {
    const void* q;
    float* out;
    const float* x;
    const float* norm;
    float* xt;
    const double* cA;
    const double* cB;
    double xc, c0;
    int n, A, tlo, W;
    int tiles_x;
    long long norm_stride; // elements between slices of norm (0 = shared)
};

template <typename QT, int MODE>
__global__ __launch_bounds__(256) void k_backproject(BpKArgs a)
{
    __shared__ QT win[BP_CHUNK * BP_SLOT];
    __shared__ double s_cA[BP_CHUNK], s_cB[BP_CHUNK];
    __shared__ int s_base[BP_CHUNK];

    const int tid = threadIdx.x;
    const int tx = tid & 15, ty = tid >> 4;
    const int wave = tid >> 6, lane = tid & 63;
    const int tile = blockIdx.x;
    const int s = blockIdx.y;
    const int i0 = (tile % a.tiles_x) * BP_TILE;
    const int j0 = (tile / a.tiles_x) * BP_TILE;

    double X[4], Y[4];
#pragma unroll
    for (int p = 0; p < 4; ++p) X[p] = (double)(i0 + tx + 16 * p) - a.xc;
#pragma unroll
    for (int q = 0; q < 4; ++q) Y[q] = (double)(j0 + ty + 16 * q) - a.xc;
    const double Xmin = (double)i0 - a.xc, Xmax = (double)(i0 + BP_TILE - 1) - a.xc;
    const double Ymin = (double)j0 - a.xc, Ymax = (double)(j0 + BP_TILE - 1) - a.xc;

    const QT* qs = static_cast<const QT*>(a.q) + (size_t)s * a.A * a.W;
    const int thi_g = a.tlo + a.W - 1;

    float tot[4][4];
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int p = 0; p < 4; ++p) tot[q][p] = 0.0f;

    for (int a0 = 0; a0 < a.A; a0 += BP_CHUNK) {
        const int nA = min(BP_CHUNK, a.A - a0);
        __syncthreads();
        // ---- stage the chunk's bin windows (wave w: angles w, w+4, ...) ----
        for (int al = wave; al < nA; al += 4) {
            const int ang = a0 + al;
            const double cA = a.cA[ang], cB = a.cB[ang];
            const double u00 = fma(Xmin, cA, fma(Ymin, cB, a.c0));
            const double u10 = fma(Xmax, cA, fma(Ymin, cB, a.c0));
            const double u01 = fma(Xmin, cA, fma(Ymax, cB, a.c0));
            const double u11 = fma(Xmax, cA, fma(Ymax, cB, a.c0));
            const double umin = fmin(fmin(u00, u10), fmin(u01, u11));
            const double umax = fmax(fmax(u00, u10), fmax(u01, u11));
            // any pixel's trunc(u) lies in [floor(umin)-1, floor(umax)+2]
            int tl = (int)floor(umin) - 1;
            int th = (int)floor(umax) + 2;
            tl = max(tl, a.tlo);
            th = min(th, thi_g);
            const int width = th - tl + 1;
            const QT* src = qs + (size_t)ang * a.W + (tl - a.tlo);
            QT* dst = win + al * BP_SLOT;
            dst[lane] = lane < width ? src[lane] : zero_q<QT>();
            if (lane + 64 < BP_SLOT) dst[lane + 64] = (lane + 64) < width ? src[lane + 64] : zero_q<QT>();
            if (lane == 0) {
                s_cA[al] = cA;
                s_cB[al] = cB;
                s_base[al] = al * BP_SLOT - tl;
            }
        }
        __syncthreads();
        // ---- sweep ----
        float part[4][4];
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int p = 0; p < 4; ++p) part[q][p] = 0.0f;
#pragma unroll 2
        for (int al = 0; al < nA; ++al) {
            const double cA = s_cA[al], cB = s_cB[al];
            const QT* wb = win + s_base[al];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const double rt = fma(Y[q], cB, a.c0);
#pragma unroll
                for (int p = 0; p < 4; ++p) {
                    const int T = (int)fma(X[p], cA, rt);
                    part[q][p] += to_f32<QT>(wb[T]);
                }
            }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int p = 0; p < 4; ++p) tot[q][p] += part[q][p];
    }

    // ---- epilogue ----
    const size_t so = (size_t)s * a.n * a.n;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int j = j0 + ty + 16 * q;
        if (j >= a.n) continue;
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            const int i = i0 + tx + 16 * p;
            if (i >= a.n) continue;
            const size_t idx = so + (size_t)j * a.n + i;
            if (MODE == 1) {
                const float nv = a.norm[(size_t)s * a.norm_stride + (size_t)j * a.n + i];
                const float xv = a.x[idx] * tot[q][p] / nv;
                a.out[idx] = xv;
                if (a.xt) a.xt[so + (size_t)i * a.n + j] = xv;
            } else {
                a.out[idx] = tot[q][p];
            }
        }
    }
}

// rows for B(1): q[a][w] = wa[a] where 0 <= tlo + w < nbins, else 0
__global__ void k_fill_rows(float* q, const float* wa, int S, int A, int W, int tlo, int nbins)
{
    const size_t total = (size_t)S * A * W;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const int w = (int)(e % W);
        const int ang = (int)((e / W) % A);
        const int T = tlo + w;
        q[e] = (T >= 0 && T < nbins) ? wa[ang] : 0.0f;
    }
}

// out[s][c][r] = in[s][r][c], 32 x 32 LDS tiles
__global__ __launch_bounds__(256) void k_transpose(const float* in, float* out, int rows, int cols)
{
    __shared__ float t[32][33];
    const int s = blockIdx.z;
    const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5; // 32 x 8
    const float* src = in + (size_t)s * rows * cols;
    float* dst = out + (size_t)s * rows * cols;
#pragma unroll
    for (int k = 0; k < 32; k += 8) {
        const int r = r0 + ty + k, c = c0 + tx;
        if (r < rows && c < cols) t[ty + k][tx] = src[(size_t)r * cols + c];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 32; k += 8) {
        const int c = c0 + ty + k, r = r0 + tx;
        if (r < rows && c < cols) dst[(size_t)c * rows + r] = t[tx][ty + k];
    }
}

} // namespace

void launch_backproject(const BpArgs& b, int store, hipStream_t st)
{
    if (b.S <= 0 || b.n <= 0) return;
    BpKArgs k;
    k.q = b.q; k.out = b.out; k.x = b.x; k.norm = b.norm; k.xt = b.xt;
    k.cA = b.cA; k.cB = b.cB; k.xc = b.xc; k.c0 = b.c0;
    k.n = b.n; k.A = b.A; k.tlo = b.tlo; k.W = b.W;
    k.tiles_x = (b.n + BP_TILE - 1) / BP_TILE;
    k.norm_stride = 0;
    dim3 grid(k.tiles_x * k.tiles_x, b.S);
    dim3 block(256);
    if (b.mode == 1) {
        if (store == RAFT_STORE_F16) hipLaunchKernelGGL((k_backproject<__half, 1>), grid, block, 0, st, k);
        else hipLaunchKernelGGL((k_backproject<float, 1>), grid, block, 0, st, k);
    } else {
        if (store == RAFT_STORE_F16) hipLaunchKernelGGL((k_backproject<__half, 0>), grid, block, 0, st, k);
        else hipLaunchKernelGGL((k_backproject<float, 0>), grid, block, 0, st, k);
    }
    RAFT_HIP(hipGetLastError());
}

void launch_fill_rows(float* q, const float* wa, int S, int A, int W, int tlo, int nbins, hipStream_t st)
{
    const size_t total = (size_t)S * A * W;
    if (!total) return;
    int blocks = (int)std::min<size_t>((total + 255) / 256, 65536);
    hipLaunchKernelGGL(k_fill_rows, dim3(blocks), dim3(256), 0, st, q, wa, S, A, W, tlo, nbins);
    RAFT_HIP(hipGetLastError());
}

void launch_transpose(const float* in, float* out, int S, int rows, int cols, hipStream_t st)
{
    if (S <= 0 || rows <= 0 || cols <= 0) return;
    dim3 grid((cols + 31) / 32, (rows + 31) / 32, S);
    hipLaunchKernelGGL(k_transpose, grid, dim3(256), 0, st, in, out, rows, cols);
    RAFT_HIP(hipGetLastError());
}

} // namespace raft
