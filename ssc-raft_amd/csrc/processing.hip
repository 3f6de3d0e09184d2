// processing.hip -- the step before the reconstruction path: flat/dark
// background correction with optional -log (background_correction.cu:13-45,
// opt.cu:196-214), in place on [nslices][nangles][nrays] frames.
#include "raft_internal.hpp"

namespace raft {

namespace {

// v = max(D - dark, 0.5) / max(F - dark, 0.5), F = flat row 0, or with two or
// more flats the blend of rows 0 and 1 by angle / nangles; then -log(v) if asked.
__device__ __forceinline__ float background_one(float D, float dk, float ft, int is_log)
{
    const float T = D - dk;
    const float Q = ft - dk;
    float v = fmaxf(T, 0.5f) / fmaxf(Q, 0.5f);
    if (is_log) v = -logf(v);
    return v;
}

// one block per frame row (slice z, angle y); 16-byte accesses when R % 4 == 0
__global__ __launch_bounds__(256) void k_background(float* data, const float* flat, const float* dark, int R, int A,
                                                    int nflats, int is_log)
{
    const long long row = blockIdx.x;
    const int z = (int)(row / A), y = (int)(row % A);
    float* d = data + row * R;
    const float* dk = dark + (size_t)z * R;
    const float* fb = flat + (size_t)z * R * nflats;
    const float* fa = fb + R;
    const float interp = (float)y / (float)A;
    auto flat_at = [&](int x) {
        float ft = fb[x];
        if (nflats > 1) ft = fb[x] * (1.0f - interp) + interp * fa[x];
        return ft;
    };
    if ((R & 3) == 0) {
        float4* d4 = reinterpret_cast<float4*>(d);
        const float4* k4 = reinterpret_cast<const float4*>(dk);
        const float4* b4 = reinterpret_cast<const float4*>(fb);
        const float4* a4 = reinterpret_cast<const float4*>(nflats > 1 ? fa : fb);
        for (int i = threadIdx.x; i < R / 4; i += blockDim.x) {
            float4 v = d4[i];
            const float4 k = k4[i], f0 = b4[i], f1 = a4[i];
            float4 ft = f0;
            if (nflats > 1) {
                ft.x = f0.x * (1.0f - interp) + interp * f1.x;
                ft.y = f0.y * (1.0f - interp) + interp * f1.y;
                ft.z = f0.z * (1.0f - interp) + interp * f1.z;
                ft.w = f0.w * (1.0f - interp) + interp * f1.w;
            }
            v.x = background_one(v.x, k.x, ft.x, is_log);
            v.y = background_one(v.y, k.y, ft.y, is_log);
            v.z = background_one(v.z, k.z, ft.z, is_log);
            v.w = background_one(v.w, k.w, ft.w, is_log);
            d4[i] = v;
        }
    } else {
        for (int x = threadIdx.x; x < R; x += blockDim.x) d[x] = background_one(d[x], dk[x], flat_at(x), is_log);
    }
}

} // namespace

void launch_background(float* data, const float* flat, const float* dark, int R, int A, int S, int nflats, int is_log,
                       hipStream_t st)
{
    const long long rows = (long long)S * A;
    if (rows <= 0 || R <= 0) return;
    if (rows > 0x7fffffffLL) fail(RAFT_EUNSUPPORTED, "too many frame rows");
    hipLaunchKernelGGL(k_background, dim3((unsigned)rows), dim3(256), 0, st, data, flat, dark, R, A, nflats, is_log);
    RAFT_HIP(hipGetLastError());
}

} // namespace raft

namespace raft {

namespace {

__global__ void k_fill(float* p, long long n, float v)
{
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x)
        p[e] = v;
}

} // namespace

// p[0 .. n) = v, as a kernel (usable inside captured graphs, unlike memset nodes)
void launch_fill(float* p, long long n, float v, hipStream_t st)
{
    if (n <= 0) return;
    const long long blocks = std::min<long long>((n + 255) / 256, 16384);
    hipLaunchKernelGGL(k_fill, dim3((unsigned)blocks), dim3(256), 0, st, p, n, v);
    RAFT_HIP(hipGetLastError());
}

} // namespace raft
