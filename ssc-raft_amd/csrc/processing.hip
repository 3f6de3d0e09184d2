// processing.hip -- the step before the reconstruction path: flat/dark
// background correction with optional -log (background_correction.cu:13-45,
// opt.cu:196-214), in place on [nslices][nangles][nrays] frames.
#include "raft_internal.hpp"

namespace raft {

namespace {

// v = max(D - dark, 0.5) / max(F - dark, 0.5), F = flat row 0, or with two or
// more flats the blend of rows 0 and 1 by angle / nangles; then -log(v) if asked.
__global__ void k_background(float* data, const float* flat, const float* dark, int R, int A, int S, int nflats,
                             int is_log)
{
    const long long total = (long long)S * A * R;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const int x = (int)(e % R);
        const int y = (int)((e / R) % A);
        const long long z = e / ((long long)R * A);
        const float dk = dark[z * R + x];
        const float fb = flat[z * R * nflats + x];
        float ft = fb;
        if (nflats > 1) {
            const float interp = (float)y / (float)A;
            const float fa = flat[z * R * nflats + R + x];
            ft = fb * (1.0f - interp) + interp * fa;
        }
        const float T = data[e] - dk;
        const float Q = ft - dk;
        float v = fmaxf(T, 0.5f) / fmaxf(Q, 0.5f);
        if (is_log) v = -logf(v);
        data[e] = v;
    }
}

} // namespace

void launch_background(float* data, const float* flat, const float* dark, int R, int A, int S, int nflats, int is_log,
                       hipStream_t st)
{
    const long long total = (long long)S * A * R;
    if (total <= 0) return;
    const long long blocks = std::min<long long>((total + 255) / 256, 16384);
    hipLaunchKernelGGL(k_background, dim3((unsigned)blocks), dim3(256), 0, st, data, flat, dark, R, A, S, nflats,
                       is_log);
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
