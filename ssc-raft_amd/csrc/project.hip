// project.hip -- ray-driven line sums for Radon-RT and the eEM forward
// operator (gfx950, wave64), plus the deterministic FP64 KL reduction.
//
// Layout: the image is read slice-interleaved, [group][y][x][4], so one
// 16-byte gather serves four slices that share the ray geometry.  For angles
// closer to the y axis than to the x axis (|sin| > |cos|) the kernel reads the
// interleaved transpose instead, so the fast-varying direction across rays is
// always the contiguous one.
//
// Work split: one workgroup per (block of PROJ_ANGLES consecutive angles,
// slice group), one wave per angle.  The waves walk the same blocks of R
// consecutive rays in lockstep (a barrier per block), so the 8 nearby angles
// read nearly the same image region at the same time and the reuse is served
// by L1/L2.  Every ray is split over Kp = 64/R lanes that take its samples
// k0+kp, k0+kp+Kp, ...; at each step a wave therefore reads an R x Kp patch of
// the rotated sampling grid, which maps to a compact patch of the image (R is
// 64 for angles within ~2 degrees of an axis and 8 near 45 degrees).  The Kp
// partial sums of a ray are combined with lane shuffles.
//
// Sample coordinates are IEEE-double affine functions of the sample index,
//     ux = fma(k - ck, -sin*ry, fma(i - ci, cos, ucx)),  X = (int)ux
//     uy = fma(k - ck,  cos,    fma(i - ci, sin*rx, ucy)), Y = (int)uy
// (bit-identical to the CPU oracle).  The range of k whose sample lies in the
// image is computed exactly per ray: both coordinates are monotone in k, so
// the valid set is an interval; a generous real-valued estimate is shrunk and
// extended by evaluating the same expressions at its ends, and the inner loop
// has no range test.
//
// eEM mode fuses kernel_radonWithDivision (em_RT_kernel.cu:188-245): it forms
// v = dt * sum, the clamped smoothed ratio r = y SIGN(v)/sqrt(e^2+v^2)
// (configs.hpp:31-32, e = 1e-4), stores r * dtheta_a into the interleaved
// staged-row layout of the backprojector, and accumulates the FP64 KL term
// v - y + y ln(y/v) (bins with v <= 0 or y < 0 are counted, not summed).
// Radon mode restates Radon_RT_version_sscRadon (radon_rt.cu:131-174).
#include "raft_internal.hpp"

namespace raft {

namespace {

struct ProjK {
    const float4* img;
    const float4* imgT;
    float* out;
    const float* y;
    const float* wa;
    double* kl_part;
    unsigned long long* ex_part;
    const double* cosv;
    const double* sinv;
    int Nx, Ny, nrays, nangles, nk, S;
    double ci, ck, ucx, ucy, ry, rx;
    float scale;
    int tlo, W;
};

__device__ __forceinline__ bool in_image(double kr, double bxk, double uxi, double cs, double uyi, int Nx, int Ny)
{
    const int X = (int)fma(kr, bxk, uxi);
    const int Y = (int)fma(kr, cs, uyi);
    return (unsigned)X < (unsigned)Nx && (unsigned)Y < (unsigned)Ny;
}

// open real interval of k where lo_b < fma(k - ck, b, u0) < hi_b (b may be 0)
__device__ __forceinline__ void k_interval(double b, double u0, double lo_b, double hi_b, double ck, double& klo,
                                           double& khi)
{
    if (b > 0.0) {
        klo = fmax(klo, ck + (lo_b - u0) / b);
        khi = fmin(khi, ck + (hi_b - u0) / b);
    } else if (b < 0.0) {
        klo = fmax(klo, ck + (hi_b - u0) / b);
        khi = fmin(khi, ck + (lo_b - u0) / b);
    } else if (!(u0 > lo_b && u0 < hi_b)) {
        khi = -1.0e300;
    }
}

__device__ __forceinline__ float4 add4(float4 a, float4 b)
{
    return make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w);
}

// 16-byte gather through a buffer resource: 32-bit byte offsets, and offsets
// at or past the end of the buffer read as zero (used for the padding samples
// of the last unrolled step).  Word 3 = 0x00020000 is the gfx9-family raw
// buffer format.
__device__ __forceinline__ float4 gather(__amdgpu_buffer_rsrc_t rs, unsigned off)
{
    const auto v = __builtin_amdgcn_raw_buffer_load_b128(rs, (int)off, 0, 0);
    return make_float4(__uint_as_float(v[0]), __uint_as_float(v[1]), __uint_as_float(v[2]), __uint_as_float(v[3]));
}
constexpr unsigned OFF_NONE = 0xFFFFFFF0u;

__device__ __forceinline__ float ratio(float yv, float v)
{
    const float e = 1.0e-4f;
    const float sg = (v > 0.0f) ? 1.0f : ((v < 0.0f) ? -1.0f : 0.0f);
    const float r = yv * (sg / sqrtf(e * e + v * v));
    return r < 0.0f ? 0.0f : r;
}

template <int EM>
__global__ __launch_bounds__(64 * PROJ_ANGLES) void k_project(ProjK a)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int a0 = blockIdx.x * PROJ_ANGLES;
    const int ang = a0 + wave;
    const bool have = ang < a.nangles;
    const int g = blockIdx.y;
    // layout and patch shape are chosen once per workgroup, from its middle angle
    const int amid = min(a0 + PROJ_ANGLES / 2, a.nangles - 1);
    const double cm = fabs(a.cosv[amid]), sm = fabs(a.sinv[amid]);
    const bool steep = sm > cm && a.imgT != nullptr;
    const double tn = steep ? cm / sm : (cm > sm ? sm / cm : 1.0);
    const int lgR = tn < 0.031 ? 6 : (tn < 0.125 ? 5 : (tn < 0.5 ? 4 : 3));
    const int R = 1 << lgR, Kp = 64 >> lgR;
    const int r = lane & (R - 1), kp = lane >> lgR;
    const int nblocks = (a.nrays + R - 1) >> lgR;

    const size_t plane = (size_t)a.Nx * a.Ny;
    const float4* __restrict__ img = (steep ? a.imgT : a.img) + (size_t)g * plane;
    const int sY = steep ? 1 : a.Nx; // element = img[Y * sY + X * sX]
    const int sX = steep ? a.Ny : 1;
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<float4*>(img), 0, (int)(plane * sizeof(float4)), 0x00020000);

    const double cs = have ? a.cosv[ang] : 1.0, sn = have ? a.sinv[ang] : 0.0;
    const double bxk = -sn * a.ry;
    const double byi = sn * a.rx;
    // slopes below 1e-9 (but not 0) make the interval estimate ill-conditioned:
    // such angles take the full sweep with a per-sample test
    const bool sweep = (bxk != 0.0 && fabs(bxk) < 1.0e-9) || (cs != 0.0 && fabs(cs) < 1.0e-9);

    double klsum[4] = {0.0, 0.0, 0.0, 0.0};
    unsigned long long exc[4] = {0, 0, 0, 0};

    for (int rb = 0; rb < nblocks; ++rb) {
        const int i = (rb << lgR) + r;
        float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
        if (have && i < a.nrays) {
            const double ir = (double)i - a.ci;
            const double uxi = fma(ir, cs, a.ucx);
            const double uyi = fma(ir, byi, a.ucy);
            if (sweep) {
                double kr = (double)kp - a.ck;
                for (int k = kp; k < a.nk; k += Kp) {
                    const int X = (int)fma(kr, bxk, uxi);
                    const int Y = (int)fma(kr, cs, uyi);
                    if ((unsigned)X < (unsigned)a.Nx && (unsigned)Y < (unsigned)a.Ny) acc = add4(acc, img[Y * sY + X * sX]);
                    kr += (double)Kp;
                }
            } else {
                double klo = 0.0, khi = (double)a.nk;
                k_interval(bxk, uxi, -1.0, (double)a.Nx, a.ck, klo, khi);
                k_interval(cs, uyi, -1.0, (double)a.Ny, a.ck, klo, khi);
                int k0 = 0, k1 = 0;
                if (khi >= klo) {
                    k0 = (int)fmax(floor(klo) - 1.0, 0.0);
                    k1 = (int)fmin(ceil(khi) + 1.0, (double)a.nk);
                    k0 = min(k0, a.nk);
                    k1 = max(k1, k0);
                }
                while (k0 < k1 && !in_image((double)k0 - a.ck, bxk, uxi, cs, uyi, a.Nx, a.Ny)) ++k0;
                while (k1 > k0 && !in_image((double)(k1 - 1) - a.ck, bxk, uxi, cs, uyi, a.Nx, a.Ny)) --k1;
                if (k1 > k0) {
                    while (k0 > 0 && in_image((double)(k0 - 1) - a.ck, bxk, uxi, cs, uyi, a.Nx, a.Ny)) --k0;
                    while (k1 < a.nk && in_image((double)k1 - a.ck, bxk, uxi, cs, uyi, a.Nx, a.Ny)) ++k1;
                }
                // samples k0+kp, k0+kp+Kp, ... < k1, four gathers in flight per step;
                // kr + u*Kp is exact, so each coordinate is the oracle's expression
                const int cnt = (k1 > k0 + kp) ? (k1 - k0 - kp + Kp - 1) / Kp : 0;
                double kr = (double)(k0 + kp) - a.ck;
                const double kstep4 = (double)(4 * Kp);
                for (int m = 0; m < cnt; m += 4) {
                    unsigned off[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const double kru = kr + (double)(u * Kp);
                        const int X = (int)fma(kru, bxk, uxi);
                        const int Y = (int)fma(kru, cs, uyi);
                        off[u] = (m + u < cnt) ? (unsigned)(Y * sY + X * sX) * 16u : OFF_NONE;
                    }
                    const float4 v0 = gather(rs, off[0]), v1 = gather(rs, off[1]);
                    const float4 v2 = gather(rs, off[2]), v3 = gather(rs, off[3]);
                    acc = add4(acc, add4(add4(v0, v1), add4(v2, v3)));
                    kr += kstep4;
                }
            }
        }
        // combine the Kp partial sums of each ray (lanes r, r+R, r+2R, ...)
        for (int off = R; off < 64; off <<= 1) {
            acc.x += __shfl_xor(acc.x, off);
            acc.y += __shfl_xor(acc.y, off);
            acc.z += __shfl_xor(acc.z, off);
            acc.w += __shfl_xor(acc.w, off);
        }
        if (have && kp == 0 && i < a.nrays) {
            const float v[4] = {acc.x * a.scale, acc.y * a.scale, acc.z * a.scale, acc.w * a.scale};
            if (!EM) {
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    const int sl = g * 4 + s;
                    if (sl < a.S) a.out[((size_t)sl * a.nangles + ang) * a.nrays + i] = v[s];
                }
            } else if (EM == 2) {
                // transmission EM (em_RT_kernel.cu:20-31 with the slice's own flat
                // row, y = flat [S][nrays]): flat * exp(-v), times dtheta
                const float wv = a.wa[ang];
                float rr[4];
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    const int sl = g * 4 + s;
                    rr[s] = (sl < a.S) ? a.y[(size_t)sl * a.nrays + i] * expf(-v[s]) * wv : 0.0f;
                }
                reinterpret_cast<float4*>(a.out)[((size_t)g * a.nangles + ang) * a.W + (i - a.tlo)] =
                    make_float4(rr[0], rr[1], rr[2], rr[3]);
            } else {
                const float wv = a.wa[ang];
                float rr[4];
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    const int sl = g * 4 + s;
                    const float yv = (sl < a.S) ? a.y[((size_t)sl * a.nangles + ang) * a.nrays + i] : 0.0f;
                    rr[s] = ratio(yv, v[s]) * wv;
                    if (a.kl_part && sl < a.S) {
                        const double vd = (double)v[s], yd = (double)yv;
                        if (vd > 0.0 && yd >= 0.0) klsum[s] += vd - yd + (yd > 0.0 ? yd * log(yd / vd) : 0.0);
                        else exc[s] += 1;
                    }
                }
                reinterpret_cast<float4*>(a.out)[((size_t)g * a.nangles + ang) * a.W + (i - a.tlo)] =
                    make_float4(rr[0], rr[1], rr[2], rr[3]);
            }
        }
        __syncthreads(); // keep the waves (angles) of the workgroup on the same ray block
    }
    if (EM != 1 || !a.kl_part || !have) return;
    // deterministic per-(slice, angle) KL partial: wave reduction in fixed order
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        double d = klsum[s];
        unsigned long long e = exc[s];
        for (int off = 32; off > 0; off >>= 1) {
            d += __shfl_xor(d, off);
            e += __shfl_xor(e, off);
        }
        const int sl = g * 4 + s;
        if (lane == 0 && sl < a.S) {
            a.kl_part[(size_t)sl * a.nangles + ang] = d;
            a.ex_part[(size_t)sl * a.nangles + ang] = e;
        }
    }
}

// kl[s*K + iter] = sum of the slice's partials, fixed order
__global__ __launch_bounds__(256) void k_kl_reduce(const double* part, const unsigned long long* expart, int nb,
                                                   double* kl, unsigned long long* exo, int K, int iter,
                                                   const int* iter_ptr)
{
    const int s = blockIdx.x;
    __shared__ double sd[256];
    __shared__ unsigned long long se[256];
    double acc = 0.0;
    unsigned long long ec = 0;
    for (int b = threadIdx.x; b < nb; b += 256) {
        acc += part[(size_t)s * nb + b];
        ec += expart[(size_t)s * nb + b];
    }
    sd[threadIdx.x] = acc;
    se[threadIdx.x] = ec;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            sd[threadIdx.x] += sd[threadIdx.x + w];
            se[threadIdx.x] += se[threadIdx.x + w];
        }
        __syncthreads();
    }
    const int it = iter_ptr ? *iter_ptr : iter;
    if (threadIdx.x == 0 && it >= 0 && it < K) {
        kl[(size_t)s * K + it] = sd[0];
        exo[(size_t)s * K + it] = se[0];
    }
}

// iteration counter for graph-replayed eEM iterations: op 0 clears, op 1 increments
__global__ void k_counter(int* c, int op)
{
    if (threadIdx.x == 0) *c = op ? *c + 1 : 0;
}

} // namespace

void launch_project(const ProjArgs& p, int em_mode, hipStream_t st)
{
    if (p.S <= 0 || p.g.nrays <= 0 || p.g.nangles <= 0) return;
    ProjK k;
    k.img = reinterpret_cast<const float4*>(p.img);
    k.imgT = reinterpret_cast<const float4*>(p.imgT);
    k.out = p.out; k.y = p.y; k.wa = p.wa;
    k.kl_part = p.kl_part; k.ex_part = p.ex_part; k.cosv = p.cosv; k.sinv = p.sinv;
    k.Nx = p.g.Nx; k.Ny = p.g.Ny; k.nrays = p.g.nrays; k.nangles = p.g.nangles; k.nk = p.g.nk; k.S = p.S;
    k.ci = p.g.ci; k.ck = p.g.ck; k.ucx = p.g.ucx; k.ucy = p.g.ucy; k.ry = p.g.ry; k.rx = p.g.rx;
    k.scale = p.g.scale;
    k.tlo = p.tlo; k.W = p.W;
    const int G = (p.S + 3) / 4;
    if (G > 65535) fail(RAFT_EUNSUPPORTED, "projector grid too large");
    dim3 grid((p.g.nangles + PROJ_ANGLES - 1) / PROJ_ANGLES, G);
    if (em_mode == 2) hipLaunchKernelGGL((k_project<2>), grid, dim3(64 * PROJ_ANGLES), 0, st, k);
    else if (em_mode) hipLaunchKernelGGL((k_project<1>), grid, dim3(64 * PROJ_ANGLES), 0, st, k);
    else hipLaunchKernelGGL((k_project<0>), grid, dim3(64 * PROJ_ANGLES), 0, st, k);
    RAFT_HIP(hipGetLastError());
}

void launch_kl_reduce(const double* part, const unsigned long long* expart, int nb, int S, double* kl,
                      unsigned long long* ex, int K, int iter, const int* iter_ptr, hipStream_t st)
{
    if (S <= 0) return;
    hipLaunchKernelGGL(k_kl_reduce, dim3(S), dim3(256), 0, st, part, expart, nb, kl, ex, K, iter, iter_ptr);
    RAFT_HIP(hipGetLastError());
}

void launch_counter(int* c, int op, hipStream_t st)
{
    hipLaunchKernelGGL(k_counter, dim3(1), dim3(64), 0, st, c, op);
    RAFT_HIP(hipGetLastError());
}

} // namespace raft
