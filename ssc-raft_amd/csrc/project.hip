// project.hip -- ray-driven line sums for Radon-RT and the eEM forward
// operator (gfx950, wave64), plus the deterministic FP64 KL reduction.
//
// One thread per ray (256 consecutive rays of one angle per workgroup).  The
// sample coordinates are IEEE-double affine functions of the sample index,
//     ux = fma(k - ck, -sin*ry, fma(i - ci, cos, ucx)),  X = (int)ux
//     uy = fma(k - ck,  cos,    fma(i - ci, sin*rx, ucy)), Y = (int)uy
// (bit-identical to the CPU oracle), and the range of k whose sample lies in
// the image is computed exactly per ray (both coordinates are monotone in k,
// so the valid set is an interval: a generous real-valued estimate is shrunk
// by evaluating the same expressions at its ends).  The inner loop then has
// no range test.  For angles closer to the y axis than to the x axis
// (|sin| > |cos|) the kernel reads a transposed copy of the image, so that
// neighbouring rays of a wave touch neighbouring addresses.
//
// eEM mode fuses kernel_radonWithDivision (em_RT_kernel.cu:188-245): it
// forms v = dt * sum, the clamped smoothed ratio r = y SIGN(v)/sqrt(e^2+v^2)
// (configs.hpp:31-32, e = 1e-4), stores r * dtheta_a into the staged-row
// layout of the backprojector, and accumulates the FP64 KL term
// v - y + y ln(y/v) per workgroup (bins with v <= 0 or y < 0 are counted,
// not summed).
//
// Radon mode restates Radon_RT_version_sscRadon (radon_rt.cu:131-174).
#include "raft_internal.hpp"

namespace raft {

namespace {

struct ProjK {
    const float* img;
    const float* imgT;
    float* out;
    const float* y;
    const float* wa;
    double* kl_part;
    unsigned long long* ex_part;
    const double* cosv;
    const double* sinv;
    int Nx, Ny, nrays, nangles, nk;
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

template <int EM>
__global__ __launch_bounds__(256) void k_project(ProjK a)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int ang = blockIdx.y;
    const int s = blockIdx.z;
    const double cs = a.cosv[ang], sn = a.sinv[ang];
    const bool steep = fabs(sn) > fabs(cs) && a.imgT != nullptr;
    const size_t img_off = (size_t)s * a.Nx * a.Ny;
    const float* __restrict__ img = steep ? a.imgT + img_off : a.img + img_off;
    // address = Y * sY + X * sX
    const int sY = steep ? 1 : a.Nx;
    const int sX = steep ? a.Ny : 1;

    float acc = 0.0f;
    if (i < a.nrays) {
        const double ir = (double)i - a.ci;
        const double bxk = -sn * a.ry;
        const double byi = sn * a.rx;
        const double uxi = fma(ir, cs, a.ucx);
        const double uyi = fma(ir, byi, a.ucy);
        // slopes below 1e-9 (but not 0) make the real-valued interval estimate
        // ill-conditioned: such angles take the tested full sweep
        const bool sweep = (bxk != 0.0 && fabs(bxk) < 1.0e-9) || (cs != 0.0 && fabs(cs) < 1.0e-9);
        if (sweep) {
            double kr = -a.ck;
            for (int k = 0; k < a.nk; ++k) {
                const int X = (int)fma(kr, bxk, uxi);
                const int Y = (int)fma(kr, cs, uyi);
                if ((unsigned)X < (unsigned)a.Nx && (unsigned)Y < (unsigned)a.Ny) acc += img[Y * sY + X * sX];
                kr += 1.0;
            }
        } else {
            // generous estimate of the valid k range: -1 < u < N on both axes
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
            // shrink to the exact interval, then extend in case the estimate was short
            while (k0 < k1 && !in_image((double)k0 - a.ck, bxk, uxi, cs, uyi, a.Nx, a.Ny)) ++k0;
            while (k1 > k0 && !in_image((double)(k1 - 1) - a.ck, bxk, uxi, cs, uyi, a.Nx, a.Ny)) --k1;
            if (k1 > k0) {
                while (k0 > 0 && in_image((double)(k0 - 1) - a.ck, bxk, uxi, cs, uyi, a.Nx, a.Ny)) --k0;
                while (k1 < a.nk && in_image((double)k1 - a.ck, bxk, uxi, cs, uyi, a.Nx, a.Ny)) ++k1;
            }
            double kr = (double)k0 - a.ck;
            for (int k = k0; k < k1; ++k) {
                const int X = (int)fma(kr, bxk, uxi);
                const int Y = (int)fma(kr, cs, uyi);
                acc += img[Y * sY + X * sX];
                kr += 1.0;
            }
        }
    }
    const float v = acc * a.scale;

    if (!EM) {
        if (i < a.nrays) a.out[((size_t)s * a.nangles + ang) * a.nrays + i] = v;
        return;
    }

    double term = 0.0;
    unsigned long long ex = 0;
    if (i < a.nrays) {
        const float yv = a.y[((size_t)s * a.nangles + ang) * a.nrays + i];
        const float e = 1.0e-4f;
        const float sg = (v > 0.0f) ? 1.0f : ((v < 0.0f) ? -1.0f : 0.0f);
        float r = yv * (sg / sqrtf(e * e + v * v));
        if (r < 0.0f) r = 0.0f;
        a.out[((size_t)s * a.nangles + ang) * a.W + (i - a.tlo)] = r * a.wa[ang];
        if (a.kl_part) {
            const double vd = (double)v, yd = (double)yv;
            if (vd > 0.0 && yd >= 0.0) term = vd - yd + (yd > 0.0 ? yd * log(yd / vd) : 0.0);
            else ex = 1;
        }
    }
    if (!a.kl_part) return;
    // workgroup reduction (fixed order -> deterministic)
    __shared__ double sd[256];
    __shared__ unsigned long long se[256];
    sd[threadIdx.x] = term;
    se[threadIdx.x] = ex;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            sd[threadIdx.x] += sd[threadIdx.x + w];
            se[threadIdx.x] += se[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const size_t pidx = ((size_t)s * a.nangles + ang) * gridDim.x + blockIdx.x;
        a.kl_part[pidx] = sd[0];
        a.ex_part[pidx] = se[0];
    }
}

// kl[s*K + iter] = sum of the slice's partials, fixed order
__global__ __launch_bounds__(256) void k_kl_reduce(const double* part, const unsigned long long* expart, int nb,
                                                   double* kl, unsigned long long* exo, int K, const int* iter_ptr,
                                                   int iter)
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
    if (threadIdx.x == 0) {
        const int it = iter_ptr ? *iter_ptr : iter;
        if (it < K) {
            kl[(size_t)s * K + it] = sd[0];
            exo[(size_t)s * K + it] = se[0];
        }
    }
}

} // namespace

int proj_blocks_per_slice(const RayGeom& g) { return ((g.nrays + 255) / 256) * g.nangles; }

void launch_project(const ProjArgs& p, int em_mode, hipStream_t st)
{
    if (p.S <= 0 || p.g.nrays <= 0 || p.g.nangles <= 0) return;
    ProjK k;
    k.img = p.img; k.imgT = p.imgT; k.out = p.out; k.y = p.y; k.wa = p.wa;
    k.kl_part = p.kl_part; k.ex_part = p.ex_part; k.cosv = p.cosv; k.sinv = p.sinv;
    k.Nx = p.g.Nx; k.Ny = p.g.Ny; k.nrays = p.g.nrays; k.nangles = p.g.nangles; k.nk = p.g.nk;
    k.ci = p.g.ci; k.ck = p.g.ck; k.ucx = p.g.ucx; k.ucy = p.g.ucy; k.ry = p.g.ry; k.rx = p.g.rx;
    k.scale = p.g.scale;
    k.tlo = p.tlo; k.W = p.W;
    if (p.S > 65535 || p.g.nangles > 65535) fail(RAFT_EUNSUPPORTED, "projector grid too large");
    dim3 grid((p.g.nrays + 255) / 256, p.g.nangles, p.S);
    if (em_mode) hipLaunchKernelGGL((k_project<1>), grid, dim3(256), 0, st, k);
    else hipLaunchKernelGGL((k_project<0>), grid, dim3(256), 0, st, k);
    RAFT_HIP(hipGetLastError());
}

void launch_kl_reduce(const double* part, const unsigned long long* expart, int nb, int S, double* kl,
                      unsigned long long* ex, int K, int iter, hipStream_t st)
{
    if (S <= 0) return;
    hipLaunchKernelGGL(k_kl_reduce, dim3(S), dim3(256), 0, st, part, expart, nb, kl, ex, K, nullptr, iter);
    RAFT_HIP(hipGetLastError());
}

} // namespace raft
