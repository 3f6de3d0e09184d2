// project.hip -- ray-driven line sums for Radon-RT and the eEM forward
// operator (gfx950, wave64), plus the deterministic FP64 KL reduction.
//
// Layout: the image is read slice-interleaved, [group][y][x][4], so one
// 16-byte gather serves four slices that share the ray geometry.  For angles
// closer to the y axis than to the x axis (|sin| > |cos|) the kernel reads the
// interleaved transpose instead, so the fast-varying direction across rays is
// always the contiguous one.
//
// Two forms share the sampling and the epilogue:
//
// Staged form (k_project_lds, dense angle sets).  One workgroup = 8 waves =
// 8 consecutive angles, one ray chunk, one slice group.  The waves walk the
// same block of PL_RAYS rays in steps of PL_K samples; at every step the
// workgroup copies the union of the 8 patches' bounding boxes (rays
// [ib, ib + 32) x samples [ks, ks + 32), mapped into the image) from HBM/L2
// into LDS with row-contiguous 16-byte loads -- zeros outside the image --
// and each wave gathers its 1024 samples from LDS.  Because out-of-image
// samples read staged zeros, the inner loop has no range test.  The gathers
// of a direct form are limited by the L1 tag rate (a wave's 64 gathers touch
// ~38 distinct 64-byte lines); staged, a step costs ~17 lines per staged row
// and the gathers become LDS reads.  Rows are staged at a pitch that is a
// multiple of 16 entries of 16 bytes whenever the box still fits the tile
// (else at the narrowest pitch of 1 mod 4, which the host admission assumes):
// the 16 lanes a 16-byte LDS read serves together hold 16 neighbouring rays
// at one sample, whose entries then fall on distinct bank groups unless two
// share a column in different rows (a pitch of 1 mod 4 also makes lanes ~11
// apart collide at slopes near 1: about twice the conflict cycles in a bank
// simulation of this pattern).  Near-vertical chunks stage from the
// transposed image, so that neighbouring rays always step mostly along a
// staged row.
// Steps whose union box is empty (no wave has samples there) are dropped from
// a batch's step list.  In the prefetch form (the default), the first PL_PF
// rows per wave of the next step's box are loaded into registers before the
// current step's gathers and written to LDS after them, so the staging
// latency overlaps the gathers; rows past 8 PL_PF and columns past 64 are
// staged after the barrier, as in the plain form.  Both forms stage the same
// values, so they agree bit for bit.
// The host admits a launch to this form only when every chunk's union box
// fits PL_CAP entries at every patch position (proj_plan); sparse angle sets
// use the direct form.
//
// Direct form (k_project): one workgroup per (block of PROJ_ANGLES
// consecutive angles, slice group), one wave per angle.  The waves walk the
// same blocks of R consecutive rays in lockstep (a barrier per block), so the
// 8 nearby angles read nearly the same image region at the same time and the
// reuse is served by L1/L2.  Every ray is split over Kp = 64/R lanes that
// take its samples k0+kp, k0+kp+Kp, ...; at each step a wave therefore reads
// an R x Kp patch of the rotated sampling grid (R is 64 for angles within ~2
// degrees of an axis and 8 near 45 degrees).  The Kp partial sums of a ray
// are combined with lane shuffles.

// Sample coordinates are IEEE-double affine functions of the sample index,
//     ux = fma(k - ck, -sin*ry, fma(i - ci, cos, ucx)),  X = (int)ux
//     uy = fma(k - ck,  cos,    fma(i - ci, sin*rx, ucy)), Y = (int)uy
// (bit-identical to the CPU oracle).  The range of k whose sample lies in the
// image is computed exactly per ray in the direct form: both coordinates are
// monotone in k, so the valid set is an interval; a generous real-valued
// estimate is shrunk and extended by evaluating the same expressions at its
// ends, and the inner loop has no range test.
//
// eEM mode fuses kernel_radonWithDivision (em_RT_kernel.cu:188-245): it forms
// v = dt * sum, the clamped smoothed ratio r = y SIGN(v)/sqrt(e^2+v^2)
// (configs.hpp:31-32, e = 1e-4), stores r * dtheta_a into the interleaved
// staged-row layout of the backprojector, and accumulates the FP64 KL term
// v - y + y ln(y/v) (bins with v <= 0 or y < 0 are counted, not summed),
// one partial per (slice, angle, ray chunk) in a fixed order.
// Radon mode restates Radon_RT_version_sscRadon (radon_rt.cu:131-174).
#include "raft_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

namespace raft {

namespace {

constexpr int PL_RAYS = 32;   // rays per patch (staged form)
constexpr int PL_K = 32;      // samples per ray per step (staged form)
constexpr int PL_BATCH = 64;  // steps whose boxes are computed together
constexpr int PL_CAP = 4480;  // staged 16-byte entries per workgroup (70 KB: two workgroups per CU)
// rows per wave staged through registers ahead of the gathers (prefetch form):
// 56 rows cover ~96% of the union boxes at 2048 angles of 2048^2
constexpr int PL_PF = 7;

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
    int RC;          // ray chunks per angle chunk (staged form; 1 for the direct form)
    int kl_stride;   // KL partials per slice = nangles * RC
    int xcd_order;   // staged form: give each XCD a contiguous range of the grid (see k_project_lds)
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

// Output of one ray for the 4 slices of group g: Radon line sums, eEM ratio
// row (+ KL terms), or tEM flat * exp(-v) row.  `part` of the `nsplit` lanes
// that hold the ray's sum (nsplit in {1, 2, 4}): part 0 writes the row; the
// FP64 KL terms of slices s = part + j * nsplit are computed by that lane and
// summed in its slot j, so each (ray, slice) term is added exactly once.
template <int EM>
__device__ __forceinline__ void ray_out(const ProjK& a, int g, int ang, int i, float4 acc, double* klsum,
                                        unsigned* exc, int part = 0, int nsplit = 1)
{
    const float v[4] = {acc.x * a.scale, acc.y * a.scale, acc.z * a.scale, acc.w * a.scale};
    if (EM == 1 && a.kl_part) {
        // the slice is picked by selects so that all lanes run one FP64 block
        // per round; slots are indexed by constants (registers, not scratch)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (j >= 4 / nsplit) break;
            const int s = part + j * nsplit;
            const int sl = g * 4 + s;
            if (sl < a.S) {
                const float vs = s == 0 ? v[0] : (s == 1 ? v[1] : (s == 2 ? v[2] : v[3]));
                const double vd = (double)vs, yd = (double)a.y[((size_t)sl * a.nangles + ang) * a.nrays + i];
                const bool ok = vd > 0.0 && yd >= 0.0;
                const double t = ok ? vd - yd + (yd > 0.0 ? yd * log(yd / vd) : 0.0) : 0.0;
                klsum[j] += t;
                exc[j] += ok ? 0u : 1u;
            }
        }
    }
    if (EM == 1 && part > 0) return;
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
    } else if (a.out) { // null: KL only, no ratio rows
        const float wv = a.wa[ang];
        float rr[4];
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const int sl = g * 4 + s;
            const float yv = (sl < a.S) ? a.y[((size_t)sl * a.nangles + ang) * a.nrays + i] : 0.0f;
            rr[s] = ratio(yv, v[s]) * wv;
        }
        reinterpret_cast<float4*>(a.out)[((size_t)g * a.nangles + ang) * a.W + (i - a.tlo)] =
            make_float4(rr[0], rr[1], rr[2], rr[3]);
    }
}

// deterministic KL partial of one wave (one angle): wave reduction in fixed
// order.  Slice s sits in slot s / nsplit of the lanes whose part is s % nsplit
// (zero elsewhere)
__device__ __forceinline__ void kl_flush(const ProjK& a, int g, int slot, int lane, const double* klsum,
                                         const unsigned* exc, int part, int nsplit)
{
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const bool mine = part < nsplit && s % nsplit == part;
        const int j = s / nsplit;
        const double kj = j == 0 ? klsum[0] : (j == 1 ? klsum[1] : (j == 2 ? klsum[2] : klsum[3]));
        const unsigned ej = j == 0 ? exc[0] : (j == 1 ? exc[1] : (j == 2 ? exc[2] : exc[3]));
        double d = mine ? kj : 0.0;
        unsigned long long e = mine ? ej : 0u;
        for (int off = 32; off > 0; off >>= 1) {
            d += __shfl_xor(d, off);
            e += __shfl_xor(e, off);
        }
        const int sl = g * 4 + s;
        if (lane == 0 && sl < a.S) {
            a.kl_part[(size_t)sl * a.kl_stride + slot] = d;
            a.ex_part[(size_t)sl * a.kl_stride + slot] = e;
        }
    }
}

// Four waves per SIMD (<= 128 VGPRs): two workgroups per CU (the eEM form
// held 141 VGPRs and ran one; it now spills two loop invariants, reloaded
// once per ray block).
template <int EM>
__global__ __launch_bounds__(64 * PROJ_ANGLES, 4) void k_project(ProjK a)
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
    unsigned exc[4] = {0, 0, 0, 0};

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
        // every lane of the ray holds its sum: up to 4 of them share the KL terms
        const int nsplit = EM == 1 ? min(Kp, 4) : 1;
        if (have && kp < nsplit && i < a.nrays) ray_out<EM>(a, g, ang, i, acc, klsum, exc, kp, nsplit);
        __syncthreads(); // keep the waves (angles) of the workgroup on the same ray block
    }
    if (EM == 1 && a.kl_part && have) kl_flush(a, g, ang, lane, klsum, exc, kp, min(Kp, 4));
}

// ---- staged form ----
// Bounding box of one wave's patch (rays [ib, ib + PL_RAYS) x samples
// [ks, ks + PL_K)) in staged (row, col) coordinates, half-open.  Each
// coordinate is affine in (ir, kr), so its extremes over the patch are
// separable sums; the margins cover the rounding of the sample expressions
// and truncation toward zero.
__device__ __forceinline__ int4 patch_box(double ir0, double kr0, double rU, double rI, double rK, double cU, double cI,
                                          double cK)
{
    const double ir1 = ir0 + (double)(PL_RAYS - 1), kr1 = kr0 + (double)(PL_K - 1);
    const double rlo = rU + fmin(ir0 * rI, ir1 * rI) + fmin(kr0 * rK, kr1 * rK);
    const double rhi = rU + fmax(ir0 * rI, ir1 * rI) + fmax(kr0 * rK, kr1 * rK);
    const double clo = cU + fmin(ir0 * cI, ir1 * cI) + fmin(kr0 * cK, kr1 * cK);
    const double chi = cU + fmax(ir0 * cI, ir1 * cI) + fmax(kr0 * cK, kr1 * cK);
    return make_int4((int)floor(rlo) - 1, (int)floor(rhi) + 3, (int)floor(clo) - 1, (int)floor(chi) + 3);
}

// byte offset ((R * pitch + C) * 16 - off16) of a staged sample in two VALU ops:
// a signed 24-bit multiply-add (R may be negative: the box extends past the image
// edge; |R|, pitch * 16 < 2^23) and a shift-add (the compiler's form takes three)
__device__ __forceinline__ unsigned tile_off(int R, int C, int pitch16, int moff16)
{
    int rp;
    unsigned e;
    asm("v_mad_i32_i24 %0, %1, %2, %3" : "=v"(rp) : "v"(R), "v"(pitch16), "v"(moff16));
    asm("v_lshl_add_u32 %0, %1, 4, %2" : "=v"(e) : "v"(C), "v"(rp));
    return e;
}

// Wait for every vector-memory load of this wave (s_waitcnt vmcnt(0); expcnt
// and lgkmcnt left free: 0xF70 in the gfx9 encoding).  The staging writes sit
// under lane masks, so on some paths the compiler sees a load whose wait was
// skipped and then waits conservatively (vmcnt(0)) at the next reuse of a
// register -- which would fall inside the gathers.  One explicit wait where
// the loads are complete anyway keeps every later count exact.
__device__ __forceinline__ void wait_loads() { __builtin_amdgcn_s_waitcnt(0xF70); }

// Union box of one step in staged (row, col) coordinates: rows
// [ux, ux + nr) x cols [uz, uz + w), staged at row pitch `pitch` entries --
// a multiple of 16 when that fits the tile, else the narrowest pitch of
// 1 mod 4.  Every field is the same for every lane (scalar registers).
struct StepBox {
    int ux, uz, nr, w, pitch;
};

__device__ __forceinline__ StepBox step_box(const int4& ub)
{
    StepBox b;
    b.ux = __builtin_amdgcn_readfirstlane(ub.x);
    const int uy = __builtin_amdgcn_readfirstlane(ub.y);
    b.uz = __builtin_amdgcn_readfirstlane(ub.z);
    const int uw = __builtin_amdgcn_readfirstlane(ub.w);
    b.nr = uy - b.ux;
    b.w = uw - b.uz;
    b.pitch = 1;
    if (b.nr > 0) {
        const int pitch16 = (b.w + 15) & ~15;
        b.pitch = b.nr * pitch16 <= PL_CAP ? pitch16 : b.w + ((1 - b.w) & 3);
        // the launch was admitted only if every union fits; the clamp keeps
        // every LDS access in bounds regardless (its division runs only when
        // it changes nr)
        if (b.nr * b.pitch > PL_CAP) b.nr = PL_CAP / b.pitch;
    }
    return b;
}

// Buffer offset of staged row 0 for this lane's column.  Row R = ux + rr of
// lane column C is base + rr * nCol * 16 (mod 2^32): rows outside the image
// fall past the buffer's end (negative ones wrap to >= 2^31), and a lane
// whose column lies outside starts at 2^31, so both read zeros with no test
// in the loops.
__device__ __forceinline__ unsigned stage_base(const StepBox& b, int nCol, int lane)
{
    const int C = b.uz + lane;
    const bool cin = lane < b.w && (unsigned)C < (unsigned)nCol;
    return cin ? (unsigned)(b.ux * nCol + C) * 16u : 0x80000000u;
}

// Rows rr = first, first + 8, ... of the box (one wave per row, eight rows'
// loads in flight before their LDS writes), then the columns past the first
// 64 of every row (boxes wider than a wave); all of them complete on return.
__device__ __forceinline__ void stage_rows(float4* tile, __amdgpu_buffer_rsrc_t rs, const StepBox& b, int nCol,
                                           int lane, int wave, int first)
{
    const unsigned row16 = (unsigned)nCol * 16u;
    const unsigned base = stage_base(b, nCol, lane);
    const bool wr = lane < b.w;
    for (int rr0 = first; rr0 < b.nr; rr0 += 8 * PROJ_ANGLES) {
        float4 v[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int rr = rr0 + q * PROJ_ANGLES;
            v[q] = rr < b.nr ? gather(rs, base + (unsigned)rr * row16) : make_float4(0.f, 0.f, 0.f, 0.f);
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int rr = rr0 + q * PROJ_ANGLES;
            if (wr && rr < b.nr) tile[rr * b.pitch + lane] = v[q];
        }
    }
    for (int cc0 = 64; cc0 < b.w; cc0 += 64) {
        // column uz + cc0 + lane of row ux + rr: inside the image, or OFF_NONE
        const int C = b.uz + cc0 + lane;
        const bool cin = cc0 + lane < b.w && (unsigned)C < (unsigned)nCol;
        for (int rr = wave; rr < b.nr; rr += PROJ_ANGLES) {
            // rows outside the image fall past the buffer's end, as in stage_base
            const unsigned off = cin ? (unsigned)((b.ux + rr) * nCol + C) * 16u : OFF_NONE;
            const float4 v = gather(rs, off);
            if (cc0 + lane < b.w) tile[rr * b.pitch + cc0 + lane] = v;
        }
    }
    wait_loads();
}

// Prefetch: rows rr = wave + 8 q, q < PF, of the box into registers ...
template <int PF>
__device__ __forceinline__ void stage_issue(__amdgpu_buffer_rsrc_t rs, const StepBox& b, int nCol, int lane, int wave,
                                            float4 (&pv)[PF])
{
    const unsigned row16 = (unsigned)nCol * 16u;
    const unsigned base = stage_base(b, nCol, lane);
#pragma unroll
    for (int q = 0; q < PF; ++q) {
        const int rr = wave + q * PROJ_ANGLES;
        pv[q] = rr < b.nr ? gather(rs, base + (unsigned)rr * row16) : make_float4(0.f, 0.f, 0.f, 0.f);
    }
}

// ... and from the registers into the tile (after the barrier that ends the
// previous step's gathers)
template <int PF>
__device__ __forceinline__ void stage_commit(float4* tile, const StepBox& b, int lane, int wave, const float4 (&pv)[PF])
{
    const bool wr = lane < b.w;
    wait_loads();
#pragma unroll
    for (int q = 0; q < PF; ++q) {
        const int rr = wave + q * PROJ_ANGLES;
        if (wr && rr < b.nr) tile[rr * b.pitch + lane] = pv[q];
    }
}

// LDS of one staged-form workgroup
struct ProjSmem {
    float4 tile[PL_CAP];
    int4 ubox[PL_BATCH]; // union boxes of the batch's steps with samples
    int kst[PL_BATCH];   // their first sample index
    int nst;             // their count
    int kr[PROJ_ANGLES][2];
};

// One workgroup's work of the staged form: angle chunk ab (angles
// [8 ab, 8 ab + 8)), ray chunk rc (ray blocks rc, rc + RC, ...), slice group
// g.  Every LDS write is preceded by a barrier, so a persistent workgroup can
// run items back to back.
template <int EM, int PF>
__device__ __forceinline__ void proj_lds_item(const ProjK& a, int ab, int rc, int g, ProjSmem& sm)
{
    float4* tile = sm.tile;
    int4* s_ubox = sm.ubox;
    int* s_kst = sm.kst;
    int* s_nst = &sm.nst;
    int(*s_kr)[2] = sm.kr;
    // per-wave boxes of a batch of steps, computed while no step is staged
    int4* s_wbox = reinterpret_cast<int4*>(tile);
    static_assert(PROJ_ANGLES * PL_BATCH * sizeof(int4) <= sizeof(float4) * PL_CAP, "box scratch exceeds the tile");

    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); // uniform: scalar row loops
    const int a0 = ab * PROJ_ANGLES;
    const int ang = a0 + wave;
    const bool have = ang < a.nangles;
    // staged orientation, chosen once per workgroup from its middle angle:
    // rows = Y, cols = X (image [Y][X]), or rows = X, cols = Y (transpose)
    const int amid = min(a0 + PROJ_ANGLES / 2, a.nangles - 1);
    const bool steep = fabs(a.sinv[amid]) > fabs(a.cosv[amid]) && a.imgT != nullptr;
    const int nRow = steep ? a.Nx : a.Ny, nCol = steep ? a.Ny : a.Nx;
    const size_t plane = (size_t)a.Nx * a.Ny;
    const float4* __restrict__ src = (steep ? a.imgT : a.img) + (size_t)g * plane;
    // the plane through a buffer resource (the host admits planes below 2^31
    // bytes): byte offsets at or past its end read as zero
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<float4*>(src), 0, (int)(plane * sizeof(float4)), 0x00020000);

    const double cs = have ? a.cosv[ang] : 1.0, sn = have ? a.sinv[ang] : 0.0;
    const double bxk = -sn * a.ry, byi = sn * a.rx;
    // X = (int)fma(kr, bxk, fma(ir, cs, ucx)), Y = (int)fma(kr, cs, fma(ir, byi, ucy)),
    // renamed to row / col: row = (int)fma(kr, rK, fma(ir, rI, rU)), col likewise
    const double rI = steep ? cs : byi, rK = steep ? bxk : cs, rU = steep ? a.ucx : a.ucy;
    const double cI = steep ? byi : cs, cK = steep ? cs : bxk, cU = steep ? a.ucy : a.ucx;

    const int r = lane & (PL_RAYS - 1), kp = lane >> 5; // 32 rays x 2 sample phases
    const int nrb = (a.nrays + PL_RAYS - 1) / PL_RAYS;

    double klsum[4] = {0.0, 0.0, 0.0, 0.0};
    unsigned exc[4] = {0, 0, 0, 0};

    for (int rb = rc; rb < nrb; rb += a.RC) {
        const int ib = rb * PL_RAYS;
        const int i = ib + r;
        const bool ray_ok = have && i < a.nrays;
        const double ir = (double)i - a.ci;
        const double rowi = fma(ir, rI, rU), coli = fma(ir, cI, cU);
        // a superset of the k with in-image samples (exact bounds are not
        // needed: out-of-image samples read zeros), clamped to [0, nk)
        int kl = 0x7fffffff, kh = -1;
        if (ray_ok) {
            double klo = 0.0, khi = (double)a.nk;
            k_interval(rK, rowi, -1.0, (double)nRow, a.ck, klo, khi);
            k_interval(cK, coli, -1.0, (double)nCol, a.ck, klo, khi);
            if (khi >= klo) {
                kl = (int)fmax(floor(klo) - 1.0, 0.0);
                kh = (int)fmin(ceil(khi) + 2.0, (double)a.nk);
            }
        }
        for (int off = 32; off > 0; off >>= 1) {
            kl = min(kl, __shfl_xor(kl, off));
            kh = max(kh, __shfl_xor(kh, off));
        }
        __syncthreads(); // the previous ray block is done with s_kr, s_ubox and the tile
        if (lane == 0) {
            s_kr[wave][0] = kl;
            s_kr[wave][1] = kh;
        }
        __syncthreads();
        int Kmin = 0x7fffffff, Kmax = -1;
#pragma unroll
        for (int w = 0; w < PROJ_ANGLES; ++w) {
            Kmin = min(Kmin, s_kr[w][0]);
            Kmax = max(Kmax, s_kr[w][1]);
        }
        const int wkl = s_kr[wave][0], wkh = s_kr[wave][1];
        float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);

        for (int kb = Kmin; kb < Kmax; kb += PL_K * PL_BATCH) {
            const int nst = min(PL_BATCH, (Kmax - kb + PL_K - 1) / PL_K);
            // boxes of this batch of steps: lane = step of its own wave, then
            // the union over the waves, one thread per step
            {
                const int ks = kb + lane * PL_K;
                int4 bx = make_int4(0, 0, 0, 0);
                if (have && lane < nst && ks < wkh && ks + PL_K > wkl)
                    bx = patch_box((double)ib - a.ci, (double)ks - a.ck, rU, rI, rK, cU, cI, cK);
                // the tile (aliased by s_wbox) is free: the last step ended with a barrier
                s_wbox[wave * PL_BATCH + lane] = bx;
                __syncthreads();
                if (threadIdx.x < 64) { // wave 0, one lane per step (nst <= 64)
                    int4 u = make_int4(0x7fffffff, -0x7fffffff, 0x7fffffff, -0x7fffffff);
                    bool any = false;
                    if (threadIdx.x < nst) {
#pragma unroll
                        for (int w = 0; w < PROJ_ANGLES; ++w) {
                            const int4 b = s_wbox[w * PL_BATCH + threadIdx.x];
                            if (b.y > b.x) {
                                u.x = min(u.x, b.x);
                                u.y = max(u.y, b.y);
                                u.z = min(u.z, b.z);
                                u.w = max(u.w, b.w);
                                any = true;
                            }
                        }
                    }
                    // steps where no wave has samples are dropped; the list keeps
                    // the others in order, with their first sample index
                    const unsigned long long m = __ballot(any);
                    const int idx = __popcll(m & ((1ull << threadIdx.x) - 1ull));
                    if (any) {
                        s_ubox[idx] = u;
                        s_kst[idx] = kb + (int)threadIdx.x * PL_K;
                    }
                    if (threadIdx.x == 0) *s_nst = __popcll(m);
                }
                __syncthreads();
            }
            const int nsv = __builtin_amdgcn_readfirstlane(*s_nst); // steps with samples
            // Steps are staged in order.  With PF > 0, the first PF rows per wave
            // of step st + 1 are loaded into registers before step st's gathers and
            // written to LDS after them, so their latency overlaps the gathers; the
            // rest of a box (rows past 8 PF, columns past 64) is staged as in the
            // plain form.  The staged values, and so every sum, are the same.
            // Every listed step has a box, so every prefetch is consumed by a
            // commit: no load is left in flight when its registers are reused.
            float4 pv[PF > 0 ? PF : 1];
            StepBox bnext{};
            if constexpr (PF > 0) {
                if (nsv > 0) {
                    bnext = step_box(s_ubox[0]);
                    stage_issue<PF>(rs, bnext, nCol, lane, wave, pv);
                }
            }
            for (int st = 0; st < nsv; ++st) {
                const int ks = __builtin_amdgcn_readfirstlane(s_kst[st]);
                const StepBox b = PF > 0 ? bnext : step_box(s_ubox[st]);
                if constexpr (PF > 0) stage_commit<PF>(tile, b, lane, wave, pv);
                stage_rows(tile, rs, b, nCol, lane, wave, wave + PF * PROJ_ANGLES);
                __syncthreads();
                if constexpr (PF > 0) {
                    if (st + 1 < nsv) {
                        bnext = step_box(s_ubox[st + 1]);
                        stage_issue<PF>(rs, bnext, nCol, lane, wave, pv);
                    }
                }
                const int pitch = b.pitch;
                if (have && ks < wkh && ks + PL_K > wkl) {
                    // byte offset of sample (R, C) in the tile: R * pitch16 - off16 + C * 16
                    // (24-bit multiply-add, then shift-add).  Every sample lies in its wave's
                    // patch_box (margins of 1 and 2 rows / columns around the FP64 extremes,
                    // far beyond their rounding difference), so in the union box: the offset
                    // is below nr * pitch * 16 <= PL_CAP * 16, the bound proj_plan admitted
                    // the launch with.  No clamp is needed; the staging clamp of nr keeps the
                    // tile's writes in bounds
                    const int pitch16 = pitch << 4;
                    const int moff16 = -((b.ux * pitch + b.uz) << 4);
                    const char* tb = reinterpret_cast<const char*>(tile);
                    double kr = (double)(ks + kp) - a.ck;
                    float4 sacc = make_float4(0.f, 0.f, 0.f, 0.f); // this step's sum (two-level summation)
                    const int jmax = ks + PL_K <= a.nk ? PL_K / 2 : (a.nk - ks - kp + 1) >> 1; // ks + kp + 2j < nk
                    if (jmax == PL_K / 2) {
                        // GQ samples per trip, all their gathers issued before the
                        // first add, so GQ LDS reads are in flight; the adds keep the
                        // order of the one-sample loop (kr + 2q is exact: kr is an
                        // integer or half-integer far below 2^52).  GQ as timed per
                        // mode at the 128-register budget: 8 for the eEM form, 4 for
                        // the others (8 made Radon ~1% slower)
                        constexpr int GQ = EM == 1 ? 8 : 4;
#pragma unroll 2
                        for (int j = 0; j < PL_K / 2; j += GQ) {
                            unsigned e[GQ];
#pragma unroll
                            for (int q = 0; q < GQ; ++q) {
                                const double kq = kr + 2.0 * q;
                                e[q] = tile_off((int)fma(kq, rK, rowi), (int)fma(kq, cK, coli), pitch16, moff16);
                            }
                            float4 x[GQ];
#pragma unroll
                            for (int q = 0; q < GQ; ++q) x[q] = *reinterpret_cast<const float4*>(tb + e[q]);
#pragma unroll
                            for (int q = 0; q < GQ; ++q) sacc = add4(sacc, x[q]);
                            kr += 2.0 * GQ;
                        }
                    } else {
                        for (int j = 0; j < jmax; ++j) {
                            const int R = (int)fma(kr, rK, rowi);
                            const int C = (int)fma(kr, cK, coli);
                            const unsigned e = tile_off(R, C, pitch16, moff16);
                            sacc = add4(sacc, *reinterpret_cast<const float4*>(tb + e));
                            kr += 2.0;
                        }
                    }
                    acc = add4(acc, sacc);
                }
                __syncthreads();
            }
        }
        // combine the two sample phases of each ray (lanes r and r + 32)
        acc.x += __shfl_xor(acc.x, 32);
        acc.y += __shfl_xor(acc.y, 32);
        acc.z += __shfl_xor(acc.z, 32);
        acc.w += __shfl_xor(acc.w, 32);
        // both phases hold the ray's sum and share its KL terms
        if (ray_ok && (EM == 1 || kp == 0)) ray_out<EM>(a, g, ang, i, acc, klsum, exc, kp, EM == 1 ? 2 : 1);
    }
    if (EM == 1 && a.kl_part && have) kl_flush(a, g, ang * a.RC + rc, lane, klsum, exc, kp, 2);
}

// The hardware deals workgroups to the 8 XCDs round-robin by linear block
// id, so in grid order XCD x would run every 8th angle chunk.  With
// xcd_order the linear id l is renumbered (l mod 8) * (total / 8) + l / 8, so
// that XCD x runs the contiguous range [x total / 8, (x + 1) total / 8) of
// (angle chunk, ray chunk, slice group) items: consecutive angle chunks of the
// same rays, whose staged boxes overlap, share that XCD's L2.  Only the order
// changes; every item computes the same values.
template <int EM, int PF>
__global__ __launch_bounds__(64 * PROJ_ANGLES, 4) void k_project_lds(ProjK a)
{
    __shared__ ProjSmem sm;
    unsigned ab = blockIdx.x, rc = blockIdx.y, g = blockIdx.z;
    const unsigned nx = gridDim.x, ny = gridDim.y, total = nx * ny * gridDim.z;
    if (a.xcd_order && (total & 7u) == 0) {
        const unsigned l = blockIdx.x + nx * (blockIdx.y + ny * blockIdx.z);
        const unsigned m = (l & 7u) * (total >> 3) + (l >> 3);
        ab = m % nx;
        rc = (m / nx) % ny;
        g = m / (nx * ny);
    }
    proj_lds_item<EM, PF>(a, (int)ab, (int)rc, (int)g, sm);
}

// Persistent work-queue form of the staged projector over several volumes:
// item i of segment b is (angle chunk, ray chunk, slice group) =
// (l % nab, (l / nab) % RC, l / (nab RC)), l = i - first[b], of descs[b].
template <int EM>
__global__ __launch_bounds__(64 * PROJ_ANGLES, 4) void k_project_lds_queue(const ProjK* __restrict__ descs,
                                                                            ProjQueue qu)
{
    __shared__ ProjSmem sm;
    __shared__ int s_item;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0)
            s_item = __hip_atomic_fetch_add(qu.counter, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        const int item = __builtin_amdgcn_readfirstlane(s_item);
        if (item >= qu.first[qu.nseg]) break;
        int b = 0;
        while (b + 1 < qu.nseg && item >= qu.first[b + 1]) ++b;
        b = __builtin_amdgcn_readfirstlane(b);
        const int local = item - qu.first[b];
        const int nab = qu.nab[b], RC = descs[b].RC;
        proj_lds_item<EM, PL_PF>(descs[b], local % nab, (local / nab) % RC, local / (nab * RC), sm);
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

static ProjK proj_kargs(const ProjArgs& p)
{
    ProjK k;
    k.img = reinterpret_cast<const float4*>(p.img);
    k.imgT = reinterpret_cast<const float4*>(p.imgT);
    k.out = p.out; k.y = p.y; k.wa = p.wa;
    k.kl_part = p.kl_part; k.ex_part = p.ex_part; k.cosv = p.cosv; k.sinv = p.sinv;
    k.Nx = p.g.Nx; k.Ny = p.g.Ny; k.nrays = p.g.nrays; k.nangles = p.g.nangles; k.nk = p.g.nk; k.S = p.S;
    k.ci = p.g.ci; k.ck = p.g.ck; k.ucx = p.g.ucx; k.ucy = p.g.ucy; k.ry = p.g.ry; k.rx = p.g.rx;
    k.scale = p.g.scale;
    {
        const char* e = std::getenv("RAFT_PROJ_XCD_ORDER");
        k.xcd_order = (e && e[0] == '0') ? 0 : 1;
    }
    k.tlo = p.tlo; k.W = p.W;
    k.RC = p.lds_rc > 0 ? p.lds_rc : 1;
    k.kl_stride = p.g.nangles * k.RC;
    return k;
}

size_t proj_desc_size() { return sizeof(ProjK); }

void proj_desc_fill(const ProjArgs& p, void* dst)
{
    if (p.lds_rc <= 0) fail(RAFT_EINVAL, "internal: the projector queue runs the staged form only");
    const ProjK k = proj_kargs(p);
    std::memcpy(dst, &k, sizeof k);
}

int proj_angle_chunks(int nangles) { return (nangles + PROJ_ANGLES - 1) / PROJ_ANGLES; }

static const void* proj_queue_kernel(int em_mode)
{
    if (em_mode == 2) return reinterpret_cast<const void*>(&k_project_lds_queue<2>);
    if (em_mode == 1) return reinterpret_cast<const void*>(&k_project_lds_queue<1>);
    return reinterpret_cast<const void*>(&k_project_lds_queue<0>);
}

int proj_queue_grid(int em_mode)
{
    int dev = 0, cus = 0, per = 0;
    RAFT_HIP(hipGetDevice(&dev));
    RAFT_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    RAFT_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, proj_queue_kernel(em_mode), 64 * PROJ_ANGLES, 0));
    return std::max(1, cus * std::max(1, per));
}

void launch_project_queue(const void* d_descs, const ProjQueue& q, int em_mode, int grid, hipStream_t st)
{
    if (q.nseg <= 0 || q.first[q.nseg] <= 0) return;
    if (q.nseg > QUEUE_MAXSEG) fail(RAFT_EINVAL, "internal: too many queue segments");
    ProjQueue qq = q;
    const ProjK* descs = static_cast<const ProjK*>(d_descs);
    void* args[] = {&descs, &qq};
    RAFT_HIP(hipLaunchKernel(proj_queue_kernel(em_mode), dim3(std::max(1, std::min(grid, q.first[q.nseg]))),
                             dim3(64 * PROJ_ANGLES), args, 0, st));
}

void launch_project(const ProjArgs& p, int em_mode, hipStream_t st)
{
    if (p.S <= 0 || p.g.nrays <= 0 || p.g.nangles <= 0) return;
    ProjK k = proj_kargs(p);
    const int G = (p.S + 3) / 4;
    if (G > 65535) fail(RAFT_EUNSUPPORTED, "projector grid too large");
    const int nab = (p.g.nangles + PROJ_ANGLES - 1) / PROJ_ANGLES;
    const dim3 block(64 * PROJ_ANGLES);
    if (p.lds_rc > 0) {
        const dim3 grid(nab, k.RC, G);
        // the prefetch form by default; RAFT_PROJ_PREFETCH=0 selects the plain
        // one (same values bit for bit)
        const char* pfe = std::getenv("RAFT_PROJ_PREFETCH");
        const bool pf = !(pfe && pfe[0] == '0');
        if (pf) {
            if (em_mode == 2) hipLaunchKernelGGL((k_project_lds<2, PL_PF>), grid, block, 0, st, k);
            else if (em_mode) hipLaunchKernelGGL((k_project_lds<1, PL_PF>), grid, block, 0, st, k);
            else hipLaunchKernelGGL((k_project_lds<0, PL_PF>), grid, block, 0, st, k);
        } else {
            if (em_mode == 2) hipLaunchKernelGGL((k_project_lds<2, 0>), grid, block, 0, st, k);
            else if (em_mode) hipLaunchKernelGGL((k_project_lds<1, 0>), grid, block, 0, st, k);
            else hipLaunchKernelGGL((k_project_lds<0, 0>), grid, block, 0, st, k);
        }
    } else {
        const dim3 grid(nab, G);
        if (em_mode == 2) hipLaunchKernelGGL((k_project<2>), grid, block, 0, st, k);
        else if (em_mode) hipLaunchKernelGGL((k_project<1>), grid, block, 0, st, k);
        else hipLaunchKernelGGL((k_project<0>), grid, block, 0, st, k);
    }
    RAFT_HIP(hipGetLastError());
}

void proj_tile_consts(int* c4)
{
    c4[0] = PL_RAYS;
    c4[1] = PL_K;
    c4[2] = PROJ_ANGLES;
    c4[3] = PL_CAP;
}

int proj_plan(const RayGeom& g, const double* c, const double* s, int S, bool have_T)
{
    const char* env = std::getenv("RAFT_PROJ_LDS");
    if (env && env[0] == '0') return 0;
    const int A = g.nangles;
    if (A <= 0 || g.nrays <= 0 || S <= 0) return 0;
    // Worst union box of a chunk.  A wave's box at patch origin (ir0, kr0) is
    // [lo, hi) with lo = f(ir0, kr0) + min(0, 31 rI) + min(0, 31 rK) and
    // hi = f + max(0, 31 rI) + max(0, 31 rK), f affine (patch_box: the extremes
    // of the patch corners).  The union's width max_a hi_a - min_a lo_a is a
    // maximum of affine functions minus a minimum of affine functions, so it
    // is convex in the origin and largest at a corner of the origins' domain:
    // ray blocks ib in [0, last block], steps ks in [0, nk - 1].  The kernel's
    // box [floor(lo) - 1, floor(hi) + 3) spans at most floor(hi - lo) + 5; the
    // 1e-9 covers the rounding of the kernel's own FP64 evaluation.
    const double irs[2] = {-g.ci, (double)(((g.nrays - 1) / PL_RAYS) * PL_RAYS) - g.ci};
    const double krs[2] = {-g.ck, (double)(g.nk - 1) - g.ck};
    for (int a0 = 0; a0 < A; a0 += PROJ_ANGLES) {
        const int a1 = std::min(A, a0 + PROJ_ANGLES);
        const int am = std::min(a0 + PROJ_ANGLES / 2, A - 1);
        const bool steep = have_T && std::fabs(s[am]) > std::fabs(c[am]);
        double wr = 0.0, wc = 0.0;
        for (double ir : irs)
            for (double kr : krs) {
                double rlo = 1e300, rhi = -1e300, clo = 1e300, chi = -1e300;
                for (int a = a0; a < a1; ++a) {
                    const double bxk = -s[a] * g.ry, byi = s[a] * g.rx;
                    const double rI = steep ? c[a] : byi, rK = steep ? bxk : c[a];
                    const double cI = steep ? byi : c[a], cK = steep ? c[a] : bxk;
                    const double fr = ir * rI + kr * rK, fc = ir * cI + kr * cK;
                    const double er = (PL_RAYS - 1) * rI, ek = (PL_K - 1) * rK;
                    const double ec = (PL_RAYS - 1) * cI, el = (PL_K - 1) * cK;
                    rlo = std::min(rlo, fr + std::min(0.0, er) + std::min(0.0, ek));
                    rhi = std::max(rhi, fr + std::max(0.0, er) + std::max(0.0, ek));
                    clo = std::min(clo, fc + std::min(0.0, ec) + std::min(0.0, el));
                    chi = std::max(chi, fc + std::max(0.0, ec) + std::max(0.0, el));
                }
                wr = std::max(wr, rhi - rlo);
                wc = std::max(wc, chi - clo);
            }
        const double rows = std::floor(wr + 1e-9) + 5.0, cols = std::floor(wc + 1e-9) + 5.0;
        if (!(rows < 4096.0 && cols < 4096.0)) return 0;
        const int ic = (int)cols;
        const long long pitch = ic + ((1 - ic) & 3); // the kernel's narrowest row pitch
        if ((long long)rows * pitch > PL_CAP) return 0;
    }
    // ray chunks: enough workgroups to fill the device several times
    const long long nrb = (g.nrays + PL_RAYS - 1) / PL_RAYS;
    const long long wgs = (long long)((A + PROJ_ANGLES - 1) / PROJ_ANGLES) * ((S + 3) / 4);
    const long long rc = std::max(1LL, std::min(nrb, (2048 + wgs - 1) / wgs));
    return (int)rc;
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
