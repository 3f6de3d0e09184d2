// fft_filter.hip -- FBP row filter on gfx950: edge-replicate pad, FFT,
// multiply by the (Hermitian) filter, inverse FFT, window store.
//
// One workgroup filters two sinogram rows at once -- the same angle of two
// consecutive slices: the padded rows u0, u1 (length n = npad) form
// z = u0 + i u1 in LDS; a mixed-radix Stockham FFT (radices 8, 4, 2, 3, 5, 7,
// 11, 13; twiddles from an FP64-built table) runs entirely in LDS with the
// butterfly inputs held in registers across one barrier per pass.  Because
// the filter H is Hermitian (H_{n-k} = conj H_k, the spectrum of a real
// kernel), conj(FFT(conj(Z H))) returns q0 + i q1, the two filtered rows.  H
// carries 1/(npad * pixel) (filters.cu:184-186); the store multiplies by
// dtheta_a * angular scale and writes only the bin window the backprojector
// reads, zeros outside [0, n), into the slice-interleaved layout
// [group][angle][bin][SB] (SB = 4 for FP32 rows, 8 for FP16 rows).
//
// Padded sizes with a prime factor > 5 or n > 16384 use k_filter_direct:
// q[m] = sum_j g[(m - j) mod n] u[j] with g = IDFT(H) (built on the host in
// FP64), the same circular convolution.
//
// Reference behaviour restated: paddR2R (padding.cu:140-178), filterFBP /
// convolution_R2C_C2R_1D / fbp_filtering_R2C2R (filters.cu:33-51,68-87,163-190),
// Filter::apply (filters.cu:395-438).
#include "raft_internal.hpp"

#include <cstdlib>

namespace raft {

namespace {

__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(fmaf(a.x, b.x, -(a.y * b.y)), fmaf(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 mul_mi(float2 a) { return make_float2(a.y, -a.x); } // * (-i)

// Twiddles W_n^m = exp(-2 pi i m / n), m < n <= 16384, from two small tables
// kept in LDS: W^m = hi[m >> 7] * lo[m & 127] (one FP32 product of two
// FP64-built values; error ~2 ulp).
struct Tw {
    const float2* lo; // [128]
    const float2* hi; // [n / 128 + 1]
    __device__ __forceinline__ float2 operator()(int m) const { return cmul(hi[m >> 7], lo[m & 127]); }
};

// LDS index with one complex of padding per 16: Stockham passes with small
// strides (Ns = 1, 8) would otherwise put consecutive lanes on the same banks.
__device__ __forceinline__ int pix(int e) { return e + (e >> 4); }
__host__ __device__ constexpr size_t lds_elems(int n) { return (size_t)n + (size_t)(n >> 4) + 1; }

// in-register forward DFTs of size R (W_R = exp(-2 pi i / R))
template <int R> struct Dft;

template <> struct Dft<2> {
    __device__ static void run(float2* v, const Tw&, int)
    {
        float2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};

template <> struct Dft<3> {
    __device__ static void run(float2* v, const Tw&, int)
    {
        const float c1 = -0.5f, s1 = 0.86602540378443864676f; // sin(2 pi/3)
        float2 a = v[0], b = v[1], c = v[2];
        float2 sp = cadd(b, c), sm = csub(b, c);
        float2 m = make_float2(a.x + c1 * sp.x, a.y + c1 * sp.y);
        // W = c1 - i s1 : y1 = m - i s1 sm, y2 = m + i s1 sm
        float2 t = make_float2(s1 * sm.y, -s1 * sm.x); // -i s1 sm
        v[0] = cadd(a, sp);
        v[1] = cadd(m, t);
        v[2] = csub(m, t);
    }
};

template <> struct Dft<4> {
    __device__ static void run(float2* v, const Tw&, int)
    {
        float2 a = v[0], b = v[1], c = v[2], d = v[3];
        float2 s0 = cadd(a, c), s1 = csub(a, c), s2 = cadd(b, d), s3 = mul_mi(csub(b, d));
        v[0] = cadd(s0, s2);
        v[1] = cadd(s1, s3);
        v[2] = csub(s0, s2);
        v[3] = csub(s1, s3);
    }
};

template <> struct Dft<8> {
    __device__ static void run(float2* v, const Tw&, int)
    {
        const float h = 0.70710678118654752440f;
        float2 e[4] = {v[0], v[2], v[4], v[6]};
        float2 o[4] = {v[1], v[3], v[5], v[7]};
        Dft<4>::run(e, Tw{}, 0);
        Dft<4>::run(o, Tw{}, 0);
        // o[k] *= W8^k
        float2 o1 = make_float2(h * (o[1].x + o[1].y), h * (o[1].y - o[1].x)); // * h(1 - i)
        float2 o2 = mul_mi(o[2]);
        float2 o3 = make_float2(h * (o[3].y - o[3].x), -h * (o[3].x + o[3].y)); // * h(-1 - i)
        v[0] = cadd(e[0], o[0]);
        v[4] = csub(e[0], o[0]);
        v[1] = cadd(e[1], o1);
        v[5] = csub(e[1], o1);
        v[2] = cadd(e[2], o2);
        v[6] = csub(e[2], o2);
        v[3] = cadd(e[3], o3);
        v[7] = csub(e[3], o3);
    }
};

// odd prime radices: direct sum with W_R^m = W_n^(m n / R)
template <int R> struct DftGeneric {
    __device__ static void run(float2* v, const Tw& tw, int n)
    {
        float2 in[R];
#pragma unroll
        for (int q = 0; q < R; ++q) in[q] = v[q];
        const int st = n / R;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            float2 acc = in[0];
#pragma unroll
            for (int q = 1; q < R; ++q) acc = cadd(acc, cmul(in[q], tw(((q * r) % R) * st)));
            v[r] = acc;
        }
    }
};
template <> struct Dft<5> : DftGeneric<5> {};

// One Stockham pass (decimation in time, autosorting):
//   butterfly j (0 <= j < n/R), jm = j mod Ns:
//   in  x_r = z[j + r n/R] * W_{Ns R}^{jm r}   (W^{jm r} = (W^{jm})^r)
//   out z[(j - jm) R + jm + r Ns] = DFT_R(x)_r
// n / THREADS <= E, so a thread holds at most ceil(E/R) butterflies.
// Ns is a power of two for every pass before the first odd radix (radices
// are ordered 8, 4, 2 first), so jm is a mask there.
template <int R, int THREADS, int E>
__device__ __forceinline__ void fft_pass(float2* z, const Tw& tw, int n, int Ns, int twstep, unsigned nsm, int nss,
                                         int tid)
{
    constexpr int MAXB = (E + R - 1) / R;
    const int nb = n / R;
    const bool pow2 = (Ns & (Ns - 1)) == 0;
    // j mod Ns for j < nb: q = (j * nsm) >> nss equals j / Ns exactly (the host
    // picks nsm = ceil(2^nss / Ns) with 2^nss >= nb * Ns and checks every j),
    // in 24-bit full-rate multiplies instead of a 32-bit division
    auto jmod = [&](int j) __attribute__((always_inline)) {
        return pow2 ? (j & (Ns - 1)) : (int)((unsigned)j - __umul24(__umul24((unsigned)j, nsm) >> nss, (unsigned)Ns));
    };
    float2 v[MAXB][R];
#pragma unroll
    for (int b = 0; b < MAXB; ++b) {
        const int j = tid + b * THREADS;
        if (j < nb) {
            const int jm = jmod(j);
            // W^(r ts) = (W^ts)^r: one table twiddle per butterfly, the rest by
            // recurrence (at most R - 2 extra roundings)
            const float2 w1 = tw(jm * twstep);
            float2 w = w1;
            v[b][0] = z[pix(j)];
#pragma unroll
            for (int r = 1; r < R; ++r) {
                v[b][r] = cmul(z[pix(j + r * nb)], w);
                if (r + 1 < R) w = cmul(w, w1);
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < MAXB; ++b) {
        const int j = tid + b * THREADS;
        if (j < nb) {
            Dft<R>::run(v[b], tw, n);
            const int jm = jmod(j);
            const int d = (j - jm) * R + jm;
#pragma unroll
            for (int r = 0; r < R; ++r) z[pix(d + r * Ns)] = v[b][r];
        }
    }
    __syncthreads();
}

// Waves per SIMD asked of the 768-thread LDS FFT kernels (n in (4096, 6144]):
// six keep them within 80 VGPRs, so two workgroups share a CU (twice ~54 KB
// of LDS at n = 6144) and one workgroup's barriers overlap the other's passes.
// Other sizes fit several workgroups per CU at their natural register count.
template <int THREADS> constexpr int fft_wpe() { return THREADS == 768 ? 6 : 1; }

struct FftK {
    int n, nstages;
    int radix[24];
    // per stage: twiddle step n / (Ns R) and the exact-division pair for Ns
    int twstep[24];
    unsigned nsm[24];
    int nss[24];
};

template <int THREADS, int E>
__device__ void fft_run(float2* z, const Tw& tw, const FftK& p, int tid = threadIdx.x)
{
    int Ns = 1;
    for (int st = 0; st < p.nstages; ++st) {
        const int R = p.radix[st];
        const int ts = p.twstep[st], ss = p.nss[st];
        const unsigned sm = p.nsm[st];
        switch (R) {
        case 8: fft_pass<8, THREADS, E>(z, tw, p.n, Ns, ts, sm, ss, tid); break;
        case 4: fft_pass<4, THREADS, E>(z, tw, p.n, Ns, ts, sm, ss, tid); break;
        case 2: fft_pass<2, THREADS, E>(z, tw, p.n, Ns, ts, sm, ss, tid); break;
        case 3: fft_pass<3, THREADS, E>(z, tw, p.n, Ns, ts, sm, ss, tid); break;
        default: fft_pass<5, THREADS, E>(z, tw, p.n, Ns, ts, sm, ss, tid); break;
        }
        Ns *= R;
    }
}

struct FilterK {
    const float* tomo;
    void* q;
    const float2* H;
    const float2* tw; // [128 + n/128 + 1]: lo table then hi table
    const double* g; // direct-convolution kernel (FP64)
    const float* wa;
    int N, A, S, npad, L, tlo, W, none;
    int SB; // slices per interleaved group
    float* rows; // k_filter_fft_entry_lds only: non-null -> windows to rows [S][A][W] instead of q
};

// block -> (group, angle, pair): rows (s0 = g*SB + 2*pair, a) and (s0 + 1, a)
struct PairIdx {
    int g, a, s0;
};
__device__ __forceinline__ PairIdx pair_index(const FilterK& k)
{
    const int pairs = k.SB / 2;
    PairIdx r;
    const int b = blockIdx.x;
    const int pr = b % pairs;
    const int ga = b / pairs;
    r.a = ga % k.A;
    r.g = ga / k.A;
    r.s0 = r.g * k.SB + 2 * pr;
    return r;
}

// padded rows of slices s0, s0+1 (missing slices read as zero) into z
template <int THREADS>
__device__ __forceinline__ void load_pair(const FilterK& k, const PairIdx& ix, float2* z)
{
    const bool v0 = ix.s0 < k.S, v1 = ix.s0 + 1 < k.S;
    const float* r0 = k.tomo + ((size_t)ix.s0 * k.A + ix.a) * k.N;
    const float* r1 = r0 + (size_t)k.A * k.N;
    for (int m = threadIdx.x; m < k.npad; m += THREADS) {
        int ii = m - k.L;
        ii = ii < 0 ? 0 : (ii > k.N - 1 ? k.N - 1 : ii);
        z[pix(m)] = make_float2(v0 ? r0[ii] : 0.0f, v1 ? r1[ii] : 0.0f);
    }
}

template <typename OT> struct Pair2;
template <> struct Pair2<float> {
    using T = float2;
    __device__ static __forceinline__ T make(float a, float b) { return make_float2(a, b); }
};
template <> struct Pair2<__half> {
    using T = __half2;
    __device__ static __forceinline__ T make(float a, float b) { return __floats2half2_rn(a, b); }
};

// window store: bins [tlo, tlo + W) of both rows, times the angle weight
template <typename OT, int THREADS>
__device__ __forceinline__ void store_pair(const FilterK& k, const PairIdx& ix, const float2* z, float sign1)
{
    using P2 = typename Pair2<OT>::T;
    P2* dst = reinterpret_cast<P2*>(static_cast<OT*>(k.q) + (((size_t)ix.g * k.A + ix.a) * k.W) * k.SB +
                                    (ix.s0 - ix.g * k.SB));
    const int stride = k.SB / 2; // in pairs
    const float wv = k.wa[ix.a];
    for (int w = threadIdx.x; w < k.W; w += THREADS) {
        const int T = k.tlo + w;
        float v0 = 0.0f, v1 = 0.0f;
        if (T >= 0 && T < k.npad) {
            const float2 zv = z[pix(T)];
            v0 = zv.x * wv;
            v1 = sign1 * zv.y * wv;
        }
        dst[(size_t)w * stride] = Pair2<OT>::make(v0, v1);
    }
}

template <typename OT, int THREADS, int E>
__global__ __launch_bounds__(THREADS, fft_wpe<THREADS>()) void k_filter_fft(FilterK a, FftK p)
{
    extern __shared__ float2 z[]; // [lds_elems(npad)] padded data, then the twiddle tables
    float2* tws = z + lds_elems(a.npad);
    const int ntw = 128 + a.npad / 128 + 1;
    for (int k = threadIdx.x; k < ntw; k += THREADS) tws[k] = a.tw[k];
    const Tw tw{tws, tws + 128};
    const PairIdx ix = pair_index(a);
    load_pair<THREADS>(a, ix, z);
    __syncthreads();
    if (a.none) {
        store_pair<OT, THREADS>(a, ix, z, 1.0f);
        return;
    }
    fft_run<THREADS, E>(z, tw, p);
    for (int k = threadIdx.x; k < a.npad; k += THREADS) {
        const float2 v = cmul(z[pix(k)], a.H[k]);
        z[pix(k)] = make_float2(v.x, -v.y);
    }
    __syncthreads();
    fft_run<THREADS, E>(z, tw, p);
    store_pair<OT, THREADS>(a, ix, z, -1.0f);
}

// Whole-entry form: one workgroup filters every slice pair of one (slice
// group, angle) in turn through the same LDS and stores full 16-byte entries
// of the interleaved rows (SB slices per bin), so no 64-byte line is written
// in halves by two workgroups.  The filtered window of the earlier pairs
// waits in registers: J window bins per thread (the host picks this form
// when W <= J * THREADS).  The arithmetic of each pair is that of
// k_filter_fft, so both forms store the same values.
template <typename OT, int THREADS, int E, int J>
__global__ __launch_bounds__(THREADS) void k_filter_fft_entry(FilterK a, FftK p)
{
    constexpr int SB = sizeof(OT) == 4 ? 4 : 8;
    constexpr int PAIRS = SB / 2;
    extern __shared__ float2 z[];
    float2* tws = z + lds_elems(a.npad);
    const int ntw = 128 + a.npad / 128 + 1;
    for (int k = threadIdx.x; k < ntw; k += THREADS) tws[k] = a.tw[k];
    const Tw tw{tws, tws + 128};
    const int ang = blockIdx.x % a.A, g = blockIdx.x / a.A;
    const float wv = a.wa[ang];
    const float sign1 = a.none ? 1.0f : -1.0f;
    float2 hold[PAIRS][J];
#pragma unroll
    for (int pr = 0; pr < PAIRS; ++pr) {
        const PairIdx ix{g, ang, g * SB + 2 * pr};
        __syncthreads(); // the twiddles are in; the previous pair's window reads are done
        load_pair<THREADS>(a, ix, z);
        __syncthreads();
        if (!a.none) {
            fft_run<THREADS, E>(z, tw, p);
            for (int k = threadIdx.x; k < a.npad; k += THREADS) {
                const float2 v = cmul(z[pix(k)], a.H[k]);
                z[pix(k)] = make_float2(v.x, -v.y);
            }
            __syncthreads();
            fft_run<THREADS, E>(z, tw, p);
        }
#pragma unroll
        for (int j = 0; j < J; ++j) {
            const int T = a.tlo + (int)threadIdx.x + j * THREADS;
            float2 v = make_float2(0.0f, 0.0f);
            if ((int)threadIdx.x + j * THREADS < a.W && T >= 0 && T < a.npad) {
                const float2 zv = z[pix(T)];
                v = make_float2(zv.x * wv, sign1 * zv.y * wv);
            }
            hold[pr][j] = v;
        }
    }
    uint4* dst = reinterpret_cast<uint4*>(static_cast<OT*>(a.q)) + ((size_t)g * a.A + ang) * a.W;
#pragma unroll
    for (int j = 0; j < J; ++j) {
        const int w = (int)threadIdx.x + j * THREADS;
        if (w >= a.W) continue;
        uint4 e;
        if constexpr (SB == 4) {
            e = make_uint4(__float_as_uint(hold[0][j].x), __float_as_uint(hold[0][j].y), __float_as_uint(hold[1][j].x),
                           __float_as_uint(hold[1][j].y));
        } else {
            unsigned u[PAIRS];
#pragma unroll
            for (int q = 0; q < PAIRS; ++q) {
                const __half2 h = __floats2half2_rn(hold[q][j].x, hold[q][j].y);
                u[q] = *reinterpret_cast<const unsigned*>(&h);
            }
            e = make_uint4(u[0], u[1], u[2], u[3]);
        }
        dst[w] = e;
    }
}

// Whole-entry form for FP32 rows with the first pair's filtered window kept
// in LDS (after the twiddle tables) instead of registers: without the held
// window and with the register budget of WPE waves per SIMD (80 VGPRs at 768
// threads), two workgroups fit a CU (twice 77 KB of LDS at n = 6144), so one
// workgroup's barriers overlap the other's passes.  The arithmetic of each
// pair is that of k_filter_fft, so the stored entries are the same bit for
// bit.
template <int THREADS, int E, int WPE, bool ROWS>
__global__ __launch_bounds__(THREADS, WPE) void k_filter_fft_entry_lds(FilterK a, FftK p)
{
    extern __shared__ float2 z[];
    float2* tws = z + lds_elems(a.npad);
    const int ntw = 128 + a.npad / 128 + 1;
    float2* win = tws + ntw; // [W]: pair 0's window
    for (int k = threadIdx.x; k < ntw; k += THREADS) tws[k] = a.tw[k];
    const Tw tw{tws, tws + 128};
    const int ang = blockIdx.x % a.A, g = blockIdx.x / a.A;
    const float wv = a.wa[ang];
    const float sign1 = a.none ? 1.0f : -1.0f;
    uint4* dst = reinterpret_cast<uint4*>(static_cast<float*>(a.q)) + ((size_t)g * a.A + ang) * a.W;
#pragma unroll
    for (int pr = 0; pr < 2; ++pr) {
        const PairIdx ix{g, ang, g * 4 + 2 * pr};
        __syncthreads(); // the twiddles are in; the previous pair's window reads are done
        load_pair<THREADS>(a, ix, z);
        __syncthreads();
        if (!a.none) {
            fft_run<THREADS, E>(z, tw, p);
            for (int k = threadIdx.x; k < a.npad; k += THREADS) {
                const float2 v = cmul(z[pix(k)], a.H[k]);
                z[pix(k)] = make_float2(v.x, -v.y);
            }
            __syncthreads();
            fft_run<THREADS, E>(z, tw, p);
        }
        for (int w = threadIdx.x; w < a.W; w += THREADS) {
            const int T = a.tlo + w;
            float2 v = make_float2(0.0f, 0.0f);
            if (T >= 0 && T < a.npad) {
                const float2 zv = z[pix(T)];
                v = make_float2(zv.x * wv, sign1 * zv.y * wv);
            }
            if (pr == 0) {
                win[w] = v; // read back by the same thread
            } else if (ROWS) {
                // the group's four rows of this angle, each written by this
                // workgroup only and after all of them were read (in place
                // when rows is tomo)
                const float2 h = win[w];
                const int s0 = g * 4;
                const size_t ss = (size_t)a.A * a.W;
                float* r0 = a.rows + ((size_t)s0 * a.A + ang) * a.W + w;
                if (s0 < a.S) r0[0] = h.x;
                if (s0 + 1 < a.S) r0[ss] = h.y;
                if (s0 + 2 < a.S) r0[2 * ss] = v.x;
                if (s0 + 3 < a.S) r0[3 * ss] = v.y;
            } else {
                const float2 h = win[w];
                dst[w] = make_uint4(__float_as_uint(h.x), __float_as_uint(h.y), __float_as_uint(v.x),
                                    __float_as_uint(v.y));
            }
        }
    }
}

// direct circular convolution for sizes the LDS FFT does not cover:
// q[m] = sum_j g[(m - j) mod n] u[j]
template <typename OT>
__global__ __launch_bounds__(256) void k_filter_direct(FilterK a)
{
    extern __shared__ float2 z[]; // u0 + i u1; the filtered values go straight to the staged rows
    const PairIdx ix = pair_index(a);
    load_pair<256>(a, ix, z);
    __syncthreads();
    if (a.none) {
        store_pair<OT, 256>(a, ix, z, 1.0f);
        return;
    }
    using P2 = typename Pair2<OT>::T;
    P2* dst = reinterpret_cast<P2*>(static_cast<OT*>(a.q) + (((size_t)ix.g * a.A + ix.a) * a.W) * a.SB +
                                    (ix.s0 - ix.g * a.SB));
    const int stride = a.SB / 2;
    const float wv = a.wa[ix.a];
    for (int w = threadIdx.x; w < a.W; w += 256) {
        const int T = a.tlo + w;
        // FP64 kernel and sums: the kernel sums to ~0 (no DC), so n FP32
        // products would lose ~sqrt(n) ulps of a cancelling sum
        double s0 = 0.0, s1 = 0.0;
        if (T >= 0 && T < a.npad) {
            int d = T; // (T - j) mod n, j = 0..n-1
            for (int j = 0; j < a.npad; ++j) {
                const double gv = a.g[d];
                const float2 zj = z[pix(j)];
                s0 = fma(gv, (double)zj.x, s0);
                s1 = fma(gv, (double)zj.y, s1);
                d = (d == 0) ? a.npad - 1 : d - 1;
            }
        }
        dst[(size_t)w * stride] = Pair2<OT>::make((float)s0 * wv, (float)s1 * wv);
    }
}

template <typename OT, int THREADS, int E>
void launch_fft_t(const FilterK& k, const FftK& p, int blocks, size_t lds, hipStream_t st)
{
    // whole-entry form when the window fits 4 bins per thread (every padding
    // of 2 or more at the benchmark sizes); blocks = (groups x angles) x pairs
    if (k.W <= 4 * THREADS && !std::getenv("RAFT_FILTER_PAIRS")) {
        const void* fn = reinterpret_cast<const void*>(&k_filter_fft_entry<OT, THREADS, E, 4>);
        RAFT_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL((k_filter_fft_entry<OT, THREADS, E, 4>), dim3(blocks / (k.SB / 2)), dim3(THREADS), lds, st,
                           k, p);
        return;
    }
    RAFT_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_filter_fft<OT, THREADS, E>),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL((k_filter_fft<OT, THREADS, E>), dim3(blocks), dim3(THREADS), lds, st, k, p);
}

template <int THREADS, int WPE, bool ROWS>
void launch_fft_lds_entry_t(const FilterK& k, const FftK& p, int blocks, size_t lds, hipStream_t st)
{
    const void* fn = reinterpret_cast<const void*>(&k_filter_fft_entry_lds<THREADS, 8, WPE, ROWS>);
    RAFT_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL((k_filter_fft_entry_lds<THREADS, 8, WPE, ROWS>), dim3(blocks / 2), dim3(THREADS), lds, st, k,
                       p);
}

// the rows form is its own instantiation, so the staged-row form's code is
// unchanged by it
template <int THREADS, int WPE>
void launch_fft_lds_entry(const FilterK& k, const FftK& p, int blocks, size_t lds, hipStream_t st)
{
    if (k.rows) launch_fft_lds_entry_t<THREADS, WPE, true>(k, p, blocks, lds, st);
    else launch_fft_lds_entry_t<THREADS, WPE, false>(k, p, blocks, lds, st);
}

// THREADS = the smallest of 64, 128, 192, 256, 384, 512, 768, 1024 that is
// >= n/8, with E = 8 elements per thread; n in (8192, 16384] runs 1024
// threads with E = 16
template <typename OT>
bool launch_fft(const FilterK& k, const FftK& p, int threads, int blocks, size_t lds, hipStream_t st)
{
    // FP32 rows, n in (2048, 6144]: the two-workgroup form when two of them
    // fit a CU's 160 KB of LDS (RAFT_FFT_FORM=regs keeps the window in
    // registers instead: one workgroup per CU at n = 6144)
    const char* fe = std::getenv("RAFT_FFT_FORM");
    const size_t lw = lds + sizeof(float2) * (size_t)k.W;
    if (sizeof(OT) == 4 && p.n > 2048 && p.n <= 6144 && 2 * lw <= 160 * 1024 && !(fe && fe[0] == 'r') &&
        !std::getenv("RAFT_FILTER_PAIRS")) {
        if (p.n <= 3072) launch_fft_lds_entry<384, 3>(k, p, blocks, lw, st);
        else launch_fft_lds_entry<768, 6>(k, p, blocks, lw, st);
        return k.rows != nullptr;
    }
    if (p.n > 8192) {
        launch_fft_t<OT, 1024, 16>(k, p, blocks, lds, st);
        return false;
    }
    switch (threads) {
    case 64: launch_fft_t<OT, 64, 8>(k, p, blocks, lds, st); break;
    case 128: launch_fft_t<OT, 128, 8>(k, p, blocks, lds, st); break;
    case 192: launch_fft_t<OT, 192, 8>(k, p, blocks, lds, st); break;
    case 256: launch_fft_t<OT, 256, 8>(k, p, blocks, lds, st); break;
    case 384: launch_fft_t<OT, 384, 8>(k, p, blocks, lds, st); break;
    case 512: launch_fft_t<OT, 512, 8>(k, p, blocks, lds, st); break;
    case 768: launch_fft_t<OT, 768, 8>(k, p, blocks, lds, st); break;
    default: launch_fft_t<OT, 1024, 8>(k, p, blocks, lds, st); break;
    }
    return false;
}

// ---------------------------------------------------------------------------
// FBP by the Backprojection Slice Theorem (bst.cu:22-140, 251-312): mirrored
// half-rows of the filtered, padded projections -> C2C FFT -> sigma weights ->
// polar-to-Cartesian gather on the frequency grid -> 2-D inverse FFT.
// ---------------------------------------------------------------------------
struct BstPolarK {
    const float* q; // filtered padded rows, interleaved [G][A][n][4]
    float2* polar;  // [S][2A][n/2]
    const float2* tw;
    int n, A, S;
};

// row r of slice sl: r < A -> q[sl][r][n/2 - 1 - idx], r >= A -> q[sl][r - A][n/2 + idx],
// idx < n/2, weight 4 (2 at idx 0), zeros for idx >= n/2; forward FFT; times
// sigma_k = 2 / (n max(min(k, n - k), 1/2)).  Both rows of an angle are real,
// so one complex FFT carries the two: z = left + i right, and with
// W_k = conj(Z_{(n - k) mod n}), LEFT_k = (Z_k + W_k) / 2, RIGHT_k = (Z_k - W_k) / 2i.
template <int THREADS, int E>
__global__ __launch_bounds__(THREADS, fft_wpe<THREADS>()) void k_bst_polar(BstPolarK a, FftK p)
{
    extern __shared__ float2 z[];
    float2* tws = z + lds_elems(a.n);
    const int ntw = 128 + a.n / 128 + 1;
    for (int k = threadIdx.x; k < ntw; k += THREADS) tws[k] = a.tw[k];
    const Tw tw{tws, tws + 128};
    const int sl = blockIdx.x / a.A, ang = blockIdx.x % a.A;
    const int center = a.n / 2;
    const float* row = a.q + ((size_t)(sl / 4) * a.A + ang) * a.n * 4 + (sl % 4);
    for (int idx = threadIdx.x; idx < a.n; idx += THREADS) {
        float vl = 0.0f, vr = 0.0f;
        if (idx < a.n / 2) {
            const float w = idx == 0 ? 2.0f : 4.0f;
            vl = row[(size_t)(center - 1 - idx) * 4] * w;
            vr = row[(size_t)(center + idx) * 4] * w;
        }
        z[pix(idx)] = make_float2(vl, vr);
    }
    __syncthreads();
    fft_run<THREADS, E>(z, tw, p);
    // only radii [0, n/2) are ever interpolated (irho + 1 < n/2 - 1 + 1): rows
    // of n/2 entries
    const int h = a.n / 2;
    float2* outl = a.polar + ((size_t)sl * 2 * a.A + ang) * h;
    float2* outr = outl + (size_t)a.A * h;
    for (int k = threadIdx.x; k < h; k += THREADS) {
        const float sig = 2.0f / ((float)a.n * fmaxf(fminf((float)k, (float)(a.n - k)), 0.5f));
        const float2 zk = z[pix(k)];
        const float2 zm = z[pix(k == 0 ? 0 : a.n - k)];
        const float hs = 0.5f * sig;
        outl[k] = make_float2((zk.x + zm.x) * hs, (zk.y - zm.y) * hs);
        outr[k] = make_float2((zk.y + zm.y) * hs, (zm.x - zk.x) * hs);
    }
}

// One polar transform of k_bst_filter_polar: FFT of z, then the two real
// transforms it carries, separated and times sigma, into row offset `row` of
// the first slice (pa) and, if present, of the second (pb)
template <int THREADS, int E>
__device__ __forceinline__ void bst_pair_rows(float2* z, const Tw& tw, const FftK& p, float2* pa, float2* pb,
                                              size_t row, int h, int n, bool two)
{
    fft_run<THREADS, E>(z, tw, p);
    for (int k = threadIdx.x; k < h; k += THREADS) {
        const float sig = 2.0f / ((float)n * fmaxf(fminf((float)k, (float)(n - k)), 0.5f));
        const float2 zk = z[pix(k)];
        const float2 zm = z[pix(k == 0 ? 0 : n - k)];
        const float hs = 0.5f * sig;
        pa[row + k] = make_float2((zk.x + zm.x) * hs, (zk.y - zm.y) * hs);
        if (two) pb[row + k] = make_float2((zk.y + zm.y) * hs, (zm.x - zk.x) * hs);
    }
    __syncthreads();
}

// Filter and polar FFTs in one pass (the default BST form): one workgroup =
// one slice pair (s0, s0 + 1) of one angle.  The pair's padded rows are
// filtered as k_filter_fft does (one complex FFT carries both real rows), and
// the filtered values q_a = Re z wv, q_b = sign Im z wv -- the values the
// staged rows would hold -- go straight into the polar transforms, so the
// filtered rows are never written.  The polar rows of both slices come from
// two complex FFTs: z1 = w (q_a + i q_b) over the left half-rows (mirrored)
// and z2 = w (q_a + i q_b) over the right half-rows; the two real transforms
// each carries are separated by Hermitian symmetry as in k_bst_polar (first
// slice: (Z_k + conj Z_-k) / 2, second: (Z_k - conj Z_-k) / 2i), times sigma.
// The right half-rows wait in LDS (rh, n/2 entries) while z1 is transformed.
template <int THREADS, int E>
__global__ __launch_bounds__(THREADS, fft_wpe<THREADS>()) void k_bst_filter_polar(FilterK a, BstPolarK b, FftK p)
{
    extern __shared__ float2 z[];
    const int n = a.npad;
    float2* tws = z + lds_elems(n);
    const int ntw = 128 + n / 128 + 1;
    float2* rh = tws + ntw;
    for (int k = threadIdx.x; k < ntw; k += THREADS) tws[k] = a.tw[k];
    const Tw tw{tws, tws + 128};
    const PairIdx ix = pair_index(a);
    // the second pair of a last group with one slice has no slice: it writes
    // nothing (uniform for the workgroup, before any barrier)
    if (ix.s0 >= a.S) return;
    load_pair<THREADS>(a, ix, z);
    __syncthreads();
    float sign1 = 1.0f;
    if (!a.none) {
        fft_run<THREADS, E>(z, tw, p);
        for (int k = threadIdx.x; k < n; k += THREADS) {
            const float2 v = cmul(z[pix(k)], a.H[k]);
            z[pix(k)] = make_float2(v.x, -v.y);
        }
        __syncthreads();
        fft_run<THREADS, E>(z, tw, p);
        sign1 = -1.0f;
    }
    const int c = n / 2, h = n / 2;
    const float wv = a.wa[ix.a];
    // mirrored left half-rows through registers back into z (zeros past h),
    // right half-rows into rh
    {
        constexpr int JL = (E + 1) / 2; // h <= E THREADS / 2
        float2 lv[JL];
#pragma unroll
        for (int j = 0; j < JL; ++j) {
            const int idx = threadIdx.x + j * THREADS;
            lv[j] = make_float2(0.0f, 0.0f);
            if (idx < h) {
                const float w = idx == 0 ? 2.0f : 4.0f;
                const float2 zl = z[pix(c - 1 - idx)], zr = z[pix(c + idx)];
                lv[j] = make_float2(zl.x * wv * w, sign1 * zl.y * wv * w);
                rh[idx] = make_float2(zr.x * wv * w, sign1 * zr.y * wv * w);
            }
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < JL; ++j) {
            const int idx = threadIdx.x + j * THREADS;
            if (idx < h) z[pix(idx)] = lv[j];
        }
    }
    for (int idx = h + threadIdx.x; idx < n; idx += THREADS) z[pix(idx)] = make_float2(0.0f, 0.0f);
    __syncthreads();
    const bool two = ix.s0 + 1 < a.S;
    float2* pa = b.polar + (size_t)ix.s0 * 2 * b.A * h;
    float2* pb = pa + (size_t)2 * b.A * h;
    bst_pair_rows<THREADS, E>(z, tw, p, pa, pb, (size_t)ix.a * h, h, n, two);
    for (int idx = threadIdx.x; idx < n; idx += THREADS) z[pix(idx)] = idx < h ? rh[idx] : make_float2(0.0f, 0.0f);
    __syncthreads();
    bst_pair_rows<THREADS, E>(z, tw, p, pa, pb, (size_t)(b.A + ix.a) * h, h, n, two);
}

// unnormalised inverse C2C FFT of `rows` rows of length n, in place
template <int THREADS, int E>
__global__ __launch_bounds__(THREADS, fft_wpe<THREADS>()) void k_ifft_rows(float2* data, const float2* twg, FftK p)
{
    extern __shared__ float2 z[];
    const int n = p.n;
    float2* tws = z + lds_elems(n);
    const int ntw = 128 + n / 128 + 1;
    for (int k = threadIdx.x; k < ntw; k += THREADS) tws[k] = twg[k];
    const Tw tw{tws, tws + 128};
    float2* row = data + (size_t)blockIdx.x * n;
    for (int k = threadIdx.x; k < n; k += THREADS) {
        const float2 v = row[k];
        z[pix(k)] = make_float2(v.x, -v.y);
    }
    __syncthreads();
    fft_run<THREADS, E>(z, tw, p);
    for (int k = threadIdx.x; k < n; k += THREADS) {
        const float2 v = z[pix(k)];
        row[k] = make_float2(v.x, -v.y);
    }
}

// polar (2A rows x n/2 radii) -> Cartesian frequency grid s x s (bst.cu:63-99),
// index arithmetic in FP64; the reference's weights (tfrac, 1 - tfrac) with
// tfrac = iarc - angle are kept as they are
__global__ void k_bst_p2c(const float2* polar, float2* cart, int n, int A2, int s, int S)
{
    const long long total = (long long)S * s * s;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const int tx = (int)(e % s), ty = (int)((e / s) % s);
        const long long sl = e / ((long long)s * s);
        const int posx = tx - s / 2, posy = ty - s / 2;
        const double rho = (double)n * hypot((double)posx, (double)posy) / (double)s;
        const double angle = (double)A2 * (0.5 * atan2((double)posy, (double)posx) / M_PI + 0.5);
        const long long irho = (long long)rho;
        int iarc = (int)angle;
        float2 val = make_float2(0.0f, 0.0f);
        if (irho < n / 2 - 1) {
            const float pfrac = (float)(rho - (double)irho);
            const float tfrac = (float)((double)iarc - angle);
            iarc = iarc % A2;
            const int uarc = (iarc + 1) % A2;
            const int h = n / 2; // polar rows hold radii [0, n/2)
            const float2* P = polar + sl * (long long)A2 * h;
            const float2 a0 = P[(size_t)iarc * h + irho], a1 = P[(size_t)iarc * h + irho + 1];
            const float2 b0 = P[(size_t)uarc * h + irho], b1 = P[(size_t)uarc * h + irho + 1];
            const float2 i0 = make_float2(a0.x * (1.0f - pfrac) + a1.x * pfrac, a0.y * (1.0f - pfrac) + a1.y * pfrac);
            const float2 i1 = make_float2(b0.x * (1.0f - pfrac) + b1.x * pfrac, b0.y * (1.0f - pfrac) + b1.y * pfrac);
            val = make_float2(i0.x * tfrac + i1.x * (1.0f - tfrac), i0.y * tfrac + i1.y * (1.0f - tfrac));
        }
        const float sg = ((tx & 1) == (ty & 1)) ? 1.0f : -1.0f; // 4 (tx%2 - 1/2)(ty%2 - 1/2)
        cart[(sl * s + (ty + s / 2) % s) * (long long)s + (tx + s / 2) % s] = make_float2(val.x * sg, val.y * sg);
    }
}

// Fused form of k_bst_p2c and the inverse FFT of the grid rows (k_ifft_rows):
// one workgroup per destination row tyd of the shifted frequency grid and
// group of up to BST_SPB slices.  The FP64 polar coordinates of the row's
// cells do not depend on the slice and are computed once; for each slice the
// interpolated cells go straight into LDS, the row is inverse-transformed
// there (conjugate, forward FFT, conjugate), and only the N columns
// [L, L + N) the crop keeps are written, as rows of out [S][s][N].  The grid
// is never written to HBM.  The expressions are those of the two unfused
// kernels, so the values are the same bit for bit.
constexpr int BST_SPB = 8;
struct BstRowsK {
    const float2* polar; // [S][A2][n/2]
    float2* out;         // [S][s][N]
    int n, A2, s, S, N, L;
};

template <int THREADS, int E, int EC>
__global__ __launch_bounds__(THREADS) void k_bst_rows(BstRowsK a, const float2* twg, FftK p)
{
    extern __shared__ float2 z[];
    const int n = a.n, s = a.s;
    float2* tws = z + lds_elems(n);
    const int ntw = 128 + n / 128 + 1;
    for (int k = threadIdx.x; k < ntw; k += THREADS) tws[k] = twg[k];
    const Tw tw{tws, tws + 128};
    const int tyd = blockIdx.x;
    const int sl0 = blockIdx.y * BST_SPB, sl1 = min(a.S, sl0 + BST_SPB);
    const int ty = (tyd + s - s / 2) % s; // (ty + s/2) % s == tyd
    // the cells txd = threadIdx.x + j THREADS of this row (THREADS EC >= s), in
    // three registers each: oc = polar offset of (iarc, irho), with bit 30 set
    // when uarc wraps to 0 (-1: zero cell or past s); ps = the radial fraction
    // with the cell's checkerboard sign in its sign bit; tf = the angular one
    int oc[EC];
    float ps[EC], tf[EC];
    const int h = n / 2;               // polar rows hold radii [0, n/2)
    const int wrapb = (a.A2 - 1) * h; // (iarc, irho) -> (0, irho)
#pragma unroll
    for (int j = 0; j < EC; ++j) {
        const int txd = threadIdx.x + j * THREADS;
        oc[j] = -1;
        ps[j] = tf[j] = 0.0f;
        if (txd < s) {
            const int tx = (txd + s - s / 2) % s;
            const int posx = tx - s / 2, posy = ty - s / 2;
            const double rho = (double)n * hypot((double)posx, (double)posy) / (double)s;
            const double angle = (double)a.A2 * (0.5 * atan2((double)posy, (double)posx) / M_PI + 0.5);
            const long long irho = (long long)rho;
            int iarc = (int)angle;
            float pf = 0.0f;
            if (irho < n / 2 - 1) {
                pf = (float)(rho - (double)irho);
                tf[j] = (float)((double)iarc - angle);
                iarc = iarc % a.A2;
                oc[j] = (iarc * h + (int)irho) | (iarc + 1 == a.A2 ? (1 << 30) : 0);
            }
            ps[j] = ((tx & 1) == (ty & 1)) ? pf : -pf;
        }
    }
    for (int sl = sl0; sl < sl1; ++sl) {
        const float2* P = a.polar + (size_t)sl * a.A2 * h;
        __syncthreads(); // the previous slice's row is read (and the twiddles are in)
#pragma unroll
        for (int j = 0; j < EC; ++j) {
            const int txd = threadIdx.x + j * THREADS;
            if (txd < s) {
                float2 val = make_float2(0.0f, 0.0f);
                const float pfrac = fabsf(ps[j]);
                if (oc[j] >= 0) {
                    const int o0 = oc[j] & ((1 << 30) - 1);
                    const int o1 = (oc[j] >> 30) ? o0 - wrapb : o0 + h;
                    const float tfrac = tf[j];
                    const float2 a0 = P[o0], a1 = P[o0 + 1];
                    const float2 b0 = P[o1], b1 = P[o1 + 1];
                    const float2 i0 = make_float2(a0.x * (1.0f - pfrac) + a1.x * pfrac,
                                                  a0.y * (1.0f - pfrac) + a1.y * pfrac);
                    const float2 i1 = make_float2(b0.x * (1.0f - pfrac) + b1.x * pfrac,
                                                  b0.y * (1.0f - pfrac) + b1.y * pfrac);
                    val = make_float2(i0.x * tfrac + i1.x * (1.0f - tfrac), i0.y * tfrac + i1.y * (1.0f - tfrac));
                }
                const float sg = __builtin_signbit(ps[j]) ? -1.0f : 1.0f;
                const float2 v = make_float2(val.x * sg, val.y * sg);
                z[pix(txd)] = make_float2(v.x, -v.y);
            }
        }
        __syncthreads();
        fft_run<THREADS, E>(z, tw, p);
        for (int c = threadIdx.x; c < a.N; c += THREADS) {
            const float2 v = z[pix(a.L + c)];
            a.out[((size_t)sl * s + tyd) * a.N + c] = make_float2(v.x, -v.y);
        }
    }
}

// out[sl][c][y] = in[sl][y][c0 + c] for the nc columns [c0, c0 + nc) of an
// s-row slice of row pitch ip: the columns the crop keeps, as rows for the
// second 1-D pass (32 x 32 LDS tiles)
__global__ __launch_bounds__(256) void k_transpose_c(const float2* in, float2* out, int s, int c0, int nc, int ip)
{
    __shared__ float2 t[32][33];
    const int sl = blockIdx.z;
    const int x0 = blockIdx.x * 32, y0 = blockIdx.y * 32;
    const float2* src = in + (size_t)sl * s * ip;
    float2* dst = out + (size_t)sl * nc * s;
    for (int k = threadIdx.x; k < 32 * 32; k += 256) {
        const int dy = k / 32, dx = k % 32;
        if (y0 + dy < s && x0 + dx < nc) t[dy][dx] = src[(size_t)(y0 + dy) * ip + c0 + x0 + dx];
    }
    __syncthreads();
    for (int k = threadIdx.x; k < 32 * 32; k += 256) {
        const int dx = k / 32, dy = k % 32;
        if (y0 + dy < s && x0 + dx < nc) dst[(size_t)(x0 + dx) * s + y0 + dy] = t[dy][dx];
    }
}

// obj[sl][j][i] = Re(R[sl][i][j + L]) / scale, R the inverse FFT along the
// rows of the kept columns ([sl][N][s])
__global__ void k_bst_extract(const float2* R, float* obj, int s, int N, int L, float scale, int S)
{
    const long long total = (long long)S * N * N;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(e % N), j = (int)((e / N) % N);
        const long long sl = e / ((long long)N * N);
        obj[e] = R[(sl * N + i) * (long long)s + j + L].x / scale;
    }
}

template <int THREADS, int E>
void launch_bst_polar_t(const BstPolarK& k, const FftK& p, size_t lds, hipStream_t st)
{
    RAFT_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_bst_polar<THREADS, E>),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL((k_bst_polar<THREADS, E>), dim3((unsigned)((long long)k.S * k.A)), dim3(THREADS), lds, st, k,
                       p);
}

template <int THREADS, int E>
void launch_bst_filter_polar_t(const FilterK& f, const BstPolarK& k, const FftK& p, size_t lds, hipStream_t st)
{
    RAFT_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_bst_filter_polar<THREADS, E>),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    const long long blocks = (long long)((f.S + f.SB - 1) / f.SB) * f.A * (f.SB / 2);
    hipLaunchKernelGGL((k_bst_filter_polar<THREADS, E>), dim3((unsigned)blocks), dim3(THREADS), lds, st, f, k, p);
}

template <int THREADS, int E, int EC = E>
void launch_bst_rows_t(const BstRowsK& k, const float2* tw, const FftK& p, size_t lds, hipStream_t st)
{
    RAFT_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_bst_rows<THREADS, E, EC>),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL((k_bst_rows<THREADS, E, EC>), dim3((unsigned)k.s, (unsigned)((k.S + BST_SPB - 1) / BST_SPB)),
                       dim3(THREADS), lds, st, k, tw, p);
}

template <int THREADS, int E>
void launch_ifft_t(float2* data, const float2* tw, const FftK& p, long long rows, size_t lds, hipStream_t st)
{
    RAFT_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_ifft_rows<THREADS, E>),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL((k_ifft_rows<THREADS, E>), dim3((unsigned)rows), dim3(THREADS), lds, st, data, tw, p);
}

FftK fftk_of(const FftPlan& plan)
{
    FftK p;
    p.n = plan.n;
    p.nstages = plan.nstages;
    for (int i = 0; i < 24; ++i) {
        p.radix[i] = plan.radix[i];
        p.twstep[i] = 0;
        p.nsm[i] = 0;
        p.nss[i] = 0;
    }
    // stage st works on Ns = radix[0] * .. * radix[st-1] and j < nb = n / radix[st]:
    // nsm = ceil(2^nss / Ns) with 2^nss >= nb * Ns makes (j * nsm) >> nss == j / Ns
    // for every j < nb (the error j * (nsm Ns - 2^nss) / (Ns 2^nss) stays below the
    // gap 1/Ns); nsm <= 2 nb + 1 < 2^24 and j * nsm < 2^32, as the 24-bit multiplies need
    long long Ns = 1;
    for (int st = 0; st < p.nstages; ++st) {
        const int R = p.radix[st];
        const long long nb = p.n / R;
        p.twstep[st] = (int)(p.n / (Ns * R));
        int k = 0;
        while ((1LL << k) < nb * Ns) ++k;
        const long long m = ((1LL << k) + Ns - 1) / Ns;
        if (m >= (1LL << 24) || (nb - 1) * m >= (1LL << 32) || nb >= (1LL << 24))
            fail(RAFT_EUNSUPPORTED, "internal: FFT stage index out of the 24-bit range");
        if ((Ns & (Ns - 1)) != 0)
            for (long long j = 0; j < nb; ++j)
                if ((long long)(((unsigned long long)j * (unsigned long long)m) >> k) != j / Ns)
                    fail(RAFT_EUNSUPPORTED, "internal: FFT stage division check failed");
        p.nsm[st] = (unsigned)m;
        p.nss[st] = k;
        Ns *= R;
    }
    return p;
}

} // namespace

FftPlan fft_plan(int n)
{
    FftPlan p;
    p.n = n;
    if (n < 2 || n > 16384) return p;
    int m = n;
    static const int order[] = {8, 4, 2, 3, 5};
    for (int r : order) {
        while (m % r == 0 && m > 1) {
            if (p.nstages >= 24) return p;
            p.radix[p.nstages++] = r;
            m /= r;
        }
    }
    if (m != 1) return p;
    // threads = n / 8 when that is an allowed block size, so every radix-8
    // butterfly of a pass has its own thread (6144 -> 768, 3072 -> 384)
    static const int tsz[] = {64, 128, 192, 256, 384, 512, 768, 1024};
    int t = 1024;
    for (int c : tsz)
        if (c * 8 >= n) {
            t = c;
            break;
        }
    p.threads = t;
    p.ok = true;
    return p;
}

bool launch_fbp_filter(const FilterArgs& a, const FftPlan& plan, int store, hipStream_t st)
{
    if (a.S <= 0 || a.A <= 0) return a.rows != nullptr && store == RAFT_STORE_F32;
    FilterK k;
    k.tomo = a.tomo; k.q = a.q; k.H = a.H; k.tw = a.tw; k.g = a.g; k.wa = a.wa;
    k.N = a.N; k.A = a.A; k.S = a.S; k.npad = a.npad; k.L = a.L; k.tlo = a.tlo; k.W = a.W; k.none = a.none;
    k.SB = (store == RAFT_STORE_F16) ? 8 : 4;
    k.rows = store == RAFT_STORE_F32 ? a.rows : nullptr;
    bool direct = false;
    const long long groups = (a.S + k.SB - 1) / k.SB;
    const long long blocks = groups * a.A * (k.SB / 2);
    if (blocks > 0x7fffffffLL) fail(RAFT_EUNSUPPORTED, "too many sinogram rows for one filter launch");
    const size_t lds = sizeof(float2) * lds_elems(a.npad);
    const size_t lds_fft = lds + sizeof(float2) * (size_t)(128 + a.npad / 128 + 1);
    if (plan.ok || a.none) {
        const FftK p = fftk_of(plan);
        const int threads = plan.ok ? plan.threads : 256;
        if (store == RAFT_STORE_F16) launch_fft<__half>(k, p, threads, (int)blocks, lds_fft, st);
        else direct = launch_fft<float>(k, p, threads, (int)blocks, lds_fft, st);
    } else {
        if (lds > 160 * 1024) fail(RAFT_EUNSUPPORTED, "padded row too long for the direct filter path");
        if (store == RAFT_STORE_F16) {
            RAFT_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_filter_direct<__half>),
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
            hipLaunchKernelGGL((k_filter_direct<__half>), dim3((int)blocks), dim3(256), lds, st, k);
        } else {
            RAFT_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_filter_direct<float>),
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
            hipLaunchKernelGGL((k_filter_direct<float>), dim3((int)blocks), dim3(256), lds, st, k);
        }
    }
    RAFT_HIP(hipGetLastError());
    return direct;
}

bool bst_filter_polar_fits(int n)
{
    return sizeof(float2) * (lds_elems(n) + (size_t)(128 + n / 128 + 1) + (size_t)(n / 2)) <= 160 * 1024;
}

void launch_bst(const BstArgs& a, const FftPlan& plan, hipStream_t st, const FilterArgs* filt)
{
    if (a.S <= 0) return;
    if (!plan.ok) fail(RAFT_EUNSUPPORTED, "BST needs a padded size n = N(1+pad) <= 16384 with factors 2, 3, 5");
    const int n = plan.n, s = n;
    const FftK p = fftk_of(plan);
    const size_t lds = sizeof(float2) * (lds_elems(n) + (size_t)(128 + n / 128 + 1));
    if ((long long)a.S * 2 * a.A > 0x7fffffffLL || (long long)a.S * s > 0x7fffffffLL)
        fail(RAFT_EUNSUPPORTED, "BST block too large for one launch");
    BstPolarK k{a.q, a.polar, a.tw, n, a.A, a.S};
    if (filt) {
        // filter + polar FFTs in one pass from the sinogram (k_bst_filter_polar)
        if (filt->npad != n || filt->S != a.S || filt->A != a.A)
            fail(RAFT_EINVAL, "BST filter arguments do not match the transform");
        FilterK f;
        f.tomo = filt->tomo; f.q = nullptr; f.H = filt->H; f.tw = filt->tw; f.g = nullptr; f.wa = filt->wa;
        f.N = filt->N; f.A = filt->A; f.S = filt->S; f.npad = n; f.L = filt->L; f.tlo = 0; f.W = n;
        f.none = filt->none;
        f.SB = 4;
        f.rows = nullptr;
        if ((long long)((f.S + f.SB - 1) / f.SB) * f.A * (f.SB / 2) > 0x7fffffffLL)
            fail(RAFT_EUNSUPPORTED, "too many sinogram rows for one BST launch");
        const size_t lds_fp = lds + sizeof(float2) * (size_t)(n / 2);
        if (lds_fp > 160 * 1024) fail(RAFT_EUNSUPPORTED, "padded row too long for the BST filter pass");
        if (n > 8192) launch_bst_filter_polar_t<1024, 16>(f, k, p, lds_fp, st);
        else {
            switch (plan.threads) {
            case 64: launch_bst_filter_polar_t<64, 8>(f, k, p, lds_fp, st); break;
            case 128: launch_bst_filter_polar_t<128, 8>(f, k, p, lds_fp, st); break;
            case 192: launch_bst_filter_polar_t<192, 8>(f, k, p, lds_fp, st); break;
            case 256: launch_bst_filter_polar_t<256, 8>(f, k, p, lds_fp, st); break;
            case 384: launch_bst_filter_polar_t<384, 8>(f, k, p, lds_fp, st); break;
            case 512: launch_bst_filter_polar_t<512, 8>(f, k, p, lds_fp, st); break;
            case 768: launch_bst_filter_polar_t<768, 8>(f, k, p, lds_fp, st); break;
            default: launch_bst_filter_polar_t<1024, 8>(f, k, p, lds_fp, st); break;
            }
        }
    } else if (n > 8192) launch_bst_polar_t<1024, 16>(k, p, lds, st);
    else {
        switch (plan.threads) {
        case 64: launch_bst_polar_t<64, 8>(k, p, lds, st); break;
        case 128: launch_bst_polar_t<128, 8>(k, p, lds, st); break;
        case 192: launch_bst_polar_t<192, 8>(k, p, lds, st); break;
        case 256: launch_bst_polar_t<256, 8>(k, p, lds, st); break;
        case 384: launch_bst_polar_t<384, 8>(k, p, lds, st); break;
        case 512: launch_bst_polar_t<512, 8>(k, p, lds, st); break;
        case 768: launch_bst_polar_t<768, 8>(k, p, lds, st); break;
        default: launch_bst_polar_t<1024, 8>(k, p, lds, st); break;
        }
    }
    RAFT_HIP(hipGetLastError());
    if (a.fused) {
        // polar -> Cartesian rows, inverse-transformed in LDS, N kept columns
        // per row -> cart [S][s][N]; then transposed to cartT [S][N][s]
        const BstRowsK rk{a.polar, a.cart, n, 2 * a.A, s, a.S, a.N, a.L};
        if (n > 8192) launch_bst_rows_t<1024, 16>(rk, a.tw, p, lds, st);
        else {
            switch (plan.threads) {
            case 64: launch_bst_rows_t<64, 8>(rk, a.tw, p, lds, st); break;
            case 128: launch_bst_rows_t<128, 8>(rk, a.tw, p, lds, st); break;
            case 192: launch_bst_rows_t<192, 8>(rk, a.tw, p, lds, st); break;
            case 256: launch_bst_rows_t<256, 8>(rk, a.tw, p, lds, st); break;
            case 384: launch_bst_rows_t<384, 8>(rk, a.tw, p, lds, st); break;
            case 512: launch_bst_rows_t<512, 8>(rk, a.tw, p, lds, st); break;
            case 768: launch_bst_rows_t<768, 8>(rk, a.tw, p, lds, st); break;
            default: launch_bst_rows_t<1024, 8>(rk, a.tw, p, lds, st); break;
            }
        }
        RAFT_HIP(hipGetLastError());
        dim3 g((a.N + 31) / 32, (s + 31) / 32, a.S);
        hipLaunchKernelGGL(k_transpose_c, g, dim3(256), 0, st, a.cart, a.cartT, s, 0, a.N, a.N);
        RAFT_HIP(hipGetLastError());
    } else {
        const long long cells = (long long)a.S * s * s;
        hipLaunchKernelGGL(k_bst_p2c, dim3((unsigned)std::min<long long>((cells + 255) / 256, 65536)), dim3(256), 0,
                           st, a.polar, a.cart, n, 2 * a.A, s, a.S);
        RAFT_HIP(hipGetLastError());
    }
    // 2-D inverse FFT: all rows (unfused form), then only the N columns
    // [L, L + N) the crop keeps (transposed to rows); the other columns are
    // never needed
    for (int pass = a.fused ? 1 : 0; pass < 2; ++pass) {
        float2* d = pass == 0 ? a.cart : a.cartT;
        const long long rows = (long long)a.S * (pass == 0 ? s : a.N);
        if (n > 8192) launch_ifft_t<1024, 16>(d, a.tw, p, rows, lds, st);
        else {
            switch (plan.threads) {
            case 64: launch_ifft_t<64, 8>(d, a.tw, p, rows, lds, st); break;
            case 128: launch_ifft_t<128, 8>(d, a.tw, p, rows, lds, st); break;
            case 192: launch_ifft_t<192, 8>(d, a.tw, p, rows, lds, st); break;
            case 256: launch_ifft_t<256, 8>(d, a.tw, p, rows, lds, st); break;
            case 384: launch_ifft_t<384, 8>(d, a.tw, p, rows, lds, st); break;
            case 512: launch_ifft_t<512, 8>(d, a.tw, p, rows, lds, st); break;
            case 768: launch_ifft_t<768, 8>(d, a.tw, p, rows, lds, st); break;
            default: launch_ifft_t<1024, 8>(d, a.tw, p, rows, lds, st); break;
            }
        }
        RAFT_HIP(hipGetLastError());
        if (pass == 0) {
            dim3 g((a.N + 31) / 32, (s + 31) / 32, a.S);
            hipLaunchKernelGGL(k_transpose_c, g, dim3(256), 0, st, a.cart, a.cartT, s, a.L, a.N, s);
            RAFT_HIP(hipGetLastError());
        }
    }
    const long long outn = (long long)a.S * a.N * a.N;
    hipLaunchKernelGGL(k_bst_extract, dim3((unsigned)std::min<long long>((outn + 255) / 256, 65536)), dim3(256), 0,
                       st, a.cartT, a.obj, s, a.N, a.L, a.scale, a.S);
    RAFT_HIP(hipGetLastError());
}

} // namespace raft
