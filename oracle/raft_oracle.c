/*
 * raft_oracle.c -- CPU restatement of ssc-raft's parallel-beam hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle (and, built
 * in FP32, the timed CPU baseline).  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it, and only as the checker.  The
 * product library (ssc-raft_amd/csrc) never links or calls it.
 *
 * Compiled twice from this one source (see oracle/Makefile):
 *   REAL=double -> symbols *_f64 : the FP64 oracle every parity test uses;
 *   REAL=float  -> symbols *_f32 : FP32 restatement, timed as the CPU baseline.
 * Bin indices are always evaluated in FP64 (both builds), with the exact
 * expressions written next to each function below.  Build with
 * -ffp-contract=off so that the compiler does not fuse or re-associate them;
 * every fused multiply-add in an index expression is an explicit fma().
 *
 * Reference files restated (paths relative to the reference checkout):
 *   pad            cuda/src/common/padding.cu:140-178 (paddR2R)
 *   crop           cuda/src/common/padding.cu:245-268 (remove_paddR2R)
 *   filter         cuda/src/processing/filters/filters.cu:33-51,68-87,163-190
 *                  (fbp_filtering_R2C2R, convolution_R2C_C2R_1D, filterFBP),
 *                  filters.cu:395-438 (Filter::apply), inc/processing/filters.hpp:21-47
 *   backprojection cuda/src/geometries/parallel/reconstruction_methods/fbp.cu:8-76
 *                  (BackProjection_SS), fbp.cu:81-122 (getFBP), fbp.cu:127-236
 *   parameters     reconstruction_methods/parameters.cu:131-219 (setFBPParameters),
 *                  parameters.cu:8-59 (setEMRTParameters)
 *   Radon          cuda/src/geometries/parallel/forward_methods/radon_rt.cu:131-174
 *                  (Radon_RT_version_sscRadon), radon_rt.cu:178-295
 *   eEM            reconstruction_methods/em_rt.cu:316-355 (get_eEM_RT),
 *                  reconstruction_methods/kernels/em_RT_kernel.cu:188-245
 *                  (kernel_radonWithDivision), :249-299
 *                  (kernel_backprojectionWithUpdate), :303-350
 *                  (kernel_backprojectionOfOnes), inc/common/configs.hpp:31-32
 *                  (SIGN, APPROXINVX)
 *   KL objective   not in the reference; defined here (see oracle_kl).
 *
 * The reference's cuFFT R2C/C2R pair (filters.cu:80,84) is restated as the
 * unnormalised DFT definition: forward e^{-2 pi i k m / n}; C2R = real part of
 * the inverse over the Hermitian extension of bins [0, n/2], the imaginary part
 * of the Nyquist bin ignored.
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "raft_oracle.h"

#ifndef REAL
#define REAL double
#endif
#ifndef SFX
#define SFX(name) name##_f64
#endif
typedef REAL real;

/* ------------------------------------------------------------------------- */
/* small parallel-for over [0, n) on nthreads pthreads                        */
/* ------------------------------------------------------------------------- */
typedef void (*range_fn)(void* ctx, long lo, long hi);
typedef struct {
    range_fn fn;
    void* ctx;
    long n;
    int nt, id;
} par_job;

static void* par_entry(void* p)
{
    par_job* j = (par_job*)p;
    long lo = j->n * j->id / j->nt, hi = j->n * (j->id + 1) / j->nt;
    if (hi > lo) j->fn(j->ctx, lo, hi);
    return NULL;
}

static void parallel_for(long n, int nthreads, range_fn fn, void* ctx)
{
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    if (nthreads > n) nthreads = (int)(n > 0 ? n : 1);
    if (nthreads == 1) { fn(ctx, 0, n); return; }
    pthread_t th[256];
    par_job jobs[256];
    for (int t = 0; t < nthreads; ++t) {
        jobs[t].fn = fn; jobs[t].ctx = ctx; jobs[t].n = n; jobs[t].nt = nthreads; jobs[t].id = t;
        pthread_create(&th[t], NULL, par_entry, &jobs[t]);
    }
    for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
}

/* ------------------------------------------------------------------------- */
/* geometry shared by FBP and eEM                                            */
/* ------------------------------------------------------------------------- */

/* dtheta_a = theta_{a+1} - theta_a, last = theta_{A-1} - theta_{A-2}
 * (fbp.cu:60-64, em_RT_kernel.cu:281-285).  Float32 angles promoted to double:
 * the differences are exact. */
static double angle_step(const float* angles, int nangles, int a)
{
    if (a == nangles - 1) return (double)angles[a] - (double)angles[a - 1];
    return (double)angles[a + 1] - (double)angles[a];
}

/* Angular correction of fbp.cu:24-28: the comparison is the float32 one the
 * reference makes (angles[A-1] > float(M_PI)); the value is pi/theta_last in
 * double. */
static double fbp_angle_scale(const float* angles, int nangles)
{
    float last = angles[nangles - 1];
    if (last > (float)M_PI) return M_PI / (double)last;
    return 1.0;
}

/* ------------------------------------------------------------------------- */
/* FFT (mixed radix, recursive decimation in time, generic prime butterflies) */
/* ------------------------------------------------------------------------- */
typedef struct { real re, im; } cpx;

typedef struct {
    int n;
    int nf;
    int factors[64];
    cpx* tw; /* tw[t] = exp(-2 pi i t / n) */
} fft_plan;

static int fft_plan_init(fft_plan* p, int n)
{
    p->n = n;
    p->nf = 0;
    int m = n;
    static const int small[3] = {4, 2, 3};
    for (int s = 0; s < 3; ++s)
        while (m % small[s] == 0 && m > 1) { p->factors[p->nf++] = small[s]; m /= small[s]; }
    for (int f = 5; m > 1; f += 2)
        while (m % f == 0) { p->factors[p->nf++] = f; m /= f; }
    p->tw = (cpx*)malloc(sizeof(cpx) * (size_t)n);
    if (!p->tw) return -1;
    for (int t = 0; t < n; ++t) {
        double ang = -2.0 * M_PI * (double)t / (double)n;
        p->tw[t].re = (real)cos(ang);
        p->tw[t].im = (real)sin(ang);
    }
    return 0;
}

static void fft_plan_free(fft_plan* p) { free(p->tw); p->tw = NULL; }

/* out[0..len) = DFT_len(in[0], in[stride], ...), forward sign.
 * twstride = n_full / len, fi = index of the first factor of len. */
static void fft_rec(const fft_plan* P, const cpx* in, cpx* out, int len, long stride, int fi, cpx* scratch)
{
    if (len == 1) { out[0] = in[0]; return; }
    int p = P->factors[fi];
    int m = len / p;
    for (int q = 0; q < p; ++q) fft_rec(P, in + q * stride, out + q * m, m, stride * p, fi + 1, scratch);
    int twstride = P->n / len;
    /* butterflies: X[k + r m] = sum_q W_len^{q k} W_p^{q r} Y_q[k] */
    cpx tmp[64];
    cpx* t = (p <= 64) ? tmp : scratch;
    for (int k = 0; k < m; ++k) {
        for (int q = 0; q < p; ++q) {
            cpx y = out[q * m + k];
            long ti = ((long)q * k * twstride) % P->n;
            cpx w = P->tw[ti];
            t[q].re = y.re * w.re - y.im * w.im;
            t[q].im = y.re * w.im + y.im * w.re;
        }
        for (int r = 0; r < p; ++r) {
            real sr = 0, si = 0;
            for (int q = 0; q < p; ++q) {
                long ti = ((long)q * r % p) * (P->n / p);
                cpx w = P->tw[ti];
                sr += t[q].re * w.re - t[q].im * w.im;
                si += t[q].re * w.im + t[q].im * w.re;
            }
            out[k + r * m].re = sr;
            out[k + r * m].im = si;
        }
    }
}

/* forward DFT of x (length n) into X, unnormalised */
static void fft_forward(const fft_plan* P, const cpx* x, cpx* X, cpx* scratch)
{
    fft_rec(P, x, X, P->n, 1, 0, scratch);
}

/* ------------------------------------------------------------------------- */
/* FBP parameter decode (parameters.cu:131-219) and filter table               */
/* ------------------------------------------------------------------------- */
#define PLANK 4.135667696E-15 /* eV s, configs.hpp */
#define VC 299792458.0        /* m/s, configs.hpp */

typedef struct {
    int nrays, nangles, nslices, objsize, pad, filter;
    int npad, L, M;
    double reg, paganin, offset, pixel;
} fbp_cfg;

static int fbp_decode(fbp_cfg* c, const float* paramf, const int* parami)
{
    c->nrays = parami[0];
    c->nangles = parami[1];
    c->nslices = parami[2];
    c->objsize = parami[3];
    c->pad = parami[4];
    c->filter = parami[7];
    if (c->nrays < 2 || c->nangles < 2 || c->nslices < 0 || c->pad < 0) return -1;
    if (c->objsize != c->nrays) return -1; /* reference Python always passes objsize = nrays */
    if (c->filter < 0 || c->filter > 7) return -1;
    c->npad = c->nrays * (1 + c->pad);
    c->L = (c->pad * c->nrays) / 2;
    c->M = c->npad / 2 + 1;
    double bd = (double)paramf[0];
    c->reg = (double)paramf[1];
    double energy = (double)paramf[2];
    double z2 = (double)paramf[3];
    c->pixel = (double)paramf[4];
    c->offset = (double)paramf[6];
    double lambda = (bd == 0.0) ? 1.0 : (PLANK * VC) / energy;
    c->paganin = lambda * z2 * M_PI * bd / (c->pixel * c->pixel);
    return 0;
}

/* Filter::apply (filters.cu:395-438) in double.  The reference's float
 * literals are used at their float32 values. */
static double filter_apply(int type, double w, double reg, double pag)
{
    double pg = 1.0 / (1.0 + pag * w * w);
    switch (type) {
    case 1: return w * exp(-(double)0.693f * reg * w * w) * pg;
    case 2: return w * (1.0 / (1.0 + reg * w * w)) * pg;
    case 3: return w * cos(M_PI * 0.5 * w) * pg;
    case 4: {
        double param = fmax(w * reg * M_PI * 0.5, (double)1E-4f);
        return w * (sin(param) / param) * pg;
    }
    case 5: return w * (0.5 + 0.5 * cos(2.0 * M_PI * w)) * pg;
    case 6: return w * ((double)0.54f + (double)0.46f * cos(2.0 * M_PI * w)) * pg;
    case 7: return w * pg;
    default: return 1.0;
    }
}

/* Full-length Hermitian filter H[0..n) including the 1/(npad*pixel) scale
 * (filters.cu:184-186):
 *   k in [0, M):  H_k = exp(-i 2 pi c k / (2M-2)) * apply(0.5 min(k, M-k) / M)
 *   k in [M, n):  H_k = conj(H_{n-k})
 *   n even:       H_{n/2} keeps only its real part (C2R ignores Im of Nyquist). */
static void fbp_filter_table(const fbp_cfg* c, double* Hre, double* Him)
{
    int n = c->npad, M = c->M;
    double scale = 1.0 / ((double)n * c->pixel);
    for (int k = 0; k < M; ++k) {
        double w = 0.5 * (double)(k < M - k ? k : M - k) / (double)M;
        double a = filter_apply(c->filter, w, c->reg, c->paganin);
        double ph = -2.0 * M_PI * c->offset * (double)k / (double)(2 * M - 2);
        Hre[k] = a * cos(ph) * scale;
        Him[k] = a * sin(ph) * scale;
    }
    for (int k = M; k < n; ++k) { Hre[k] = Hre[n - k]; Him[k] = -Him[n - k]; }
    if (n % 2 == 0) Him[n / 2] = 0.0;
}

/* ------------------------------------------------------------------------- */
/* pad + filter stage: q[s][a][0..npad)                                        */
/* ------------------------------------------------------------------------- */
typedef struct {
    const fbp_cfg* c;
    const float* tomo;
    real* q;
    const double *Hre, *Him;
    fft_plan plan;
} filt_ctx;

static void filter_rows(void* vctx, long lo, long hi)
{
    filt_ctx* X = (filt_ctx*)vctx;
    const fbp_cfg* c = X->c;
    int n = c->npad, N = c->nrays, L = c->L;
    cpx* z = (cpx*)malloc(sizeof(cpx) * n * 3);
    cpx* Z = z + n;
    cpx* scratch = z + 2 * n;
    for (long row = lo; row < hi; ++row) {
        const float* p = X->tomo + (size_t)row * N;
        real* q = X->q + (size_t)row * n;
        /* edge-replicate pad: u[m] = p[clamp(m - L, 0, N-1)] (padding.cu:140-178) */
        for (int m = 0; m < n; ++m) {
            int ii = m - L;
            if (ii < 0) ii = 0;
            if (ii > N - 1) ii = N - 1;
            z[m].re = (real)p[ii];
            z[m].im = 0;
        }
        if (c->filter == 0) { /* 'none': no filtering and no scaling (fbp.cu:102-103) */
            for (int m = 0; m < n; ++m) q[m] = z[m].re;
            continue;
        }
        fft_forward(&X->plan, z, Z, scratch);
        /* multiply by H, then inverse via conj(DFT(conj(.))) */
        for (int k = 0; k < n; ++k) {
            real hr = (real)X->Hre[k], hi_ = (real)X->Him[k];
            real re = Z[k].re * hr - Z[k].im * hi_;
            real im = Z[k].re * hi_ + Z[k].im * hr;
            Z[k].re = re;
            Z[k].im = -im;
        }
        fft_forward(&X->plan, Z, z, scratch);
        for (int m = 0; m < n; ++m) q[m] = z[m].re; /* real part; 1/(npad h) is in H */
    }
    free(z);
}

static int fbp_filter_stage(const fbp_cfg* c, const float* tomo, real* q, int nthreads)
{
    int n = c->npad;
    filt_ctx X;
    X.c = c; X.tomo = tomo; X.q = q;
    double* H = (double*)malloc(sizeof(double) * 2 * (size_t)n);
    if (!H) return -2;
    X.Hre = H; X.Him = H + n;
    fbp_filter_table(c, H, H + n);
    if (fft_plan_init(&X.plan, n) != 0) { free(H); return -2; }
    parallel_for((long)c->nslices * c->nangles, nthreads, filter_rows, &X);
    fft_plan_free(&X.plan);
    free(H);
    return 0;
}

int SFX(oracle_fbp_filter)(real* q, const float* tomo, const float* paramf, const int* parami, int nthreads)
{
    fbp_cfg c;
    if (fbp_decode(&c, paramf, parami) != 0) return -1;
    return fbp_filter_stage(&c, tomo, q, nthreads);
}

/* ------------------------------------------------------------------------- */
/* FBP backprojection, crop only                                              */
/*   half = npad/2 (double);  xr = (i + L) - half;  yr = (j + L) - half;      */
/*   u = fma(xr, cos t, fma(-yr, sin t, half));  T = (int)u  (trunc)          */
/*   f[j][i] = scale * sum_a [0 <= T < npad] q_a[T] dtheta_a                  */
/* This is fbp.cu:38-74 with x = xmin + I dx, t = x cos - y sin,              */
/* T = (int)((t - tmin)/dt): the pixel size cancels, I = i + L, J = j + L.    */
/* ------------------------------------------------------------------------- */
typedef struct {
    const fbp_cfg* c;
    const real* q;
    real* obj;
    const double *cs, *sn, *dth;
    double scale;
} bp_ctx;

static void fbp_bp_rows(void* vctx, long lo, long hi)
{
    bp_ctx* X = (bp_ctx*)vctx;
    const fbp_cfg* c = X->c;
    int N = c->nrays, A = c->nangles, n = c->npad, L = c->L;
    double half = (double)n / 2.0;
    for (long r = lo; r < hi; ++r) {
        long s = r / N;
        int j = (int)(r % N);
        const real* qs = X->q + (size_t)s * A * n;
        real* out = X->obj + ((size_t)s * N + j) * N;
        double yr = (double)(j + L) - half;
        for (int i = 0; i < N; ++i) {
            double xr = (double)(i + L) - half;
            real acc = 0;
            for (int a = 0; a < A; ++a) {
                double u = fma(xr, X->cs[a], fma(-yr, X->sn[a], half));
                int T = (int)u;
                if (T > -1 && T < n) acc += qs[(size_t)a * n + T] * (real)X->dth[a];
            }
            out[i] = acc * (real)X->scale;
        }
    }
}

/* Padded grid then crop, the way the reference runs it (fbp.cu:38-74 over the
 * whole npad x npad grid, padding.cu:245-268 crop), with the pixel size kept:
 *   x = -h npad/2 + I h, t = x cos - y sin, T = (int)((t + h npad/2) / h).
 * Only for small sizes: it exists to pin the crop-only form above. */
static void fbp_bp_padded_rows(void* vctx, long lo, long hi)
{
    bp_ctx* X = (bp_ctx*)vctx;
    const fbp_cfg* c = X->c;
    int N = c->nrays, A = c->nangles, n = c->npad, L = c->L;
    double h = c->pixel;
    double tmin = -h * (double)n / 2.0;
    for (long r = lo; r < hi; ++r) {
        long s = r / n;
        int J = (int)(r % n);
        const real* qs = X->q + (size_t)s * A * n;
        double y = tmin + (double)J * h;
        for (int I = 0; I < n; ++I) {
            double x = tmin + (double)I * h;
            real acc = 0;
            for (int a = 0; a < A; ++a) {
                double t = x * X->cs[a] - y * X->sn[a];
                int T = (int)((t - tmin) / h);
                if (T > -1 && T < n) acc += qs[(size_t)a * n + T] * (real)X->dth[a];
            }
            int ii = I - L, jj = J - L; /* remove_paddR2R */
            if (ii >= 0 && ii < N && jj >= 0 && jj < N) X->obj[((size_t)s * N + jj) * N + ii] = acc * (real)X->scale;
        }
    }
}

static int fbp_run(real* obj, const float* tomo, const float* angles, const float* paramf, const int* parami,
                   int nthreads, int padded)
{
    fbp_cfg c;
    if (fbp_decode(&c, paramf, parami) != 0) return -1;
    int A = c.nangles, n = c.npad;
    real* q = (real*)malloc(sizeof(real) * (size_t)c.nslices * A * n);
    double* tab = (double*)malloc(sizeof(double) * 3 * (size_t)A);
    if (!q || !tab) { free(q); free(tab); return -2; }
    int rc = fbp_filter_stage(&c, tomo, q, nthreads);
    if (rc == 0) {
        bp_ctx X;
        X.c = &c; X.q = q; X.obj = obj;
        X.cs = tab; X.sn = tab + A; X.dth = tab + 2 * A;
        for (int a = 0; a < A; ++a) {
            tab[a] = cos((double)angles[a]);
            tab[A + a] = sin((double)angles[a]);
            tab[2 * A + a] = angle_step(angles, A, a);
        }
        X.scale = fbp_angle_scale(angles, A);
        if (padded)
            parallel_for((long)c.nslices * n, nthreads, fbp_bp_padded_rows, &X);
        else
            parallel_for((long)c.nslices * c.nrays, nthreads, fbp_bp_rows, &X);
    }
    free(q);
    free(tab);
    return rc;
}

int SFX(oracle_fbp)(real* obj, const float* tomo, const float* angles, const float* paramf, const int* parami,
                    int nthreads)
{
    return fbp_run(obj, tomo, angles, paramf, parami, nthreads, 0);
}

int SFX(oracle_fbp_padded)(real* obj, const float* tomo, const float* angles, const float* paramf,
                           const int* parami, int nthreads)
{
    return fbp_run(obj, tomo, angles, paramf, parami, nthreads, 1);
}

/* ------------------------------------------------------------------------- */
/* ray-sampled line sums (Radon-RT and the eEM forward operator)              */
/*                                                                           */
/* Ray i at angle theta, samples k = 0..nk-1 (radon_rt.cu:150-173):           */
/*   t = -ax + i dx,  s_k = -ay + k dy,  x = t cos - s_k sin,                  */
/*   y = t sin + s_k cos,  X = (int)((x+1)/dx),  Y = (int)((y+1)/dy),          */
/*   dx = 2 ax/(Nx-1), dy = 2 ay/(Ny-1).                                       */
/* Written about the rotation centre (t = 0 at i = ci = (Nx-1)/2, s = 0 at     */
/* k = ck = (Ny-1)/2), the bin coordinates are                                 */
/*   ux = (i-ci) cos - (k-ck) sin (dy/dx) + 1/dx                               */
/*   uy = (i-ci) sin (dx/dy) + (k-ck) cos + 1/dy                               */
/* and are evaluated in double exactly as                                     */
/*   ux = fma(k-ck, -sin*ry, fma(i-ci, cos, ucx))                              */
/*   uy = fma(k-ck,  cos,    fma(i-ci, sin*rx, ucy))                           */
/* with ucx = (Nx-1)/(2 ax), ucy = (Ny-1)/(2 ay), ry = dy/dx and rx = dx/dy    */
/* written from the integer sizes (ry = rx = 1 exactly for a square image      */
/* with ax = ay).  At theta = 0, and at the centre sample of the centre ray,  */
/* every coordinate sits exactly on a bin edge; this form is exact there, as  */
/* exact arithmetic is.  X = (int)ux, Y = (int)uy (truncation); the sample    */
/* counts iff 0<=X<Nx and 0<=Y<Ny.                                             */
/* ------------------------------------------------------------------------- */
typedef struct {
    int Nx, Ny, nrays, nangles, nk;
    double ci, ck, ucx, ucy, ry, rx, dy;
} ray_geom;

/* Radon geometry: nk = Ny samples per ray over an Ny x Nx image */
static void ray_geom_radon(ray_geom* g, int Nx, int Ny, double ax, double ay)
{
    g->Nx = Nx; g->Ny = Ny; g->nk = Ny;
    g->ci = (double)(Nx - 1) / 2.0;
    g->ck = (double)(Ny - 1) / 2.0;
    g->ucx = (double)(Nx - 1) / (2.0 * ax);
    g->ucy = (double)(Ny - 1) / (2.0 * ay);
    g->ry = (ay * (double)(Nx - 1)) / (ax * (double)(Ny - 1));
    g->rx = (ax * (double)(Ny - 1)) / (ay * (double)(Nx - 1));
    g->dy = 2.0 * ay / (double)(Ny - 1);
}

/* eEM forward geometry (em_RT_kernel.cu:202-222): a = 1, dx = dy = dt =
 * 2/(nrays-1), nrays samples per ray, image n x n */
static void ray_geom_em(ray_geom* g, int nrays, int n)
{
    g->Nx = n; g->Ny = n; g->nk = nrays;
    g->ci = (double)(nrays - 1) / 2.0;
    g->ck = g->ci;
    g->ucx = g->ci;
    g->ucy = g->ci;
    g->ry = 1.0;
    g->rx = 1.0;
    g->dy = 2.0 / (double)(nrays - 1);
}

static real ray_sum(const ray_geom* g, const real* img, double cs, double sn, int i)
{
    double ir = (double)i - g->ci;
    double bxk = -sn * g->ry;
    double byi = sn * g->rx;
    double uxi = fma(ir, cs, g->ucx);
    double uyi = fma(ir, byi, g->ucy);
    real acc = 0;
    for (int k = 0; k < g->nk; ++k) {
        double kr = (double)k - g->ck;
        int X = (int)fma(kr, bxk, uxi);
        int Y = (int)fma(kr, cs, uyi);
        if (X >= 0 && X < g->Nx && Y >= 0 && Y < g->Ny) acc += img[(size_t)Y * g->Nx + X];
    }
    return acc;
}

typedef struct {
    const ray_geom* g;
    const float* img_f; /* float input (Radon) */
    const real* img_r;  /* real input (EM)     */
    real* out;
    const double *cs, *sn;
    real* tmp_img;      /* per-slice promoted image for float input */
} radon_ctx;

static void radon_rows(void* vctx, long lo, long hi)
{
    radon_ctx* X = (radon_ctx*)vctx;
    const ray_geom* g = X->g;
    for (long r = lo; r < hi; ++r) {
        long s = r / g->nangles;
        int a = (int)(r % g->nangles);
        const real* img = X->img_r + (size_t)s * g->Nx * g->Ny;
        real* out = X->out + ((size_t)s * g->nangles + a) * g->nrays;
        for (int i = 0; i < g->nrays; ++i) out[i] = ray_sum(g, img, X->cs[a], X->sn[a], i) * (real)g->dy;
    }
}

/* getRadonRTMultiGPU(..., sizeImagex=Nx, sizeImagey=Ny, nrays, nangles, nslices, ax, ay)
 * proj[s][a][i] = dy * sum_k phantom[s][Y][X]   (radon_rt.cu:131-174).
 * The Python wrapper's *nrays*pixel/2 (radon_methods.py:58) is not applied here. */
int SFX(oracle_radon)(real* proj, const float* phantom, const float* angles, int Nx, int Ny, int nrays, int nangles,
                      int nslices, float ax, float ay, int nthreads)
{
    if (Nx < 2 || Ny < 2 || nrays < 1 || nangles < 1 || nslices < 0) return -1;
    ray_geom g;
    ray_geom_radon(&g, Nx, Ny, (double)ax, (double)ay);
    g.nrays = nrays; g.nangles = nangles;
    size_t npx = (size_t)nslices * Nx * Ny;
    real* img = (real*)malloc(sizeof(real) * (npx ? npx : 1));
    double* tab = (double*)malloc(sizeof(double) * 2 * (size_t)nangles);
    if (!img || !tab) { free(img); free(tab); return -2; }
    for (size_t p = 0; p < npx; ++p) img[p] = (real)phantom[p];
    for (int a = 0; a < nangles; ++a) { tab[a] = cos((double)angles[a]); tab[nangles + a] = sin((double)angles[a]); }
    radon_ctx X;
    X.g = &g; X.img_f = phantom; X.img_r = img; X.out = proj; X.cs = tab; X.sn = tab + nangles;
    parallel_for((long)nslices * nangles, nthreads, radon_rows, &X);
    free(img);
    free(tab);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* eEM (emission MLEM) with ray-tracing operators                              */
/*                                                                           */
/* dt = 2/(nrays-1), dxy = 2/(n-1) (n = objsize):                              */
/*   forward  v[a][r] = dt * sum_{k<nrays} x[Y][X]   (ray sums above with       */
/*            ax = ay = 1, dx = dy = dt, Nx = Ny = n)   em_RT_kernel.cu:188-231 */
/*   ratio    rr = y * SIGN(v) / sqrt(e^2 + v^2), e = 1e-4f; rr<0 -> 0  (:233-238) */
/*   backprojection, pixel (i,j):  x_i = -1 + i dxy, y_j = -1 + j dxy,         */
/*            u = (x_i cos + y_j sin + 1)/dt; about the image centre           */
/*            c = (n-1)/2 this is (i-c) cos rr + (j-c) sin rr + (nrays-1)/2,   */
/*            rr = dxy/dt = (nrays-1)/(n-1), evaluated as                      */
/*            u = fma(i-c, cos*rr, fma(j-c, sin*rr, (nrays-1)/2)), T = (int)u  */
/*            B(q)[j][i] = sum_a [0<=T<nrays] q[a][T] dtheta_a   (:272-291)     */
/*   backones = B(1)                                    (:303-350)             */
/*   update   x <- x * B(rr) / backones                 (:293-297)             */
/* KL(y || v) per slice, per iteration, on v = forward(x) before the update:  */
/*   sum over bins with v > 0 and y >= 0 of  v - y + (y > 0 ? y ln(y/v) : 0);  */
/*   other bins are excluded and counted.                                      */
/* ------------------------------------------------------------------------- */
double SFX(oracle_kl)(const real* y, const real* v, long nbins, long* excluded)
{
    /* Neumaier-compensated sum so the value does not depend on bin order */
    double sum = 0.0, comp = 0.0;
    long ex = 0;
    for (long b = 0; b < nbins; ++b) {
        double yy = (double)y[b], vv = (double)v[b];
        if (!(vv > 0.0) || !(yy >= 0.0)) { ++ex; continue; }
        double term = vv - yy + (yy > 0.0 ? yy * log(yy / vv) : 0.0);
        double t = sum + term;
        if (fabs(sum) >= fabs(term)) comp += (sum - t) + term;
        else comp += (term - t) + sum;
        sum = t;
    }
    if (excluded) *excluded = ex;
    return sum + comp;
}

typedef struct {
    int n, nrays, nangles, niter;
    const float* tomo;
    real* obj;
    const double *cs, *sn, *dth;
    const double *bi, *bj, *a0; /* per-angle backprojection coefficients */
    double dt;
    ray_geom g;
    double* kl;
    long* kl_excluded;
} em_ctx;

/* per-angle coefficients of u = fma(i-c, bi, fma(j-c, bj, a0)) */
static void em_bp_coeffs(int nangles, int nrays, int n, const double* cs, const double* sn, double* bi, double* bj,
                         double* a0)
{
    double rr = (double)(nrays - 1) / (double)(n - 1);
    for (int a = 0; a < nangles; ++a) {
        bi[a] = cs[a] * rr;
        bj[a] = sn[a] * rr;
        a0[a] = (double)(nrays - 1) / 2.0;
    }
}

/* rows [jlo, jhi) of B(q) (or of B(1) when ones) */
static void em_backproject_rows(const em_ctx* E, const real* q, real* out, int ones, long jlo, long jhi)
{
    int n = E->n, A = E->nangles, R = E->nrays;
    double c = (double)(n - 1) / 2.0;
    for (long j = jlo; j < jhi; ++j) {
        for (int i = 0; i < n; ++i) {
            real acc = 0;
            for (int a = 0; a < A; ++a) {
                double u = fma((double)i - c, E->bi[a], fma((double)j - c, E->bj[a], E->a0[a]));
                int T = (int)u;
                if (T > -1 && T < R) acc += (ones ? (real)1 : q[(size_t)a * R + T]) * (real)E->dth[a];
            }
            out[(size_t)j * n + i] = acc;
        }
    }
}

static void em_backproject(const em_ctx* E, const real* q, real* out, int ones)
{
    em_backproject_rows(E, q, out, ones, 0, E->n);
}

/* angles [alo, ahi) of the forward projection v = dt * A x and the clamped
 * ratio rr = y SIGN(v) / sqrt(e^2 + v^2) */
static void em_forward_ratio(const em_ctx* E, const real* x, const real* y, real* v, real* rr, long alo, long ahi)
{
    int R = E->nrays;
    const double e = (double)0.0001f;
    for (long a = alo; a < ahi; ++a)
        for (int r = 0; r < R; ++r) {
            real vv = ray_sum(&E->g, x, E->cs[a], E->sn[a], r) * (real)E->dt;
            v[(size_t)a * R + r] = vv;
            real sg = (vv > 0) ? (real)1 : ((vv < 0) ? (real)-1 : (real)0);
            real q = y[(size_t)a * R + r] * (sg / (real)sqrt(e * e + (double)vv * (double)vv));
            rr[(size_t)a * R + r] = (q < 0) ? (real)0 : q;
        }
}

/* one slice's work split over threads: every element is computed exactly as
 * in the serial loops (the split changes no arithmetic) */
typedef struct {
    const em_ctx* E;
    const real *x, *y, *q;
    real *v, *rr, *out;
    int ones;
} em_part_ctx;

static void em_forward_part(void* vctx, long lo, long hi)
{
    em_part_ctx* P = (em_part_ctx*)vctx;
    em_forward_ratio(P->E, P->x, P->y, P->v, P->rr, lo, hi);
}

static void em_backproject_part(void* vctx, long lo, long hi)
{
    em_part_ctx* P = (em_part_ctx*)vctx;
    em_backproject_rows(P->E, P->q, P->out, P->ones, lo, hi);
}

static void em_run_slice(const em_ctx* E, long s, real* back1, real* bq, real* v, real* y, real* rr, int nthreads)
{
    int n = E->n, A = E->nangles, R = E->nrays;
    real* x = E->obj + (size_t)s * n * n;
    const float* ys = E->tomo + (size_t)s * A * R;
    for (long b = 0; b < (long)A * R; ++b) y[b] = (real)ys[b];
    em_part_ctx P = {E, x, y, NULL, v, rr, back1, 1};
    parallel_for(n, nthreads, em_backproject_part, &P);
    for (int it = 0; it < E->niter; ++it) {
        parallel_for(A, nthreads, em_forward_part, &P);
        if (E->kl) {
            long ex = 0;
            E->kl[(size_t)s * E->niter + it] = SFX(oracle_kl)(y, v, (long)A * R, &ex);
            if (E->kl_excluded) E->kl_excluded[(size_t)s * E->niter + it] = ex;
        }
        em_part_ctx B = {E, x, y, rr, v, rr, bq, 0};
        parallel_for(n, nthreads, em_backproject_part, &B);
        for (long p = 0; p < (long)n * n; ++p) x[p] = x[p] * bq[p] / back1[p];
    }
}

static int em_slices_run(em_ctx* E, long lo, long hi, int inner_threads)
{
    int n = E->n, A = E->nangles, R = E->nrays;
    real* back1 = (real*)malloc(sizeof(real) * (size_t)n * n);
    real* bq = (real*)malloc(sizeof(real) * (size_t)n * n);
    real* v = (real*)malloc(sizeof(real) * (size_t)A * R);
    real* y = (real*)malloc(sizeof(real) * (size_t)A * R);
    real* rr = (real*)malloc(sizeof(real) * (size_t)A * R);
    int rc = (back1 && bq && v && y && rr) ? 0 : -2;
    for (long s = lo; s < hi && rc == 0; ++s) em_run_slice(E, s, back1, bq, v, y, rr, inner_threads);
    free(back1); free(bq); free(v); free(y); free(rr);
    return rc;
}

static void em_slices(void* vctx, long lo, long hi)
{
    (void)em_slices_run((em_ctx*)vctx, lo, hi, 1);
}

/* get_eEM_RT_MultiGPU(..., obj (in/out initial guess), tomogram, angles, paramf, parami)
 * parami = [nrays, nangles, nslices, objsize, 0, 0, 0, nflats, iterations, blocksize]
 * (parameters.cu:8-59, em_methods.py:55-59).  kl/kl_excluded: [nslices][iterations] or NULL. */
int SFX(oracle_eem)(real* obj, const float* tomo, const float* angles, const int* parami, double* kl,
                    long* kl_excluded, int nthreads)
{
    em_ctx E;
    E.nrays = parami[0];
    E.nangles = parami[1];
    int nslices = parami[2];
    E.n = parami[3];
    E.niter = parami[8];
    if (E.nrays < 2 || E.nangles < 2 || E.n < 2 || nslices < 0 || E.niter < 0) return -1;
    E.tomo = tomo; E.obj = obj; E.kl = kl; E.kl_excluded = kl_excluded;
    E.dt = 2.0 / (double)(E.nrays - 1);
    ray_geom_em(&E.g, E.nrays, E.n);
    E.g.nrays = E.nrays; E.g.nangles = E.nangles;
    int A = E.nangles;
    double* tab = (double*)malloc(sizeof(double) * 6 * (size_t)A);
    if (!tab) return -2;
    for (int a = 0; a < A; ++a) {
        tab[a] = cos((double)angles[a]);
        tab[A + a] = sin((double)angles[a]);
        tab[2 * A + a] = angle_step(angles, A, a);
    }
    E.cs = tab; E.sn = tab + A; E.dth = tab + 2 * A;
    em_bp_coeffs(A, E.nrays, E.n, tab, tab + A, tab + 3 * A, tab + 4 * A, tab + 5 * A);
    E.bi = tab + 3 * A; E.bj = tab + 4 * A; E.a0 = tab + 5 * A;
    /* threads over slices when there are enough of them, else over the angles
     * and image rows of one slice at a time */
    int rc = 0;
    if (nslices >= nthreads) parallel_for(nslices, nthreads, em_slices, &E);
    else rc = em_slices_run(&E, 0, nslices, nthreads);
    free(tab);
    return rc;
}

/* B(q) alone and the forward operator alone, for stage-level checks. */
int SFX(oracle_em_backproject)(real* out, const real* q, const float* angles, int nrays, int nangles, int n, int ones)
{
    if (nrays < 2 || nangles < 2 || n < 2) return -1;
    em_ctx E;
    memset(&E, 0, sizeof(E));
    E.nrays = nrays; E.nangles = nangles; E.n = n;
    E.dt = 2.0 / (double)(nrays - 1);
    int A = nangles;
    double* tab = (double*)malloc(sizeof(double) * 6 * (size_t)A);
    if (!tab) return -2;
    for (int a = 0; a < A; ++a) {
        tab[a] = cos((double)angles[a]);
        tab[A + a] = sin((double)angles[a]);
        tab[2 * A + a] = angle_step(angles, A, a);
    }
    E.cs = tab; E.sn = tab + A; E.dth = tab + 2 * A;
    em_bp_coeffs(A, nrays, n, tab, tab + A, tab + 3 * A, tab + 4 * A, tab + 5 * A);
    E.bi = tab + 3 * A; E.bj = tab + 4 * A; E.a0 = tab + 5 * A;
    em_backproject(&E, q, out, ones);
    free(tab);
    return 0;
}

int SFX(oracle_em_forward)(real* v, const real* x, const float* angles, int nrays, int nangles, int n)
{
    if (nrays < 2 || nangles < 1 || n < 2) return -1;
    ray_geom g;
    ray_geom_em(&g, nrays, n);
    g.nrays = nrays; g.nangles = nangles;
    for (int a = 0; a < nangles; ++a) {
        double cs = cos((double)angles[a]), sn = sin((double)angles[a]);
        for (int r = 0; r < nrays; ++r) v[(size_t)a * nrays + r] = ray_sum(&g, x, cs, sn, r) * (real)g.dy;
    }
    return 0;
}
