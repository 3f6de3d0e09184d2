// raft_internal.hpp -- shared host/device declarations of libraft (gfx950).
//
// Error handling: internal code throws raft::Error; every extern "C" entry
// point catches it and stores code + text in thread-local state
// (raft_last_error / raft_last_error_string).  Nothing calls exit().
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <string>
#include <vector>

#include "../../include/raft.h"

namespace raft {

struct Error {
    int code;
    std::string msg;
};

[[noreturn]] void fail(int code, const std::string& msg);
// sets this thread's error state (raft_last_error) without throwing
void record_error(int code, const std::string& msg);
void hip_check(hipError_t e, const char* what);
#define RAFT_HIP(call) ::raft::hip_check((call), #call)

// device allocation owned by the caller's scope
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    explicit DevBuf(size_t n) { alloc(n); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
    ~DevBuf() { release(); }
    void alloc(size_t n);
    void release();
    template <typename T> T* as() const { return static_cast<T*>(p); }
};

// ---------------------------------------------------------------------------
// Backprojection geometry shared by FBP and the eEM update.
// For pixel (i, j) of an n x n image and angle a:
//     X_i = i - xc,  Y_j = j - xc         (exact: small integers or halves)
//     u   = fma(X_i, cA[a], fma(Y_j, cB[a], c0))     (IEEE double)
//     T   = (int)u (truncation); the sample counts iff 0 <= T < nbins.
// FBP-RT:  xc = npad/2 - L, cA = cos, cB = -sin, c0 = npad/2, nbins = npad
//          (fbp.cu:38-69 with the pixel size cancelled and the crop folded in).
// eEM:     xc = (n-1)/2, cA = cos*rr, cB = sin*rr, c0 = (nrays-1)/2, nbins = nrays,
//          rr = (nrays-1)/(n-1)   (em_RT_kernel.cu:262-291 about the image centre).
// Rows staged for the kernel hold the bins [tlo, tlo + width) of each angle;
// bins outside [0, nbins) are stored as zeros so the kernel needs no range test.
// ---------------------------------------------------------------------------
struct BpGeom {
    int n = 0;        // image side
    int nangles = 0;
    int nbins = 0;    // valid bin range [0, nbins)
    double xc = 0;    // pixel-centre offset
    double c0 = 0;
    int tlo = 0;      // first staged bin
    int width = 0;    // staged bins per angle row
    std::vector<double> cA, cB; // per angle
};

// Tile of the backprojection kernel: 64 x 32 pixels, 256 threads, 4 x 2 pixels each.
constexpr int BP_TILE_X = 64;
constexpr int BP_TILE_Y = 32;
constexpr int BP_CHUNK = 32;   // angles staged per LDS round (default; 16 and 48 are autotuner variants)
// variant flag: FP32 rows, two 4-slice groups per workgroup (8 slices per bin
// lookup); BP_PAIR | 16 and BP_PAIR | 32 are the autotuner's pair variants
constexpr int BP_PAIR = 1024;
// | BP_SPF: the single form whose sweep loads the next 4 angles' coefficients
// one trip ahead (fewer registers; same sums bit for bit)
constexpr int BP_SPF = 2048;
constexpr int BP_GROUP = 16;   // angles per partial sum of the FP32 two-level summation
constexpr int BP_SLOT = 80;    // LDS entries per staged angle window (>= |(63, 31)| + 5)
// the pair form's window (two workgroups per CU fit 160 KB of LDS at 32
// angles per round): takes geometries whose window rr |(63, 31)| + 5 <= 76,
// rr = bin steps per pixel step (1 for FBP; (nrays - 1)/(n - 1) for eEM)
constexpr int BP_SLOT_PAIR = 76;
// staged rows hold the bins of BP_SLICES_F32 (FP32) or BP_SLICES_F16 (FP16)
// slices side by side: [group][angle][bin][slice]
constexpr int BP_SLICES_F32 = 4;
constexpr int BP_SLICES_F16 = 8;
// projector: consecutive angles per workgroup (one wave each)
constexpr int PROJ_ANGLES = 8;

// global bin window covering every pixel of the tile grid
void bp_window(BpGeom& g);

// ---------------------------------------------------------------------------
// Ray geometry shared by Radon-RT and the eEM forward operator.
// Ray i at angle a, sample k (centred forms, see oracle/raft_oracle.c):
//     ir = i - ci, kr = k - ck
//     ux = fma(kr, -sin*ry, fma(ir, cos, ucx)),  X = (int)ux
//     uy = fma(kr,  cos,    fma(ir, sin*rx, ucy)), Y = (int)uy
// sample counts iff 0 <= X < Nx and 0 <= Y < Ny; result = scale * sum.
// ---------------------------------------------------------------------------
struct RayGeom {
    int Nx = 0, Ny = 0, nrays = 0, nangles = 0, nk = 0;
    double ci = 0, ck = 0, ucx = 0, ucy = 0, ry = 1, rx = 1;
    float scale = 1;
};
RayGeom ray_geom_radon(int Nx, int Ny, int nrays, int nangles, double ax, double ay);
RayGeom ray_geom_em(int nrays, int nangles, int n);

// per-angle tables on the device
struct AngleTables {
    DevBuf cosv, sinv; // double[A]
};

// ---------------------------------------------------------------------------
// kernel launchers (defined in the .hip files)
// ---------------------------------------------------------------------------
struct FftPlan {
    int n = 0;
    int nstages = 0;
    int radix[24] = {0};
    int threads = 256;
    bool ok = false;   // false -> direct-convolution path
};
FftPlan fft_plan(int n);

struct FilterArgs {
    const float* tomo;  // [S][A][N]
    void* q;            // [G][A][W][SB] float (SB = 4) or __half (SB = 8)
    const float2* H;    // [npad] Hermitian filter incl. 1/(npad*pixel)
    const float2* tw;   // two-level twiddles: lo[128] = W^k, hi[npad/128 + 1] = W^(128k), W = exp(-2 pi i/npad)
    const double* g;    // direct-convolution kernel (real, [npad], FP64) or null
    const float* wa;    // [A] per-angle weight (dtheta * angular scale)
    int N, A, S, npad, L, tlo, W;
    int none;           // filter 'none': no filtering, no 1/(npad h)
    float* rows = nullptr; // FP32, W = N: where the kernel form allows, the window is written
                           // straight to rows [S][A][W] (may be tomo itself) instead of q
};
// true when the filtered windows went straight to a.rows (else to a.q)
bool launch_fbp_filter(const FilterArgs& a, const FftPlan& plan, int store, hipStream_t st);

// BST after the row filter: q = filtered padded rows (interleaved [G][A][n][4],
// window [0, n)), polar [S][2A][n/2], obj [S][N][N]; tw the twiddle tables of n.
// fused: cart [S][n][N] (the kept columns of the inverse-transformed rows) and
// cartT [S][N][n]; unfused: cart [S][n][n] (the whole frequency grid) and cartT
// [S][N][n]
struct BstArgs {
    const float* q;
    float2* polar;
    float2* cart;
    float2* cartT;
    float* obj;
    const float2* tw;
    int N, A, S, L;
    float scale;        // 4 n pixel
    int fused = 1;      // gather + row inverse FFT in one kernel (k_bst_rows)
};
// filt != nullptr: the filter runs inside the polar pass from filt->tomo
// (a.q unused); otherwise a.q holds the filtered rows
void launch_bst(const BstArgs& a, const FftPlan& plan, hipStream_t st, const FilterArgs* filt = nullptr);
// whether the one-pass filter + polar form's LDS fits at padded size n
bool bst_filter_polar_fits(int n);

struct BpArgs {
    const void* q;      // [G][A][W][SB] staged rows (float SB = 4, or half SB = 8)
    float* out;         // mode 0: planar [S][n][n]; mode 1: interleaved [G][n][n][4]
    const float* x;     // mode 1: interleaved current image (may alias out)
    const float* norm;  // mode 1: backones [n][n] shared by all slices, or per slice interleaved (norm_il)
    float* xt;          // mode 1: interleaved transposed copy of the updated image (or null)
    const double* cA;
    const double* cB;
    double xc, c0;
    int n, A, S, tlo, W;
    int mode;           // 0 = store the sums (FBP, backones), 1 = EM update x * bp / norm
    int norm_il = 0;    // mode 1: norm is per slice, interleaved [G][n][n][4] (tEM)
    int chunk = 0;      // variant: angles per LDS round (16, 32 or 48; 0 = 32), | BP_PAIR for the pair form
    size_t q_bytes = 0; // bytes of the buffer behind q when known: launches that would read past it are refused
};
void launch_backproject(const BpArgs& a, int store, hipStream_t st);
// whether variant (see BpArgs::chunk) exists for rows of this store
bool bp_variant_supported(int variant, int store);
// slices one workgroup (one queue item) of this variant covers
int bp_slices_per_item(int store, int mode, int variant);
// bytes per second of conflict-free 16-byte LDS reads by every CU of the
// current device (raft_lds_read_probe)
double lds_read_probe(int* cus, int* clock_khz);
// tiles per slice group of an n x n image
int bp_tiles(int n);

// ---------------------------------------------------------------------------
// Persistent work-queue launches: one resident grid takes items from a
// device counter over several volumes ("segments", e.g. the scans of a batch
// grouped by geometry).  Segment b owns items [first[b], first[b+1]); the
// kernels decode an item into (volume b, its work unit).  counter points at
// two ints that are zero between launches (the last workgroup clears them).
// ---------------------------------------------------------------------------
constexpr int QUEUE_MAXSEG = 32;
struct BpQueue {
    int* counter;
    int nseg;
    int first[QUEUE_MAXSEG + 1];
    int tiles[QUEUE_MAXSEG]; // tiles per slice group of segment b
};
// device descriptor of one volume for launch_backproject_queue (bp_desc_size() bytes)
size_t bp_desc_size();
void bp_desc_fill(const BpArgs& b, int store, void* dst);
// resident workgroups of the queue kernel (CUs x blocks per CU)
int bp_queue_grid(int store, int mode, int chunk);
void launch_backproject_queue(const void* d_descs, const BpQueue& q, int store, int mode, int chunk, int grid,
                              hipStream_t st);

struct ProjArgs {
    const float* img;   // interleaved [G][Ny][Nx][4]
    const float* imgT;  // interleaved transpose [G][Nx][Ny][4] (null: no transposed path)
    float* out;         // Radon: planar [S][A][nrays]; eEM: interleaved ratio rows [G][A][W][4] at (i - tlo)
    const float* y;     // eEM: measured sinogram, planar [S][A][nrays] (null for Radon); tEM: flat [S][nrays]
    const float* wa;    // eEM: per-angle dtheta
    double* kl_part;    // eEM: per-(slice, angle) KL partial sums [S][A] (null = skip)
    unsigned long long* ex_part; // eEM: per-(slice, angle) excluded counts
    const double* cosv;
    const double* sinv;
    RayGeom g;
    int S;
    int tlo, W;         // eEM ratio-row layout
    int lds_rc = 0;     // > 0: staged form with this many ray chunks (from proj_plan); 0: direct form
};
// em_mode 0 = Radon, 1 = eEM ratio (+ KL), 2 = tEM flat * exp(-v).
// KL partials: kl_part[s * proj_kl_slots(a) + slot], reduced by launch_kl_reduce.
void launch_project(const ProjArgs& a, int em_mode, hipStream_t st);
// Staged-form plan for S slices and host angle tables c, s: the number of ray
// chunks, or 0 when a chunk of 8 angles cannot be staged (sparse angles) or
// RAFT_PROJ_LDS=0.  have_T: the transposed image will be passed as imgT.
int proj_plan(const RayGeom& g, const double* c, const double* s, int S, bool have_T);
// persistent work-queue form of the staged projector (see BpQueue); segment b
// has nab[b] angle chunks x RC ray chunks x slice groups items
struct ProjQueue {
    int* counter;
    int nseg;
    int first[QUEUE_MAXSEG + 1];
    int nab[QUEUE_MAXSEG];
};
size_t proj_desc_size();
void proj_desc_fill(const ProjArgs& p, void* dst);
int proj_angle_chunks(int nangles);
int proj_queue_grid(int em_mode);
void launch_project_queue(const void* d_descs, const ProjQueue& q, int em_mode, int grid, hipStream_t st);
// the staged form's tile constants: rays per patch, samples per step, angles
// per workgroup, staged 16-byte entries per workgroup
void proj_tile_consts(int* c4);
inline int proj_kl_slots(int nangles, int lds_rc) { return nangles * (lds_rc > 0 ? lds_rc : 1); }

// planar [S][rows][cols] -> interleaved [G][rows][cols][4] (or its transpose [G][cols][rows][4])
void launch_to_interleaved(const float* in, float* out, int S, int rows, int cols, bool transpose, hipStream_t st);
// interleaved [G][rows][cols][4] -> planar [S][rows][cols]
void launch_from_interleaved(const float* in, float* out, int S, int rows, int cols, hipStream_t st);
// rows for B(1): interleaved [A][W][4], value wa[a] on valid bins, 0 elsewhere
void launch_fill_rows(float* q, const float* wa, int A, int W, int tlo, int nbins, hipStream_t st);
// rows from planar data: q[g][a][w][s] = data[4g+s][a][tlo+w] * wa[a] (0 outside [0, nbins) and past S)
void launch_rows_from_planar(float* q, const float* data, const float* wa, int S, int A, int W, int tlo, int nbins,
                             hipStream_t st);
// kl[s*K + it] for it = *iter_ptr (if iter_ptr) or iter
void launch_kl_reduce(const double* part, const unsigned long long* expart, int nb, int S, double* kl,
                      unsigned long long* ex, int K, int iter, const int* iter_ptr, hipStream_t st);
// device iteration counter: op 0 clears, op 1 increments (a kernel, so it may be graph-captured)
void launch_counter(int* c, int op, hipStream_t st);

// flat/dark background correction (+ optional -log), in place on [S][A][R]
void launch_background(float* data, const float* flat, const float* dark, int R, int A, int S, int nflats, int is_log,
                       hipStream_t st);

// p[0 .. n) = v (a kernel, so it may be captured into a graph)
void launch_fill(float* p, long long n, float v, hipStream_t st);

// device address of this device's copy of the fixed scratch array of
// RAFT_NOALLOC_SCRATCH_BYTES (noalloc.hip); dev must be the current device
void* noalloc_scratch(int dev);

// autotuner (autotune.cpp): the cached choice for key, else the fastest of
// cands timed with launch(c) on stream st (or cands[0] when tuning is off)
bool tune_enabled();
int tune_pick(const std::string& key, const std::vector<int>& cands, hipStream_t st,
              const std::function<void(int)>& launch);
// the cached choice for key, or dflt
int tune_cached(const std::string& key, int dflt);

// profiling hooks (events around kernel families)
enum KernelFamily { KF_FILTER = 0, KF_BP = 1, KF_PROJ = 2, KF_EMBP = 3, KF_COUNT = 4 };
struct ProfScope {
    int fam;
    hipStream_t st;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ProfScope(int family, hipStream_t stream);
    ~ProfScope();
};

} // namespace raft
