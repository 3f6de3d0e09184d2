// raft_host.cpp -- host side of libraft: error state, parameter decode,
// geometry tables, per-device caches, the reference C ABI
// (getFBPMultiGPU / get_eEM_RT_MultiGPU / getRadonRTMultiGPU) and the
// device-pointer entry points.
//
// Design (MI355X, one process, one host thread per device in the call):
//   * the reference's slab split over gpus[] (fbp.cu:267-300, em_rt.cu:221-259,
//     radon_rt.cu:259-294) is kept: contiguous slabs of slices, one std::thread
//     per device, each binding its own device for the whole slab;
//   * inside a device, slices are processed in blocks sized from free HBM
//     (or parami[9] when the caller sets it); device buffers live in a
//     per-device grow-only workspace instead of being allocated per call;
//   * geometry (cos/sin of the float32 angles, bin windows, the filter
//     spectrum, FFT twiddles) is built on the host in FP64.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <condition_variable>
#include <map>
#include <memory>
#include <mutex>
#include <thread>

#include <sys/mman.h>

#include "raft_internal.hpp"

namespace raft {

// ---------------------------------------------------------------------------
// error state
// ---------------------------------------------------------------------------
static thread_local int g_err = RAFT_OK;
static thread_local std::string g_msg;

void fail(int code, const std::string& msg) { throw Error{code, msg}; }

void record_error(int code, const std::string& msg)
{
    g_err = code;
    g_msg = msg;
}

void hip_check(hipError_t e, const char* what)
{
    if (e != hipSuccess) fail(RAFT_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}

void DevBuf::alloc(size_t n)
{
    release();
    if (n == 0) return;
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, n);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        fail(RAFT_EHIP, "hipMalloc(" + std::to_string(n) + " bytes): " + hipGetErrorString(e));
    }
    p = q;
    bytes = n;
}

void DevBuf::release()
{
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
}

template <typename F> static int guarded(F&& f)
{
    try {
        f();
        g_err = RAFT_OK;
        g_msg.clear();
        return RAFT_OK;
    } catch (const Error& e) {
        g_err = e.code;
        g_msg = e.msg;
        return e.code;
    } catch (const std::exception& e) {
        g_err = RAFT_EHIP;
        g_msg = e.what();
        return RAFT_EHIP;
    }
}

// ---------------------------------------------------------------------------
// profiling (events around kernel families, resolved lazily)
// ---------------------------------------------------------------------------
static thread_local bool g_prof = false;
static bool prof_enabled() { return g_prof; }
struct ProfRec {
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev[KF_COUNT];
};
static thread_local ProfRec g_prof_rec;

static void prof_clear()
{
    for (auto& v : g_prof_rec.ev) {
        for (auto& pr : v) {
            (void)hipEventDestroy(pr.first);
            (void)hipEventDestroy(pr.second);
        }
        v.clear();
    }
}

ProfScope::ProfScope(int family, hipStream_t stream) : fam(family), st(stream)
{
    if (!g_prof) return;
    RAFT_HIP(hipEventCreate(&e0));
    RAFT_HIP(hipEventCreate(&e1));
    RAFT_HIP(hipEventRecord(e0, st));
}

ProfScope::~ProfScope()
{
    if (!e0) return;
    (void)hipEventRecord(e1, st);
    g_prof_rec.ev[fam].push_back({e0, e1});
}

// ---------------------------------------------------------------------------
// geometry
// ---------------------------------------------------------------------------
static double angle_step(const float* ang, int A, int a)
{
    if (a == A - 1) return (double)ang[a] - (double)ang[a - 1];
    return (double)ang[a + 1] - (double)ang[a];
}

// fbp.cu:24-28: float32 comparison, value pi/theta_last
static double fbp_angle_scale(const float* ang, int A)
{
    const float last = ang[A - 1];
    return (last > (float)M_PI) ? M_PI / (double)last : 1.0;
}

void bp_window(BpGeom& g)
{
    const int tx = (g.n + BP_TILE_X - 1) / BP_TILE_X, ty = (g.n + BP_TILE_Y - 1) / BP_TILE_Y;
    const double xlo = -g.xc, xhi = (double)(tx * BP_TILE_X - 1) - g.xc;
    const double ylo = -g.xc, yhi = (double)(ty * BP_TILE_Y - 1) - g.xc;
    long long tmin = 0x7fffffffLL, tmax = -0x7fffffffLL;
    for (int a = 0; a < g.nangles; ++a) {
        const double u[4] = {std::fma(xlo, g.cA[a], std::fma(ylo, g.cB[a], g.c0)),
                             std::fma(xhi, g.cA[a], std::fma(ylo, g.cB[a], g.c0)),
                             std::fma(xlo, g.cA[a], std::fma(yhi, g.cB[a], g.c0)),
                             std::fma(xhi, g.cA[a], std::fma(yhi, g.cB[a], g.c0))};
        const double umin = std::min(std::min(u[0], u[1]), std::min(u[2], u[3]));
        const double umax = std::max(std::max(u[0], u[1]), std::max(u[2], u[3]));
        tmin = std::min(tmin, (long long)std::floor(umin) - 1);
        tmax = std::max(tmax, (long long)std::floor(umax) + 2);
    }
    // the per-tile window of the kernel may reach two bins past these bounds
    // only if rounding disagrees; keep a margin of 2 more on each side
    tmin -= 2;
    tmax += 2;
    g.tlo = (int)tmin;
    g.width = (int)(tmax - tmin + 1);
}

RayGeom ray_geom_radon(int Nx, int Ny, int nrays, int nangles, double ax, double ay)
{
    RayGeom g;
    g.Nx = Nx; g.Ny = Ny; g.nrays = nrays; g.nangles = nangles; g.nk = Ny;
    g.ci = (double)(Nx - 1) / 2.0;
    g.ck = (double)(Ny - 1) / 2.0;
    g.ucx = (double)(Nx - 1) / (2.0 * ax);
    g.ucy = (double)(Ny - 1) / (2.0 * ay);
    g.ry = (ay * (double)(Nx - 1)) / (ax * (double)(Ny - 1));
    g.rx = (ax * (double)(Ny - 1)) / (ay * (double)(Nx - 1));
    g.scale = (float)(2.0 * ay / (double)(Ny - 1));
    return g;
}

RayGeom ray_geom_em(int nrays, int nangles, int n)
{
    RayGeom g;
    g.Nx = n; g.Ny = n; g.nrays = nrays; g.nangles = nangles; g.nk = nrays;
    g.ci = (double)(nrays - 1) / 2.0;
    g.ck = g.ci;
    g.ucx = g.ci;
    g.ucy = g.ci;
    g.ry = 1.0;
    g.rx = 1.0;
    g.scale = (float)(2.0 / (double)(nrays - 1));
    return g;
}

// ---------------------------------------------------------------------------
// autotuning keys and candidates
// ---------------------------------------------------------------------------
static std::string bp_tune_key(int n, int A, int W, int store, int mode, int groups)
{
    char k[160];
    snprintf(k, sizeof k, "bp/n=%d/A=%d/W=%d/store=%d/mode=%d/groups=%d", n, A, W, store, mode, groups);
    return k;
}
// backprojector variants the tuner times: angle chunks, the scalar-prefetch
// sweep; for FP32 rows also the pair form (two slice groups per workgroup)
// where its narrower window holds the tile's bins (rr = bin steps per pixel
// step); every variant gives the same sums
static const std::vector<int>& bp_cands(int store, double rr = 1.0)
{
    static const std::vector<int> f32 = {BP_CHUNK, 16, 48, BP_PAIR | 32, BP_PAIR | 16, BP_SPF | 16, BP_SPF | 32};
    static const std::vector<int> f32_nopair = {BP_CHUNK, 16, 48, BP_SPF | 16, BP_SPF | 32};
    static const std::vector<int> f16 = {BP_CHUNK, 16, 48, BP_SPF | 16, BP_SPF | 32};
    const bool pair_fits = rr * std::hypot((double)(BP_TILE_X - 1), (double)(BP_TILE_Y - 1)) + 5.0 <= BP_SLOT_PAIR;
    if (store == RAFT_STORE_F16) return f16;
    return pair_fits ? f32 : f32_nopair;
}

static std::string proj_tune_key(const RayGeom& g, int em, int groups)
{
    char k[200];
    snprintf(k, sizeof k, "proj/Nx=%d/Ny=%d/nrays=%d/A=%d/em=%d/groups=%d/ry=%.6f", g.Nx, g.Ny, g.nrays, g.nangles, em,
             groups, g.ry);
    return k;
}
// ray-chunk counts of the staged projector: the plan's, then powers of two up to the ray blocks
static std::vector<int> proj_rc_cands(const RayGeom& g, int planned)
{
    std::vector<int> c = {planned};
    const int nrb = (g.nrays + 31) / 32;
    for (int r = 1; r <= nrb && r <= 32; r *= 2)
        if (r != planned) c.push_back(r);
    return c;
}

// ---------------------------------------------------------------------------
// FBP parameters (parameters.cu:131-219, Filter::apply filters.cu:395-438)
// ---------------------------------------------------------------------------
struct FbpCfg {
    int N, A, S, P, filter, blocksize;
    int npad, L, M;
    double pixel, reg, paganin, offset;
};

static FbpCfg decode_fbp(const float* paramf, const int* parami)
{
    if (!paramf || !parami) fail(RAFT_EINVAL, "null parameter array");
    FbpCfg c;
    c.N = parami[0];
    c.A = parami[1];
    c.S = parami[2];
    const int objsize = parami[3];
    c.P = parami[4];
    c.filter = parami[7];
    c.blocksize = parami[9];
    if (c.N < 2) fail(RAFT_EINVAL, "nrays must be >= 2");
    if (c.A < 2) fail(RAFT_EINVAL, "nangles must be >= 2 (the angle step of the last angle needs two angles)");
    if (c.S < 0) fail(RAFT_EINVAL, "nslices must be >= 0");
    if (c.P < 0) fail(RAFT_EINVAL, "padding must be >= 0");
    if (objsize != c.N) fail(RAFT_EUNSUPPORTED, "objsize must equal nrays (as sscRaft.fbpGPU passes it)");
    if (c.filter < 0 || c.filter > 7) fail(RAFT_EINVAL, "filter type must be in 0..7");
    if ((long long)c.N * (1 + c.P) > (1 << 26)) fail(RAFT_EINVAL, "padded size too large");
    c.npad = c.N * (1 + c.P);
    c.L = (c.P * c.N) / 2;
    c.M = c.npad / 2 + 1;
    const double bd = (double)paramf[0];
    c.reg = (double)paramf[1];
    const double energy = (double)paramf[2];
    const double z2 = (double)paramf[3];
    c.pixel = (double)paramf[4];
    c.offset = (double)paramf[6];
    if (!(c.pixel != 0.0) && c.filter != 0) fail(RAFT_EINVAL, "pixel size must be non-zero");
    const double plank = 4.135667696E-15, vc = 299792458.0;
    const double lambda = (bd == 0.0) ? 1.0 : (plank * vc) / energy;
    c.paganin = lambda * z2 * M_PI * bd / (c.pixel * c.pixel);
    return c;
}

static double filter_apply(int type, double w, double reg, double pag)
{
    const double pg = 1.0 / (1.0 + pag * w * w);
    switch (type) {
    case 1: return w * std::exp(-(double)0.693f * reg * w * w) * pg;
    case 2: return w * (1.0 / (1.0 + reg * w * w)) * pg;
    case 3: return w * std::cos(M_PI * 0.5 * w) * pg;
    case 4: {
        const double prm = std::max(w * reg * M_PI * 0.5, (double)1E-4f);
        return w * (std::sin(prm) / prm) * pg;
    }
    case 5: return w * (0.5 + 0.5 * std::cos(2.0 * M_PI * w)) * pg;
    case 6: return w * ((double)0.54f + (double)0.46f * std::cos(2.0 * M_PI * w)) * pg;
    case 7: return w * pg;
    default: return 1.0;
    }
}

// ---------------------------------------------------------------------------
// per-device state: lock, cached filter tables, grow-only workspace
// ---------------------------------------------------------------------------
struct FilterTables {
    DevBuf H, tw, g; // float2[n], float2[n], double[n]
    FftPlan plan;
};

// page-locked host buffer (hipHostMalloc), grow-only
struct PinnedBuf {
    void* p = nullptr;
    size_t bytes = 0;
    PinnedBuf() = default;
    PinnedBuf(const PinnedBuf&) = delete;
    PinnedBuf& operator=(const PinnedBuf&) = delete;
    ~PinnedBuf() { release(); }
    void release()
    {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        bytes = 0;
    }
    void* ensure(size_t n)
    {
        if (bytes >= n) return p;
        release();
        void* q = nullptr;
        const hipError_t e = hipHostMalloc(&q, n, hipHostMallocDefault);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            fail(RAFT_EHIP, "hipHostMalloc(" + std::to_string(n) + " bytes): " + hipGetErrorString(e));
        }
        p = q;
        bytes = n;
        return p;
    }
};

struct Device {
    static constexpr int kSlots = 14;
    std::mutex mu;
    std::map<std::string, std::unique_ptr<FilterTables>> filters;
    DevBuf ws[kSlots];
    DevBuf tab; // per-call tables (see TableSet)
    PinnedBuf pin[6];                     // host-pointer entry points: staging of up to three chunks in and out
    std::vector<std::vector<char>> keep;  // host bytes of table uploads in flight (TableSet::keep_alive)
    // the slot is a template argument so that every index is checked at compile time
    template <int SLOT> void* workspace(size_t bytes)
    {
        static_assert(SLOT >= 0 && SLOT < kSlots, "workspace slot out of range");
        if (ws[SLOT].bytes < bytes) {
            ws[SLOT].release();
            ws[SLOT].alloc(bytes);
        }
        return ws[SLOT].p;
    }
    char* tables(size_t bytes)
    {
        if (tab.bytes < bytes) {
            tab.release();
            tab.alloc(std::max(bytes, (size_t)64 << 10));
        }
        return static_cast<char*>(tab.p);
    }
};

// A call's small tables (angle cosines and sines, bin coefficients, weights)
// packed into one host buffer and copied with one transfer into the device's
// grow-only table arena: no allocation (and no implicitly synchronising free)
// per call.  Calls on a device run one at a time under its lock and
// synchronise their stream before returning, so the next call may overwrite
// the arena.
class TableSet;
struct TabRef {
    const TableSet* set;
    size_t off;
    template <typename T> T* as() const;
};
class TableSet {
  public:
    explicit TableSet(Device& D) : D_(D) {}
    template <typename T> TabRef add(const std::vector<T>& v)
    {
        const size_t off = (host_.size() + 255) & ~(size_t)255;
        host_.resize(off + sizeof(T) * v.size());
        if (!v.empty()) std::memcpy(host_.data() + off, v.data(), sizeof(T) * v.size());
        return TabRef{this, off};
    }
    void upload(hipStream_t st)
    {
        base_ = D_.tables(host_.size());
        if (!host_.empty())
            RAFT_HIP(hipMemcpyAsync(base_, host_.data(), host_.size(), hipMemcpyHostToDevice, st));
    }
    // for calls that return before their stream is synchronised: the host
    // bytes of the upload stay alive in the device state until the caller
    // synchronises and clears it
    void keep_alive() { D_.keep.push_back(std::move(host_)); }
    char* base() const
    {
        if (!base_) fail(RAFT_EINVAL, "internal: table set used before upload");
        return base_;
    }

  private:
    Device& D_;
    std::vector<char> host_;
    char* base_ = nullptr;
};
template <typename T> T* TabRef::as() const { return reinterpret_cast<T*>(set->base() + off); }

static std::mutex g_dev_mu;
static std::map<int, std::unique_ptr<Device>> g_devices;

static Device& device_state(int dev)
{
    std::lock_guard<std::mutex> lk(g_dev_mu);
    auto& d = g_devices[dev];
    if (!d) d.reset(new Device());
    return *d;
}

static void bind_device(int dev)
{
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count <= 0) {
        (void)hipGetLastError();
        fail(RAFT_EDEVICE, "no HIP device available");
    }
    if (dev < 0 || dev >= count) fail(RAFT_EDEVICE, "invalid device id " + std::to_string(dev));
    RAFT_HIP(hipSetDevice(dev));
}

// bytes of device memory a block of slices may use
static size_t block_budget()
{
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
        (void)hipGetLastError();
        return (size_t)8 << 30;
    }
    return std::max<size_t>(fr / 2, (size_t)256 << 20);
}

template <typename T> static void upload(DevBuf& b, const std::vector<T>& v, hipStream_t st)
{
    b.alloc(sizeof(T) * v.size());
    RAFT_HIP(hipMemcpyAsync(b.p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, st));
}

// host copies of the filter tables: the spectrum H (filter, offset phase and
// 1/(npad*pixel)), the FFT twiddles and, for sizes the LDS FFT does not take,
// the real convolution kernel g
struct FilterHost {
    std::vector<float2> H, tw;
    std::vector<double> g;
    FftPlan plan;
};

static void filter_host_tables(const FbpCfg& c, FilterHost& F)
{
    const int n = c.npad, M = c.M;
    std::vector<double> Hr(n), Hi(n);
    const double scale = 1.0 / ((double)n * c.pixel);
    for (int k = 0; k < M; ++k) {
        const double w = 0.5 * (double)std::min(k, M - k) / (double)M;
        const double amp = filter_apply(c.filter, w, c.reg, c.paganin);
        const double ph = -2.0 * M_PI * c.offset * (double)k / (double)(2 * M - 2);
        Hr[k] = amp * std::cos(ph) * scale;
        Hi[k] = amp * std::sin(ph) * scale;
    }
    for (int k = M; k < n; ++k) {
        Hr[k] = Hr[n - k];
        Hi[k] = -Hi[n - k];
    }
    if (n % 2 == 0) Hi[n / 2] = 0.0;
    F.H.resize(n);
    for (int k = 0; k < n; ++k) F.H[k] = make_float2((float)Hr[k], (float)Hi[k]);
    // two-level twiddle tables: W^m = hi[m >> 7] * lo[m & 127], W = exp(-2 pi i / n)
    const int nhi = n / 128 + 1;
    F.tw.resize(128 + nhi);
    for (int k = 0; k < 128; ++k) {
        const double ang = -2.0 * M_PI * (double)k / (double)n;
        F.tw[k] = make_float2((float)std::cos(ang), (float)std::sin(ang));
    }
    for (int k = 0; k < nhi; ++k) {
        const double ang = -2.0 * M_PI * (double)(((long long)k * 128) % n) / (double)n;
        F.tw[128 + k] = make_float2((float)std::cos(ang), (float)std::sin(ang));
    }
    F.plan = fft_plan(n);
    F.g.clear();
    if (!F.plan.ok && c.filter != 0) {
        // g[d] = sum_k H_k exp(2 pi i k d / n): real because H is Hermitian
        F.g.resize(n);
        for (int d = 0; d < n; ++d) {
            double acc = 0.0;
            for (int k = 0; k < n; ++k) {
                const long long kd = ((long long)k * d) % n;
                const double ang = 2.0 * M_PI * (double)kd / (double)n;
                acc += Hr[k] * std::cos(ang) - Hi[k] * std::sin(ang);
            }
            F.g[d] = acc;
        }
    }
}

// the key of a filter table set: (n, filter, reg, paganin, offset, pixel)
static std::string filter_key(const FbpCfg& c)
{
    char key[256];
    snprintf(key, sizeof key, "%d/%d/%.17g/%.17g/%.17g/%.17g", c.npad, c.filter, c.reg, c.paganin, c.offset,
             c.pixel);
    return key;
}

// the filter tables on the device, cached per filter_key
static FilterTables& filter_tables(Device& D, const FbpCfg& c, hipStream_t st)
{
    const std::string key = filter_key(c);
    auto it = D.filters.find(key);
    if (it != D.filters.end()) return *it->second;
    std::unique_ptr<FilterTables> F(new FilterTables());
    FilterHost h;
    filter_host_tables(c, h);
    F->plan = h.plan;
    upload(F->H, h.H, st);
    upload(F->tw, h.tw, st);
    if (!h.g.empty()) upload(F->g, h.g, st);
    RAFT_HIP(hipStreamSynchronize(st));
    auto& ref = *F;
    D.filters[key] = std::move(F);
    return ref;
}

// ---------------------------------------------------------------------------
// FBP on device pointers
// ---------------------------------------------------------------------------
// backprojection geometry and per-angle weights dtheta_a * angular scale
static void fbp_geometry(const FbpCfg& c, const float* angles, BpGeom& g, std::vector<float>& wa)
{
    const int A = c.A;
    g.n = c.N;
    g.nangles = A;
    g.nbins = c.npad;
    g.c0 = (double)c.npad / 2.0;
    g.xc = g.c0 - (double)c.L;
    g.cA.resize(A);
    g.cB.resize(A);
    wa.resize(A);
    const double ascale = fbp_angle_scale(angles, A);
    for (int a = 0; a < A; ++a) {
        const double th = (double)angles[a];
        g.cA[a] = std::cos(th);
        g.cB[a] = -std::sin(th);
        wa[a] = (float)(angle_step(angles, A, a) * ascale);
    }
    bp_window(g);
}

// FBP work of one call on stream st; the caller holds D.mu with the device
// bound.  sync: wait for the stream before returning (the per-call tables in
// the arena are then free for the next call).
static void fbp_run(Device& D, float* d_obj, const float* d_tomo, const float* angles, const FbpCfg& c, int store,
                    hipStream_t st, bool sync)
{
    const int A = c.A, N = c.N;
    BpGeom g;
    std::vector<float> wa;
    fbp_geometry(c, angles, g, wa);

    FilterTables& F = filter_tables(D, c, st);
    TableSet tb(D);
    const TabRef dcA = tb.add(g.cA), dcB = tb.add(g.cB), dwa = tb.add(wa);
    tb.upload(st);
    if (!sync) tb.keep_alive();

    const int SB = (store == RAFT_STORE_F16) ? BP_SLICES_F16 : BP_SLICES_F32;
    const size_t per_group = (size_t)A * g.width * 16; // SB values per bin, 16 bytes
    // blocks are whole slice groups (except the last)
    int B = c.blocksize > 0 ? std::min(c.blocksize, c.S) : c.S;
    B = std::min(c.S, ((B + SB - 1) / SB) * SB);
    const size_t budget = block_budget();
    while (B > SB && (size_t)((B + SB - 1) / SB) * per_group > budget) B = (((B + 1) / 2 + SB - 1) / SB) * SB;
    void* q = D.workspace<0>((size_t)((B + SB - 1) / SB) * per_group);

    for (int s0 = 0; s0 < c.S; s0 += B) {
        const int nb = std::min(B, c.S - s0);
        FilterArgs fa;
        fa.tomo = d_tomo + (size_t)s0 * A * N;
        fa.q = q;
        fa.H = F.H.as<float2>();
        fa.tw = F.tw.as<float2>();
        fa.g = F.g.as<double>();
        fa.wa = dwa.as<float>();
        fa.N = N; fa.A = A; fa.S = nb; fa.npad = c.npad; fa.L = c.L; fa.tlo = g.tlo; fa.W = g.width;
        fa.none = (c.filter == 0);
        {
            ProfScope ps(KF_FILTER, st);
            launch_fbp_filter(fa, F.plan, store, st);
        }
        BpArgs ba;
        ba.q = q;
        ba.out = d_obj + (size_t)s0 * N * N;
        ba.x = nullptr; ba.norm = nullptr; ba.xt = nullptr;
        ba.cA = dcA.as<double>();
        ba.cB = dcB.as<double>();
        ba.xc = g.xc; ba.c0 = g.c0;
        ba.n = N; ba.A = A; ba.S = nb; ba.tlo = g.tlo; ba.W = g.width; ba.mode = 0;
        ba.q_bytes = (size_t)((B + SB - 1) / SB) * per_group;
        // angle chunk: tuned on this block's rows (the timed launches write
        // the same output as the final one)
        ba.chunk = tune_pick(bp_tune_key(N, A, g.width, store, 0, (nb + SB - 1) / SB), bp_cands(store), st,
                             [&](int c) {
                                 BpArgs t = ba;
                                 t.chunk = c;
                                 launch_backproject(t, store, st);
                             });
        {
            ProfScope ps(KF_BP, st);
            launch_backproject(ba, store, st);
        }
    }
    // the small tables above are freed when this call returns
    if (sync) RAFT_HIP(hipStreamSynchronize(st));
}

static void fbp_device(int dev, float* d_obj, const float* d_tomo, const float* angles, const float* paramf,
                       const int* parami, int store, hipStream_t st)
{
    const FbpCfg c = decode_fbp(paramf, parami);
    if (!angles) fail(RAFT_EINVAL, "null angles");
    if (store != RAFT_STORE_F32 && store != RAFT_STORE_F16) fail(RAFT_EINVAL, "store must be F32 or F16");
    if (c.S == 0) return;
    if (!d_obj || !d_tomo) fail(RAFT_EINVAL, "null data pointer");
    bind_device(dev);
    Device& D = device_state(dev);
    std::lock_guard<std::mutex> lk(D.mu);
    fbp_run(D, d_obj, d_tomo, angles, c, store, st, true);
}

// FBP with every intermediate in the library's fixed device scratch array
// (noalloc.hip) instead of the workspace: the filter tables, the angle
// cosines, sines and weights, then as many slice groups of filtered rows as
// fit.  The volume runs in blocks of that many groups through the same
// kernels with the same arguments as fbp_run, so the output is fbp_run's bit
// for bit (slice groups are independent, and every angle chunk of the
// backprojector gives the same sums).  Nothing here allocates device memory:
// no workspace, no table arena, no cached device filter tables, no tuning
// runs (the backprojector takes the tuner's cached chunk or its default).
// Nothing in the scratch must start at zero: the filter writes every entry of
// the rows the backprojector reads, and the tables are copied in whole.
static std::mutex g_host_filters_mu;
static std::map<std::string, std::unique_ptr<FilterHost>> g_host_filters; // host tables per filter_key

static void fbp_device_noalloc(int dev, float* d_obj, const float* d_tomo, const float* angles, const float* paramf,
                               const int* parami, int store, hipStream_t st)
{
    const FbpCfg c = decode_fbp(paramf, parami);
    if (!angles) fail(RAFT_EINVAL, "null angles");
    if (store != RAFT_STORE_F32 && store != RAFT_STORE_F16) fail(RAFT_EINVAL, "store must be F32 or F16");
    if (c.S == 0) return;
    if (!d_obj || !d_tomo) fail(RAFT_EINVAL, "null data pointer");
    const int A = c.A, N = c.N;
    BpGeom g;
    std::vector<float> wa;
    fbp_geometry(c, angles, g, wa);
    const FilterHost* F = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_host_filters_mu);
        auto& slot = g_host_filters[filter_key(c)];
        if (!slot) {
            std::unique_ptr<FilterHost> h(new FilterHost());
            filter_host_tables(c, *h);
            slot = std::move(h);
        }
        F = slot.get();
    }

    // scratch layout: tables (256-byte aligned pieces), then the rows
    std::vector<char> host;
    auto put = [&](const void* src, size_t bytes) {
        const size_t off = (host.size() + 255) & ~(size_t)255;
        host.resize(off + bytes);
        if (bytes) std::memcpy(host.data() + off, src, bytes);
        return off;
    };
    const size_t oH = put(F->H.data(), F->H.size() * sizeof(float2));
    const size_t oTw = put(F->tw.data(), F->tw.size() * sizeof(float2));
    const size_t oG = put(F->g.data(), F->g.size() * sizeof(double));
    const size_t oA = put(g.cA.data(), (size_t)A * sizeof(double));
    const size_t oB = put(g.cB.data(), (size_t)A * sizeof(double));
    const size_t oW = put(wa.data(), (size_t)A * sizeof(float));
    const size_t oRows = (host.size() + 255) & ~(size_t)255;

    const int SB = (store == RAFT_STORE_F16) ? BP_SLICES_F16 : BP_SLICES_F32;
    const size_t per_group = (size_t)A * g.width * 16;
    const size_t cap = (size_t)RAFT_NOALLOC_SCRATCH_BYTES;
    if (oRows > cap || cap - oRows < per_group)
        fail(RAFT_EUNSUPPORTED, "raft_fbp_device_noalloc: tables (" + std::to_string(oRows) +
                                    " bytes) and one slice group of filtered rows (" + std::to_string(per_group) +
                                    " bytes) exceed the fixed scratch of " + std::to_string(cap) + " bytes");
    const int groups_fit = (int)std::min<size_t>((cap - oRows) / per_group, 1 << 20);
    int B = c.blocksize > 0 ? std::min(c.blocksize, c.S) : c.S;
    B = std::min(c.S, ((B + SB - 1) / SB) * SB);
    B = std::min(B, groups_fit * SB);

    bind_device(dev);
    Device& D = device_state(dev);
    std::lock_guard<std::mutex> lk(D.mu);
    char* base = static_cast<char*>(noalloc_scratch(dev));
    RAFT_HIP(hipMemcpyAsync(base, host.data(), host.size(), hipMemcpyHostToDevice, st));
    void* q = base + oRows;
    const size_t q_bytes = (size_t)((B + SB - 1) / SB) * per_group;
    const int chunk = tune_cached(bp_tune_key(N, A, g.width, store, 0, (B + SB - 1) / SB), BP_CHUNK);

    for (int s0 = 0; s0 < c.S; s0 += B) {
        const int nb = std::min(B, c.S - s0);
        FilterArgs fa;
        fa.tomo = d_tomo + (size_t)s0 * A * N;
        fa.q = q;
        fa.H = reinterpret_cast<const float2*>(base + oH);
        fa.tw = reinterpret_cast<const float2*>(base + oTw);
        fa.g = F->g.empty() ? nullptr : reinterpret_cast<const double*>(base + oG);
        fa.wa = reinterpret_cast<const float*>(base + oW);
        fa.N = N; fa.A = A; fa.S = nb; fa.npad = c.npad; fa.L = c.L; fa.tlo = g.tlo; fa.W = g.width;
        fa.none = (c.filter == 0);
        launch_fbp_filter(fa, F->plan, store, st);
        BpArgs ba;
        ba.q = q;
        ba.out = d_obj + (size_t)s0 * N * N;
        ba.x = nullptr; ba.norm = nullptr; ba.xt = nullptr;
        ba.cA = reinterpret_cast<const double*>(base + oA);
        ba.cB = reinterpret_cast<const double*>(base + oB);
        ba.xc = g.xc; ba.c0 = g.c0;
        ba.n = N; ba.A = A; ba.S = nb; ba.tlo = g.tlo; ba.W = g.width; ba.mode = 0;
        ba.q_bytes = q_bytes;
        ba.chunk = bp_variant_supported(chunk, store) ? chunk : BP_CHUNK;
        launch_backproject(ba, store, st);
    }
    // the host copy of the tables stays alive until the stream is done
    RAFT_HIP(hipStreamSynchronize(st));
}

// ---------------------------------------------------------------------------
// Radon on device pointers
// ---------------------------------------------------------------------------
// FBP by the Backprojection Slice Theorem (bst.cu:251-312, 536-736): same
// parameter layout as getFBPMultiGPU; the padded rows (edge replicate, pad
// L = pad N / 2) are filtered in frequency by apply(min(k, n-k)/n) times
// exp(-2 pi i c k / n), c = (int) axis offset (filters.cu:295-307, 369-390,
// unnormalised inverse), then launch_bst; the image is cropped from the
// n x n plane at offset L and divided by 4 n pixel.  The reference assumes
// angles uniform on [0, pi) (the Python front end passes linspace).
static void bst_device(int dev, float* d_obj, const float* d_tomo, const float* paramf, const int* parami,
                       hipStream_t st)
{
    const FbpCfg c = decode_fbp(paramf, parami);
    if (c.S == 0) return;
    if (!d_obj || !d_tomo) fail(RAFT_EINVAL, "null data pointer");
    bind_device(dev);
    Device& D = device_state(dev);
    std::lock_guard<std::mutex> lk(D.mu);
    const int n = c.npad, A = c.A, N = c.N;
    FilterTables& F = filter_tables(D, c, st);
    if (!F.plan.ok) fail(RAFT_EUNSUPPORTED, "BST needs a padded size n = N(1+pad) <= 16384 with factors 2, 3, 5");
    std::vector<float2> H(n);
    const int center = (int)c.offset;
    for (int k = 0; k < n; ++k) {
        const double w = (double)std::min(k, n - k) / (double)n;
        const double amp = filter_apply(c.filter, w, c.reg, c.paganin);
        const double ph = -2.0 * M_PI * (double)center * (double)k / (double)n;
        H[k] = make_float2((float)(amp * std::cos(ph)), (float)(amp * std::sin(ph)));
    }
    const std::vector<float> ones(A, 1.0f);
    TableSet tb(D);
    const TabRef dH = tb.add(H), dwa = tb.add(ones);
    tb.upload(st);
    // the fused form (default) never holds the whole n x n grid: the kept N
    // columns of every row, then their transpose; RAFT_BST_FUSED=0 runs the
    // gather and the row transforms as separate passes over the grid
    const char* fenv = std::getenv("RAFT_BST_FUSED");
    const bool fused = !(fenv && fenv[0] == '0');
    // the filter runs inside the polar pass (default, where its LDS fits), so
    // the filtered padded rows are never written; RAFT_BST_FILTER_FUSED=0
    // filters into staged rows first, as a separate pass
    const char* ffenv = std::getenv("RAFT_BST_FILTER_FUSED");
    const bool ffused = !(ffenv && ffenv[0] == '0') && bst_filter_polar_fits(n);
    const size_t s2 = (size_t)n * n;
    const size_t cart_elems = fused ? (size_t)n * N : s2;
    const size_t per_slice = (ffused ? 0 : (size_t)A * n * 4) + (size_t)2 * A * (n / 2) * 8 +
                             (cart_elems + (size_t)n * N) * 8;
    int B = c.blocksize > 0 ? std::min(c.blocksize, c.S) : c.S;
    B = ((B + 3) / 4) * 4;
    const size_t budget = block_budget();
    while (B > 4 && (size_t)B * per_slice > budget) B = (((B + 1) / 2 + 3) / 4) * 4;
    float* q = ffused ? nullptr : static_cast<float*>(D.workspace<0>((size_t)B * A * n * 4));
    float2* polar = static_cast<float2*>(D.workspace<1>((size_t)B * 2 * A * (n / 2) * 8));
    float2* cart = static_cast<float2*>(D.workspace<2>((size_t)B * cart_elems * 8));
    float2* cartT = static_cast<float2*>(D.workspace<3>((size_t)B * n * N * 8));
    for (int s0 = 0; s0 < c.S; s0 += B) {
        const int nb = std::min(B, c.S - s0);
        FilterArgs fa;
        fa.tomo = d_tomo + (size_t)s0 * A * N;
        fa.q = q;
        fa.H = dH.as<float2>();
        fa.tw = F.tw.as<float2>();
        fa.g = nullptr;
        fa.wa = dwa.as<float>();
        fa.N = N; fa.A = A; fa.S = nb; fa.npad = n; fa.L = c.L; fa.tlo = 0; fa.W = n;
        fa.none = (c.filter == 0);
        if (!ffused) launch_fbp_filter(fa, F.plan, RAFT_STORE_F32, st);
        BstArgs ba;
        ba.q = q; ba.polar = polar; ba.cart = cart; ba.cartT = cartT;
        ba.obj = d_obj + (size_t)s0 * N * N;
        ba.tw = F.tw.as<float2>();
        ba.N = N; ba.A = A; ba.S = nb; ba.L = c.L;
        ba.scale = (float)((double)n * c.pixel * 4.0);
        ba.fused = fused ? 1 : 0;
        launch_bst(ba, F.plan, st, ffused ? &fa : nullptr);
    }
    RAFT_HIP(hipStreamSynchronize(st));
}

// Filter-only export (low_pass.cu:9-141 / filters.cu:120-161 filterFBPpad):
// edge-replicate pad, the FBP filter with no axis offset, crop back to the N
// original bins, / (npad * pixel); in place.  Filter 'none' leaves the data as is.
static void lowpass_device(int dev, float* d_tomo, const float* paramf, const int* parami, hipStream_t st)
{
    if (!paramf) fail(RAFT_EINVAL, "null parameter array");
    float pf[7];
    std::memcpy(pf, paramf, sizeof(pf));
    pf[6] = 0.0f; // getFilterLowPass builds its Filter with axis_offset = 0
    const FbpCfg c = decode_fbp(pf, parami);
    if (c.S == 0 || c.filter == 0) return;
    if (!d_tomo) fail(RAFT_EINVAL, "null data pointer");
    bind_device(dev);
    Device& D = device_state(dev);
    std::lock_guard<std::mutex> lk(D.mu);
    const int A = c.A, N = c.N;
    FilterTables& F = filter_tables(D, c, st);
    std::vector<float> ones(A, 1.0f);
    TableSet tb(D);
    const TabRef dwa = tb.add(ones);
    tb.upload(st);
    const int SB = BP_SLICES_F32;
    const size_t per_group = (size_t)A * N * 16;
    int B = c.blocksize > 0 ? std::min(c.blocksize, c.S) : c.S;
    B = std::min(c.S, ((B + SB - 1) / SB) * SB);
    const size_t budget = block_budget();
    while (B > SB && (size_t)((B + SB - 1) / SB) * per_group > budget) B = (((B + 1) / 2 + SB - 1) / SB) * SB;
    float* q = static_cast<float*>(D.workspace<0>((size_t)((B + SB - 1) / SB) * per_group));
    for (int s0 = 0; s0 < c.S; s0 += B) {
        const int nb = std::min(B, c.S - s0);
        FilterArgs fa;
        fa.tomo = d_tomo + (size_t)s0 * A * N;
        fa.q = q;
        fa.H = F.H.as<float2>();
        fa.tw = F.tw.as<float2>();
        fa.g = F.g.as<double>();
        fa.wa = dwa.as<float>();
        fa.N = N; fa.A = A; fa.S = nb; fa.npad = c.npad; fa.L = c.L; fa.tlo = c.L; fa.W = N;
        fa.none = 0;
        // in place where the filter form writes rows directly (FP32 whole
        // entries with the window in LDS), else through q
        fa.rows = d_tomo + (size_t)s0 * A * N;
        bool direct = false;
        {
            ProfScope ps(KF_FILTER, st);
            direct = launch_fbp_filter(fa, F.plan, RAFT_STORE_F32, st);
        }
        if (!direct) launch_from_interleaved(q, d_tomo + (size_t)s0 * A * N, nb, A, N, st);
    }
    RAFT_HIP(hipStreamSynchronize(st));
}

static void radon_device(int dev, float* d_proj, const float* d_obj, const float* angles, int Nx, int Ny, int nrays,
                         int nangles, int nslices, float ax, float ay, hipStream_t st)
{
    if (Nx < 2 || Ny < 2) fail(RAFT_EINVAL, "image sizes must be >= 2");
    if (nrays < 1 || nangles < 1 || nslices < 0) fail(RAFT_EINVAL, "invalid projection sizes");
    if (!(ax != 0.0f) || !(ay != 0.0f)) fail(RAFT_EINVAL, "ax, ay must be non-zero");
    if ((double)Nx * (double)Ny * 16.0 >= 2147483648.0) fail(RAFT_EUNSUPPORTED, "image too large for 32-bit offsets");
    if (!angles) fail(RAFT_EINVAL, "null angles");
    if (nslices == 0) return;
    if (!d_proj || !d_obj) fail(RAFT_EINVAL, "null data pointer");
    bind_device(dev);
    Device& D = device_state(dev);
    std::lock_guard<std::mutex> lk(D.mu);

    std::vector<double> cv(nangles), sv(nangles);
    for (int a = 0; a < nangles; ++a) {
        cv[a] = std::cos((double)angles[a]);
        sv[a] = std::sin((double)angles[a]);
    }
    TableSet tb(D);
    const TabRef dc = tb.add(cv), ds = tb.add(sv);
    tb.upload(st);
    RayGeom g = ray_geom_radon(Nx, Ny, nrays, nangles, (double)ax, (double)ay);
    const size_t img = (size_t)Nx * Ny;
    int B = ((nslices + 3) / 4) * 4;
    const size_t budget = block_budget();
    while (B > 4 && (size_t)B * img * 8 > budget) B = (((B + 1) / 2 + 3) / 4) * 4;
    float* il = static_cast<float*>(D.workspace<1>((size_t)B * img * 4));
    float* ilT = static_cast<float*>(D.workspace<2>((size_t)B * img * 4));
    const int prc = proj_plan(g, cv.data(), sv.data(), B, true);
    for (int s0 = 0; s0 < nslices; s0 += B) {
        const int nb = std::min(B, nslices - s0);
        const float* src = d_obj + (size_t)s0 * img;
        launch_to_interleaved(src, il, nb, Ny, Nx, false, st);
        launch_to_interleaved(src, ilT, nb, Ny, Nx, true, st);
        ProjArgs p;
        p.img = il; p.imgT = ilT; p.out = d_proj + (size_t)s0 * nangles * nrays;
        p.y = nullptr; p.wa = nullptr; p.kl_part = nullptr; p.ex_part = nullptr;
        p.cosv = dc.as<double>(); p.sinv = ds.as<double>();
        p.g = g; p.S = nb; p.tlo = 0; p.W = nrays; p.lds_rc = prc;
        if (prc > 0)
            p.lds_rc = tune_pick(proj_tune_key(g, 0, (nb + 3) / 4), proj_rc_cands(g, prc), st, [&](int c) {
                ProjArgs t = p;
                t.lds_rc = c;
                launch_project(t, 0, st);
            });
        ProfScope ps(KF_PROJ, st);
        launch_project(p, 0, st);
    }
    RAFT_HIP(hipStreamSynchronize(st));
}

// ---------------------------------------------------------------------------
// eEM on device pointers
// ---------------------------------------------------------------------------
// kl_only: KL(y || A x) of the given images (one projection, x unchanged), kl[S]
// kl_iters >= 0: the KL of the first kl_iters iterations only, kl[S][kl_iters]
// (the others run without the FP64 terms); < 0: of every iteration, kl[S][K]
// sched (host, [S], non-increasing): slice s runs sched[s] iterations (the
// iteration count of parami is ignored; no KL); iteration t updates only the
// slices with sched[s] > t, and the launches cover only their slice groups
static void eem_device(int dev, float* d_obj, const float* d_tomo, const float* angles, const int* parami,
                       double* kl, long long* kl_ex, hipStream_t st, bool kl_only = false, int kl_iters = -1,
                       const int* sched = nullptr)
{
    if (!parami) fail(RAFT_EINVAL, "null parameter array");
    if (sched && parami[2] > 0) {
        for (int s = 0; s < parami[2]; ++s)
            if (sched[s] < 0 || (s > 0 && sched[s] > sched[s - 1]))
                fail(RAFT_EINVAL, "the per-slice iteration counts must be >= 0 and non-increasing");
        if (kl || kl_ex || kl_only) fail(RAFT_EINVAL, "a scheduled eEM run records no KL");
    }
    const int R = parami[0], A = parami[1], S = parami[2], n = parami[3];
    const int K = kl_only ? 1 : (sched ? (S > 0 ? sched[0] : 0) : parami[8]);
    if (kl_only && !kl && !kl_ex) fail(RAFT_EINVAL, "no KL output");
    int blocksize = parami[9];
    if (R < 2) fail(RAFT_EINVAL, "nrays must be >= 2");
    if (A < 2) fail(RAFT_EINVAL, "nangles must be >= 2");
    if (n < 2) fail(RAFT_EINVAL, "objsize must be >= 2");
    if (S < 0 || K < 0) fail(RAFT_EINVAL, "nslices and iterations must be >= 0");
    if (!angles) fail(RAFT_EINVAL, "null angles");
    const double rr = (double)(R - 1) / (double)(n - 1);
    if ((double)n * (double)n * 16.0 >= 2147483648.0) fail(RAFT_EUNSUPPORTED, "image too large for 32-bit offsets");
    if (rr * std::hypot((double)(BP_TILE_X - 1), (double)(BP_TILE_Y - 1)) + 5.0 > (double)BP_SLOT)
        fail(RAFT_EUNSUPPORTED, "nrays larger than objsize beyond the backprojection tile window");
    if (S == 0) return;
    if (!d_obj || !d_tomo) fail(RAFT_EINVAL, "null data pointer");
    bind_device(dev);
    Device& D = device_state(dev);
    std::lock_guard<std::mutex> lk(D.mu);

    std::vector<double> cv(A), sv(A);
    std::vector<float> wa(A);
    BpGeom g;
    g.n = n;
    g.nangles = A;
    g.nbins = R;
    g.xc = (double)(n - 1) / 2.0;
    g.c0 = (double)(R - 1) / 2.0;
    g.cA.resize(A);
    g.cB.resize(A);
    for (int a = 0; a < A; ++a) {
        cv[a] = std::cos((double)angles[a]);
        sv[a] = std::sin((double)angles[a]);
        g.cA[a] = cv[a] * rr;
        g.cB[a] = sv[a] * rr;
        wa[a] = (float)angle_step(angles, A, a);
    }
    bp_window(g);
    // the ratio rows are written at offset (i - tlo): the window must cover [0, R)
    if (g.tlo > 0 || g.tlo + g.width < R) {
        const int lo = std::min(g.tlo, 0), hi = std::max(g.tlo + g.width, R);
        g.tlo = lo;
        g.width = hi - lo;
    }
    TableSet tb(D);
    const TabRef dc = tb.add(cv), ds = tb.add(sv), dcA = tb.add(g.cA), dcB = tb.add(g.cB), dwa = tb.add(wa);
    tb.upload(st);
    RayGeom rg = ray_geom_em(R, A, n);

    const size_t img = (size_t)n * n;
    // per slice: interleaved image + transpose, staged ratio rows, KL partials
    const size_t per_slice = img * 8 + (size_t)A * g.width * 4 + (size_t)A * 16;
    int B = blocksize > 0 ? std::min(blocksize, S) : S;
    B = ((B + 3) / 4) * 4;
    const size_t budget = block_budget();
    while (B > 4 && (size_t)B * per_slice > budget) B = (((B + 1) / 2 + 3) / 4) * 4;
    const int Gb = B / 4;

    float* rows = static_cast<float*>(D.workspace<0>((size_t)Gb * A * g.width * 16));
    float* xil = static_cast<float*>(D.workspace<1>((size_t)Gb * img * 16));
    float* xT = static_cast<float*>(D.workspace<2>((size_t)Gb * img * 16));
    float* ones = static_cast<float*>(D.workspace<3>(img * 4 + (size_t)A * g.width * 16));
    float* onerows = ones + img;
    // staged-projector ray chunks: the plan's, or the autotuner's choice among
    // powers of two (made on the first block, below); the KL partial slots
    // are sized for the largest candidate
    int prc = proj_plan(rg, cv.data(), sv.data(), B, true);
    const std::vector<int> rc_cands = prc > 0 ? proj_rc_cands(rg, prc) : std::vector<int>{prc};
    int kslots = proj_kl_slots(A, prc);
    size_t kslots_cap = (size_t)kslots;
    for (int c : rc_cands) kslots_cap = std::max(kslots_cap, (size_t)proj_kl_slots(A, c));
    double* part = static_cast<double*>(D.workspace<4>((size_t)B * kslots_cap * 16));
    unsigned long long* expart = reinterpret_cast<unsigned long long*>(part + (size_t)B * kslots_cap);
    int bp_chunk = BP_CHUNK;
    const bool want_kl = (kl != nullptr) || (kl_ex != nullptr);
    // iterations whose KL is recorded (the first Kkl), the row stride of kl
    const int Kkl = want_kl ? (kl_iters < 0 ? K : std::min(kl_iters, K)) : 0;
    double* dkl = nullptr;
    unsigned long long* dex = nullptr;
    if (Kkl > 0) {
        double* buf = static_cast<double*>(D.workspace<5>((size_t)S * Kkl * 16));
        dkl = buf;
        dex = reinterpret_cast<unsigned long long*>(buf + (size_t)S * Kkl);
    }

    // backones = B(1): the same for every slice
    if (!kl_only) {
        launch_fill_rows(onerows, dwa.as<float>(), A, g.width, g.tlo, R, st);
        BpArgs b1;
        b1.q = onerows; b1.out = ones; b1.x = nullptr; b1.norm = nullptr; b1.xt = nullptr;
        b1.cA = dcA.as<double>(); b1.cB = dcB.as<double>(); b1.xc = g.xc; b1.c0 = g.c0;
        b1.n = n; b1.A = A; b1.S = 1; b1.tlo = g.tlo; b1.W = g.width; b1.mode = 0;
        b1.q_bytes = (size_t)A * g.width * 16;
        launch_backproject(b1, RAFT_STORE_F32, st);
    }
    // ratio rows: margins stay zero, the projector writes [0, R) (none for the KL alone)
    if (!kl_only) RAFT_HIP(hipMemsetAsync(rows, 0, (size_t)Gb * A * g.width * 16, st));

    // one iteration: projector + ratio (+ KL partials), KL reduction, backprojection + update
    int* counter = reinterpret_cast<int*>(static_cast<char*>(D.workspace<6>(256)));
    auto iteration = [&](int it, int nb, const float* y, double* klo, unsigned long long* exo, bool graph,
                         hipStream_t s, bool with_kl) {
        ProjArgs p;
        p.img = xil; p.imgT = xT; p.out = kl_only ? nullptr : rows; p.y = y; p.wa = dwa.as<float>();
        p.kl_part = with_kl ? part : nullptr;
        p.ex_part = with_kl ? expart : nullptr;
        p.cosv = dc.as<double>(); p.sinv = ds.as<double>();
        p.g = rg; p.S = nb; p.tlo = g.tlo; p.W = g.width; p.lds_rc = prc;
        {
            ProfScope ps(KF_PROJ, s);
            launch_project(p, 1, s);
        }
        if (with_kl) launch_kl_reduce(part, expart, kslots, nb, klo, exo, Kkl, it, graph ? counter : nullptr, s);
        if (kl_only) return;
        BpArgs b;
        b.q = rows; b.out = xil; b.x = xil; b.norm = ones; b.xt = xT;
        b.cA = dcA.as<double>(); b.cB = dcB.as<double>(); b.xc = g.xc; b.c0 = g.c0;
        b.n = n; b.A = A; b.S = nb; b.tlo = g.tlo; b.W = g.width; b.mode = 1; b.chunk = bp_chunk;
        b.q_bytes = (size_t)Gb * A * g.width * 16;
        {
            ProfScope ps(KF_EMBP, s);
            launch_backproject(b, RAFT_STORE_F32, s);
        }
        if (graph) launch_counter(counter, 1, s);
    };
    // tile variants, timed on the first block's data: the projector writes
    // only ratio rows (rewritten by the first iteration), the update reads
    // those rows (nb slices, all their groups) and writes a scratch image
    auto tune = [&](int nb, const float* y) {
        if (prc > 0) {
            ProjArgs tp;
            tp.img = xil; tp.imgT = xT; tp.out = kl_only ? nullptr : rows; tp.y = y; tp.wa = dwa.as<float>();
            tp.kl_part = nullptr; tp.ex_part = nullptr;
            tp.cosv = dc.as<double>(); tp.sinv = ds.as<double>();
            tp.g = rg; tp.S = nb; tp.tlo = g.tlo; tp.W = g.width;
            prc = tune_pick(proj_tune_key(rg, 1, (nb + 3) / 4), rc_cands, st, [&](int c) {
                tp.lds_rc = c;
                launch_project(tp, 1, st);
            });
            kslots = proj_kl_slots(A, prc);
        }
        if (kl_only) return;
        const std::string key = bp_tune_key(n, A, g.width, RAFT_STORE_F32, 1, (nb + 3) / 4);
        float* tmp = nullptr;
        if (tune_enabled() && tune_cached(key, -1) < 0)
            tmp = static_cast<float*>(D.workspace<7>((size_t)((nb + 3) / 4) * img * 16));
        BpArgs tb;
        tb.q = rows; tb.out = tmp; tb.x = xil; tb.norm = ones; tb.xt = nullptr;
        tb.cA = dcA.as<double>(); tb.cB = dcB.as<double>(); tb.xc = g.xc; tb.c0 = g.c0;
        tb.n = n; tb.A = A; tb.S = nb; tb.tlo = g.tlo; tb.W = g.width; tb.mode = 1;
        tb.q_bytes = (size_t)Gb * A * g.width * 16;
        bp_chunk = tune_pick(key, bp_cands(RAFT_STORE_F32, rr), st, [&](int c) {
            tb.chunk = c;
            launch_backproject(tb, RAFT_STORE_F32, st);
        });
    };
    // Iterations are replayed from a captured hipGraph unless per-kernel timing is on
    // (RAFT_EM_GRAPH=0 disables it).  The graph holds kernels only: the counter it
    // advances is cleared by a kernel outside it.
    const char* genv = std::getenv("RAFT_EM_GRAPH");
    // iterations [0, it0) run as direct launches with the KL terms; the rest
    // [it0, K) with the KL terms when kl_rest, replayed from a graph when there
    // are at least two of them
    const bool kl_rest = Kkl > 0 && Kkl == K;
    const int it0 = kl_rest ? 0 : Kkl;
    const bool use_graph = !kl_only && !sched && K - it0 >= 2 && !prof_enabled() && !(genv && genv[0] == '0');

    for (int s0 = 0; s0 < S; s0 += B) {
        const int nb = std::min(B, S - s0);
        float* x = d_obj + (size_t)s0 * img;
        const float* y = d_tomo + (size_t)s0 * A * R;
        double* klo = Kkl > 0 ? dkl + (size_t)s0 * Kkl : nullptr;
        unsigned long long* exo = Kkl > 0 ? dex + (size_t)s0 * Kkl : nullptr;
        launch_to_interleaved(x, xil, nb, n, n, false, st);
        launch_to_interleaved(x, xT, nb, n, n, true, st);
        if (s0 == 0) tune(nb, y);
        for (int it = 0; it < it0; ++it) iteration(it, nb, y, klo, exo, false, st, true);
        if (use_graph) {
            launch_counter(counter, 0, st);
            hipStream_t cs = nullptr;
            RAFT_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
            hipGraph_t graph = nullptr;
            hipGraphExec_t exec = nullptr;
            hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
            if (e == hipSuccess) {
                try {
                    iteration(0, nb, y, klo, exo, true, cs, kl_rest);
                } catch (...) {
                    (void)hipStreamEndCapture(cs, &graph);
                    if (graph) (void)hipGraphDestroy(graph);
                    (void)hipStreamDestroy(cs);
                    throw;
                }
                e = hipStreamEndCapture(cs, &graph);
            }
            if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            if (e == hipSuccess)
                for (int it = it0; it < K && e == hipSuccess; ++it) e = hipGraphLaunch(exec, st);
            if (exec) (void)hipGraphExecDestroy(exec);
            if (graph) (void)hipGraphDestroy(graph);
            (void)hipStreamDestroy(cs);
            RAFT_HIP(e);
        } else if (sched) {
            // slices still running at iteration it: a prefix of the block
            int active = nb;
            for (int it = 0; it < K; ++it) {
                while (active > 0 && sched[s0 + active - 1] <= it) --active;
                if (active == 0) break;
                iteration(it, active, y, klo, exo, false, st, false);
            }
        } else {
            for (int it = it0; it < K; ++it) iteration(it, nb, y, klo, exo, false, st, kl_rest);
        }
        if (!kl_only) launch_from_interleaved(xil, x, nb, n, n, st);
    }
    if (Kkl > 0) {
        std::vector<double> hk((size_t)S * Kkl);
        std::vector<unsigned long long> he((size_t)S * Kkl);
        RAFT_HIP(hipMemcpyAsync(hk.data(), dkl, hk.size() * 8, hipMemcpyDeviceToHost, st));
        RAFT_HIP(hipMemcpyAsync(he.data(), dex, he.size() * 8, hipMemcpyDeviceToHost, st));
        RAFT_HIP(hipStreamSynchronize(st));
        if (kl) std::memcpy(kl, hk.data(), hk.size() * 8);
        if (kl_ex)
            for (size_t e = 0; e < he.size(); ++e) kl_ex[e] = (long long)he[e];
    }
    RAFT_HIP(hipStreamSynchronize(st));
}

// Transmission EM (em_rt.cu:155-192, em_RT_kernel.cu:20-31,52-65,92-184):
//   backcounts = B(count);  K x { x <- x * B(flat * exp(-A x)) / backcounts }
// with the eEM projector / backprojector geometry.  The flat field is the
// slice's own row flat[s][nrays] (the reference indexes it by angle,
// em_RT_kernel.cu:29, reading rows it never copied).
static void tem_device(int dev, float* d_obj, const float* d_count, const float* d_flat, const float* angles,
                       const int* parami, hipStream_t st)
{
    if (!parami) fail(RAFT_EINVAL, "null parameter array");
    const int R = parami[0], A = parami[1], S = parami[2], n = parami[3], K = parami[8];
    const int blocksize = parami[9];
    if (R < 2) fail(RAFT_EINVAL, "nrays must be >= 2");
    if (A < 2) fail(RAFT_EINVAL, "nangles must be >= 2");
    if (n < 2) fail(RAFT_EINVAL, "objsize must be >= 2");
    if (S < 0 || K < 0) fail(RAFT_EINVAL, "nslices and iterations must be >= 0");
    if (!angles) fail(RAFT_EINVAL, "null angles");
    const double rr = (double)(R - 1) / (double)(n - 1);
    if ((double)n * (double)n * 16.0 >= 2147483648.0) fail(RAFT_EUNSUPPORTED, "image too large for 32-bit offsets");
    if (rr * std::hypot((double)(BP_TILE_X - 1), (double)(BP_TILE_Y - 1)) + 5.0 > (double)BP_SLOT)
        fail(RAFT_EUNSUPPORTED, "nrays larger than objsize beyond the backprojection tile window");
    if (S == 0) return;
    if (!d_obj || !d_count || !d_flat) fail(RAFT_EINVAL, "null data pointer");
    bind_device(dev);
    Device& D = device_state(dev);
    std::lock_guard<std::mutex> lk(D.mu);

    std::vector<double> cv(A), sv(A);
    std::vector<float> wa(A);
    BpGeom g;
    g.n = n;
    g.nangles = A;
    g.nbins = R;
    g.xc = (double)(n - 1) / 2.0;
    g.c0 = (double)(R - 1) / 2.0;
    g.cA.resize(A);
    g.cB.resize(A);
    for (int a = 0; a < A; ++a) {
        cv[a] = std::cos((double)angles[a]);
        sv[a] = std::sin((double)angles[a]);
        g.cA[a] = cv[a] * rr;
        g.cB[a] = sv[a] * rr;
        wa[a] = (float)angle_step(angles, A, a);
    }
    bp_window(g);
    if (g.tlo > 0 || g.tlo + g.width < R) {
        const int lo = std::min(g.tlo, 0), hi = std::max(g.tlo + g.width, R);
        g.tlo = lo;
        g.width = hi - lo;
    }
    TableSet tb(D);
    const TabRef dc = tb.add(cv), ds = tb.add(sv), dcA = tb.add(g.cA), dcB = tb.add(g.cB), dwa = tb.add(wa);
    tb.upload(st);
    RayGeom rg = ray_geom_em(R, A, n);

    const size_t img = (size_t)n * n;
    const size_t per_slice = img * 16 + (size_t)A * g.width * 4;
    int B = blocksize > 0 ? std::min(blocksize, S) : S;
    B = ((B + 3) / 4) * 4;
    const size_t budget = block_budget();
    while (B > 4 && (size_t)B * per_slice > budget) B = (((B + 1) / 2 + 3) / 4) * 4;
    const int Gb = B / 4;
    float* rows = static_cast<float*>(D.workspace<0>((size_t)Gb * A * g.width * 16));
    float* xil = static_cast<float*>(D.workspace<1>((size_t)Gb * img * 16));
    float* xT = static_cast<float*>(D.workspace<2>((size_t)Gb * img * 16));
    float* normil = static_cast<float*>(D.workspace<3>((size_t)Gb * img * 16));
    float* tmp = static_cast<float*>(D.workspace<5>((size_t)B * img * 4));
    const int prc = proj_plan(rg, cv.data(), sv.data(), B, true);

    for (int s0 = 0; s0 < S; s0 += B) {
        const int nb = std::min(B, S - s0);
        float* x = d_obj + (size_t)s0 * img;
        // backcounts = B(count), per slice, into the interleaved layout of the iterate
        launch_rows_from_planar(rows, d_count + (size_t)s0 * A * R, dwa.as<float>(), nb, A, g.width, g.tlo, R, st);
        BpArgs b0;
        b0.q = rows; b0.out = tmp; b0.x = nullptr; b0.norm = nullptr; b0.xt = nullptr;
        b0.cA = dcA.as<double>(); b0.cB = dcB.as<double>(); b0.xc = g.xc; b0.c0 = g.c0;
        b0.n = n; b0.A = A; b0.S = nb; b0.tlo = g.tlo; b0.W = g.width; b0.mode = 0;
        b0.q_bytes = (size_t)Gb * A * g.width * 16;
        launch_backproject(b0, RAFT_STORE_F32, st);
        launch_to_interleaved(tmp, normil, nb, n, n, false, st);
        // the rows' margins (bins outside [0, R)) are zero from here on: the
        // projector rewrites only [0, R)
        launch_to_interleaved(x, xil, nb, n, n, false, st);
        launch_to_interleaved(x, xT, nb, n, n, true, st);
        for (int it = 0; it < K; ++it) {
            ProjArgs p;
            p.img = xil; p.imgT = xT; p.out = rows; p.y = d_flat + (size_t)s0 * R; p.wa = dwa.as<float>();
            p.kl_part = nullptr; p.ex_part = nullptr;
            p.cosv = dc.as<double>(); p.sinv = ds.as<double>();
            p.g = rg; p.S = nb; p.tlo = g.tlo; p.W = g.width; p.lds_rc = prc;
            {
                ProfScope ps(KF_PROJ, st);
                launch_project(p, 2, st);
            }
            BpArgs b;
            b.q = rows; b.out = xil; b.x = xil; b.norm = normil; b.norm_il = 1; b.xt = xT;
            b.cA = dcA.as<double>(); b.cB = dcB.as<double>(); b.xc = g.xc; b.c0 = g.c0;
            b.n = n; b.A = A; b.S = nb; b.tlo = g.tlo; b.W = g.width; b.mode = 1;
            b.q_bytes = (size_t)Gb * A * g.width * 16;
            {
                ProfScope ps(KF_EMBP, st);
                launch_backproject(b, RAFT_STORE_F32, st);
            }
        }
        launch_from_interleaved(xil, x, nb, n, n, st);
    }
    RAFT_HIP(hipStreamSynchronize(st));
}

// ---------------------------------------------------------------------------
// Batched uneven scans: FBP and eEM of many independent scans of different
// sizes in ONE graph launch.  Scans with the same geometry (size, angles, FBP
// parameters) form a bucket; each bucket is gathered into contiguous slices,
// reconstructed with the same kernels as fbp_device / eem_device, and
// scattered back.  All buckets are captured into one hipGraph on forked
// streams (independent branches), then launched once.
// ---------------------------------------------------------------------------
struct BatchBucket {
    std::vector<int> scans;
    int N = 0, A = 0;
    std::vector<float> angles, paramf;
    // FBP
    FbpCfg cf;
    BpGeom gf;
    FilterTables* F = nullptr;
    DevBuf fcA, fcB, fwa, frows, tomo, fout;
    // eEM
    BpGeom ge;
    RayGeom rg;
    int prc = 0; // staged projector plan (ray chunks), 0 = direct form
    DevBuf dc, ds, ecA, ecB, ewa, erows, xil, xT, ones, y, x;
    hipStream_t s = nullptr;
    hipEvent_t done = nullptr;
};

static void batch_device(int dev, int nscans, const int* nrays, const int* nangles, const float* const* angles,
                         const float* const* d_tomo, const float* const* d_y, float* const* d_fbp, float* const* d_em,
                         const float* paramf, int pad, int filter, int store, int K, hipStream_t st,
                         float* graph_ms)
{
    if (nscans < 0 || !nrays || !nangles || !angles || !paramf) fail(RAFT_EINVAL, "null batch description");
    if (store != RAFT_STORE_F32 && store != RAFT_STORE_F16) fail(RAFT_EINVAL, "store must be F32 or F16");
    if (K < 0) fail(RAFT_EINVAL, "iterations must be >= 0");
    const bool do_fbp = d_fbp != nullptr && d_tomo != nullptr;
    const bool do_em = d_em != nullptr && d_y != nullptr && K > 0;
    if (nscans == 0 || (!do_fbp && !do_em)) return;
    bind_device(dev);
    Device& D = device_state(dev);
    std::lock_guard<std::mutex> lk(D.mu);

    // buckets: same size, angles and FBP parameters
    std::map<std::string, std::unique_ptr<BatchBucket>> bmap;
    std::vector<BatchBucket*> buckets;
    for (int i = 0; i < nscans; ++i) {
        const int N = nrays[i], A = nangles[i];
        if (N < 2 || A < 2) fail(RAFT_EINVAL, "each scan needs nrays >= 2 and nangles >= 2");
        if (!angles[i]) fail(RAFT_EINVAL, "null angles");
        if ((do_fbp && (!d_tomo[i] || !d_fbp[i])) || (do_em && (!d_y[i] || !d_em[i])))
            fail(RAFT_EINVAL, "null scan pointer");
        std::string key(reinterpret_cast<const char*>(&N), sizeof N);
        key.append(reinterpret_cast<const char*>(&A), sizeof A);
        key.append(reinterpret_cast<const char*>(angles[i]), sizeof(float) * A);
        key.append(reinterpret_cast<const char*>(paramf + 7 * (size_t)i), sizeof(float) * 7);
        auto& b = bmap[key];
        if (!b) {
            b.reset(new BatchBucket());
            b->N = N;
            b->A = A;
            b->angles.assign(angles[i], angles[i] + A);
            b->paramf.assign(paramf + 7 * (size_t)i, paramf + 7 * (size_t)i + 7);
            buckets.push_back(b.get());
        }
        b->scans.push_back(i);
    }

    // set-up outside the capture: tables, uploads, buffers
    for (BatchBucket* b : buckets) {
        const int N = b->N, A = b->A, S = (int)b->scans.size();
        const size_t img = (size_t)N * N, sino = (size_t)A * N;
        if (do_fbp) {
            const int pi[10] = {N, A, S, N, pad, 0, 0, filter, 0, 0};
            b->cf = decode_fbp(b->paramf.data(), pi);
            std::vector<float> wa;
            fbp_geometry(b->cf, b->angles.data(), b->gf, wa);
            b->F = &filter_tables(D, b->cf, st);
            upload(b->fcA, b->gf.cA, st);
            upload(b->fcB, b->gf.cB, st);
            upload(b->fwa, wa, st);
            const int SB = (store == RAFT_STORE_F16) ? BP_SLICES_F16 : BP_SLICES_F32;
            b->frows.alloc((size_t)((S + SB - 1) / SB) * A * b->gf.width * 16);
            b->tomo.alloc((size_t)S * sino * 4);
            b->fout.alloc((size_t)S * img * 4);
        }
        if (do_em) {
            const double rr = 1.0;
            std::vector<double> cv(A), sv(A);
            std::vector<float> wa(A);
            BpGeom& g = b->ge;
            g.n = N;
            g.nangles = A;
            g.nbins = N;
            g.xc = (double)(N - 1) / 2.0;
            g.c0 = (double)(N - 1) / 2.0;
            g.cA.resize(A);
            g.cB.resize(A);
            for (int a = 0; a < A; ++a) {
                cv[a] = std::cos((double)b->angles[a]);
                sv[a] = std::sin((double)b->angles[a]);
                g.cA[a] = cv[a] * rr;
                g.cB[a] = sv[a] * rr;
                wa[a] = (float)angle_step(b->angles.data(), A, a);
            }
            bp_window(g);
            if (g.tlo > 0 || g.tlo + g.width < N) {
                const int lo = std::min(g.tlo, 0), hi = std::max(g.tlo + g.width, N);
                g.tlo = lo;
                g.width = hi - lo;
            }
            if ((double)N * (double)N * 16.0 >= 2147483648.0) fail(RAFT_EUNSUPPORTED, "image too large");
            upload(b->dc, cv, st);
            upload(b->ds, sv, st);
            upload(b->ecA, g.cA, st);
            upload(b->ecB, g.cB, st);
            upload(b->ewa, wa, st);
            b->rg = ray_geom_em(N, A, N);
            b->prc = proj_plan(b->rg, cv.data(), sv.data(), S, true);
            const int G = (S + 3) / 4;
            b->erows.alloc((size_t)G * A * g.width * 16);
            b->xil.alloc((size_t)G * img * 16);
            b->xT.alloc((size_t)G * img * 16);
            b->ones.alloc(img * 4 + (size_t)A * g.width * 16);
            b->y.alloc((size_t)S * sino * 4);
            b->x.alloc((size_t)S * img * 4);
        }
    }
    RAFT_HIP(hipStreamSynchronize(st));

    // Work of one bucket in phases.  With the work queue (default) the
    // backprojections and projections of all buckets run as single persistent
    // launches over every bucket (one launch per stage and iteration); the
    // per-bucket phases (copies, filter, eEM set-up) run as graph branches.
    // RAFT_BATCH_QUEUE=0 runs every stage per bucket, all inside its branch.
    const char* qenv = std::getenv("RAFT_BATCH_QUEUE");
    const bool use_queue = !(qenv && qenv[0] == '0');
    auto fbp_bp_args = [&](BatchBucket* b) {
        BpArgs ba;
        ba.q = b->frows.p; ba.out = b->fout.as<float>(); ba.x = nullptr; ba.norm = nullptr; ba.xt = nullptr;
        ba.cA = b->fcA.as<double>(); ba.cB = b->fcB.as<double>(); ba.xc = b->gf.xc; ba.c0 = b->gf.c0;
        ba.n = b->N; ba.A = b->A; ba.S = (int)b->scans.size(); ba.tlo = b->gf.tlo; ba.W = b->gf.width; ba.mode = 0;
        ba.q_bytes = b->frows.bytes;
        return ba;
    };
    auto em_proj_args = [&](BatchBucket* b) {
        ProjArgs p;
        p.img = b->xil.as<float>(); p.imgT = b->xT.as<float>(); p.out = b->erows.as<float>();
        p.y = b->y.as<float>(); p.wa = b->ewa.as<float>(); p.kl_part = nullptr; p.ex_part = nullptr;
        p.cosv = b->dc.as<double>(); p.sinv = b->ds.as<double>();
        p.g = b->rg; p.S = (int)b->scans.size(); p.tlo = b->ge.tlo; p.W = b->ge.width; p.lds_rc = b->prc;
        return p;
    };
    auto em_bp_args = [&](BatchBucket* b) {
        BpArgs bu;
        bu.q = b->erows.as<float>(); bu.out = b->xil.as<float>(); bu.x = b->xil.as<float>();
        bu.norm = b->ones.as<float>(); bu.xt = b->xT.as<float>();
        bu.cA = b->ecA.as<double>(); bu.cB = b->ecB.as<double>(); bu.xc = b->ge.xc; bu.c0 = b->ge.c0;
        bu.n = b->N; bu.A = b->A; bu.S = (int)b->scans.size(); bu.tlo = b->ge.tlo; bu.W = b->ge.width; bu.mode = 1;
        bu.q_bytes = b->erows.bytes;
        return bu;
    };
    // phase 1 (branch): gathers, FBP filter, eEM set-up
    auto enqueue_setup = [&](BatchBucket* b, hipStream_t s) {
        const int N = b->N, A = b->A, S = (int)b->scans.size();
        const size_t img = (size_t)N * N, sino = (size_t)A * N;
        if (do_fbp) {
            for (int k = 0; k < S; ++k)
                RAFT_HIP(hipMemcpyAsync(b->tomo.as<float>() + k * sino, d_tomo[b->scans[k]], sino * 4,
                                        hipMemcpyDeviceToDevice, s));
            FilterArgs fa;
            fa.tomo = b->tomo.as<float>();
            fa.q = b->frows.p;
            fa.H = b->F->H.as<float2>();
            fa.tw = b->F->tw.as<float2>();
            fa.g = b->F->g.as<double>();
            fa.wa = b->fwa.as<float>();
            fa.N = N; fa.A = A; fa.S = S; fa.npad = b->cf.npad; fa.L = b->cf.L; fa.tlo = b->gf.tlo; fa.W = b->gf.width;
            fa.none = (b->cf.filter == 0);
            launch_fbp_filter(fa, b->F->plan, store, s);
        }
        if (do_em) {
            const BpGeom& g = b->ge;
            for (int k = 0; k < S; ++k) {
                RAFT_HIP(hipMemcpyAsync(b->y.as<float>() + k * sino, d_y[b->scans[k]], sino * 4,
                                        hipMemcpyDeviceToDevice, s));
                RAFT_HIP(hipMemcpyAsync(b->x.as<float>() + k * img, d_em[b->scans[k]], img * 4,
                                        hipMemcpyDeviceToDevice, s));
            }
            float* ones = b->ones.as<float>();
            float* onerows = ones + img;
            launch_fill_rows(onerows, b->ewa.as<float>(), A, g.width, g.tlo, N, s);
            BpArgs b1;
            b1.q = onerows; b1.out = ones; b1.x = nullptr; b1.norm = nullptr; b1.xt = nullptr;
            b1.cA = b->ecA.as<double>(); b1.cB = b->ecB.as<double>(); b1.xc = g.xc; b1.c0 = g.c0;
            b1.n = N; b1.A = A; b1.S = 1; b1.tlo = g.tlo; b1.W = g.width; b1.mode = 0;
            b1.q_bytes = (size_t)A * g.width * 16;
            launch_backproject(b1, RAFT_STORE_F32, s);
            const int G = (S + 3) / 4;
            launch_fill(b->erows.as<float>(), (long long)G * A * g.width * 4, 0.0f, s);
            launch_to_interleaved(b->x.as<float>(), b->xil.as<float>(), S, N, N, false, s);
            launch_to_interleaved(b->x.as<float>(), b->xT.as<float>(), S, N, N, true, s);
        }
    };
    auto enqueue_fbp_scatter = [&](BatchBucket* b, hipStream_t s) {
        const size_t img = (size_t)b->N * b->N;
        for (size_t k = 0; k < b->scans.size(); ++k)
            RAFT_HIP(hipMemcpyAsync(d_fbp[b->scans[k]], b->fout.as<float>() + k * img, img * 4,
                                    hipMemcpyDeviceToDevice, s));
    };
    auto enqueue_em_finish = [&](BatchBucket* b, hipStream_t s) {
        const int N = b->N, S = (int)b->scans.size();
        const size_t img = (size_t)N * N;
        launch_from_interleaved(b->xil.as<float>(), b->x.as<float>(), S, N, N, s);
        for (int k = 0; k < S; ++k)
            RAFT_HIP(hipMemcpyAsync(d_em[b->scans[k]], b->x.as<float>() + k * img, img * 4,
                                    hipMemcpyDeviceToDevice, s));
    };
    // the whole bucket in its branch (RAFT_BATCH_QUEUE=0)
    auto enqueue = [&](BatchBucket* b) {
        hipStream_t s = b->s;
        enqueue_setup(b, s);
        if (do_fbp) {
            launch_backproject(fbp_bp_args(b), store, s);
            enqueue_fbp_scatter(b, s);
        }
        if (do_em) {
            for (int it = 0; it < K; ++it) {
                launch_project(em_proj_args(b), 1, s);
                launch_backproject(em_bp_args(b), RAFT_STORE_F32, s);
            }
            enqueue_em_finish(b, s);
        }
    };

    // queue descriptors: segments in decreasing cost per item (the longest
    // items are taken first), at most QUEUE_MAXSEG per launch
    struct QLaunch {
        std::vector<BatchBucket*> seg;
        size_t desc_off = 0;
        int counter = 0; // index of its counter pair
    };
    std::vector<QLaunch> q_fbp, q_proj, q_embp;
    std::vector<BatchBucket*> em_direct; // buckets whose projection takes the direct form
    std::vector<char> qdesc;
    int ncounters = 0;
    auto plan_queue = [&](std::vector<BatchBucket*> v, auto cost, std::vector<QLaunch>& out, size_t dsize,
                          auto fill) {
        std::stable_sort(v.begin(), v.end(), [&](BatchBucket* x, BatchBucket* y) { return cost(x) > cost(y); });
        for (size_t i = 0; i < v.size(); i += QUEUE_MAXSEG) {
            QLaunch L;
            L.seg.assign(v.begin() + i, v.begin() + std::min(v.size(), i + QUEUE_MAXSEG));
            L.desc_off = (qdesc.size() + 255) & ~(size_t)255;
            qdesc.resize(L.desc_off + dsize * L.seg.size());
            for (size_t k = 0; k < L.seg.size(); ++k) fill(L.seg[k], qdesc.data() + L.desc_off + k * dsize);
            L.counter = ncounters++;
            out.push_back(std::move(L));
        }
    };
    DevBuf qbuf, qctr;
    if (use_queue) {
        std::vector<BatchBucket*> all(buckets.begin(), buckets.end()), staged;
        for (BatchBucket* b : buckets) (b->prc > 0 ? staged : em_direct).push_back(b);
        if (do_fbp)
            plan_queue(all, [](BatchBucket* b) { return (double)b->A; }, q_fbp, bp_desc_size(),
                       [&](BatchBucket* b, void* d) { bp_desc_fill(fbp_bp_args(b), store, d); });
        if (do_em) {
            plan_queue(staged,
                       [](BatchBucket* b) {
                           const int nrb = (b->N + 31) / 32;
                           return (double)b->N * (double)((nrb + b->prc - 1) / b->prc);
                       },
                       q_proj, proj_desc_size(), [&](BatchBucket* b, void* d) { proj_desc_fill(em_proj_args(b), d); });
            plan_queue(all, [](BatchBucket* b) { return (double)b->A; }, q_embp, bp_desc_size(),
                       [&](BatchBucket* b, void* d) { bp_desc_fill(em_bp_args(b), RAFT_STORE_F32, d); });
        }
        if (!qdesc.empty()) {
            qbuf.alloc(qdesc.size());
            RAFT_HIP(hipMemcpyAsync(qbuf.p, qdesc.data(), qdesc.size(), hipMemcpyHostToDevice, st));
        }
        if (ncounters) {
            qctr.alloc((size_t)ncounters * 2 * sizeof(int));
            RAFT_HIP(hipMemsetAsync(qctr.p, 0, (size_t)ncounters * 2 * sizeof(int), st));
        }
        RAFT_HIP(hipStreamSynchronize(st));
    }
    auto bp_queue_of = [&](const QLaunch& L, int per_group_slices) {
        BpQueue q;
        q.counter = qctr.as<int>() + 2 * L.counter;
        q.nseg = (int)L.seg.size();
        q.first[0] = 0;
        for (int k = 0; k < q.nseg; ++k) {
            BatchBucket* b = L.seg[k];
            const int G = ((int)b->scans.size() + per_group_slices - 1) / per_group_slices;
            q.tiles[k] = bp_tiles(b->N);
            q.first[k + 1] = q.first[k] + q.tiles[k] * G;
        }
        return q;
    };
    auto proj_queue_of = [&](const QLaunch& L) {
        ProjQueue q;
        q.counter = qctr.as<int>() + 2 * L.counter;
        q.nseg = (int)L.seg.size();
        q.first[0] = 0;
        for (int k = 0; k < q.nseg; ++k) {
            BatchBucket* b = L.seg[k];
            const int G = ((int)b->scans.size() + 3) / 4;
            q.nab[k] = proj_angle_chunks(b->A);
            q.first[k + 1] = q.first[k] + q.nab[k] * b->prc * G;
        }
        return q;
    };
    const int sb_fbp = (store == RAFT_STORE_F16) ? BP_SLICES_F16 : BP_SLICES_F32;
    const int grid_fbp = use_queue && do_fbp ? bp_queue_grid(store, 0, 32) : 0;
    // the eEM update takes the pair form (two slice groups per item: one FP64
    // bin index per 8 slices; a bucket's scans fill many groups)
    const int embp_variant = BP_PAIR | 32;
    const int grid_embp = use_queue && do_em ? bp_queue_grid(RAFT_STORE_F32, 1, embp_variant) : 0;
    const int grid_proj = use_queue && do_em ? proj_queue_grid(1) : 0;
    // the stages after the per-bucket set-up, on the capture stream
    auto enqueue_queued = [&](hipStream_t s, hipEvent_t evf) {
        if (do_fbp) {
            for (const QLaunch& L : q_fbp)
                launch_backproject_queue(qbuf.as<char>() + L.desc_off, bp_queue_of(L, sb_fbp), store, 0, 32, grid_fbp,
                                         s);
            for (BatchBucket* b : buckets) enqueue_fbp_scatter(b, s);
        }
        if (do_em) {
            for (int it = 0; it < K; ++it) {
                // direct-form buckets (sparse angle sets) on their branches, beside the queue
                if (!em_direct.empty()) {
                    RAFT_HIP(hipEventRecord(evf, s));
                    for (BatchBucket* b : em_direct) {
                        RAFT_HIP(hipStreamWaitEvent(b->s, evf, 0));
                        launch_project(em_proj_args(b), 1, b->s);
                        RAFT_HIP(hipEventRecord(b->done, b->s));
                    }
                }
                for (const QLaunch& L : q_proj)
                    launch_project_queue(qbuf.as<char>() + L.desc_off, proj_queue_of(L), 1, grid_proj, s);
                for (BatchBucket* b : em_direct) RAFT_HIP(hipStreamWaitEvent(s, b->done, 0));
                for (const QLaunch& L : q_embp)
                    launch_backproject_queue(qbuf.as<char>() + L.desc_off,
                                             bp_queue_of(L, bp_slices_per_item(RAFT_STORE_F32, 1, embp_variant)),
                                             RAFT_STORE_F32, 1, embp_variant, grid_embp, s);
            }
            for (BatchBucket* b : buckets) enqueue_em_finish(b, s);
        }
    };

    // capture every bucket as a branch of one graph
    hipStream_t cs = nullptr;
    hipEvent_t fork = nullptr, t0 = nullptr, t1 = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    auto cleanup = [&] {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (fork) (void)hipEventDestroy(fork);
        if (t0) (void)hipEventDestroy(t0);
        if (t1) (void)hipEventDestroy(t1);
        for (BatchBucket* b : buckets) {
            if (b->done) (void)hipEventDestroy(b->done);
            if (b->s) (void)hipStreamDestroy(b->s);
        }
        if (cs) (void)hipStreamDestroy(cs);
    };
    try {
        // streams and events only once every bucket is validated and set up,
        // and inside the block whose cleanup destroys them
        for (BatchBucket* b : buckets) {
            RAFT_HIP(hipStreamCreateWithFlags(&b->s, hipStreamNonBlocking));
            RAFT_HIP(hipEventCreateWithFlags(&b->done, hipEventDisableTiming));
        }
        RAFT_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
        RAFT_HIP(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
        RAFT_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
        try {
            RAFT_HIP(hipEventRecord(fork, cs));
            for (BatchBucket* b : buckets) {
                RAFT_HIP(hipStreamWaitEvent(b->s, fork, 0));
                if (use_queue) enqueue_setup(b, b->s);
                else enqueue(b);
                RAFT_HIP(hipEventRecord(b->done, b->s));
                RAFT_HIP(hipStreamWaitEvent(cs, b->done, 0));
            }
            if (use_queue) enqueue_queued(cs, fork);
        } catch (...) {
            (void)hipStreamEndCapture(cs, &graph);
            throw;
        }
        RAFT_HIP(hipStreamEndCapture(cs, &graph));
        RAFT_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        if (graph_ms) {
            RAFT_HIP(hipEventCreate(&t0));
            RAFT_HIP(hipEventCreate(&t1));
            RAFT_HIP(hipEventRecord(t0, st));
        }
        RAFT_HIP(hipGraphLaunch(exec, st));
        if (graph_ms) RAFT_HIP(hipEventRecord(t1, st));
        RAFT_HIP(hipStreamSynchronize(st));
        if (graph_ms) RAFT_HIP(hipEventElapsedTime(graph_ms, t0, t1));
    } catch (...) {
        cleanup();
        throw;
    }
    cleanup();
}

// ---------------------------------------------------------------------------
// host-pointer wrappers: slab split over gpus[], one host thread per device
// ---------------------------------------------------------------------------
static void check_gpus(const int* gpus, int ngpus)
{
    if (!gpus || ngpus < 1) fail(RAFT_EDEVICE, "empty GPU list");
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count <= 0) {
        (void)hipGetLastError();
        fail(RAFT_EDEVICE, "no HIP device available");
    }
    for (int i = 0; i < ngpus; ++i)
        if (gpus[i] < 0 || gpus[i] >= count) fail(RAFT_EDEVICE, "invalid device number " + std::to_string(gpus[i]));
}

// run fn(device, first_slice, count) on contiguous slabs, one thread per device
template <typename F> static void run_slabs(const int* gpus, int ngpus, int nslices, F fn)
{
    const int per = (nslices + ngpus - 1) / ngpus;
    if (ngpus == 1) {
        fn(gpus[0], 0, nslices);
        return;
    }
    std::vector<std::thread> th;
    std::vector<Error> errs(ngpus, Error{RAFT_OK, ""});
    int ptr = 0;
    for (int i = 0; i < ngpus; ++i) {
        const int cnt = std::max(0, std::min(nslices - ptr, per));
        const int first = ptr;
        ptr += cnt;
        th.emplace_back([&, i, first, cnt]() {
            try {
                if (cnt > 0) fn(gpus[i], first, cnt);
            } catch (const Error& e) {
                errs[i] = e;
            } catch (const std::exception& e) {
                errs[i] = Error{RAFT_EHIP, e.what()};
            }
        });
    }
    for (auto& t : th) t.join();
    for (auto& e : errs)
        if (e.code != RAFT_OK) throw e;
}

struct Stream {
    hipStream_t s = nullptr;
    Stream() { RAFT_HIP(hipStreamCreate(&s)); }
    ~Stream() { if (s) (void)hipStreamDestroy(s); }
};

// Host memcpy split over a small pool of threads (pageable caller arrays to
// and from page-locked staging: one thread moves ~10 GB/s, the pool several
// times that).  copy() blocks until every part is done.
class CopyPool {
  public:
    explicit CopyPool(int nthreads) : n_(std::max(1, nthreads))
    {
        const char* e = std::getenv("RAFT_COPY_NT");
        nt_ = !(e && e[0] == '0');
        for (int i = 1; i < n_; ++i) th_.emplace_back([this, i] { worker(i); });
    }
    ~CopyPool()
    {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        go_.notify_all();
        for (auto& t : th_) t.join();
    }
    void copy(void* dst, const void* src, size_t bytes)
    {
        {
            std::lock_guard<std::mutex> lk(mu_);
            dst_ = static_cast<char*>(dst);
            src_ = static_cast<const char*>(src);
            bytes_ = bytes;
            left_ = n_ - 1;
            ++gen_;
        }
        go_.notify_all();
        part(0);
        std::unique_lock<std::mutex> lk(mu_);
        done_.wait(lk, [&] { return left_ == 0; });
    }

  private:
    // Copy with non-temporal 16-byte stores: the destination lines are not
    // read first (a plain copy reads each one for ownership), so a copy moves
    // two bytes of host memory per byte instead of three; host memory
    // bandwidth bounds these entry points.  The store fence makes them
    // visible before the part is reported done.
    static void copy_nt(char* dst, const char* src, size_t n)
    {
        typedef long long v2i __attribute__((vector_size(16)));
        size_t head = (16 - ((uintptr_t)dst & 15)) & 15;
        if (head > n) head = n;
        std::memcpy(dst, src, head);
        dst += head;
        src += head;
        n -= head;
        const size_t body = n & ~(size_t)63;
        for (size_t i = 0; i < body; i += 64) {
            v2i a, b, c, d;
            std::memcpy(&a, src + i, 16);
            std::memcpy(&b, src + i + 16, 16);
            std::memcpy(&c, src + i + 32, 16);
            std::memcpy(&d, src + i + 48, 16);
            __builtin_nontemporal_store(a, reinterpret_cast<v2i*>(dst + i));
            __builtin_nontemporal_store(b, reinterpret_cast<v2i*>(dst + i + 16));
            __builtin_nontemporal_store(c, reinterpret_cast<v2i*>(dst + i + 32));
            __builtin_nontemporal_store(d, reinterpret_cast<v2i*>(dst + i + 48));
        }
        std::memcpy(dst + body, src + body, n - body);
        __builtin_ia32_sfence(); // non-temporal stores are ordered only by a fence
    }
    void part(int i)
    {
        // 4 KB-aligned split points
        const size_t lo = ((bytes_ * (size_t)i / n_) >> 12) << 12;
        const size_t hi = (i + 1 == n_) ? bytes_ : ((bytes_ * (size_t)(i + 1) / n_) >> 12) << 12;
        if (hi > lo) {
            if (nt_) copy_nt(dst_ + lo, src_ + lo, hi - lo);
            else std::memcpy(dst_ + lo, src_ + lo, hi - lo);
        }
    }
    void worker(int i)
    {
        long seen = 0;
        for (;;) {
            std::unique_lock<std::mutex> lk(mu_);
            go_.wait(lk, [&] { return stop_ || gen_ != seen; });
            if (stop_) return;
            seen = gen_;
            lk.unlock();
            part(i);
            lk.lock();
            if (--left_ == 0) done_.notify_one();
        }
    }
    int n_;
    bool nt_ = true;
    std::vector<std::thread> th_;
    std::mutex mu_;
    std::condition_variable go_, done_;
    char* dst_ = nullptr;
    const char* src_ = nullptr;
    size_t bytes_ = 0;
    long gen_ = 0;
    int left_ = 0;
    bool stop_ = false;
};

// The caller's output array is usually freshly allocated, so the copies out
// fault its pages in; 2 MB pages (where the system grants them on request)
// take 512 times fewer faults.  Advisory only: the memory's contents and
// mapping are unchanged, and a refusal is ignored.
static void advise_huge(void* p, size_t bytes)
{
    const uintptr_t a = ((uintptr_t)p + ((1u << 21) - 1)) & ~(uintptr_t)((1u << 21) - 1);
    const uintptr_t b = ((uintptr_t)p + bytes) & ~(uintptr_t)((1u << 21) - 1);
    if (b > a) (void)madvise(reinterpret_cast<void*>(a), b - a, MADV_HUGEPAGE);
}

static int copy_threads()
{
    const char* e = std::getenv("RAFT_COPY_THREADS");
    if (e && std::atoi(e) > 0) return std::min(64, std::atoi(e));
    return (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
}

static void fbp_host(int* gpus, int ngpus, float* obj, float* tomo, float* angles, float* paramf, int* parami)
{
    const FbpCfg c = decode_fbp(paramf, parami);
    if (!obj || !tomo || !angles) fail(RAFT_EINVAL, "null data pointer");
    check_gpus(gpus, ngpus);
    run_slabs(gpus, ngpus, c.S, [&](int dev, int first, int cnt) {
        if (cnt <= 0) return;
        bind_device(dev);
        // Pipeline over chunks of whole slice groups, NB in flight: the copy
        // pool moves chunk k from the caller's array into page-locked staging
        // while the GPU filters and backprojects the chunks before it and the
        // DMA engines move chunks in and out on their own streams.  The pool
        // copies chunk k - (NB - 1) out after issuing chunk k, so with NB = 3
        // it waits for a chunk whose copies and FBP had two chunk periods to
        // finish (RAFT_COPY_DEPTH=2: one period).
        const size_t tin = (size_t)c.A * c.N, oout = (size_t)c.N * c.N;
        const size_t target = (size_t)64 << 20; // ~64 MB of sinogram per chunk
        int C = (int)std::max<size_t>(1, target / (tin * 4));
        if (c.blocksize > 0) C = std::min(C, c.blocksize);
        C = std::max(BP_SLICES_F32, (C / BP_SLICES_F32) * BP_SLICES_F32);
        C = std::min(C, cnt);
        const int nch = (cnt + C - 1) / C;
        const char* de = std::getenv("RAFT_COPY_DEPTH");
        const int NB = (de && std::atoi(de) == 2) ? 2 : 3;
        Device& D = device_state(dev);
        std::lock_guard<std::mutex> lk(D.mu);
        void* pin_in[3] = {D.pin[0].ensure((size_t)C * tin * 4), D.pin[1].ensure((size_t)C * tin * 4),
                           NB > 2 ? D.pin[4].ensure((size_t)C * tin * 4) : nullptr};
        void* pin_out[3] = {D.pin[2].ensure((size_t)C * oout * 4), D.pin[3].ensure((size_t)C * oout * 4),
                            NB > 2 ? D.pin[5].ensure((size_t)C * oout * 4) : nullptr};
        float* d_in[3] = {static_cast<float*>(D.workspace<8>((size_t)C * tin * 4)),
                          static_cast<float*>(D.workspace<9>((size_t)C * tin * 4)),
                          NB > 2 ? static_cast<float*>(D.workspace<12>((size_t)C * tin * 4)) : nullptr};
        float* d_out[3] = {static_cast<float*>(D.workspace<10>((size_t)C * oout * 4)),
                           static_cast<float*>(D.workspace<11>((size_t)C * oout * 4)),
                           NB > 2 ? static_cast<float*>(D.workspace<13>((size_t)C * oout * 4)) : nullptr};
        Stream s_in, s_comp, s_out;
        std::vector<hipEvent_t> ev;
        auto mk = [&]() {
            hipEvent_t e = nullptr;
            RAFT_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            ev.push_back(e);
            return e;
        };
        struct Cleanup {
            Device& D;
            std::vector<hipEvent_t>& ev;
            hipStream_t s[3];
            ~Cleanup()
            {
                for (hipStream_t x : s) (void)hipStreamSynchronize(x);
                for (hipEvent_t e : ev) (void)hipEventDestroy(e);
                D.keep.clear();
            }
        } cleanup{D, ev, {s_in.s, s_comp.s, s_out.s}};
        hipEvent_t ev_h2d[3], ev_comp[3], ev_d2h[3];
        for (int b = 0; b < NB; ++b) {
            ev_h2d[b] = mk();
            ev_comp[b] = mk();
            ev_d2h[b] = mk();
        }
        CopyPool pool(copy_threads());
        advise_huge(obj + (size_t)first * oout, (size_t)cnt * oout * 4);
        auto rows_of = [&](int k) { return std::min(C, cnt - k * C); };
        auto host_out = [&](int k) {
            const int b = k % NB;
            RAFT_HIP(hipEventSynchronize(ev_d2h[b]));
            pool.copy(obj + (size_t)(first + k * C) * oout, pin_out[b], (size_t)rows_of(k) * oout * 4);
        };
        for (int k = 0; k < nch; ++k) {
            const int b = k % NB, nb = rows_of(k);
            if (k >= NB) RAFT_HIP(hipEventSynchronize(ev_h2d[b])); // pin_in[b] read by the copy of chunk k-NB
            pool.copy(pin_in[b], tomo + (size_t)(first + k * C) * tin, (size_t)nb * tin * 4);
            if (k >= NB) RAFT_HIP(hipStreamWaitEvent(s_in.s, ev_comp[b], 0)); // d_in[b] read by chunk k-NB's FBP
            RAFT_HIP(hipMemcpyAsync(d_in[b], pin_in[b], (size_t)nb * tin * 4, hipMemcpyHostToDevice, s_in.s));
            RAFT_HIP(hipEventRecord(ev_h2d[b], s_in.s));
            RAFT_HIP(hipStreamWaitEvent(s_comp.s, ev_h2d[b], 0));
            if (k >= NB) RAFT_HIP(hipStreamWaitEvent(s_comp.s, ev_d2h[b], 0)); // d_out[b] read by chunk k-NB's copy
            FbpCfg ck = c;
            ck.S = nb;
            fbp_run(D, d_out[b], d_in[b], angles, ck, RAFT_STORE_F32, s_comp.s, false);
            RAFT_HIP(hipEventRecord(ev_comp[b], s_comp.s));
            RAFT_HIP(hipStreamWaitEvent(s_out.s, ev_comp[b], 0));
            // pin_out[b] was copied out (host_out(k - NB)) before this iteration
            RAFT_HIP(hipMemcpyAsync(pin_out[b], d_out[b], (size_t)nb * oout * 4, hipMemcpyDeviceToHost, s_out.s));
            RAFT_HIP(hipEventRecord(ev_d2h[b], s_out.s));
            if (k >= NB - 1) host_out(k - (NB - 1)); // pin_out of chunk k-(NB-1) -> the caller's array
        }
        for (int k = std::max(0, nch - (NB - 1)); k < nch; ++k) host_out(k);
    });
}

static void radon_host(int* gpus, int ngpus, float* proj, float* obj, float* angles, int Nx, int Ny, int nrays,
                       int nangles, int nslices, float ax, float ay)
{
    if (!proj || !obj || !angles) fail(RAFT_EINVAL, "null data pointer");
    if (Nx < 2 || Ny < 2 || nrays < 1 || nangles < 1 || nslices < 0) fail(RAFT_EINVAL, "invalid sizes");
    check_gpus(gpus, ngpus);
    run_slabs(gpus, ngpus, nslices, [&](int dev, int first, int cnt) {
        bind_device(dev);
        Stream st;
        const size_t iin = (size_t)Nx * Ny, pout = (size_t)nangles * nrays;
        int B = cnt;
        const size_t budget = block_budget();
        while (B > 1 && (size_t)B * (2 * iin + pout) * 4 > budget) B = (B + 1) / 2;
        DevBuf di((size_t)B * iin * 4), dp((size_t)B * pout * 4);
        for (int s0 = 0; s0 < cnt; s0 += B) {
            const int nb = std::min(B, cnt - s0);
            RAFT_HIP(hipMemcpyAsync(di.p, obj + (size_t)(first + s0) * iin, (size_t)nb * iin * 4,
                                    hipMemcpyHostToDevice, st.s));
            radon_device(dev, dp.as<float>(), di.as<float>(), angles, Nx, Ny, nrays, nangles, nb, ax, ay, st.s);
            RAFT_HIP(hipMemcpyAsync(proj + (size_t)(first + s0) * pout, dp.p, (size_t)nb * pout * 4,
                                    hipMemcpyDeviceToHost, st.s));
        }
        RAFT_HIP(hipStreamSynchronize(st.s));
    });
}

static void eem_host(int* gpus, int ngpus, float* obj, float* tomo, float* angles, int* parami, double* kl,
                     long long* kl_ex)
{
    if (!parami) fail(RAFT_EINVAL, "null parameter array");
    if (!obj || !tomo || !angles) fail(RAFT_EINVAL, "null data pointer");
    const int R = parami[0], A = parami[1], S = parami[2], n = parami[3], K = parami[8];
    if (R < 2 || A < 2 || n < 2 || S < 0 || K < 0) fail(RAFT_EINVAL, "invalid eEM parameters");
    check_gpus(gpus, ngpus);
    run_slabs(gpus, ngpus, S, [&](int dev, int first, int cnt) {
        bind_device(dev);
        Stream st;
        const size_t tin = (size_t)A * R, img = (size_t)n * n;
        int B = parami[9] > 0 ? std::min(parami[9], cnt) : cnt;
        const size_t budget = block_budget();
        const size_t need = (tin + 2 * img) * 4 + (size_t)A * (size_t)(1.5 * R + 64) * 4;
        while (B > 1 && (size_t)B * need > budget) B = (B + 1) / 2;
        DevBuf dt((size_t)B * tin * 4), dx((size_t)B * img * 4);
        std::vector<int> pi(parami, parami + 10);
        std::vector<double> hk;
        std::vector<long long> he;
        for (int s0 = 0; s0 < cnt; s0 += B) {
            const int nb = std::min(B, cnt - s0);
            RAFT_HIP(hipMemcpyAsync(dt.p, tomo + (size_t)(first + s0) * tin, (size_t)nb * tin * 4,
                                    hipMemcpyHostToDevice, st.s));
            RAFT_HIP(hipMemcpyAsync(dx.p, obj + (size_t)(first + s0) * img, (size_t)nb * img * 4,
                                    hipMemcpyHostToDevice, st.s));
            pi[2] = nb;
            pi[9] = 0;
            double* klp = kl ? kl + (size_t)(first + s0) * K : nullptr;
            long long* exp_ = kl_ex ? kl_ex + (size_t)(first + s0) * K : nullptr;
            eem_device(dev, dx.as<float>(), dt.as<float>(), angles, pi.data(), klp, exp_, st.s);
            RAFT_HIP(hipMemcpyAsync(obj + (size_t)(first + s0) * img, dx.p, (size_t)nb * img * 4,
                                    hipMemcpyDeviceToHost, st.s));
        }
        RAFT_HIP(hipStreamSynchronize(st.s));
    });
}

} // namespace raft

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
using namespace raft;

extern "C" {

static void background_host(int* gpus, int ngpus, float* frames, float* flat, float* dark, int R, int A, int S,
                            int nflats, int is_log, int blocksize)
{
    if (R < 1 || A < 1 || S < 0 || nflats < 1) fail(RAFT_EINVAL, "invalid sizes");
    if (S == 0) return;
    if (!frames || !flat || !dark) fail(RAFT_EINVAL, "null data pointer");
    check_gpus(gpus, ngpus);
    run_slabs(gpus, ngpus, S, [&](int dev, int first, int cnt) {
        bind_device(dev);
        Stream st;
        const size_t per = ((size_t)A * R + (size_t)R * nflats + R) * 4;
        int B = blocksize > 0 ? std::min(blocksize, cnt) : cnt;
        const size_t budget = block_budget();
        while (B > 1 && (size_t)B * per > budget) B = (B + 1) / 2;
        DevBuf dd((size_t)B * A * R * 4), df((size_t)B * R * nflats * 4), dk((size_t)B * R * 4);
        for (int s0 = 0; s0 < cnt; s0 += B) {
            const int nb = std::min(B, cnt - s0);
            const size_t z = (size_t)(first + s0);
            float* h = frames + z * A * R;
            RAFT_HIP(hipMemcpyAsync(dd.p, h, (size_t)nb * A * R * 4, hipMemcpyHostToDevice, st.s));
            RAFT_HIP(hipMemcpyAsync(df.p, flat + z * R * nflats, (size_t)nb * R * nflats * 4, hipMemcpyHostToDevice,
                                    st.s));
            RAFT_HIP(hipMemcpyAsync(dk.p, dark + z * R, (size_t)nb * R * 4, hipMemcpyHostToDevice, st.s));
            launch_background(dd.as<float>(), df.as<float>(), dk.as<float>(), R, A, nb, nflats, is_log, st.s);
            RAFT_HIP(hipMemcpyAsync(h, dd.p, (size_t)nb * A * R * 4, hipMemcpyDeviceToHost, st.s));
        }
        RAFT_HIP(hipStreamSynchronize(st.s));
    });
}

void getBackgroundCorrectionMultiGPU(int* gpus, int ngpus, float* frames, float* flat, float* dark, int nrays,
                                     int nangles, int nslices, int numflats, int is_log, int blocksize)
{
    guarded([&] {
        background_host(gpus, ngpus, frames, flat, dark, nrays, nangles, nslices, numflats, is_log, blocksize);
    });
}

int raft_background_device(int device, float* d_frames, const float* d_flat, const float* d_dark, int nrays,
                           int nangles, int nslices, int numflats, int is_log, void* stream)
{
    return guarded([&] {
        if (nrays < 1 || nangles < 1 || nslices < 0 || numflats < 1) fail(RAFT_EINVAL, "invalid sizes");
        if (nslices == 0) return;
        if (!d_frames || !d_flat || !d_dark) fail(RAFT_EINVAL, "null data pointer");
        bind_device(device);
        launch_background(d_frames, d_flat, d_dark, nrays, nangles, nslices, numflats, is_log,
                          static_cast<hipStream_t>(stream));
    });
}

static void tem_host(int* gpus, int ngpus, float* obj, float* count, float* flat, float* angles, int* parami)
{
    if (!parami) fail(RAFT_EINVAL, "null parameter array");
    if (!obj || !count || !flat || !angles) fail(RAFT_EINVAL, "null data pointer");
    const int R = parami[0], A = parami[1], S = parami[2], n = parami[3], K = parami[8];
    if (R < 2 || A < 2 || n < 2 || S < 0 || K < 0) fail(RAFT_EINVAL, "invalid tEM parameters");
    check_gpus(gpus, ngpus);
    run_slabs(gpus, ngpus, S, [&](int dev, int first, int cnt) {
        bind_device(dev);
        Stream st;
        const size_t tin = (size_t)A * R, img = (size_t)n * n;
        int B = parami[9] > 0 ? std::min(parami[9], cnt) : cnt;
        const size_t budget = block_budget();
        const size_t need = (tin + R + 5 * img) * 4 + (size_t)A * (size_t)(1.5 * R + 64) * 4;
        while (B > 1 && (size_t)B * need > budget) B = (B + 1) / 2;
        DevBuf dt((size_t)B * tin * 4), df((size_t)B * R * 4), dx((size_t)B * img * 4);
        std::vector<int> pi(parami, parami + 10);
        for (int s0 = 0; s0 < cnt; s0 += B) {
            const int nb = std::min(B, cnt - s0);
            RAFT_HIP(hipMemcpyAsync(dt.p, count + (size_t)(first + s0) * tin, (size_t)nb * tin * 4,
                                    hipMemcpyHostToDevice, st.s));
            RAFT_HIP(hipMemcpyAsync(df.p, flat + (size_t)(first + s0) * R, (size_t)nb * R * 4, hipMemcpyHostToDevice,
                                    st.s));
            RAFT_HIP(hipMemcpyAsync(dx.p, obj + (size_t)(first + s0) * img, (size_t)nb * img * 4,
                                    hipMemcpyHostToDevice, st.s));
            pi[2] = nb;
            pi[9] = 0;
            tem_device(dev, dx.as<float>(), dt.as<float>(), df.as<float>(), angles, pi.data(), st.s);
            RAFT_HIP(hipMemcpyAsync(obj + (size_t)(first + s0) * img, dx.p, (size_t)nb * img * 4,
                                    hipMemcpyDeviceToHost, st.s));
        }
        RAFT_HIP(hipStreamSynchronize(st.s));
    });
}

void get_tEM_RT_MultiGPU(int* gpus, int ngpus, float* obj, float* count, float* flat, float* angles, float* paramf,
                         int* parami)
{
    (void)paramf;
    guarded([&] { tem_host(gpus, ngpus, obj, count, flat, angles, parami); });
}

int raft_tem_device(int device, float* d_obj, const float* d_count, const float* d_flat, const float* angles,
                    const int* parami, void* stream)
{
    return guarded(
        [&] { tem_device(device, d_obj, d_count, d_flat, angles, parami, static_cast<hipStream_t>(stream)); });
}

static void bst_host(int* gpus, int ngpus, float* obj, float* tomo, float* paramf, int* parami)
{
    const FbpCfg c = decode_fbp(paramf, parami);
    if (!obj || !tomo) fail(RAFT_EINVAL, "null data pointer");
    check_gpus(gpus, ngpus);
    run_slabs(gpus, ngpus, c.S, [&](int dev, int first, int cnt) {
        bind_device(dev);
        Stream st;
        const size_t tin = (size_t)c.A * c.N, oout = (size_t)c.N * c.N;
        const size_t n = (size_t)c.npad;
        int B = c.blocksize > 0 ? std::min(c.blocksize, cnt) : cnt;
        const size_t need = (tin + oout) * 4 + (size_t)c.A * n * 20 + n * n * 16;
        const size_t budget = block_budget();
        while (B > 1 && (size_t)B * need > budget) B = (B + 1) / 2;
        DevBuf dt((size_t)B * tin * 4), dob((size_t)B * oout * 4);
        std::vector<int> pi(parami, parami + 10);
        for (int s0 = 0; s0 < cnt; s0 += B) {
            const int nb = std::min(B, cnt - s0);
            RAFT_HIP(hipMemcpyAsync(dt.p, tomo + (size_t)(first + s0) * tin, (size_t)nb * tin * 4,
                                    hipMemcpyHostToDevice, st.s));
            pi[2] = nb;
            pi[9] = 0;
            bst_device(dev, dob.as<float>(), dt.as<float>(), paramf, pi.data(), st.s);
            RAFT_HIP(hipMemcpyAsync(obj + (size_t)(first + s0) * oout, dob.p, (size_t)nb * oout * 4,
                                    hipMemcpyDeviceToHost, st.s));
        }
        RAFT_HIP(hipStreamSynchronize(st.s));
    });
}

void getBSTMultiGPU(int* gpus, int ngpus, float* obj, float* tomogram, float* angles, float* paramf, int* parami,
                    int nstreams)
{
    (void)angles;
    (void)nstreams;
    guarded([&] { bst_host(gpus, ngpus, obj, tomogram, paramf, parami); });
}

int raft_bst_device(int device, float* d_obj, const float* d_tomo, const float* paramf, const int* parami,
                    void* stream)
{
    return guarded([&] { bst_device(device, d_obj, d_tomo, paramf, parami, static_cast<hipStream_t>(stream)); });
}

static void lowpass_host(int* gpus, int ngpus, float* tomo, float* paramf, int* parami)
{
    if (!paramf) fail(RAFT_EINVAL, "null parameter array");
    float pf[7];
    std::memcpy(pf, paramf, sizeof(pf));
    pf[6] = 0.0f;
    const FbpCfg c = decode_fbp(pf, parami);
    if (!tomo) fail(RAFT_EINVAL, "null data pointer");
    check_gpus(gpus, ngpus);
    if (c.filter == 0) return;
    run_slabs(gpus, ngpus, c.S, [&](int dev, int first, int cnt) {
        bind_device(dev);
        Stream st;
        const size_t tin = (size_t)c.A * c.N;
        int B = c.blocksize > 0 ? std::min(c.blocksize, cnt) : cnt;
        const size_t need = tin * 4 * 2;
        const size_t budget = block_budget();
        while (B > 1 && (size_t)B * need > budget) B = (B + 1) / 2;
        DevBuf dt((size_t)B * tin * 4);
        std::vector<int> pi(parami, parami + 10);
        for (int s0 = 0; s0 < cnt; s0 += B) {
            const int nb = std::min(B, cnt - s0);
            float* h = tomo + (size_t)(first + s0) * tin;
            RAFT_HIP(hipMemcpyAsync(dt.p, h, (size_t)nb * tin * 4, hipMemcpyHostToDevice, st.s));
            pi[2] = nb;
            pi[9] = 0;
            lowpass_device(dev, dt.as<float>(), pf, pi.data(), st.s);
            RAFT_HIP(hipMemcpyAsync(h, dt.p, (size_t)nb * tin * 4, hipMemcpyDeviceToHost, st.s));
        }
        RAFT_HIP(hipStreamSynchronize(st.s));
    });
}

void getFilterLowPassMultiGPU(int* gpus, int ngpus, float* tomogram, float* paramf, int* parami)
{
    guarded([&] { lowpass_host(gpus, ngpus, tomogram, paramf, parami); });
}

int raft_lowpass_device(int device, float* d_tomo, const float* paramf, const int* parami, void* stream)
{
    return guarded([&] { lowpass_device(device, d_tomo, paramf, parami, static_cast<hipStream_t>(stream)); });
}

void getFBPMultiGPU(int* gpus, int ngpus, float* obj, float* tomogram, float* angles, float* paramf, int* parami)
{
    guarded([&] { fbp_host(gpus, ngpus, obj, tomogram, angles, paramf, parami); });
}

void get_eEM_RT_MultiGPU(int* gpus, int ngpus, float* obj, float* tomogram, float* angles, float* paramf,
                         int* parami)
{
    (void)paramf;
    guarded([&] { eem_host(gpus, ngpus, obj, tomogram, angles, parami, nullptr, nullptr); });
}

void get_eEM_RT_KL_MultiGPU(int* gpus, int ngpus, float* obj, float* tomogram, float* angles, float* paramf,
                            int* parami, double* kl, long long* kl_excluded)
{
    (void)paramf;
    guarded([&] { eem_host(gpus, ngpus, obj, tomogram, angles, parami, kl, kl_excluded); });
}

void getRadonRTMultiGPU(int* gpus, int ngpus, float* projection, float* obj, float* angles, int sizeImagex,
                        int sizeImagey, int nrays, int nangles, int nslices, float ax, float ay)
{
    guarded([&] {
        radon_host(gpus, ngpus, projection, obj, angles, sizeImagex, sizeImagey, nrays, nangles, nslices, ax, ay);
    });
}

int raft_last_error(void) { return g_err; }
const char* raft_last_error_string(void) { return g_msg.c_str(); }
void raft_clear_error(void)
{
    g_err = RAFT_OK;
    g_msg.clear();
}

int raft_fbp_device(int device, float* d_obj, const float* d_tomo, const float* angles, const float* paramf,
                    const int* parami, int store, void* stream)
{
    return guarded([&] {
        fbp_device(device, d_obj, d_tomo, angles, paramf, parami, store, static_cast<hipStream_t>(stream));
    });
}

int raft_fbp_device_noalloc(int device, float* d_obj, const float* d_tomo, const float* angles, const float* paramf,
                            const int* parami, int store, void* stream)
{
    return guarded([&] {
        fbp_device_noalloc(device, d_obj, d_tomo, angles, paramf, parami, store, static_cast<hipStream_t>(stream));
    });
}

int raft_radon_device(int device, float* d_proj, const float* d_obj, const float* angles, int Nx, int Ny, int nrays,
                      int nangles, int nslices, float ax, float ay, void* stream)
{
    return guarded([&] {
        radon_device(device, d_proj, d_obj, angles, Nx, Ny, nrays, nangles, nslices, ax, ay,
                     static_cast<hipStream_t>(stream));
    });
}

int raft_eem_device(int device, float* d_obj, const float* d_tomo, const float* angles, const int* parami,
                    double* kl, long long* kl_excluded, void* stream)
{
    return guarded([&] {
        eem_device(device, d_obj, d_tomo, angles, parami, kl, kl_excluded, static_cast<hipStream_t>(stream));
    });
}

int raft_batch_device(int device, int nscans, const int* nrays, const int* nangles, const float* const* angles,
                      const float* const* d_tomo, const float* const* d_y, float* const* d_fbp, float* const* d_em,
                      const float* paramf, int pad, int filter, int store, int iterations, void* stream,
                      float* graph_ms)
{
    return guarded([&] {
        batch_device(device, nscans, nrays, nangles, angles, d_tomo, d_y, d_fbp, d_em, paramf, pad, filter, store,
                     iterations, static_cast<hipStream_t>(stream), graph_ms);
    });
}

int raft_eem_device_kl(int device, float* d_obj, const float* d_tomo, const float* angles, const int* parami,
                       int kl_iterations, double* kl, long long* kl_excluded, void* stream)
{
    return guarded([&] {
        if (kl_iterations < 0) fail(RAFT_EINVAL, "kl_iterations must be >= 0");
        eem_device(device, d_obj, d_tomo, angles, parami, kl_iterations > 0 ? kl : nullptr,
                   kl_iterations > 0 ? kl_excluded : nullptr, static_cast<hipStream_t>(stream), false, kl_iterations);
    });
}

int raft_eem_device_sched(int device, float* d_obj, const float* d_tomo, const float* angles, const int* parami,
                          const int* iterations, void* stream)
{
    return guarded([&] {
        if (!iterations) fail(RAFT_EINVAL, "null iteration counts");
        eem_device(device, d_obj, d_tomo, angles, parami, nullptr, nullptr, static_cast<hipStream_t>(stream), false,
                   -1, iterations);
    });
}

int raft_eem_kl_device(int device, const float* d_obj, const float* d_tomo, const float* angles, const int* parami,
                       double* kl, long long* kl_excluded, void* stream)
{
    return guarded([&] {
        eem_device(device, const_cast<float*>(d_obj), d_tomo, angles, parami, kl, kl_excluded,
                   static_cast<hipStream_t>(stream), true);
    });
}

int raft_fbp_plan_info(const float* angles, const float* paramf, const int* parami, int* tlo, int* width,
                       int* fft_path)
{
    return guarded([&] {
        const FbpCfg c = decode_fbp(paramf, parami);
        if (!angles) fail(RAFT_EINVAL, "null angles");
        BpGeom g;
        std::vector<float> wa;
        fbp_geometry(c, angles, g, wa);
        if (tlo) *tlo = g.tlo;
        if (width) *width = g.width;
        if (fft_path) *fft_path = fft_plan(c.npad).ok ? 1 : 0;
    });
}

int raft_proj_plan_info(const float* angles, int nangles, int Nx, int Ny, int nrays, int nslices, float ax, float ay,
                        int em, int* ray_chunks, int* tile_consts)
{
    return guarded([&] {
        if (!angles) fail(RAFT_EINVAL, "null angles");
        if (nangles < 1 || nrays < 2 || nslices < 1) fail(RAFT_EINVAL, "invalid projection sizes");
        if (Nx < 2 || Ny < 2) fail(RAFT_EINVAL, "image sizes must be >= 2");
        if (em && Nx != Ny) fail(RAFT_EINVAL, "the eEM image is square");
        if (!em && (!(ax != 0.0f) || !(ay != 0.0f))) fail(RAFT_EINVAL, "ax, ay must be non-zero");
        std::vector<double> cv(nangles), sv(nangles);
        for (int a = 0; a < nangles; ++a) {
            cv[a] = std::cos((double)angles[a]);
            sv[a] = std::sin((double)angles[a]);
        }
        const RayGeom g = em ? ray_geom_em(nrays, nangles, Nx)
                             : ray_geom_radon(Nx, Ny, nrays, nangles, (double)ax, (double)ay);
        if (ray_chunks) *ray_chunks = proj_plan(g, cv.data(), sv.data(), nslices, true);
        if (tile_consts) proj_tile_consts(tile_consts);
    });
}

int raft_lds_read_probe(int device, double* bytes_per_s, int* cus, int* clock_khz)
{
    return guarded([&] {
        bind_device(device);
        const double r = lds_read_probe(cus, clock_khz);
        if (bytes_per_s) *bytes_per_s = r;
    });
}

void raft_set_profiling(int on)
{
    prof_clear();
    g_prof = (on != 0);
}

double raft_last_kernel_ms(int which)
{
    if (which < 0 || which >= KF_COUNT) return -1.0;
    auto& v = g_prof_rec.ev[which];
    if (v.empty()) return -1.0;
    double tot = 0.0;
    for (auto& pr : v) {
        if (hipEventSynchronize(pr.second) != hipSuccess) {
            (void)hipGetLastError();
            return -1.0;
        }
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, pr.first, pr.second) != hipSuccess) {
            (void)hipGetLastError();
            return -1.0;
        }
        tot += ms;
    }
    return tot / (double)v.size();
}

} // extern "C"
