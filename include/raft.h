/*
 * raft.h -- C ABI of libraft.so, the MI355X (gfx950) implementation of
 * ssc-raft's parallel-beam reconstruction hot path: FBP (ramp-type filter +
 * backprojection), the Radon ray-tracing projector and the emission EM
 * (eEM) iteration.
 *
 * The first three functions keep the exact signatures that sscRaft's ctypes
 * layer binds, so the reference's Python API and array layouts stay a
 * drop-in.  All of their arguments are HOST pointers to C-contiguous arrays
 * (float = float32, int = int32); the calls are synchronous.  Unlike the
 * reference (which prints and exit()s on any error), failures are reported
 * through raft_last_error()/raft_last_error_string() and the call returns.
 *
 * The raft_*_device entry points take DEVICE pointers (data already resident
 * in HBM) and a hipStream_t (passed as void*, NULL = the default stream);
 * they are asynchronous with respect to the host except where noted.  They
 * have no counterpart in the reference: they exist for callers that keep data
 * on the GPU (the benchmark, pipelines, the EM warm-start work).
 *
 * Layouts: tomogram/sinogram [slices][angles][rays]; object [slices][y][x];
 * angles [angles] in radians.
 */
#ifndef RAFT_H
#define RAFT_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- reference entry points (host pointers) ------------------------------ */

/* Filtered backprojection, RT method.
 * Replaces getFBPMultiGPU, cuda/src/geometries/parallel/reconstruction_methods/fbp.cu:238-302
 * (declared in cuda/inc/geometries/parallel/fbp.hpp; bound at sscRaft/rafttypes.py:177-183,
 * called from sscRaft/geometries/parallel/fbp_methods.py:99-101).
 * parami = [nrays, nangles, nslices, objsize(=nrays), pad, pady(ignored), padz(ignored),
 *           filter_type (0 none,1 gaussian,2 lorentz,3 cosine,4 rectangle,5 hann,6 hamming,7 ramp),
 *           0, blocksize (0 = automatic)]
 * paramf = [beta/delta inverse or 0, regularization, energy[eV], z2[m], pixel_x, pixel_y, axis_offset]
 * obj: output [nslices][objsize][objsize]. */
void getFBPMultiGPU(int* gpus, int ngpus, float* obj, float* tomogram, float* angles, float* paramf, int* parami);

/* Emission EM with ray-tracing operators.
 * Replaces get_eEM_RT_MultiGPU, cuda/src/geometries/parallel/reconstruction_methods/em_rt.cu:201-262
 * (declared in cuda/inc/geometries/parallel/em.hpp:17-19; bound at sscRaft/rafttypes.py:139-144,
 * called from sscRaft/geometries/parallel/em_methods.py:66-68).
 * parami = [nrays, nangles, nslices, objsize, 0, 0, 0, nflats, iterations, blocksize]
 * paramf = unused.  obj: in/out [nslices][objsize][objsize], the initial guess on entry. */
void get_eEM_RT_MultiGPU(int* gpus, int ngpus, float* obj, float* tomogram, float* angles, float* paramf,
                         int* parami);

/* Radon transform, ray tracing.
 * Replaces getRadonRTMultiGPU, cuda/src/geometries/parallel/forward_methods/radon_rt.cu:252-295
 * (bound at sscRaft/rafttypes.py:312-320, called from sscRaft/geometries/parallel/radon_methods.py:48-52).
 * obj [nslices][sizeImagey][sizeImagex] -> projection [nslices][nangles][nrays]; ax, ay = half
 * extents of the image square (the Python wrapper passes 1.0).  The wrapper's nrays*pixel/2
 * rescale (radon_methods.py:58) stays on the Python side, as in the reference. */
void getRadonRTMultiGPU(int* gpus, int ngpus, float* projection, float* obj, float* angles, int sizeImagex,
                        int sizeImagey, int nrays, int nangles, int nslices, float ax, float ay);

/* eEM as get_eEM_RT_MultiGPU, plus the FP64 Kullback-Leibler objective
 * KL(y || A x_k) of every slice at every iteration k (before its update):
 * kl[nslices][iterations], kl_excluded[nslices][iterations] = bins left out
 * (A x <= 0 or y < 0).  Either output may be NULL.  No counterpart in the
 * reference (which computes no objective). */
void get_eEM_RT_KL_MultiGPU(int* gpus, int ngpus, float* obj, float* tomogram, float* angles, float* paramf,
                            int* parami, double* kl, long long* kl_excluded);

/* ---- error state (thread-local) -------------------------------------------- */

/* 0 after a successful call; otherwise a RAFT_E* code. */
int raft_last_error(void);
/* Human-readable text of the last error of this thread ("" if none). */
const char* raft_last_error_string(void);
/* Clears this thread's error state. */
void raft_clear_error(void);

#define RAFT_OK 0
#define RAFT_EINVAL 1    /* invalid parameters */
#define RAFT_EDEVICE 2   /* invalid device id / no device */
#define RAFT_EHIP 3      /* HIP runtime error (allocation, launch, copy) */
#define RAFT_EUNSUPPORTED 4

/* ---- device-resident entry points (no reference counterpart) ---------------- */

/* storage precision of the filtered rows / ratio rows staged for backprojection */
#define RAFT_STORE_F32 0
#define RAFT_STORE_F16 1

/* FBP on one device: d_tomo [nslices][nangles][nrays] -> d_obj [nslices][nrays][nrays],
 * both device pointers; angles is a HOST array; parameters as getFBPMultiGPU.
 * store = RAFT_STORE_F32 or RAFT_STORE_F16.  Returns 0 or an error code. */
int raft_fbp_device(int device, float* d_obj, const float* d_tomo, const float* angles, const float* paramf,
                    const int* parami, int store, void* stream);

/* raft_fbp_device that allocates no device memory.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, same kernels and the same output bit for bit; the filter and
 * angle tables and the filtered rows live in one fixed device array of
 * RAFT_NOALLOC_SCRATCH_BYTES that each GPU the library runs on holds until the
 * process ends, and the slices run in blocks of as many slice groups (4 slices
 * for F32 rows, 8 for F16) as fit there.  A geometry whose tables plus one
 * slice group of rows (nangles x staged width x 16 bytes) exceed the array is
 * refused with RAFT_EUNSUPPORTED (2048 angles x 2048 rays at pad 2 needs
 * 90.75 MiB per group).  Synchronous: it waits for its stream before returning.
 * Calls on one device share its array and are serialised by the library. */
#define RAFT_NOALLOC_SCRATCH_BYTES (256ull << 20)
int raft_fbp_device_noalloc(int device, float* d_obj, const float* d_tomo, const float* angles,
                            const float* paramf, const int* parami, int store, void* stream);

/* Radon on one device: d_obj [nslices][Ny][Nx] -> d_proj [nslices][nangles][nrays]. */
int raft_radon_device(int device, float* d_proj, const float* d_obj, const float* angles, int Nx, int Ny,
                      int nrays, int nangles, int nslices, float ax, float ay, void* stream);

/* eEM on one device: d_obj in/out [nslices][objsize][objsize], d_tomo [nslices][nangles][nrays];
 * parami as get_eEM_RT_MultiGPU.  kl/kl_excluded: HOST arrays [nslices][iterations] or NULL.
 * Synchronises the stream before returning when kl or kl_excluded is given. */
int raft_eem_device(int device, float* d_obj, const float* d_tomo, const float* angles, const int* parami,
                    double* kl, long long* kl_excluded, void* stream);

/* raft_eem_device recording the KL objective of the first kl_iterations
 * iterations only: kl/kl_excluded are HOST arrays [nslices][min(kl_iterations,
 * iterations)]; the other iterations run without the FP64 KL terms.  No
 * counterpart in the reference; for callers that test a starting image by its
 * KL (the learned warm start), which the first iteration computes anyway. */
int raft_eem_device_kl(int device, float* d_obj, const float* d_tomo, const float* angles, const int* parami,
                       int kl_iterations, double* kl, long long* kl_excluded, void* stream);

/* raft_eem_device with an iteration count per slice: iterations[s] (HOST
 * array [nslices], non-increasing -- sort the slices by it) replaces
 * parami[8]; iteration t updates only the slices with iterations[s] > t, and
 * its launches cover only their slice groups, so slices that are done cost
 * nothing.  No KL is recorded.  No counterpart in the reference (which runs
 * every slice for the same count); for batches whose slices need different
 * counts (warm starts, early stopping).  Synchronises before returning. */
int raft_eem_device_sched(int device, float* d_obj, const float* d_tomo, const float* angles, const int* parami,
                          const int* iterations, void* stream);

/* FBP by the Backprojection Slice Theorem (replaces getBSTMultiGPU,
 * cuda/src/geometries/parallel/reconstruction_methods/bst.cu:674-736, bound in
 * rafttypes.py, called from sscRaft/geometries/parallel/fbp_methods.py:200):
 * obj [nslices][nrays][nrays] out, tomogram [nslices][nangles][nrays]; paramf /
 * parami as getFBPMultiGPU (the axis offset is used as an integer shift, as the
 * reference's Python front end passes int(offset)).  Angles are assumed uniform
 * on [0, pi) (the array is not read, as in the reference); nstreams is ignored.
 * Padded size n = nrays (1 + pad) must factor into 2, 3, 5 and be <= 16384. */
void getBSTMultiGPU(int* gpus, int ngpus, float* obj, float* tomogram, float* angles, float* paramf, int* parami,
                    int nstreams);

/* Same on one device (device pointers, stream-ordered, synchronises before returning). */
int raft_bst_device(int device, float* d_obj, const float* d_tomo, const float* paramf, const int* parami,
                    void* stream);

/* Flat/dark background correction (replaces getBackgroundCorrectionMultiGPU,
 * cuda/src/processing/background_correction.cu:144-178, called from
 * sscRaft/processing/background_correction.py:105-108): in place on host
 * frames[nslices][nangles][nrays], flat[nslices][numflats][nrays],
 * dark[nslices][nrays]:  v = max(D - dark, 0.5) / max(F - dark, 0.5), F the
 * first flat or, with two or more flats, the first two blended by angle/nangles;
 * then v = -log(v) when is_log. */
void getBackgroundCorrectionMultiGPU(int* gpus, int ngpus, float* frames, float* flat, float* dark, int nrays,
                                     int nangles, int nslices, int numflats, int is_log, int blocksize);

/* Same on one device, in place on d_frames (stream-ordered, asynchronous). */
int raft_background_device(int device, float* d_frames, const float* d_flat, const float* d_dark, int nrays,
                           int nangles, int nslices, int numflats, int is_log, void* stream);

/* Transmission EM (replaces get_tEM_RT_MultiGPU, cuda/inc/geometries/parallel/
 * em.hpp:13-15, em_rt.cu:12-80; bound at rafttypes.py:128-133, called from
 * sscRaft/geometries/parallel/em_methods.py:141): obj [nslices][objsize][objsize]
 * in/out (initial guess), count [nslices][nangles][nrays] (= exp(-data), as the
 * Python wrapper passes it), flat [nslices][nrays] (one row per slice), angles
 * [nangles]; parami as eEM; paramf unused.  The flat row of slice s is used
 * for slice s (the reference indexes the flat buffer by angle). */
void get_tEM_RT_MultiGPU(int* gpus, int ngpus, float* obj, float* count, float* flat, float* angles, float* paramf,
                         int* parami);

/* Same on one device (device pointers, stream-ordered, synchronises before returning). */
int raft_tem_device(int device, float* d_obj, const float* d_count, const float* d_flat, const float* angles,
                    const int* parami, void* stream);

/* Filter-only export (replaces the reference's getFilterLowPassMultiGPU,
 * cuda/src/processing/filters/low_pass.cu:87-141, bound at rafttypes.py:190-195,
 * called from sscRaft/processing/filters.py:84): in place on host
 * tomogram[nslices][nangles][nrays]; paramf/parami as getFBPMultiGPU; the
 * axis offset paramf[6] is ignored (the reference filters with offset 0). */
void getFilterLowPassMultiGPU(int* gpus, int ngpus, float* tomogram, float* paramf, int* parami);

/* Same on one device, in place on d_tomo (stream-ordered, synchronises before returning). */
int raft_lowpass_device(int device, float* d_tomo, const float* paramf, const int* parami, void* stream);

/* Batched uneven scans in one graph launch (no counterpart in the reference,
 * which reconstructs one volume of equal-sized slices per call).  For scan i:
 *   FBP of d_tomo[i] [nangles[i]][nrays[i]] -> d_fbp[i] [nrays[i]]^2, with
 *     paramf[7 i .. 7 i + 6] (getFBPMultiGPU layout), padding `pad`, filter
 *     `filter`, filtered rows stored as `store` (RAFT_STORE_F32 / _F16);
 *   `iterations` eEM iterations on d_y[i] [nangles[i]][nrays[i]], d_em[i] the
 *     in/out iterate [nrays[i]]^2 (objsize = nrays).
 * Pass d_fbp = NULL to skip FBP, d_em = NULL (or iterations 0) to skip eEM.
 * Scans with equal geometry are batched together; every batch is a branch of
 * one captured hipGraph, launched once on `stream`.  All pointers in the
 * arrays are device pointers; the arrays themselves are host arrays.
 * *graph_ms (if non-null) receives the device time of the graph launch.
 * Synchronises the stream before returning. */
int raft_batch_device(int device, int nscans, const int* nrays, const int* nangles, const float* const* angles,
                      const float* const* d_tomo, const float* const* d_y, float* const* d_fbp, float* const* d_em,
                      const float* paramf, int pad, int filter, int store, int iterations, void* stream,
                      float* graph_ms);

/* KL(y || A x) per slice of the eEM objective for the given images (one
 * projection with the eEM projector; d_obj is not modified): kl/kl_excluded
 * are HOST arrays [nslices]; iterations in parami are ignored.  No
 * counterpart in the reference; used to compare candidate eEM starting images.
 * Synchronises the stream before returning. */
int raft_eem_kl_device(int device, const float* d_obj, const float* d_tomo, const float* angles, const int* parami,
                       double* kl, long long* kl_excluded, void* stream);

/* Host-only planning query (no device work): the bin window the FBP
 * backprojector stages per angle row for these parameters -- first bin *tlo
 * and row width *width -- and whether the filter runs on the LDS FFT path
 * (*fft_path = 1) or the direct-convolution path (0).  Returns 0 or an error. */
int raft_fbp_plan_info(const float* angles, const float* paramf, const int* parami, int* tlo, int* width,
                       int* fft_path);

/* Host-only planning query of the forward projector (no device work): for
 * the Radon geometry (em = 0: an Ny x Nx image, half extents ax, ay) or the
 * eEM one (em = 1: an Nx x Nx image, Ny = Nx, ax and ay ignored), the number
 * of ray chunks *ray_chunks of the staged (LDS) form, or 0 when the direct
 * form runs (a chunk's staged box would not fit the tile, or RAFT_PROJ_LDS=0).
 * tile_consts[4] (if non-null) receives the staged form's constants: rays per
 * patch, samples per step, angles per workgroup, staged 16-byte entries.
 * Returns 0 or an error code. */
int raft_proj_plan_info(const float* angles, int nangles, int Nx, int Ny, int nrays, int nslices, float ax, float ay,
                        int em, int* ray_chunks, int* tile_consts);

/* The autotuner's choices (tile variants per kernel family and geometry) as
 * JSON text, written into buf (NUL-terminated, truncated to len); returns the
 * length the whole text needs, NUL included.  The choices are made by the
 * first call that needs them (each candidate timed on the call's own data)
 * and kept for the process; RAFT_AUTOTUNE_CACHE=<file> keeps them across
 * processes in that JSON file, RAFT_AUTOTUNE=0 turns tuning off.  No
 * counterpart in the reference (fixed 16x16x4 blocks, configs.hpp:22-24). */
int raft_autotune_dump(char* buf, int len);
/* Forgets the choices held in memory (the next call re-reads the cache file,
 * if any, and tunes what it does not hold). */
void raft_autotune_clear(void);

/* Average device time (ms) of the last launch of each kernel family on this
 * thread's last call, for roofline reporting: which = 0 filter, 1 FBP
 * backprojection, 2 projector, 3 EM backprojection/update.  Returns -1 if
 * not measured.  Timing is recorded only when raft_set_profiling(1). */
void raft_set_profiling(int on);
double raft_last_kernel_ms(int which);

/* LDS read rate of the device: conflict-free 16-byte LDS reads (the access of
 * the backprojector's and staged projector's gathers) by every CU, timed with
 * HIP events.  *bytes_per_s = LDS bytes delivered per second (whole device),
 * *cus = compute units, *clock_khz = the device's peak clock attribute.  For
 * the roofline of the gather kernels, whose bound is LDS read bandwidth.  No
 * counterpart in the reference.  Synchronous. */
int raft_lds_read_probe(int device, double* bytes_per_s, int* cus, int* clock_khz);

#ifdef __cplusplus
}
#endif

#endif /* RAFT_H */
