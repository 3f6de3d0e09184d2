/*
 * raft_stubs.h -- exports of libraft.so for the symbols the reference's
 * ctypes loader binds outside the parallel-beam hot path.
 *
 * sscRaft's rafttypes.py sets argtypes/restype on 23 functions of
 * libraft.so.  Seven of them are built here (include/raft.h).  The other 16
 * belong to subsystems this library does not provide (cone beam, rings,
 * rotation axis search, stitching, phase retrieval, frequency-domain EM, the
 * C pipeline).  They are exported with the reference's C signatures so the
 * reference loader resolves every name -- its conebeam_tEM_gpu binding raises
 * NotImplementedError when the symbol is missing (rafttypes.py:375-391) --
 * and each one only records RAFT_EUNSUPPORTED ("not implemented: <name>") in
 * the thread-local error state (raft_last_error / raft_last_error_string)
 * and returns: void functions return nothing, int functions return -1, the
 * float one returns NaN.  Output arrays are left untouched.
 */
#ifndef RAFT_STUBS_H
#define RAFT_STUBS_H

#ifdef __cplusplus
extern "C" {
#endif

/* Struct arguments passed by value, laid out as the ctypes.Structure the
 * reference Python passes (rafttypes.py:323-350 Lab, 361-372 Lab_EM,
 * 394-409 Lab_CB). */
struct RaftStubLabFDK {
    float x, y, z, dx, dy, dz;
    int nx, ny, nz;
    float h, v, dh, dv;
    int nh, nv;
    float D, Dsd, beta_max, dbeta;
    int nbeta, fourier, filter_type;
    float reg;
    int is_slice, slice_recon_start, slice_recon_end, slice_tomo_start, slice_tomo_end, nph, padh;
    float energy;
    int rotation_axis_offset, blocksize;
};
struct RaftStubLabEM {
    float Lx, Ly, Lz, x0, y0, z0;
    int nx, ny, nz;
    float sx, sy, sz;
    int nbeta, ndetc, n_ray_points;
};
struct RaftStubLabCB {
    float x, y, z, x0, y0, z0;
    int nx, ny, nz;
    float sx, sy, sz;
    int nbeta, n_detector, n_ray_points;
};

/* opt::flip_x, cuda/src/common/opt.cu:22 (bound at rafttypes.py:104-108) */
void flip_x(float* data, int sizex, int sizey, int sizez);
/* opt::transpose_cpu_zyx2xyz, cuda/src/common/opt.cu:46 (rafttypes.py:115-119) */
void transpose_cpu_zyx2xyz(float* data, int sizex, int sizey, int sizez);
/* frequency-domain transmission EM, cuda/src/geometries/parallel/reconstruction_methods/
 * em_frequency.cu:339 (rafttypes.py:151-156) */
void get_tEM_FQ_MultiGPU(int* gpus, int ngpus, float* count, float* obj, float* angles, float* flat, float* paramf,
                         int* parami);
/* rotation axis search and correction, cuda/src/processing/rotation_axis.cu:402,437,444,587
 * (rafttypes.py:202-239) */
float findcentersino_subpixel(float* frame0, float* frame180, float* dark, float* flat, int sizex, int sizey);
int findcentersino(float* frame0, float* frame180, float* dark, float* flat, int sizex, int sizey);
void correctRotationAxis(float* tomoin, float* tomoout, int sizex, int sizey, int sizez, int deviation);
void getRotAxisCorrectionMultiGPU(int* gpus, int ngpus, float* tomogram, float axis_offset, int nrays, int nangles,
                                  int nslices, int blocksize);
/* Titarenko ring removal, cuda/src/processing/rings.cu:574 (rafttypes.py:246-252) */
void getTitarenkoRingsMultiGPU(int* gpus, int ngpus, float* data, int nrays, int nangles, int nslices,
                               float lambda_rings, int ring_blocks, int blocksize);
/* 360-degree (excentric) stitching, cuda/src/processing/excentric_tomo.cu:469,521
 * (rafttypes.py:274-293) */
int getOffsetExcentricTomoGPU(int ngpu, float* sinogram, int sizex, int sizey, int sizez);
void getExcentricTomoMultiGPU(int* gpus, int ngpus, float* data, int nrays, int nangles, int nslices, int offset);
/* phase retrieval, cuda/src/processing/phase/phase_retrieval.cu:237 (rafttypes.py:300-305) */
void getPhaseMultiGPU(int* gpus, int ngpus, float* projections, float* paramf, int* parami);
/* cone-beam FDK, cuda/src/geometries/conebeam/reconstruction_methods/fdk.cu:23 (rafttypes.py:352-357) */
void gpu_fdk(struct RaftStubLabFDK lab, float* recon, float* proj, float* angles, int* gpus, int ndevs,
             double* time);
/* cone-beam transmission EM, cuda/src/geometries/conebeam/reconstruction_methods/em.cu:29
 * (rafttypes.py:375-391) */
int conebeam_tEM_gpu(struct RaftStubLabEM lab, float* flat, float* px, float* py, float* pz, float* beta,
                     float* recon, float* tomo, int ngpus, int niter, float tv, float max_val);
/* cone-beam Radon, cuda/src/geometries/conebeam/forward_methods/radon.cu:15,47 (rafttypes.py:411-429) */
int cbradon(struct RaftStubLabCB lab, float* px, float* py, float* pz, float* beta, float* sample, float* tomo,
            int gpu_id);
int cbradon_MultiGPU(int* gpus, int ngpus, struct RaftStubLabCB lab, float* px, float* py, float* pz, float* beta,
                     float* sample, float* tomo);
/* C reconstruction pipeline, cuda/src/pipelines/pipeline.cu:17 (rafttypes.py:436-443) */
void ReconstructionPipeline(float* obj, float* data, float* flats, float* darks, float* angles,
                            float* parameters_float, int* parameters_int, int* flags, int* gpus, int ngpus);

#ifdef __cplusplus
}
#endif

#endif /* RAFT_STUBS_H */
