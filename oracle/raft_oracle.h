/*
 * raft_oracle.h -- declarations of the CPU oracle (test infrastructure only;
 * see raft_oracle.c).  Every function exists as *_f64 (FP64 oracle) and
 * *_f32 (FP32 restatement used as the timed CPU baseline).  Arrays are
 * C-contiguous; sinograms are [slices][angles][rays], images [slices][y][x].
 * Return 0 on success, -1 on invalid parameters, -2 on allocation failure.
 */
#ifndef RAFT_ORACLE_H
#define RAFT_ORACLE_H

#define RAFT_ORACLE_DECL(real, sfx)                                                                                 \
    int oracle_fbp_##sfx(real* obj, const float* tomo, const float* angles, const float* paramf, const int* parami, \
                         int nthreads);                                                                             \
    int oracle_fbp_padded_##sfx(real* obj, const float* tomo, const float* angles, const float* paramf,           \
                                const int* parami, int nthreads);                                                  \
    int oracle_fbp_filter_##sfx(real* q, const float* tomo, const float* paramf, const int* parami, int nthreads); \
    int oracle_radon_##sfx(real* proj, const float* phantom, const float* angles, int Nx, int Ny, int nrays,       \
                           int nangles, int nslices, float ax, float ay, int nthreads);                            \
    double oracle_kl_##sfx(const real* y, const real* v, long nbins, long* excluded);                              \
    int oracle_eem_##sfx(real* obj, const float* tomo, const float* angles, const int* parami, double* kl,         \
                         long* kl_excluded, int nthreads);                                                         \
    int oracle_em_backproject_##sfx(real* out, const real* q, const float* angles, int nrays, int nangles, int n,  \
                                    int ones);                                                                     \
    int oracle_em_forward_##sfx(real* v, const real* x, const float* angles, int nrays, int nangles, int n);

RAFT_ORACLE_DECL(double, f64)
RAFT_ORACLE_DECL(float, f32)

#endif
