// raft_stubs.cpp -- the symbols of the reference's ctypes loader that lie
// outside the parallel-beam hot path.  Each records RAFT_EUNSUPPORTED in the
// thread-local error state and returns without touching its arguments.
#include <cmath>

#include "../../include/raft_stubs.h"
#include "raft_internal.hpp"

namespace {
void unsupported(const char* name)
{
    raft::record_error(RAFT_EUNSUPPORTED,
                       std::string("not implemented: ") + name + " (outside the parallel-beam FBP/Radon/EM path)");
}
} // namespace

extern "C" {

void flip_x(float*, int, int, int) { unsupported("flip_x"); }
void transpose_cpu_zyx2xyz(float*, int, int, int) { unsupported("transpose_cpu_zyx2xyz"); }
void get_tEM_FQ_MultiGPU(int*, int, float*, float*, float*, float*, float*, int*)
{
    unsupported("get_tEM_FQ_MultiGPU");
}
float findcentersino_subpixel(float*, float*, float*, float*, int, int)
{
    unsupported("findcentersino_subpixel");
    return NAN;
}
int findcentersino(float*, float*, float*, float*, int, int)
{
    unsupported("findcentersino");
    return -1;
}
void correctRotationAxis(float*, float*, int, int, int, int) { unsupported("correctRotationAxis"); }
void getRotAxisCorrectionMultiGPU(int*, int, float*, float, int, int, int, int)
{
    unsupported("getRotAxisCorrectionMultiGPU");
}
void getTitarenkoRingsMultiGPU(int*, int, float*, int, int, int, float, int, int)
{
    unsupported("getTitarenkoRingsMultiGPU");
}
int getOffsetExcentricTomoGPU(int, float*, int, int, int)
{
    unsupported("getOffsetExcentricTomoGPU");
    return -1;
}
void getExcentricTomoMultiGPU(int*, int, float*, int, int, int, int) { unsupported("getExcentricTomoMultiGPU"); }
void getPhaseMultiGPU(int*, int, float*, float*, int*) { unsupported("getPhaseMultiGPU"); }
void gpu_fdk(struct RaftStubLabFDK, float*, float*, float*, int*, int, double*) { unsupported("gpu_fdk"); }
int conebeam_tEM_gpu(struct RaftStubLabEM, float*, float*, float*, float*, float*, float*, float*, int, int, float,
                     float)
{
    unsupported("conebeam_tEM_gpu");
    return -1;
}
int cbradon(struct RaftStubLabCB, float*, float*, float*, float*, float*, float*, int)
{
    unsupported("cbradon");
    return -1;
}
int cbradon_MultiGPU(int*, int, struct RaftStubLabCB, float*, float*, float*, float*, float*, float*)
{
    unsupported("cbradon_MultiGPU");
    return -1;
}
void ReconstructionPipeline(float*, float*, float*, float*, float*, float*, int*, int*, int*, int)
{
    unsupported("ReconstructionPipeline");
}

} // extern "C"
