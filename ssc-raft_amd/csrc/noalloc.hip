// noalloc.hip -- the fixed device scratch array of raft_fbp_device_noalloc.
//
// `scratch` is a file-scope __device__ array: every GPU the library runs on
// holds its own copy, at its own address, from the moment the library's code
// is loaded there until the process ends.  In host code the array's name is a
// host-side stand-in and not a device address, so host code never hands it to
// a kernel or a copy.  noalloc_scratch() asks the runtime for the device
// address once per device, with that device current, checks the result and
// keeps the pointer per device; every host use of the scratch memory goes
// through that pointer.
//
// Every call on a GPU shares that GPU's array and finds the previous call's
// contents there.  Calls on one device run one at a time under the device's
// lock and synchronise their stream before they return; calls that overlapped
// on one GPU (two processes excepted: each has its own copy) would race.
#include <map>
#include <mutex>

#include "raft_internal.hpp"

static __device__ __align__(256) unsigned char scratch[RAFT_NOALLOC_SCRATCH_BYTES];

namespace raft {

namespace {
std::mutex g_scratch_mu;
std::map<int, void*> g_scratch_ptr; // device id -> device address of that device's copy of scratch
} // namespace

void* noalloc_scratch(int dev)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch_ptr.find(dev);
    if (it != g_scratch_ptr.end()) return it->second;
    int cur = -1;
    RAFT_HIP(hipGetDevice(&cur));
    if (cur != dev) fail(RAFT_EINVAL, "internal: scratch address asked for a device that is not current");
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        (void)hipGetLastError();
        fail(RAFT_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    }
    g_scratch_ptr[dev] = p;
    return p;
}

} // namespace raft
