// le_geneig_noalloc: the FP64 generalized eigensolve of le_geneig on device buffers, with
// its workspace in a file-scope device array instead of device allocations, for callers that
// cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses
// that forbid hipMalloc).  No reference counterpart.
//
// The array `scratch` holds, per device, the workspace of one problem of dimension up to
// LE_GENEIG_NOALLOC_MAX_DIM (2 (8 d^2 + 4 d) doubles: L, C, W, T and the tridiagonal /
// inverse-iteration arrays of the eigensolver kernels), the problem descriptor and a status
// word.  Each device's copy lives at its own address, taken once per device with
// hipGetSymbolAddress on the first call with that device current; the host uses only that
// address.  Every call on one device shares the array and leaves its contents behind: calls
// that overlap on one device (other streams, other host threads) race.  The descriptor is
// written by a one-thread kernel and the status by the eigensolver itself on every path, so
// nothing needs clearing and the call records no memset.
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <string>

#include "../../include/laughlinedge.h"
#include "le_kernels.hpp"

namespace le {
void set_error(const std::string& msg);
}

namespace {

constexpr long kMaxDim = LE_GENEIG_NOALLOC_MAX_DIM;
constexpr size_t kWorkDoubles = 2 * (8 * kMaxDim * kMaxDim + 4 * kMaxDim);
constexpr size_t kWorkBytes = kWorkDoubles * sizeof(double);
constexpr size_t kDescOffset = kWorkBytes;              // one le::GenEigProblem
constexpr size_t kStatusOffset = kWorkBytes + 256;      // one int
constexpr size_t kScratchBytes = kWorkBytes + 512;
static_assert(sizeof(le::GenEigProblem) <= 256, "descriptor slot");

__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
std::map<int, unsigned char*> g_addr;   // device ordinal -> its copy of `scratch`

__global__ void k_noalloc_problem(le::GenEigProblem* p, int d)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        p->d = d;
        p->in_off = 0;
        p->eval_off = 0;
        p->work_off = 0;
    }
}

int scratch_address(unsigned char** out)
{
    int dev = -1;
    hipError_t h = hipGetDevice(&dev);
    if (h != hipSuccess) {
        le::set_error(std::string("le_geneig_noalloc: hipGetDevice: ") + hipGetErrorString(h));
        return LE_ERR_HIP;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_addr.find(dev);
    if (it != g_addr.end()) {
        *out = it->second;
        return LE_OK;
    }
    void* p = nullptr;
    h = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (h != hipSuccess || !p) {
        le::set_error(std::string("le_geneig_noalloc: hipGetSymbolAddress(scratch): ") + hipGetErrorString(h));
        return LE_ERR_HIP;
    }
    g_addr[dev] = static_cast<unsigned char*>(p);
    *out = static_cast<unsigned char*>(p);
    return LE_OK;
}

}  // namespace

extern "C" int le_geneig_noalloc(int d, const double* S_dev, const double* H_dev, double* evals_dev,
                                 double* evecs_dev, int* status_dev, void* stream)
{
    if (d <= 0 || !S_dev || !H_dev || !evals_dev) {
        le::set_error("le_geneig_noalloc: need d > 0 and device S, H and evals");
        return LE_ERR_ARG;
    }
    if (d > kMaxDim) {
        le::set_error("le_geneig_noalloc: d = " + std::to_string(d) + " exceeds the maximum of " +
                      std::to_string(kMaxDim) + " (its workspace is a fixed device array)");
        return LE_ERR_ARG;
    }
    unsigned char* base = nullptr;
    int rc = scratch_address(&base);
    if (rc != LE_OK) return rc;
    auto* work = reinterpret_cast<double*>(base);
    auto* desc = reinterpret_cast<le::GenEigProblem*>(base + kDescOffset);
    int* status = status_dev ? status_dev : reinterpret_cast<int*>(base + kStatusOffset);
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_noalloc_problem, dim3(1), dim3(64), 0, s, desc, d);
    hipError_t h = hipGetLastError();
    if (h == hipSuccess) h = le::launch_geneig(S_dev, H_dev, work, desc, 1, d, evals_dev, evecs_dev, status, s);
    if (h != hipSuccess) {
        le::set_error(std::string("le_geneig_noalloc: ") + hipGetErrorString(h));
        return LE_ERR_HIP;
    }
    return LE_OK;
}
