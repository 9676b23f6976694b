// lbib_noalloc.hip -- the allocation-free variant of lbib_domain_concentrations.
//
// lbib_domain_concentrations accumulates the per-domain sums in a device
// buffer the context allocates (and grows) on demand, clears it with a memset
// and copies the result to host memory.  Callers that cannot allocate or wait
// (code under HIP graph capture, serving loops) use
// lbib_domain_concentrations_noalloc instead: the sums accumulate in `scratch`,
// a static device array of kMaxDomains doubles (one copy per GPU, held until
// the process ends), cleared by a fill kernel, and are written to the caller's
// device array by a second kernel, all on the context's stream.  Nothing is
// allocated, copied to the host or waited for.  Every call on a GPU shares
// that GPU's scratch, so calls that overlap on one GPU race.
#include <climits>
#include <mutex>
#include <string>

#include "lbib_internal.h"

namespace lbib {
namespace {

constexpr int kMaxDomains = 65536;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[kMaxDomains];

// the address of this GPU's copy of `scratch`, looked up once per device with
// that device current (the array's name is not a device address on the host)
double* g_scratch[kMaxDevices];
std::once_flag g_once[kMaxDevices];
hipError_t g_lookup[kMaxDevices];

__global__ void k_zero(double* p, int count) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) p[i] = 0.0;
}
__global__ void k_publish(const double* __restrict__ src, double* __restrict__ dst, int count) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) dst[i] = src[i];
}

double* scratch_of_current_device(lbib_ctx* c) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        c->err = "lbib_domain_concentrations_noalloc: no usable current device";
        return nullptr;
    }
    std::call_once(g_once[dev], [dev] {
        void* p = nullptr;
        g_lookup[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        g_scratch[dev] = g_lookup[dev] == hipSuccess ? static_cast<double*>(p) : nullptr;
    });
    if (!g_scratch[dev]) {
        c->err = std::string("lbib_domain_concentrations_noalloc: hipGetSymbolAddress: ") +
                 hipGetErrorString(g_lookup[dev]);
        return nullptr;
    }
    return g_scratch[dev];
}

}  // namespace
}  // namespace lbib

using namespace lbib;

extern "C" int lbib_domain_concentrations_noalloc(lbib_ctx* c, int species, int ndom, double* d_out) {
    if (ctx_ready(c)) return -1;
    if (species < 0 || species >= c->ns || ndom < 1 || !d_out) {
        c->err = "lbib_domain_concentrations_noalloc: bad species index, domain count or output pointer";
        return -1;
    }
    if (ndom > kMaxDomains) {
        c->err = "lbib_domain_concentrations_noalloc: at most " + std::to_string(kMaxDomains) +
                 " domains (the static scratch size); use lbib_domain_concentrations";
        return -1;
    }
    double* acc = scratch_of_current_device(c);
    if (!acc) return -1;
    const int blocks = (ndom + 255) / 256;
    k_zero<<<blocks, 256, 0, c->stream>>>(acc, ndom);
    LBIB_HIP(hipGetLastError());
    if (launch_domain_concentrations(c, species, ndom, acc)) return -1;
    k_publish<<<blocks, 256, 0, c->stream>>>(acc, d_out, ndom);
    LBIB_HIP(hipGetLastError());
    return 0;
}
