// J/K build that allocates nothing: qcg_build_jk_device with its accumulators in a static device
// array instead of the context's buffers, for callers that may not allocate device memory (code under
// HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// The array holds, for one molecule of at most kNoAllocMaxNbf functions, the fixed-point J and K
// accumulators (nbf^2 words each), the four scale doubles and 2 nbf doubles of row sums: 4.2 MB,
// resident on every GPU the library runs on until the process ends.  Every call on a GPU shares that
// GPU's copy; the kernels clear the accumulators at the start of each call (fill kernels, no
// hipMemsetAsync), and calls that overlap on one GPU race on it.
#include <mutex>

#include "context.h"

namespace {

constexpr int kNoAllocMaxNbf = 512;
constexpr size_t kScratchWords = 2 * (size_t)kNoAllocMaxNbf * kNoAllocMaxNbf + 4 + 2 * (size_t)kNoAllocMaxNbf;

__device__ __align__(256) unsigned long long scratch[kScratchWords];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};
bool g_done[kMaxDevices] = {};

}  // namespace

// Device address of this GPU's copy of the scratch array (the symbol's name is only a host-side
// stand-in).  Called once per device from qcg_create, with that device current.
hipError_t qcg::noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_done[device]) return g_scratch[device] ? hipSuccess : hipErrorInvalidSymbol;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_done[device] = true;
    g_scratch[device] = e == hipSuccess ? p : nullptr;
    return e;
}

extern "C" int qcg_build_jk_device_noalloc(qcg_context* c, const double* dD, double* dJ, double* dK, void* stream) {
    if (!c || !dD) return QCG_INVALID_ARG;
    if (!c->has_basis) return fail(c, QCG_NO_BASIS, "no basis");
    if (c->nmol != 1) return fail(c, QCG_INVALID_ARG, "qcg_build_jk_device_noalloc: single molecules only");
    const int nbf = c->basis.nbf;
    if (nbf > kNoAllocMaxNbf)
        return fail(c, QCG_INVALID_ARG,
                    "qcg_build_jk_device_noalloc: " + std::to_string(nbf) + " basis functions, at most " +
                        std::to_string(kNoAllocMaxNbf) + " fit its static scratch");
    void* base = (c->device >= 0 && c->device < kMaxDevices) ? g_scratch[c->device] : nullptr;
    if (!base) return fail(c, QCG_DEVICE, "qcg_build_jk_device_noalloc: no scratch address for this device");
    HIPCHK(c, hipSetDevice(c->device));
    fix_t* w = static_cast<fix_t*>(base);
    const size_t n2 = (size_t)nbf * nbf;
    const JkAcc acc{w, w + n2, reinterpret_cast<double*>(w + 2 * n2), reinterpret_cast<double*>(w + 2 * n2 + 4)};
    return run_jk(c, dD, dJ, dK, stream ? (hipStream_t)stream : c->stream, &acc, true);
}
