// noalloc.hip -- gn_invert_noalloc: point inversion that allocates nothing.
//
// The persistent inversion kernels need device memory beyond the caller's inputs and outputs:
// claim counters (one slot of kMaxSlices 128-byte lines per launch) and, for the mixed mode's
// streamed closing step, a continuation list of point indices.  gn_invert takes them from a
// per-device ring that the library allocates on first use.  Callers that may not allocate
// (code under graph capture, tight serving loops, harnesses that forbid hipMalloc) use this entry
// point instead: its counters and list live in `scratch`, a fixed device array of this file.
//
// `scratch` names the array in device code only.  Host code takes its device address once per
// device with hipGetSymbolAddress (with that device current) and uses that address wherever it
// hands the array to a kernel.  Every call on a device shares that device's array: calls that
// overlap on one GPU (on different streams, or from different threads) race on it.
#include <mutex>

#include "invert_common.hpp"

namespace gn {
namespace {

constexpr int kNaSlots = 2;  // mixed: FP32 bulk + FP64 closing; plain FP64: one
constexpr size_t kNaCtrWords = (size_t)kNaSlots * kMaxSlices * kCounterStride;  // 64-bit words
constexpr size_t kNaListWords = kListCap / 2;  // kListCap 32-bit indices
constexpr size_t kNaWords = kNaCtrWords + kNaListWords;

// 16 KB of claim counters, then a 512 KB continuation list (528 KB per device)
__device__ __align__(256) unsigned long long scratch[kNaWords];

constexpr int kNaMaxDevices = 64;
std::mutex g_na_mu;
unsigned long long *g_na_ptr[kNaMaxDevices] = {};

// The current device's address of `scratch`, looked up once per device with it current.
int scratch_address(unsigned long long **out) {
    int dev = 0;
    GN_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kNaMaxDevices) return fail(GN_EDEVICE, "device index out of range");
    std::lock_guard<std::mutex> lk(g_na_mu);
    if (!g_na_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
        if (!p) return fail(GN_EDEVICE, "hipGetSymbolAddress(scratch) returned null");
        g_na_ptr[dev] = (unsigned long long *)p;
    }
    *out = g_na_ptr[dev];
    return GN_OK;
}

// Clears the counter words (the kernels leave them zeroed, but a call must not depend on how the
// last one ended); a kernel rather than hipMemsetAsync, so the call can be captured.
__global__ __launch_bounds__(256) void k_zero_words(unsigned long long *__restrict__ p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0ull;
}

}  // namespace
}  // namespace gn

extern "C" {

int gn_invert_noalloc(const gn_surface *s, const double *x, const double *y, const double *z,
                      size_t n, const gn_invert_opts *opts, double *u_out, double *v_out,
                      uint8_t *status, gn_stream stream) {
    using namespace gn;
    // the continuation list holds 32-bit point indices
    if (n > (size_t)UINT32_MAX)
        return fail(GN_EINVAL, "n exceeds the no-allocation maximum of 2^32-1 points per call");
    gn_invert_opts o;
    if (opts) o = *opts;
    else gn_invert_opts_default(&o);
    int rc = check_invert(s, x, y, z, n, o, u_out, v_out, status);
    if (rc) return rc;
    if (o.fallback)
        return fail(GN_EINVAL, "the global fallback allocates its point list: not available "
                               "without allocation (fallback must be 0)");
    if (n == 0) return GN_OK;
    unsigned long long *base = nullptr;
    if ((rc = scratch_address(&base))) return rc;
    hipStream_t st = (hipStream_t)stream;
    k_zero_words<<<(unsigned)((kNaCtrWords + 255) / 256), 256, 0, st>>>(base, kNaCtrWords);
    GN_HIP(hipGetLastError());
    SlotPool pool;
    pool.ctr = base;
    pool.nslots = kNaSlots;
    pool.list = reinterpret_cast<uint32_t *>(base + kNaCtrWords);
    pool.nlists = 1;
    PoolScope scope(&pool);
    return launch_invert(s->dev(), s->search(), s->d_tree, s->d_valid, s->tree_levels, x, y, z, n,
                         o, u_out, v_out, status, nullptr, nullptr, st);
}

}  // extern "C"
