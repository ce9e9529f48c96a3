// runtime.hip -- small per-device runtime services of libgnurbs:
//   - claim counters for the persistent (lane-refill) kernels: a ring of slots per device, each
//     slot kMaxSlices 64-bit counters on their own 128-byte lines plus an exit counter; every
//     launch takes the next slot.  Slots start zeroed and every kernel that uses one leaves it
//     zeroed (the last wave to retire clears it), so a launch needs no clearing kernel or memset
//     and a captured hipGraph can be replayed.  Two eager launches share a slot only if more than
//     kCounterSlots launches are in flight at once.  Captured graphs and the no-allocation entry
//     point take their slots from pools of their own (SlotPool), never from the ring.
//   - beside each slot, a continuation list of kListCap point indices (128 MB per device for the
//     ring, allocated with the counters).
//   - a copy kernel for small host-to-device images (the server's step tables).
//   - the compute-unit count used to size persistent grids.
#include <atomic>
#include <mutex>
#include <unordered_map>

#include "invert_common.hpp"

namespace gn {

constexpr int kCounterSlots = 256;
constexpr int kMaxDevices = 64;
constexpr size_t kSlotWords = (size_t)kMaxSlices * kCounterStride;

namespace {
std::mutex g_mu;
unsigned long long *g_counters[kMaxDevices] = {};
uint32_t *g_lists[kMaxDevices] = {};
int g_cus[kMaxDevices] = {};
std::atomic<unsigned> g_next{0};
thread_local SlotPool *t_pool = nullptr;
}  // namespace

size_t slot_words() { return kSlotWords; }
size_t list_entries() { return kListCap; }

PoolScope::PoolScope(SlotPool *pool) : prev_(t_pool) { t_pool = pool; }
PoolScope::~PoolScope() { t_pool = prev_; }
bool slot_pool_active() { return t_pool != nullptr; }

int ensure_runtime() {
    int dev = 0;
    GN_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(GN_EDEVICE, "device index out of range");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_counters[dev]) {
        void *p = nullptr;
        GN_HIP(hipMalloc(&p, sizeof(unsigned long long) * kCounterSlots * kSlotWords));
        GN_HIP(hipMemset(p, 0, sizeof(unsigned long long) * kCounterSlots * kSlotWords));
        GN_HIP(hipDeviceSynchronize());
        g_counters[dev] = (unsigned long long *)p;
    }
    if (!g_lists[dev]) {  // with the counters, so one inversion before a graph capture readies both
        void *p = nullptr;
        GN_HIP(hipMalloc(&p, sizeof(uint32_t) * kCounterSlots * kListCap));  // 128 MB
        g_lists[dev] = (uint32_t *)p;
    }
    return GN_OK;
}

int work_counters(hipStream_t st, unsigned long long **out) {
    if (SlotPool *p = t_pool) {
        if (p->used >= p->nslots) return fail(GN_ESTATE, "counter slot pool used up");
        *out = p->ctr + (size_t)p->used++ * kSlotWords;
        return GN_OK;
    }
    int rc = ensure_runtime();
    if (rc) return rc;
    int dev = 0;
    GN_HIP(hipGetDevice(&dev));
    const unsigned slot = g_next.fetch_add(1, std::memory_order_relaxed) % kCounterSlots;
    (void)st;
    *out = g_counters[dev] + (size_t)slot * kSlotWords;
    return GN_OK;
}

int work_list(const unsigned long long *counters, uint32_t **list) {
    if (SlotPool *p = t_pool) {
        if (counters < p->ctr || counters >= p->ctr + (size_t)p->nslots * kSlotWords)
            return fail(GN_EDEVICE, "counter slot outside the thread's pool");
        if (p->lists_used >= p->nlists) return fail(GN_ESTATE, "continuation list pool used up");
        *list = p->list + (size_t)p->lists_used++ * kListCap;
        return GN_OK;
    }
    int dev = 0;
    GN_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(GN_EDEVICE, "device index out of range");
    std::lock_guard<std::mutex> lk(g_mu);
    const unsigned long long *base = g_counters[dev];
    if (!base || !g_lists[dev] || counters < base ||
        counters >= base + (size_t)kCounterSlots * kSlotWords)
        return fail(GN_EDEVICE, "counter slot of another device");
    const size_t slot = (size_t)(counters - base) / kSlotWords;
    *list = g_lists[dev] + slot * kListCap;
    return GN_OK;
}

__global__ __launch_bounds__(256) void k_stage_copy(const uint4 *__restrict__ src,
                                                    uint4 *__restrict__ dst, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

int launch_stage_copy(const void *src, void *dst, size_t bytes, hipStream_t st) {
    const size_t n = (bytes + 15) / 16;
    if (n == 0) return GN_OK;
    const size_t blocks = (n + 255) / 256;
    k_stage_copy<<<(unsigned)(blocks < 1024 ? blocks : 1024), 256, 0, st>>>(
        (const uint4 *)src, (uint4 *)dst, n);
    GN_HIP(hipGetLastError());
    return GN_OK;
}

int blocks_per_cu(const void *kernel, int block) {
    static std::mutex mu;
    static std::unordered_map<const void *, int> cache;
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(kernel);
    if (it != cache.end()) return it->second;
    int b = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, kernel, block, 0) != hipSuccess || b <= 0)
        b = 1;
    cache[kernel] = b;
    return b;
}

int cu_count() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return 256;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_cus[dev]) {
        int n = 0;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0)
            n = 256;
        g_cus[dev] = n;
    }
    return g_cus[dev];
}

}  // namespace gn
