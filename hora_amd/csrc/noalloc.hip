// noalloc.hip -- the search entry point that allocates nothing
// (hm_index_search_device_noalloc, include/hora_mi355x.h).
//
// Callers under HIP graph capture, in tight serving loops or in harnesses
// that forbid hipMalloc cannot let the library grow its working buffers.
// This entry runs BruteForceIndex::node_search_k in the exact form (the same
// kernels, and so the same bits, as hm_index_search_device in
// HM_SEARCH_EXACT mode) with every working buffer carved from `scratch`, a
// file-scope device array of HM_NOALLOC_SCRATCH_BYTES that each GPU holds
// from the library's load until the process ends.  Queries run in chunks
// sized so that their buffers fit; an index whose single query does not fit
// is rejected with HM_ERR_CAPACITY.
//
// The array's name is a host-side stand-in, not a device address: its
// address on each device comes from hipGetSymbolAddress, called once per
// device with that device current, and only that address is passed to the
// kernels.  Every call on a GPU shares that GPU's array, which keeps the
// last call's contents (the error word is cleared by a kernel at the start
// of each call).  Calls that overlapped on one GPU would race on it, so the
// entry serialises its calls per device and returns only after its stream
// has drained.
#include <mutex>

#include "index.h"
#include "kernels.h"

using namespace hm;

namespace {

constexpr size_t SCRATCH_BYTES = HM_NOALLOC_SCRATCH_BYTES;
constexpr int MAX_DEVICES = 64;

static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

std::mutex g_setup_mu;
void *g_scratch_ptr[MAX_DEVICES] = {};
std::mutex g_call_mu[MAX_DEVICES];

// The array's address on the current device (looked up once per device).
void *scratch_on(int dev)
{
    if (dev < 0 || dev >= MAX_DEVICES) fail(HM_ERR_INVALID_ARG, "device ordinal out of range");
    std::lock_guard<std::mutex> lk(g_setup_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        HM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) fail(HM_ERR_HIP, "hipGetSymbolAddress returned a null scratch address");
        g_scratch_ptr[dev] = p;
    }
    return g_scratch_ptr[dev];
}

struct CurrentDevice {
    int prev = -1;
    explicit CurrentDevice(int dev)
    {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != dev) HM_HIP(hipSetDevice(dev));
    }
    ~CurrentDevice()
    {
        int cur = -1;
        if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
    }
};

}  // namespace

extern "C" {

HM_API int32_t hm_index_search_device_noalloc(hm_index *h, const float *d_queries, uint64_t nq,
                                              uint32_t k, uint64_t *d_idx, float *d_dist,
                                              uint32_t *d_count, void *stream)
{
    try {
        if (!h) fail(HM_ERR_INVALID_ARG, "index is null");
        std::lock_guard<std::mutex> lk(h->mu);
        CurrentDevice cd(h->device);
        void *p = scratch_on(h->device);
        std::lock_guard<std::mutex> dev_lk(g_call_mu[h->device]);
        h->impl->search_device_scratch(d_queries, nq, k, d_idx, d_dist, d_count,
                                       static_cast<hipStream_t>(stream), p, SCRATCH_BYTES);
        return HM_OK;
    } catch (const Error &e) {
        set_error(e.msg);
        return e.status;
    } catch (const std::exception &e) {
        set_error(e.what());
        return HM_ERR_INVALID_ARG;
    }
}

}  // extern "C"
