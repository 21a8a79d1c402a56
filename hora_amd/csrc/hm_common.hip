// hm_common.hip -- error state and device buffers of the runtime.
#include "hm_common.h"

#include <algorithm>

namespace hm {

namespace {
thread_local std::string g_last_error;
}

void set_error(const std::string &msg) { g_last_error = msg; }
const char *last_error() { return g_last_error.c_str(); }

void fail(int32_t status, const std::string &msg) { throw Error{status, msg}; }

DeviceBuffer::~DeviceBuffer() { release(); }

void DeviceBuffer::release()
{
    if (ptr_) (void)hipFree(ptr_);
    ptr_ = nullptr;
    bytes_ = 0;
}

PinnedBuffer::~PinnedBuffer()
{
    if (ptr_) (void)hipHostFree(ptr_);
}

void PinnedBuffer::reserve(size_t bytes)
{
    if (bytes <= bytes_) return;
    if (ptr_) (void)hipHostFree(ptr_);
    ptr_ = nullptr;
    bytes_ = 0;
    const size_t want = std::max<size_t>(bytes, 256);
    void *p = nullptr;
    hipError_t e = hipHostMalloc(&p, want, hipHostMallocDefault);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        fail(HM_ERR_OUT_OF_MEMORY, "hipHostMalloc of " + std::to_string(want) +
                                       " bytes failed: " + hipGetErrorString(e));
    }
    ptr_ = p;
    bytes_ = want;
}

void DeviceBuffer::reserve(size_t bytes)
{
    if (bytes <= bytes_) return;
    release();
    size_t want = std::max<size_t>(bytes, 256);
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        fail(HM_ERR_OUT_OF_MEMORY,
             "hipMalloc of " + std::to_string(want) + " bytes failed: " + hipGetErrorString(e));
    }
    ptr_ = p;
    bytes_ = want;
}

void DeviceBuffer::grow_keep(size_t bytes, size_t keep, hipStream_t s)
{
    if (bytes <= bytes_) return;
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        fail(HM_ERR_OUT_OF_MEMORY,
             "hipMalloc of " + std::to_string(bytes) + " bytes failed: " + hipGetErrorString(e));
    }
    if (ptr_ && keep) {
        HM_HIP(hipMemcpyAsync(p, ptr_, std::min(keep, bytes_), hipMemcpyDeviceToDevice, s));
        HM_HIP(hipStreamSynchronize(s));
    }
    release();
    ptr_ = p;
    bytes_ = bytes;
}

}  // namespace hm
