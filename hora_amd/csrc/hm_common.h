// hm_common.h -- shared host/device definitions of the MI355X runtime.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

#include "../../include/hora_mi355x.h"

namespace hm {

// Thread-local last error message (hm_last_error()).
void set_error(const std::string &msg);
const char *last_error();

struct Error {
    int32_t status;
    std::string msg;
};

[[noreturn]] void fail(int32_t status, const std::string &msg);

#define HM_HIP(expr)                                                                   \
    do {                                                                               \
        hipError_t _e = (expr);                                                        \
        if (_e != hipSuccess)                                                          \
            ::hm::fail(HM_ERR_HIP, std::string(#expr " failed: ") + hipGetErrorString(_e)); \
    } while (0)

#define HM_CHECK_LAUNCH() HM_HIP(hipGetLastError())

// Owning device allocation (hipMalloc'd), grown on demand.
class DeviceBuffer {
  public:
    DeviceBuffer() = default;
    ~DeviceBuffer();
    DeviceBuffer(const DeviceBuffer &) = delete;
    DeviceBuffer &operator=(const DeviceBuffer &) = delete;
    DeviceBuffer(DeviceBuffer &&o) noexcept : ptr_(o.ptr_), bytes_(o.bytes_) { o.ptr_ = nullptr; o.bytes_ = 0; }

    // Ensure at least `bytes` are allocated; contents are discarded when grown.
    void reserve(size_t bytes);
    // Grow to at least `bytes`, preserving the first `keep` bytes.
    void grow_keep(size_t bytes, size_t keep, hipStream_t s);
    void release();
    template <class T> T *as() const { return static_cast<T *>(ptr_); }
    void *get() const { return ptr_; }
    size_t bytes() const { return bytes_; }

  private:
    void *ptr_ = nullptr;
    size_t bytes_ = 0;
};

// Owning pinned host allocation (hipHostMalloc), grown on demand: small
// device-to-host status copies go straight to it by DMA.
class PinnedBuffer {
  public:
    PinnedBuffer() = default;
    ~PinnedBuffer();
    PinnedBuffer(const PinnedBuffer &) = delete;
    PinnedBuffer &operator=(const PinnedBuffer &) = delete;
    // Ensure at least `bytes`; contents are discarded when grown.
    void reserve(size_t bytes);
    template <class T> T *as() const { return static_cast<T *>(ptr_); }

  private:
    void *ptr_ = nullptr;
    size_t bytes_ = 0;
};

inline size_t round_up(size_t x, size_t m) { return (x + m - 1) / m * m; }
inline unsigned ceil_div(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

}  // namespace hm
