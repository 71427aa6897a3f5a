// Owning device allocations and HIP error checking for host code.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdint>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "capi_util.h"

namespace nsparse_amd {

inline void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw device_error(std::string(what) + ": " + hipGetErrorString(e));
}

// Raw device bytes. `pad` extra zeroed bytes let vector loads run past the end
// of an array without leaving the allocation.
class DeviceBuffer {
public:
    DeviceBuffer() = default;
    explicit DeviceBuffer(size_t bytes, size_t pad = 16) { allocate(bytes, pad); }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    DeviceBuffer(DeviceBuffer&& o) noexcept : ptr_(std::exchange(o.ptr_, nullptr)), bytes_(std::exchange(o.bytes_, 0)) {}
    DeviceBuffer& operator=(DeviceBuffer&& o) noexcept {
        if (this != &o) {
            release();
            ptr_ = std::exchange(o.ptr_, nullptr);
            bytes_ = std::exchange(o.bytes_, 0);
        }
        return *this;
    }
    ~DeviceBuffer() { release(); }

    void allocate(size_t bytes, size_t pad = 16) {
        release();
        const size_t total = bytes + pad;
        hip_check(hipMalloc(&ptr_, total == 0 ? 16 : total), "hipMalloc");
        bytes_ = bytes;
        if (pad > 0) hip_check(hipMemset(static_cast<char*>(ptr_) + bytes, 0, pad), "hipMemset");
    }
    void release() noexcept {
        if (ptr_ != nullptr) (void)hipFree(ptr_);
        ptr_ = nullptr;
        bytes_ = 0;
    }
    template <class T = void>
    T* get() const {
        return static_cast<T*>(ptr_);
    }
    size_t bytes() const { return bytes_; }
    bool empty() const { return ptr_ == nullptr; }

    void upload(const void* src, size_t bytes, size_t pad = 16) {
        allocate(bytes, pad);
        if (bytes > 0) hip_check(hipMemcpy(ptr_, src, bytes, hipMemcpyHostToDevice), "hipMemcpy H2D");
    }
    template <class T>
    void upload(const std::vector<T>& v, size_t pad = 16) {
        upload(v.data(), v.size() * sizeof(T), pad);
    }
    void download(void* dst, size_t bytes, size_t offset = 0) const {
        if (bytes > 0)
            hip_check(hipMemcpy(dst, static_cast<const char*>(ptr_) + offset, bytes, hipMemcpyDeviceToHost),
                      "hipMemcpy D2H");
    }
    template <class T>
    std::vector<T> to_host(size_t count) const {
        std::vector<T> v(count);
        download(v.data(), count * sizeof(T));
        return v;
    }

private:
    void* ptr_ = nullptr;
    size_t bytes_ = 0;
};

// Device scratch for launches that return before their kernels finish: a
// buffer is handed out again only once the event recorded after its last
// launch has completed, so callers on different streams never share one in
// flight. Buffers are kept until the pool is destroyed.
class ScratchPool {
public:
    ScratchPool() = default;
    ScratchPool(const ScratchPool&) = delete;
    ScratchPool& operator=(const ScratchPool&) = delete;
    ~ScratchPool() {
        for (auto& s : slots_)
            if (s.done != nullptr) (void)hipEventDestroy(s.done);
    }
    // A free buffer of at least `bytes`; its index goes to `slot`.
    void* acquire(size_t bytes, size_t& slot);
    // Records, on `stream`, the point after which buffer `slot` is free again.
    void release(size_t slot, hipStream_t stream);

private:
    struct Slot {
        DeviceBuffer buf;
        hipEvent_t done = nullptr;
        bool busy = false;
    };
    std::mutex mu_;
    std::vector<Slot> slots_;
};

// Makes sure a HIP device exists and is usable; throws device_error otherwise.
void require_device();

}  // namespace nsparse_amd
