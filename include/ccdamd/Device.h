// Process-wide device context and device buffers for the C++ interface (over include/ccd_amd.h).
#pragma once

#include <cstdint>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../ccd_amd.h"

namespace ccdamd {

// The reference's functions have no error channel (void / bool results), so failures of the device
// path surface as std::runtime_error.
inline void check(ccd_ctx *ctx, int rc, const char *what)
{
    if (rc != CCD_OK)
        throw std::runtime_error(std::string(what) + ": " + (ctx ? ccd_last_error(ctx) : "no device context") +
                                 " (status " + std::to_string(rc) + ")");
}

class Device {
public:
    static Device &get()
    {
        static Device d;
        return d;
    }
    ccd_ctx *ctx() { return ctx_; }
    std::mutex &mutex() { return mu_; }

private:
    Device()
    {
        int rc = ccd_create(&ctx_, 0);
        if (rc != CCD_OK) throw std::runtime_error("ccd_create failed: no usable HIP device");
    }
    // The process-wide context lives until exit: destroying it from a static destructor could run after
    // the HIP runtime has shut down.
    ~Device() {}
    ccd_ctx *ctx_ = nullptr;
    std::mutex mu_;
};

// Owning device allocation filled from / read back to host vectors.
class DeviceBuffer {
public:
    DeviceBuffer() {}
    explicit DeviceBuffer(int64_t bytes) { resize(bytes); }
    template <class T>
    explicit DeviceBuffer(const std::vector<T> &host)
    {
        upload(host.data(), (int64_t)(host.size() * sizeof(T)));
    }
    DeviceBuffer(const void *host, int64_t bytes) { upload(host, bytes); }
    ~DeviceBuffer()
    {
        if (ptr_) ccd_device_free(Device::get().ctx(), ptr_);
    }
    DeviceBuffer(const DeviceBuffer &) = delete;
    DeviceBuffer &operator=(const DeviceBuffer &) = delete;
    void resize(int64_t bytes)
    {
        if (bytes <= bytes_) return;
        ccd_ctx *c = Device::get().ctx();
        if (ptr_) ccd_device_free(c, ptr_);
        ptr_ = nullptr;
        bytes_ = 0;
        check(c, ccd_device_alloc(c, bytes, &ptr_), "device allocation");
        bytes_ = bytes;
    }
    void upload(const void *host, int64_t bytes)
    {
        resize(bytes);
        ccd_ctx *c = Device::get().ctx();
        check(c, ccd_copy_to_device(c, ptr_, host, bytes), "copy to device");
    }
    void download(void *host, int64_t bytes) const
    {
        ccd_ctx *c = Device::get().ctx();
        check(c, ccd_copy_to_host(c, host, ptr_, bytes), "copy to host");
    }
    template <class T>
    T *as() const
    {
        return (T *)ptr_;
    }

private:
    void *ptr_ = nullptr;
    int64_t bytes_ = 0;
};

}   // namespace ccdamd
