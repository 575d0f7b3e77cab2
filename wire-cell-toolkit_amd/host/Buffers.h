// Device and pinned host allocations owned by a component, freed on
// release or destruction; ensure() grows them (contents not kept).
#pragma once
#include <cstddef>
#include <stdexcept>
#include <string>

#include "../../include/wct_hip.h"

namespace WireCell {
namespace Hip {

struct DeviceError : public std::runtime_error {
    using std::runtime_error::runtime_error;
};

inline void device_check(int rc, const std::string& what)
{
    if (rc != WCT_OK) throw DeviceError(what + ": " + wct_last_error());
}

struct DevBuf {
    void* p{nullptr};
    size_t bytes{0};
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void ensure(size_t n)
    {
        if (n <= bytes && p) return;
        release();
        device_check(wct_malloc(&p, n), "device allocation");
        bytes = n;
    }
    void release()
    {
        if (p) wct_free(p);
        p = nullptr;
        bytes = 0;
    }
};

struct HostBuf {
    void* p{nullptr};
    size_t bytes{0};
    HostBuf() = default;
    HostBuf(const HostBuf&) = delete;
    HostBuf& operator=(const HostBuf&) = delete;
    ~HostBuf() { release(); }
    void ensure(size_t n)
    {
        if (n <= bytes && p) return;
        release();
        device_check(wct_host_malloc(&p, n), "pinned allocation");
        bytes = n;
    }
    void release()
    {
        if (p) wct_host_free(p);
        p = nullptr;
        bytes = 0;
    }
};

}  // namespace Hip
}  // namespace WireCell
