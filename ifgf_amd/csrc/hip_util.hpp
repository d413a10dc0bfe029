// hip_util.hpp -- host-side helpers shared by the HIP translation units: the error check that
// turns a failed HIP call into an exception (caught at the C ABI), and an owning device buffer
// for the init-time builders.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <stdexcept>
#include <string>
#include <vector>

#define HIP_OK(x)                                                                          \
    do                                                                                     \
    {                                                                                      \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess)                                                              \
            throw std::runtime_error(std::string(#x) + ": " + hipGetErrorString(e_));     \
    } while (0)

namespace ifgf
{
// a device allocation freed with its owner; grows geometrically (at least 256 bytes)
struct DBuf
{
    void* p = nullptr;
    size_t bytes = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { release(); }
    void release()
    {
        if (p)
            (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <class T> T* as() { return static_cast<T*>(p); }
    // at least b bytes; the first `keep` bytes survive a growth
    void reserve(size_t b, size_t keep = 0)
    {
        if (b <= bytes)
            return;
        const size_t nb = std::max<size_t>({b, bytes + bytes / 2, size_t(256)});
        void* q = nullptr;
        HIP_OK(hipMalloc(&q, nb));
        if (p && keep)
            HIP_OK(hipMemcpy(q, p, keep, hipMemcpyDeviceToDevice));
        release();
        p = q;
        bytes = nb;
    }
};

// copy count host values into d (grown as needed)
template <class T> void h2d(DBuf& d, const T* h, size_t count)
{
    d.reserve(count * sizeof(T));
    if (count)
        HIP_OK(hipMemcpy(d.p, h, count * sizeof(T), hipMemcpyHostToDevice));
}
template <class T> void h2d(DBuf& d, const std::vector<T>& h) { h2d(d, h.data(), h.size()); }
} // namespace ifgf
