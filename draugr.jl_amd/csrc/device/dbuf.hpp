// Device buffers (HBM-resident, owned) and HIP error checking.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "../host/types.hpp"

#define DRAUGR_HIP_CHECK(expr)                                                                        \
    do {                                                                                              \
        hipError_t e_ = (expr);                                                                       \
        if (e_ != hipSuccess)                                                                         \
            throw ::draugr::Error(std::string("HIP error in ") + #expr + ": " + hipGetErrorString(e_)); \
    } while (0)

namespace draugr {

template <typename T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    DBuf(DBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DBuf& operator=(DBuf&& o) noexcept {
        if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    ~DBuf() { release(); }
    // keeps the allocation when the size is unchanged: captured graphs hold
    // these addresses, so a same-size re-upload must not move the buffer
    void alloc(size_t count) {
        if (p && count == n) return;
        release();
        n = count;
        if (count) DRAUGR_HIP_CHECK(hipMalloc((void**)&p, count * sizeof(T)));
    }
    void upload(const T* src, size_t count, hipStream_t st) {
        alloc(count);
        if (count) DRAUGR_HIP_CHECK(hipMemcpyAsync(p, src, count * sizeof(T), hipMemcpyHostToDevice, st));
    }
    void upload(const std::vector<T>& v, hipStream_t st) { upload(v.data(), v.size(), st); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    size_t bytes() const { return n * sizeof(T); }
};

}  // namespace draugr
