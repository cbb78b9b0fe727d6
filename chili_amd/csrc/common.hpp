// common.hpp -- error handling and device buffers shared by the C-ABI sources.
#pragma once
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/chili_gpu.h"

struct chili_error : std::runtime_error {
    int code;
    chili_error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void chili_set_error(const std::string& msg);

#define HIP_CHECK(expr)                                                                                   \
    do {                                                                                                  \
        hipError_t e_ = (expr);                                                                           \
        if (e_ != hipSuccess)                                                                             \
            throw chili_error(CHILI_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_));             \
    } while (0)

// Runs f, converting every exception into a status code (no exception ever
// crosses the C boundary).
template <class F>
int guarded(F&& f) {
    try {
        f();
        return CHILI_OK;
    } catch (const chili_error& e) {
        chili_set_error(e.what());
        return e.code;
    } catch (const std::bad_alloc&) {
        chili_set_error("out of memory");
        return CHILI_ENOMEM;
    } catch (const std::invalid_argument& e) {
        chili_set_error(e.what());
        return CHILI_EINVAL;
    } catch (const std::out_of_range& e) {
        chili_set_error(e.what());
        return CHILI_EINVAL;
    } catch (const std::exception& e) {
        chili_set_error(e.what());
        return CHILI_EINVAL;
    } catch (...) {
        chili_set_error("unknown error");
        return CHILI_EINVAL;
    }
}

void select_device(int device);

template <class T>
struct DevBuf {
    T* ptr = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : ptr(o.ptr), n(o.n) { o.ptr = nullptr; o.n = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) {
            if (ptr) (void)hipFree(ptr);
            ptr = o.ptr; n = o.n; o.ptr = nullptr; o.n = 0;
        }
        return *this;
    }
    ~DevBuf() { if (ptr) (void)hipFree(ptr); }
    void resize(size_t cnt) {
        if (cnt <= n && ptr) return;
        if (ptr) { (void)hipFree(ptr); ptr = nullptr; n = 0; }
        if (cnt == 0) return;
        HIP_CHECK(hipMalloc((void**)&ptr, cnt * sizeof(T)));
        n = cnt;
    }
    void upload(const std::vector<T>& v, hipStream_t s) {
        resize(v.size());
        if (!v.empty()) HIP_CHECK(hipMemcpyAsync(ptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void download(T* dst, size_t cnt, hipStream_t s) const {
        if (cnt) HIP_CHECK(hipMemcpyAsync(dst, ptr, cnt * sizeof(T), hipMemcpyDeviceToHost, s));
    }
};
