// Host-side runtime pieces of the C++ drop-in layer: the library context,
// error handling and an owning device array. Header-only; links against
// libdcpse_hip.so and the HIP runtime.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "dcpse_hip.h"

namespace openfpm_dcpse {

class dcpse_error : public std::runtime_error {
   public:
    using std::runtime_error::runtime_error;
};

inline void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw dcpse_error(std::string(what) + ": " + hipGetErrorString(e));
}

// One library context per host thread and device (the C ABI asks for one
// context per host thread); destroyed at thread exit.
class Context {
   public:
    explicit Context(int device) {
        if (dcpse_ctx_create(&ctx_, device) != DCPSE_OK)
            throw dcpse_error("dcpse_ctx_create failed: no usable GPU (the DC-PSE HIP path has no CPU fallback)");
    }
    ~Context() { dcpse_ctx_destroy(ctx_); }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    dcpse_ctx* get() const { return ctx_; }
    void check(dcpse_status st, const char* what) const {
        if (st != DCPSE_OK) throw dcpse_error(std::string(what) + ": " + dcpse_last_error(ctx_));
    }

   private:
    dcpse_ctx* ctx_ = nullptr;
};

inline Context& context(int device = 0) {
    thread_local std::vector<std::pair<int, Context*>> ctxs;
    for (auto& c : ctxs)
        if (c.first == device) return *c.second;
    thread_local struct Reaper {
        std::vector<std::pair<int, Context*>>* v;
        ~Reaper() {
            for (auto& c : *v) delete c.second;
        }
    } reaper{&ctxs};
    (void)reaper;
    ctxs.emplace_back(device, new Context(device));
    return *ctxs.back().second;
}

// A linear operator on device fields of the particle count:
// out = alpha * D f + beta * out (beta == 0 never reads out).
template <typename P>
class field_operator {
   public:
    virtual ~field_operator() = default;
    virtual void apply(const P* f, P* out, P alpha, P beta) = 0;
};

// Owning device array (hipMalloc'd, copy = deep copy on the device).
template <typename T>
class device_array {
   public:
    device_array() = default;
    explicit device_array(size_t n) { resize(n); }
    ~device_array() { release(); }
    device_array(const device_array& o) { *this = o; }
    device_array& operator=(const device_array& o) {
        if (this != &o) {
            resize(o.n_);
            if (n_) hip_check(hipMemcpy(p_, o.p_, n_ * sizeof(T), hipMemcpyDeviceToDevice), "hipMemcpy D2D");
        }
        return *this;
    }
    device_array(device_array&& o) noexcept : p_(o.p_), n_(o.n_) {
        o.p_ = nullptr;
        o.n_ = 0;
    }
    device_array& operator=(device_array&& o) noexcept {
        std::swap(p_, o.p_);
        std::swap(n_, o.n_);
        return *this;
    }
    void resize(size_t n) {
        if (n == n_) return;
        release();
        if (n) hip_check(hipMalloc(&p_, n * sizeof(T)), "hipMalloc");
        n_ = n;
    }
    void upload(const T* h, size_t n) {
        resize(n);
        if (n) hip_check(hipMemcpy(p_, h, n * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy H2D");
    }
    void upload(const std::vector<T>& h) { upload(h.data(), h.size()); }
    void download(T* h) const {
        if (n_) hip_check(hipMemcpy(h, p_, n_ * sizeof(T), hipMemcpyDeviceToHost), "hipMemcpy D2H");
    }
    std::vector<T> to_host() const {
        std::vector<T> h(n_);
        download(h.data());
        return h;
    }
    T* data() { return p_; }
    const T* data() const { return p_; }
    size_t size() const { return n_; }

   private:
    void release() {
        if (p_) (void)hipFree(p_);
        p_ = nullptr;
        n_ = 0;
    }
    T* p_ = nullptr;
    size_t n_ = 0;
};

}  // namespace openfpm_dcpse
