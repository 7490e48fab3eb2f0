// Host helpers shared by the C ABI translation units: error capture, device
// allocation and launch sizing.
#pragma once

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/gpusim2grid_amd.h"

namespace g2g {

// message of the last failed call on this thread (g2g_last_error)
inline std::string& last_error()
{
    static thread_local std::string s;
    return s;
}

struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define HIPCHK(call)                                                                                    \
    do {                                                                                                \
        hipError_t e_ = (call);                                                                         \
        if (e_ != hipSuccess) throw ::g2g::HipError(std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

// runs f, mapping exceptions to the ABI's status codes
template <typename F>
int guarded(F&& f)
{
    try {
        f();
        return G2G_OK;
    } catch (const std::invalid_argument& e) {
        last_error() = e.what();
        return G2G_ERR_INVALID;
    } catch (const HipError& e) {
        last_error() = e.what();
        return G2G_ERR_HIP;
    } catch (const std::exception& e) {
        last_error() = e.what();
        return G2G_ERR_RUNTIME;
    }
}

template <typename T>
T* dalloc(size_t count)
{
    void* p = nullptr;
    if (count == 0) count = 1;
    HIPCHK(hipMalloc(&p, count * sizeof(T)));
    return static_cast<T*>(p);
}

template <typename T>
T* dupload(const std::vector<T>& v)
{
    T* p = dalloc<T>(v.size());
    if (!v.empty()) HIPCHK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return p;
}

inline void dfree(void* p)
{
    if (p) (void)hipFree(p);
}

inline unsigned blocks_for(long long n, int bs) { return static_cast<unsigned>((n + bs - 1) / bs); }

// device buffers freed together (at scope exit or on release())
struct DeviceScratch {
    std::vector<void*> p;
    DeviceScratch() = default;
    DeviceScratch(const DeviceScratch&) = delete;
    DeviceScratch& operator=(const DeviceScratch&) = delete;
    ~DeviceScratch() { release(); }
    void release()
    {
        for (void* q : p) dfree(q);
        p.clear();
    }
    template <typename T>
    T* get(size_t n)
    {
        T* q = dalloc<T>(n);
        p.push_back(q);
        return q;
    }
    template <typename T>
    static void free_one(T*& q)
    {
        dfree(q);
        q = nullptr;
    }
};

}  // namespace g2g
