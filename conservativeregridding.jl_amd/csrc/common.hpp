// common.hpp -- shared definitions of the HIP library (host + device).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/crhip.h"

#define CR_HD __host__ __device__ __forceinline__

namespace cr {

// ------------------------------------------------------------- errors --
void set_error(const std::string &msg);

struct Error {
    cr_status status;
    std::string msg;
};

#define CR_HIP(expr)                                                                     \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            throw ::cr::Error{e_ == hipErrorOutOfMemory ? CR_OUT_OF_MEMORY : CR_HIP_ERROR, \
                              std::string(#expr) + ": " + hipGetErrorString(e_)};        \
    } while (0)

#define CR_REQUIRE(cond, status, msg)                               \
    do {                                                            \
        if (!(cond)) throw ::cr::Error{(status), std::string(msg)}; \
    } while (0)

// Run a body and turn exceptions into a status + thread-local message.
template <class F>
cr_status guard(F &&f) {
    try {
        f();
        return CR_OK;
    } catch (const Error &e) {
        set_error(e.msg);
        return e.status;
    } catch (const std::bad_alloc &) {
        set_error("host allocation failed");
        return CR_OUT_OF_MEMORY;
    } catch (const std::exception &e) {
        set_error(e.what());
        return CR_INTERNAL;
    }
}

// True while `s` is being captured into a HIP graph.  Lazy per-regridder set-up
// (tile plans, long-row lists, the default variant) allocates, copies to the
// host and synchronises: none of that may run inside a capture.
inline bool stream_capturing(hipStream_t s) {
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &st) != hipSuccess) {
        (void)hipGetLastError();  // e.g. the legacy null stream during a global capture
        return true;
    }
    return st != hipStreamCaptureStatusNone;
}

#define CR_REQUIRE_NOT_CAPTURING(s, what)                                                                \
    CR_REQUIRE(!::cr::stream_capturing(s), CR_UNSUPPORTED,                                               \
               std::string(what) + " is not made yet and cannot be made during stream capture: call "   \
                                   "cr_regrid_prepare (or one cr_regrid) with the same arguments first")

// Device memory.  DevBuf(n) uses hipMalloc (long-lived buffers owned by a
// handle); DevBuf(n, stream) takes stream-ordered memory from the device's
// default pool (build temporaries), which keeps freed blocks for the next
// build instead of returning them to the driver.
void configure_pool();

template <class T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    hipStream_t st = nullptr;
    bool pooled = false;
    DevBuf() = default;
    explicit DevBuf(size_t count) { alloc(count); }
    DevBuf(size_t count, hipStream_t s) { alloc_async(count, s); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), n(o.n), st(o.st), pooled(o.pooled) {
        o.p = nullptr;
        o.n = 0;
    }
    DevBuf &operator=(DevBuf &&o) noexcept {
        if (this != &o) {
            release();
            p = o.p; n = o.n; st = o.st; pooled = o.pooled;
            o.p = nullptr; o.n = 0;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(size_t count) {
        release();
        n = count;
        pooled = false;
        if (count) CR_HIP(hipMalloc(&p, count * sizeof(T)));
    }
    void alloc_async(size_t count, hipStream_t s) {
        release();
        n = count;
        st = s;
        pooled = true;
        configure_pool();
        if (count) CR_HIP(hipMallocAsync((void **)&p, count * sizeof(T), s));
    }
    // Pool memory that outlives the call that allocated it (owned by a handle):
    // allocated in order on `s`, released on the null stream.
    void alloc_kept(size_t count, hipStream_t s) {
        alloc_async(count, s);
        st = nullptr;
    }
    void release() {
        if (p) {
            if (pooled) (void)hipFreeAsync(p, st);
            else (void)hipFree(p);
        }
        p = nullptr;
        n = 0;
    }
    size_t bytes() const { return n * sizeof(T); }
};

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

inline void check_device() {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) throw Error{CR_NO_DEVICE, "no HIP device is visible"};
}

// Widest cell polygon a grid may hold (vertices per cell).
constexpr int kMaxCellVertices = 64;

inline unsigned grid_for(int64_t n, int block) {
    int64_t g = (n + block - 1) / block;
    return (unsigned)(g < 1 ? 1 : g);
}

}  // namespace cr

// ------------------------------------------------------------- handles --
struct cr_grid {
    int64_t ncells = 0;
    int32_t nv = 0;
    int32_t convex = 0;
    int32_t flipped = 0;
    int64_t nverts = 0;          // caller-visible vertices (points of a point grid)
    cr::DevBuf<double> cells;    // ncells * nv * 3, xyz interleaved per vertex
    cr::DevBuf<int32_t> vidx;    // ncells * nv: caller vertex of each cell slot
    int32_t private_vertices = 0;  // 1: no vertex index is shared by two cells (tables of cells)
    cr::DevBuf<int32_t> fold;    // 2 * nfold: (primary, partner) cells of a folded row
    int64_t nfold = 0;
    // Planar manifold: the caller's vertices (nverts * 2, x and y); `cells`
    // then holds their gnomonic images on the unit sphere, which only the
    // candidate search reads (straight edges map to great-circle arcs)
    int32_t planar = 0;
    cr::DevBuf<double> xy;
};
