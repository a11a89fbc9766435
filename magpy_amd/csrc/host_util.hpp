// host_util.hpp -- host-side helpers shared by the translation units of
// libmagpy_amd.so: constants, HIP error checks, owned device buffers, the
// launch stream and the exception-to-status wrapper of the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "simulation.hpp"

namespace mp_host {

// include/constants.hpp:10-12
constexpr double kKB = 1.38064852e-23;
constexpr double kMU0 = 1.25663706e-6;
constexpr double kGYROMAG = 1.76086e11;
constexpr double kPiHost = 3.14159265358979323846;

// Text of the last error of this thread (mp_last_error).
std::string &last_error();

#define MP_HIP(call)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess)                                                               \
            throw simulation::Error(MP_ERR_HIP, std::string(#call) + ": " +                 \
                                                    hipGetErrorString(e_));                 \
    } while (0)

inline void require(bool ok, const std::string &msg)
{
    if (!ok) throw simulation::Error(MP_ERR_ARG, msg);
}

// Bump allocator over device memory the library does not own.  While an
// ArenaScope is open on a thread, DevBuf takes its memory from that arena
// instead of hipMalloc and never frees it: the arena's whole span is reused
// by the next scope.  A request beyond the span throws MP_ERR_ALLOC before
// anything is written.
struct Arena {
    char *base = nullptr;
    size_t cap = 0, used = 0;
    const char *who = "";
};

inline Arena *&current_arena()
{
    static thread_local Arena *a = nullptr;
    return a;
}

struct ArenaScope {
    Arena *prev;
    explicit ArenaScope(Arena *a) : prev(current_arena()) { current_arena() = a; }
    ~ArenaScope() { current_arena() = prev; }
    ArenaScope(const ArenaScope &) = delete;
    ArenaScope &operator=(const ArenaScope &) = delete;
};

// Device allocation owned by its holder (or carved from the thread's arena).
template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    bool owned = false;
    DevBuf() = default;
    explicit DevBuf(size_t count) { alloc(count); }
    void alloc(size_t count)
    {
        release();
        if (count == 0) return;
        if (Arena *ar = current_arena()) {
            const size_t off = (ar->used + 255) & ~(size_t)255;
            if (off > ar->cap || count > (ar->cap - off) / sizeof(T))
                throw simulation::Error(
                    MP_ERR_ALLOC, std::string(ar->who) + ": the run needs more device memory than "
                                  "the library's fixed scratch of " + std::to_string(ar->cap) +
                                  " bytes (" + std::to_string(off) + " bytes already placed, " +
                                  std::to_string(count) + " more elements of " +
                                  std::to_string(sizeof(T)) + " bytes asked for)");
            p = reinterpret_cast<T *>(ar->base + off);
            n = count;
            ar->used = off + count * sizeof(T);
            return;
        }
        hipError_t e = hipMalloc(&p, count * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            throw simulation::Error(MP_ERR_ALLOC, "hipMalloc of " +
                                                      std::to_string(count * sizeof(T)) +
                                                      " bytes failed");
        }
        n = count;
        owned = true;
    }
    void upload(const std::vector<T> &h)
    {
        alloc(h.size());
        if (!h.empty()) MP_HIP(hipMemcpy(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    }
    void release()
    {
        if (p && owned) (void)hipFree(p);
        p = nullptr;
        n = 0;
        owned = false;
    }
    ~DevBuf() { release(); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
};

// The launch stream and its timing events; without `own`, the null stream
// and no events (nothing is created).
struct StreamGuard {
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    explicit StreamGuard(bool own = true)
    {
        if (!own) return;
        MP_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        MP_HIP(hipEventCreate(&e0));
        MP_HIP(hipEventCreate(&e1));
    }
    ~StreamGuard()
    {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (s) (void)hipStreamDestroy(s);
    }
};


// Run f, turning exceptions into MP_* status codes and last_error().
template <typename F>
int guarded(F &&f)
{
    try {
        f();
        return MP_OK;
    } catch (const simulation::Error &e) {
        last_error() = e.what();
        return e.code;
    } catch (const std::bad_alloc &) {
        last_error() = "host allocation failed";
        return MP_ERR_ALLOC;
    } catch (const std::exception &e) {
        last_error() = e.what();
        return MP_ERR_ARG;
    }
}

}  // namespace mp_host
