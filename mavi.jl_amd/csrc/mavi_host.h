// mavi_host.h -- host-side helpers shared by the C entry points of single
// systems (mavi_capi.hip) and replica batches (mavi_batch.hip).
#pragma once
#include <string>

#include "mavi_internal.h"

inline int set_err(mavi_ctx *ctx, int code, const std::string &msg) {
    if (ctx) ctx->err = msg;
    return code;
}

#define HIPCHK(ctx, expr)                                                                                  \
    do {                                                                                                   \
        hipError_t e_ = (expr);                                                                            \
        if (e_ != hipSuccess)                                                                              \
            return set_err((ctx), MAVI_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));           \
    } while (0)

// Allocate and zero on the context's stream: the library's stream is
// non-blocking, so a zero fill on the null stream would not be ordered
// before the stream's own work on the buffer.
template <typename T>
hipError_t dalloc(T **p, int64_t count, hipStream_t st) {
    *p = nullptr;
    if (count <= 0) count = 1;
    hipError_t e = hipMalloc((void **)p, sizeof(T) * (size_t)count);
    if (e == hipSuccess) e = hipMemsetAsync(*p, 0, sizeof(T) * (size_t)count, st);
    return e;
}

template <typename T>
void dfree(T *&p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

inline int nblk(int64_t n, int t) { return (int)((n + t - 1) / t); }

inline bool is_pow2_wg(int32_t w) { return w == 64 || w == 128 || w == 256 || w == 512 || w == 1024; }

// Kernel timing of one launch class with HIP events (mavi_prof_enable): the
// launches inside the scope are bracketed on the stream they run on.
struct ProfScope {
    mavi_sys *s;
    int which;
    hipStream_t st;
    ProfScope(mavi_sys *s_, int w, hipStream_t st_) : s(s_), which(w), st(st_) {
        if (s->prof) (void)hipEventRecord(s->ev0, st);
    }
    ~ProfScope() {
        if (!s->prof) return;
        (void)hipEventRecord(s->ev1, st);
        if (hipEventSynchronize(s->ev1) == hipSuccess) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, s->ev0, s->ev1) == hipSuccess) {
                s->profs[which].ms += ms;
                s->profs[which].launches += 1;
            }
        }
    }
};
