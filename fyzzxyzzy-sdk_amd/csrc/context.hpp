// Library context: device, stream and the per-context error message that
// the C-ABI reports through fyzz_hip_last_error().
#pragma once
#include <exception>
#include <map>
#include <vector>
#include <new>
#include <string>

#include "common.hpp"

namespace fyzz_hip {
// Optional per-stage timing with HIP events recorded on the context's
// stream around named launches; read back (and reset) on demand.
struct Profiler {
    struct Span {
        hipEvent_t a, b;
    };
    bool enabled = false;
    std::map<std::string, std::vector<Span>> spans;
    std::vector<hipEvent_t> pool;
    hipEvent_t take() {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e;
        FYZZ_HIP_CHECK(hipEventCreate(&e));
        return e;
    }
    // Returns the span index to close with end(), or -1 when disabled.
    long begin(const std::string& name, hipStream_t st) {
        if (!enabled) return -1;
        Span s{take(), take()};
        FYZZ_HIP_CHECK(hipEventRecord(s.a, st));
        auto& v = spans[name];
        v.push_back(s);
        return static_cast<long>(v.size() - 1);
    }
    void end(const std::string& name, long idx, hipStream_t st) {
        if (idx < 0) return;
        FYZZ_HIP_CHECK(hipEventRecord(spans[name][static_cast<std::size_t>(idx)].b, st));
    }
    // Sum of elapsed ms and span count for `name`; releases its events.
    void read(const std::string& name, double* total_ms, std::uint64_t* count) {
        double t = 0.0;
        std::uint64_t c = 0;
        auto it = spans.find(name);
        if (it != spans.end()) {
            for (auto& s : it->second) {
                FYZZ_HIP_CHECK(hipEventSynchronize(s.b));
                float ms = 0.f;
                FYZZ_HIP_CHECK(hipEventElapsedTime(&ms, s.a, s.b));
                t += ms;
                ++c;
                pool.push_back(s.a);
                pool.push_back(s.b);
            }
            spans.erase(it);
        }
        *total_ms = t;
        *count = c;
    }
    ~Profiler() {
        for (auto& kv : spans)
            for (auto& s : kv.second) {
                (void)hipEventDestroy(s.a);
                (void)hipEventDestroy(s.b);
            }
        for (auto e : pool) (void)hipEventDestroy(e);
    }
};
struct ProfScope {
    Profiler* p;
    std::string name;
    long idx;
    hipStream_t st;
    ProfScope(Profiler* prof, const char* n, hipStream_t s) : p(prof), name(n), idx(-1), st(s) {
        if (p && p->enabled) idx = p->begin(name, st);
    }
    ~ProfScope() {
        try {
            if (p && idx >= 0) p->end(name, idx, st);
        } catch (...) {
        }
    }
};
}  // namespace fyzz_hip

struct fyzz_hip_context {
    int device = 0;
    hipStream_t stream = nullptr;  // null = the device's default stream
    bool cooperative = true;       // grid-barrier kernels: cooperative or plain launches
    std::string last_error;
    fyzz_hip::Profiler prof;
};

namespace fyzz_hip {

// Runs f, converting exceptions into a status code and the context's
// last-error string; no exception crosses the C-ABI.
template <class F>
int guarded(fyzz_hip_context* ctx, F&& f) {
    if (!ctx) return kInvalidArgument;
    try {
        FYZZ_HIP_CHECK(hipSetDevice(ctx->device));
        f();
        ctx->last_error.clear();
        return kOk;
    } catch (const Error& e) {
        ctx->last_error = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        ctx->last_error = "host allocation failed";
        return kOutOfMemory;
    } catch (const std::exception& e) {
        ctx->last_error = e.what();
        return kLogicError;
    } catch (...) {
        ctx->last_error = "unknown failure";
        return kLogicError;
    }
}

}  // namespace fyzz_hip
