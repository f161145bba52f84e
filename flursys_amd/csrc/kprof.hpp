// Per-kernel device timing with HIP events on the solver's stream.
//
// A timed launch is bracketed by an event pair.  Each record carries the
// solver iteration that issued it: the solvers launch iterations in batches
// ahead of the device-side stop test, and launches after the test has fired
// exit at once, so when a batch is read back the records of iterations at
// or beyond the number actually run are dropped.  Records of work outside
// an iteration loop carry the tag -1 and always count.  Per kernel kind the
// profiler accumulates launches, device milliseconds and the algorithmic
// HBM bytes those launches moved.
#pragma once

#include <hip/hip_runtime.h>

#include <climits>
#include <cstring>
#include <vector>

namespace fl {

struct KProf {
    struct Rec {
        int kind;
        long long iter;
        double bytes;
    };
    bool on = false;
    int only = -1;  // time only this kind; -1 times every kind
    int kinds = 0;
    const char *const *names = nullptr;
    std::vector<hipEvent_t> pool;  // pair i = pool[2i], pool[2i + 1]
    std::vector<Rec> recs;         // pending records (events not yet read)
    std::vector<double> ms, bytes;
    std::vector<long long> launches;

    void init(const char *const *kernel_names, int count) {
        names = kernel_names;
        kinds = count;
        clear();
    }
    // kind of a kernel name, -1 when unknown
    int find(const char *name) const {
        for (int k = 0; k < kinds; ++k)
            if (std::strcmp(names[k], name) == 0) return k;
        return -1;
    }
    void clear() {
        recs.clear();
        ms.assign(kinds, 0.0);
        bytes.assign(kinds, 0.0);
        launches.assign(kinds, 0);
    }
    bool want(int kind) const { return on && (only < 0 || only == kind); }
    hipError_t start(hipStream_t s, int kind, long long iter, double b) {
        const size_t i = recs.size();
        while (pool.size() < 2 * i + 2) {
            hipEvent_t e;
            const hipError_t err = hipEventCreate(&e);
            if (err != hipSuccess) return err;
            pool.push_back(e);
        }
        recs.push_back(Rec{kind, iter, b});
        return hipEventRecord(pool[2 * i], s);
    }
    hipError_t stop(hipStream_t s) { return hipEventRecord(pool[2 * recs.size() - 1], s); }
    // Call after the stream has drained.
    hipError_t flush(long long iter_limit = LLONG_MAX) {
        for (size_t i = 0; i < recs.size(); ++i) {
            const Rec &r = recs[i];
            if (r.iter >= iter_limit) continue;
            float t = 0.0f;
            const hipError_t err = hipEventElapsedTime(&t, pool[2 * i], pool[2 * i + 1]);
            if (err != hipSuccess) {
                recs.clear();
                return err;
            }
            ms[r.kind] += t;
            bytes[r.kind] += r.bytes;
            launches[r.kind] += 1;
        }
        recs.clear();
        return hipSuccess;
    }
    void destroy() {
        for (hipEvent_t e : pool) (void)hipEventDestroy(e);
        pool.clear();
        recs.clear();
    }
};

// Launch `...` (a kernel launch statement) bracketed by the profiler's
// events when its kind is being timed.  S is a solver with members `prof`
// and `ctx`; the enclosing function returns fl_status.
#define FL_TIMED(S, KIND, ITER, BYTES, ...)                                                    \
    do {                                                                                       \
        const bool tm_ = (S)->prof.want(KIND);                                                 \
        if (tm_)                                                                               \
            FL_HIP_TRY((S)->ctx, (S)->prof.start((S)->ctx->stream, (KIND), (ITER),             \
                                                 (double)(BYTES)));                            \
        __VA_ARGS__;                                                                           \
        if (tm_) FL_HIP_TRY((S)->ctx, (S)->prof.stop((S)->ctx->stream));                       \
    } while (0)

}  // namespace fl
