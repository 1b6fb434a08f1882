// common.hpp -- shared host/device helpers of libcoolpdlp_hip.so (gfx950).
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <map>
#include <mutex>
#include <vector>

#include "../../include/coolpdlp_hip.h"

namespace cpdlp {

// ---------------------------------------------------------------- errors
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const char* msg);

#define HIP_CHECK(expr)                                                                  \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            throw ::cpdlp::Error(CPDLP_ERR_HIP, std::string(#expr) + ": " +              \
                                                    hipGetErrorString(e_) + " at " +     \
                                                    __FILE__ + ":" + std::to_string(__LINE__)); \
    } while (0)

// catch-all wrapper turning exceptions into status codes + thread-local message
template <class F>
int guarded(F&& f) {
    try {
        f();
        set_last_error("");
        return CPDLP_OK;
    } catch (const Error& e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::bad_alloc&) {
        set_last_error("host allocation failed");
        return CPDLP_ERR_MEMORY;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return CPDLP_ERR_ARGUMENT;
    }
}

// ---------------------------------------------------------------- device buffers
// Process-wide cache of device blocks.  hipFree synchronises the whole device, which
// would stall every other stream (a served request's teardown would wait for all
// running solves), so released blocks go to a per-device free list and are handed out
// again; hipFree runs only to make room when hipMalloc fails.  Sizes are rounded to
// powers of two below 1 MiB and to 2 MiB multiples above.  Every buffer here is
// released only after the work that uses it has completed (its stream was
// synchronised), so a reused block is never still in use.
struct BlockCache {
    std::mutex mu;
    std::map<int, std::multimap<size_t, void*>> free_;   // device -> size -> block
    std::unordered_map<void*, std::pair<int, size_t>> live;   // block -> (device, size)
};
inline BlockCache& block_cache() {
    static BlockCache* c = new BlockCache();   // never destroyed: used until exit
    return *c;
}
inline size_t block_round(size_t b) {
    if (b <= 256) return 256;
    if (b < (size_t(1) << 20)) {
        size_t r = 256;
        while (r < b) r <<= 1;
        return r;
    }
    const size_t q = size_t(2) << 20;
    return (b + q - 1) / q * q;
}
inline void block_flush(int dev) {   // caller holds the lock
    BlockCache& c = block_cache();
    auto& fl = c.free_[dev];
    for (auto& kv : fl) (void)hipFree(kv.second);
    fl.clear();
}
inline void* dev_alloc(size_t bytes) {
    int dev = 0;
    HIP_CHECK(hipGetDevice(&dev));
    const size_t sz = block_round(bytes);
    BlockCache& c = block_cache();
    std::lock_guard<std::mutex> g(c.mu);
    auto& fl = c.free_[dev];
    auto it = fl.lower_bound(sz);
    if (it != fl.end() && it->first <= 2 * sz) {
        void* p = it->second;
        c.live[p] = {dev, it->first};
        fl.erase(it);
        return p;
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, sz);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        block_flush(dev);
        e = hipMalloc(&p, sz);
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        throw Error(CPDLP_ERR_MEMORY, "hipMalloc of " + std::to_string(bytes) + " bytes failed: " +
                                          hipGetErrorString(e));
    }
    c.live[p] = {dev, sz};
    return p;
}
inline void dev_free(void* p) {
    if (!p) return;
    BlockCache& c = block_cache();
    std::lock_guard<std::mutex> g(c.mu);
    auto it = c.live.find(p);
    if (it == c.live.end()) {
        (void)hipFree(p);
        return;
    }
    c.free_[it->second.first].emplace(it->second.second, p);
    c.live.erase(it);
}

template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    explicit DevBuf(size_t count) { alloc(count); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(size_t count) {
        release();
        if (count == 0) return;
        p = (T*)dev_alloc(count * sizeof(T));
        n = count;
    }
    void release() {
        dev_free(p);
        p = nullptr;
        n = 0;
    }
    void upload(const T* h, size_t count, hipStream_t s) {
        if (count > n) throw Error(CPDLP_ERR_ARGUMENT, "upload larger than buffer");
        if (count) HIP_CHECK(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void upload(const std::vector<T>& v, hipStream_t s) {
        if (v.size() > n) alloc(v.size());
        upload(v.data(), v.size(), s);
    }
    void download(T* h, size_t count, hipStream_t s) const {
        if (count) HIP_CHECK(hipMemcpyAsync(h, p, count * sizeof(T), hipMemcpyDeviceToHost, s));
    }
};

// ---------------------------------------------------------------- scalar math
// The reference's elementwise helpers (src/utils/linalg.jl).  Julia's min/max on
// floats propagate NaN; these do the same.
template <class T>
__host__ __device__ __forceinline__ T jmax(T a, T b) {
    return (a != a || b != b) ? a + b : (a > b ? a : b);
}
template <class T>
__host__ __device__ __forceinline__ T jmin(T a, T b) {
    return (a != a || b != b) ? a + b : (a < b ? a : b);
}
// linalg.jl:4-5
template <class T>
__host__ __device__ __forceinline__ T pos_part(T a) { return a > T(0) ? a : T(0); }
template <class T>
__host__ __device__ __forceinline__ T neg_part(T a) { return -(a < T(0) ? a : T(0)); }
// linalg.jl:19: proj_box(x, l, u) = min(u, max(l, x))
template <class T>
__host__ __device__ __forceinline__ T proj_box(T x, T l, T u) { return jmin(u, jmax(l, x)); }
// linalg.jl:17: safeprod_left(l, r) = isinf(l) ? r : l*r
template <class T>
__host__ __device__ __forceinline__ T safeprod_left(T l, T r) { return isinf(l) ? r : l * r; }
// linalg.jl:26-42
template <class T>
__host__ __device__ __forceinline__ T proj_multiplier(T lam, T l, T u) {
    bool lmin = (l == -INFINITY), umax = (u == INFINITY);
    if (lmin) return umax ? T(0) : -neg_part(lam);
    return umax ? pos_part(lam) : lam;
}
// linalg.jl:49-53
template <class T>
__host__ __device__ __forceinline__ T combine(T l, T u) {
    T ls = isfinite(l) ? fabs(l) : T(0);
    T us = isfinite(u) ? fabs(u) : T(0);
    return jmax(jmax(T(0), ls), us);
}

// ---------------------------------------------------------------- wave/block reductions
constexpr int BLOCK = 256;      // 4 wavefronts of 64
constexpr int WAVE = 64;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
    return v;
}
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
    return v;
}

// Sequential sum p[lo] + p[lo+1] + ... + p[hi-1] from zero, in that order (bit-identical
// to the plain loop), with the LDS reads of each group of 8 issued together so that their
// latencies overlap instead of adding up.
template <class T>
__device__ __forceinline__ T seq_sum(const T* p, int lo, int hi) {
    T sum = T(0);
    int e = lo;
    for (; e + 8 <= hi; e += 8) {
        T v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = p[e + u];
#pragma unroll
        for (int u = 0; u < 8; ++u) sum += v[u];
    }
    if (e + 4 <= hi) {
        T v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = p[e + u];
#pragma unroll
        for (int u = 0; u < 4; ++u) sum += v[u];
        e += 4;
    }
    for (; e < hi; ++e) sum += p[e];
    return sum;
}

// Deterministic block reduction of NQ doubles per thread; thread q < NQ of the block
// returns the block total of quantity q (others: unspecified).  Fixed order: butterfly
// within each wave, then the THREADS / 64 wave sums in wave order.
template <int NQ, int THREADS = BLOCK>
__device__ __forceinline__ void block_sum(double (&q)[NQ], double* out_partial) {
    __shared__ double red[THREADS / WAVE][NQ > 0 ? NQ : 1];
    const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x / WAVE;
#pragma unroll
    for (int i = 0; i < NQ; ++i) {
        double v = wave_sum(q[i]);
        if (lane == 0) red[wave][i] = v;
    }
    __syncthreads();
    if ((int)threadIdx.x < NQ) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < THREADS / WAVE; ++w) s += red[w][threadIdx.x];
        out_partial[threadIdx.x] = s;
    }
}

inline int ceil_div(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }

}  // namespace cpdlp
