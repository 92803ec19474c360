// Shared HIP plumbing: error propagation, device buffers, launch shapes and
// deterministic wave64 + LDS reductions.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>

namespace cfd {

std::string &last_error();

struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define CFD_HIP(call)                                                           \
    do {                                                                        \
        hipError_t e_ = (call);                                                 \
        if (e_ != hipSuccess)                                                   \
            throw ::cfd::HipError(std::string(#call " failed: ") +              \
                                  hipGetErrorString(e_) + " at " __FILE__ ":" + \
                                  std::to_string(__LINE__));                    \
    } while (0)

#define CFD_LAUNCH_CHECK() CFD_HIP(hipGetLastError())

// Owning device buffer (hipMalloc'd, freed on destruction).
template <class T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    DBuf() = default;
    explicit DBuf(size_t count) { alloc(count); }
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    DBuf(DBuf &&o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DBuf &operator=(DBuf &&o) noexcept {
        if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    ~DBuf() { release(); }
    void alloc(size_t count) {
        release();
        if (count) CFD_HIP(hipMalloc(&p, count * sizeof(T)));
        n = count;
    }
    void ensure(size_t count) {
        if (count > n) alloc(count);
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    // Host copies are set-up / I-O paths: they complete before returning, so
    // pageable host buffers may be released right after the call.
    void upload(const T *h, size_t count, hipStream_t s = 0) {
        if (count > n) throw HipError("upload larger than buffer");
        if (!count) return;
        CFD_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
        CFD_HIP(hipStreamSynchronize(s));
    }
    void download(T *h, size_t count, hipStream_t s = 0) const {
        if (count > n) throw HipError("download larger than buffer");
        if (!count) return;
        CFD_HIP(hipMemcpyAsync(h, p, count * sizeof(T), hipMemcpyDeviceToHost, s));
        CFD_HIP(hipStreamSynchronize(s));
    }
    void zero(hipStream_t s = 0) {
        if (n) CFD_HIP(hipMemsetAsync(p, 0, n * sizeof(T), s));
    }
    operator T *() const { return p; }
};

constexpr int kBlock = 256;          // 4 wave64 per workgroup
constexpr int kRedGrid = 2048;       // fixed grid of reducing kernels: 8 WG/CU
inline int grid_for(long long n, int block = kBlock) {
    long long g = (n + block - 1) / block;
    if (g < 1) g = 1;
    return (int)g;
}
inline int red_grid(long long n) {
    long long g = (n + kBlock - 1) / kBlock;
    if (g > kRedGrid) g = kRedGrid;
    if (g < 1) g = 1;
    return (int)g;
}

// ---------------------------------------------------------------- device --
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Block-wide sum of NV values (blockDim == kBlock); result valid in thread 0.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV]) {
    __shared__ double red[NV][kBlock / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = wave_sum(v[k]);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) red[k][wid] = v[k];
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            double s = 0.;
            for (int w = 0; w < kBlock / 64; ++w) s += red[k][w];
            v[k] = s;
        }
}

// Grid-wide deterministic reduction: each block writes its partials; the
// last block to arrive (atomic ticket) sums them in block order and returns
// true in thread 0 with the totals in v. part must hold NV*gridDim.x doubles.
template <int NV>
__device__ __forceinline__ bool grid_sum(double (&v)[NV], double *part, unsigned *ticket) {
    block_sum<NV>(v);
    __shared__ bool last;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) part[k * gridDim.x + blockIdx.x] = v[k];
        __threadfence();
        unsigned t = atomicAdd(ticket, 1u);
        last = (t == gridDim.x - 1);
    }
    __syncthreads();
    if (!last) return false;
    __threadfence();
    double acc[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) acc[k] = 0.;
    // all threads of the last block cooperate: strided partial sums, then a
    // fixed-order block reduction (deterministic for a fixed grid size)
    // agent-scope loads bypass this CU's L1, which may not hold other
    // blocks' freshly written partials
    for (int b = threadIdx.x; b < (int)gridDim.x; b += blockDim.x)
#pragma unroll
        for (int k = 0; k < NV; ++k)
            acc[k] += __hip_atomic_load(&part[k * gridDim.x + b], __ATOMIC_RELAXED,
                                        __HIP_MEMORY_SCOPE_AGENT);
    block_sum<NV>(acc);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) v[k] = acc[k];
        *ticket = 0u;
    }
    return threadIdx.x == 0;
}

}  // namespace cfd
