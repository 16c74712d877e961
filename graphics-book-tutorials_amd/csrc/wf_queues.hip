// Scan, stable compaction and the queue-push test hook, written for 64-wide waves.
// Contracts follow src/gpu/queues.cuh:29-54; the algorithm is reduce-then-scan:
//   1. each block reduces one tile to a partial sum (tile = blockDim * items per thread),
//   2. one block scans the partial sums (looping over 1024-entry chunks with a carry),
//   3. each block rescans its tile with its partial as the offset and writes the result.
// Inside a tile every thread owns `items` consecutive elements (one 16-byte load per 4 items),
// a wave scans the per-thread totals with shuffles, and LDS combines the waves. No block waits
// on another block, so no launch order or co-residency is assumed.
#include <mutex>

#include "wf_queues.h"

namespace wf::gpu {

// Partial sums of the allocation-free compaction: one word per tile of the smallest tile size for
// up to kNoAllocMaxN elements (512 KiB per device, held until the process ends). Each device has
// its own copy at its own address; host code reaches it only through the address that
// hipGetSymbolAddress returns on that device (noalloc_partials), never through this name.
constexpr int kNoAllocPartials = kNoAllocMaxN / kMinScanTile + 1;
static __device__ __align__(256) uint32_t scratch[kNoAllocPartials];

namespace {

constexpr int kScanThreads = 256;
constexpr int kMaxWaves = kScanThreads / 64;

__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t x) {
    const int lane = (int)lane_id();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    return x;
}

// Exclusive prefix of `t` over the block (in thread order) and the block total.
__device__ __forceinline__ uint32_t block_exclusive(uint32_t t, uint32_t* total) {
    __shared__ uint32_t wave_tot[kMaxWaves];
    const int w = threadIdx.x >> 6;
    const uint32_t incl = wave_inclusive_scan(t);
    if (lane_id() == 63) wave_tot[w] = incl;
    __syncthreads();
    uint32_t off = 0, tot = 0;
    for (int k = 0; k < kMaxWaves; ++k) {
        uint32_t v = wave_tot[k];
        if (k < w) off += v;
        tot += v;
    }
    __syncthreads();
    *total = tot;
    return off + incl - t;
}

template <int ITEMS>
__device__ __forceinline__ void load_tile(const uint32_t* in, int n, long base, uint32_t (&v)[ITEMS]) {
    const long i0 = base + (long)threadIdx.x * ITEMS;
    if (ITEMS % 4 == 0 && i0 + ITEMS <= n && ((uintptr_t)(in + i0) & 15) == 0) {
#pragma unroll
        for (int k = 0; k < ITEMS; k += 4) {
            uint4 q = *reinterpret_cast<const uint4*>(in + i0 + k);
            v[k] = q.x; v[k + 1] = q.y; v[k + 2] = q.z; v[k + 3] = q.w;
        }
    } else {
#pragma unroll
        for (int k = 0; k < ITEMS; ++k) v[k] = (i0 + k < n) ? in[i0 + k] : 0u;
    }
}

template <int ITEMS, bool INDICATOR>
__global__ __launch_bounds__(kScanThreads) void scan_reduce_kernel(const uint32_t* in, int n, uint32_t* partial) {
    const long base = (long)blockIdx.x * kScanThreads * ITEMS;
    uint32_t v[ITEMS];
    load_tile<ITEMS>(in, n, base, v);
    uint32_t t = 0;
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) t += INDICATOR ? (v[k] != 0u ? 1u : 0u) : v[k];
    uint32_t total;
    block_exclusive(t, &total);
    if (threadIdx.x == 0) partial[blockIdx.x] = total;
}

// One block of 1024 threads scans partial[0, nb) in place (exclusive) and writes the total.
__global__ __launch_bounds__(1024) void scan_partials_kernel(uint32_t* partial, int nb, uint32_t* d_total) {
    __shared__ uint32_t wtot[16];
    __shared__ uint32_t carry;
    if (threadIdx.x == 0) carry = 0u;
    __syncthreads();
    const int w = threadIdx.x >> 6;
    for (int base = 0; base < nb; base += 1024) {
        const int i = base + threadIdx.x;
        const uint32_t v = i < nb ? partial[i] : 0u;
        const uint32_t incl = wave_inclusive_scan(v);
        if (lane_id() == 63) wtot[w] = incl;
        __syncthreads();
        uint32_t off = 0, tot = 0;
        for (int k = 0; k < 16; ++k) {
            uint32_t x = wtot[k];
            if (k < w) off += x;
            tot += x;
        }
        const uint32_t c = carry;
        if (i < nb) partial[i] = c + off + incl - v;
        __syncthreads();
        if (threadIdx.x == 0) carry = c + tot;
        __syncthreads();
    }
    if (threadIdx.x == 0 && d_total) *d_total = carry;
}

template <int ITEMS>
__global__ __launch_bounds__(kScanThreads) void scan_down_kernel(const uint32_t* in, uint32_t* out, int n,
                                                                 const uint32_t* partial) {
    const long base = (long)blockIdx.x * kScanThreads * ITEMS;
    uint32_t v[ITEMS];
    load_tile<ITEMS>(in, n, base, v);
    uint32_t t = 0;
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) t += v[k];
    uint32_t total;
    uint32_t run = block_exclusive(t, &total) + partial[blockIdx.x];
    const long i0 = base + (long)threadIdx.x * ITEMS;
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        if (i0 + k < n) out[i0 + k] = run;
        run += v[k];
    }
}

template <int ITEMS>
__global__ __launch_bounds__(kScanThreads) void compact_down_kernel(const uint32_t* values, const uint32_t* flags,
                                                                    uint32_t* out, int n, const uint32_t* partial) {
    const long base = (long)blockIdx.x * kScanThreads * ITEMS;
    uint32_t f[ITEMS];
    load_tile<ITEMS>(flags, n, base, f);
    uint32_t t = 0;
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) t += f[k] != 0u ? 1u : 0u;
    uint32_t total;
    uint32_t pos = block_exclusive(t, &total) + partial[blockIdx.x];
    const long i0 = base + (long)threadIdx.x * ITEMS;
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        if (i0 + k < n && f[k] != 0u) out[pos++] = values[i0 + k];
    }
}

__global__ void test_queue_push_kernel(const uint32_t* flags, int n, uint32_t* items, uint32_t capacity,
                                       uint32_t* counter, uint32_t* overflow) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool pred = i < n && flags[i] != 0u;
    wave_push(items, counter, capacity, (uint32_t)i, pred, overflow);
}

__global__ void set_u32_kernel(uint32_t* p, uint32_t v) { *p = v; }

}  // namespace

int scan_items_for_tile(int tile) {
    int items = tile / kScanThreads;
    return (items == 2 || items == 4 || items == 8 || items == 16) ? items : 0;
}

hipError_t scan_workspace_create(ScanWorkspace* ws, int capacity) {
    ws->capacity = 0;
    ws->partial = nullptr;
    if (capacity < 0) return hipErrorInvalidValue;
    const int nb = (capacity + kMinScanTile - 1) / kMinScanTile + 1;
    hipError_t e = hipMalloc(&ws->partial, (size_t)nb * sizeof(uint32_t));
    if (e != hipSuccess) return e;
    ws->capacity = capacity;
    return hipSuccess;
}

void scan_workspace_destroy(ScanWorkspace* ws) {
    if (ws->partial) (void)hipFree(ws->partial);
    ws->partial = nullptr;
    ws->capacity = 0;
}

template <int ITEMS>
static hipError_t scan_impl(const uint32_t* d_in, uint32_t* d_out, int n, uint32_t* d_total, ScanWorkspace& ws,
                            hipStream_t stream) {
    const int tile = kScanThreads * ITEMS;
    const int nb = (n + tile - 1) / tile;
    scan_reduce_kernel<ITEMS, false><<<nb, kScanThreads, 0, stream>>>(d_in, n, ws.partial);
    scan_partials_kernel<<<1, 1024, 0, stream>>>(ws.partial, nb, d_total);
    scan_down_kernel<ITEMS><<<nb, kScanThreads, 0, stream>>>(d_in, d_out, n, ws.partial);
    return hipGetLastError();
}

template <int ITEMS>
static hipError_t compact_impl(const uint32_t* d_values, const uint32_t* d_flags, uint32_t* d_out, uint32_t* d_count,
                               int n, uint32_t* partial, hipStream_t stream) {
    const int tile = kScanThreads * ITEMS;
    const int nb = (n + tile - 1) / tile;
    scan_reduce_kernel<ITEMS, true><<<nb, kScanThreads, 0, stream>>>(d_flags, n, partial);
    scan_partials_kernel<<<1, 1024, 0, stream>>>(partial, nb, d_count);
    compact_down_kernel<ITEMS><<<nb, kScanThreads, 0, stream>>>(d_values, d_flags, d_out, n, partial);
    return hipGetLastError();
}

static hipError_t compact_dispatch(const uint32_t* d_values, const uint32_t* d_flags, uint32_t* d_out,
                                   uint32_t* d_count, int n, uint32_t* partial, hipStream_t stream, int tile) {
    if (n == 0) {
        if (d_count) set_u32_kernel<<<1, 1, 0, stream>>>(d_count, 0u);
        return hipGetLastError();
    }
    switch (scan_items_for_tile(tile)) {
        case 2: return compact_impl<2>(d_values, d_flags, d_out, d_count, n, partial, stream);
        case 4: return compact_impl<4>(d_values, d_flags, d_out, d_count, n, partial, stream);
        case 16: return compact_impl<16>(d_values, d_flags, d_out, d_count, n, partial, stream);
        case 8: return compact_impl<8>(d_values, d_flags, d_out, d_count, n, partial, stream);
        default: return hipErrorInvalidValue;
    }
}

// Device address of `scratch` on the current device, looked up once per device (with that device
// current) and kept; null if the lookup fails.
static uint32_t* noalloc_partials(hipError_t* err) {
    static std::mutex mu;
    static uint32_t* addr[64] = {};
    static bool looked_up[64] = {};
    int dev = 0;
    *err = hipGetDevice(&dev);
    if (*err != hipSuccess) return nullptr;
    if (dev < 0 || dev >= 64) {
        *err = hipErrorInvalidDevice;
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(mu);
    if (!looked_up[dev]) {
        void* p = nullptr;
        *err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (*err != hipSuccess) return nullptr;
        addr[dev] = static_cast<uint32_t*>(p);
        looked_up[dev] = true;
    }
    return addr[dev];
}

hipError_t scan_exclusive_u32(const uint32_t* d_in, uint32_t* d_out, int n, uint32_t* d_total, ScanWorkspace& ws,
                              hipStream_t stream, int tile) {
    if (n < 0 || n > ws.capacity) return hipErrorInvalidValue;
    if (n == 0) {
        if (d_total) set_u32_kernel<<<1, 1, 0, stream>>>(d_total, 0u);
        return hipGetLastError();
    }
    switch (scan_items_for_tile(tile)) {
        case 2: return scan_impl<2>(d_in, d_out, n, d_total, ws, stream);
        case 4: return scan_impl<4>(d_in, d_out, n, d_total, ws, stream);
        case 16: return scan_impl<16>(d_in, d_out, n, d_total, ws, stream);
        case 8: return scan_impl<8>(d_in, d_out, n, d_total, ws, stream);
        default: return hipErrorInvalidValue;
    }
}

hipError_t compact_u32(const uint32_t* d_values, const uint32_t* d_flags, uint32_t* d_out, uint32_t* d_count, int n,
                       ScanWorkspace& ws, hipStream_t stream, int tile) {
    if (n < 0 || n > ws.capacity) return hipErrorInvalidValue;
    return compact_dispatch(d_values, d_flags, d_out, d_count, n, ws.partial, stream, tile);
}

// Every partial that a call reads is written by its own reduce kernel first (tiles [0, nb)), so the
// array needs no clearing between calls.
hipError_t compact_u32_noalloc(const uint32_t* d_values, const uint32_t* d_flags, uint32_t* d_out, uint32_t* d_count,
                               int n, hipStream_t stream, int tile) {
    if (n < 0 || n > kNoAllocMaxN || scan_items_for_tile(tile) == 0) return hipErrorInvalidValue;
    hipError_t err = hipSuccess;
    uint32_t* partial = noalloc_partials(&err);
    if (!partial) return err != hipSuccess ? err : hipErrorInvalidSymbol;
    return compact_dispatch(d_values, d_flags, d_out, d_count, n, partial, stream, tile);
}

hipError_t test_queue_push(const uint32_t* d_flags, int n, uint32_t* d_items, uint32_t capacity, uint32_t* d_counter,
                           uint32_t* d_overflow, hipStream_t stream) {
    if (n < 0) return hipErrorInvalidValue;
    if (n == 0) return hipSuccess;
    const int nb = (n + 255) / 256;
    test_queue_push_kernel<<<nb, 256, 0, stream>>>(d_flags, n, d_items, capacity, d_counter, d_overflow);
    return hipGetLastError();
}

}  // namespace wf::gpu
