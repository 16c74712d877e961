// Scan / compaction / queue-push primitives (contracts of src/gpu/queues.cuh:29-54).
#pragma once
#include "wf_gpu.h"

namespace wf::gpu {

// Smallest supported tile; the workspace holds one partial per tile of this size.
constexpr int kMinScanTile = 512;
constexpr int kDefaultScanTile = 2048;

struct ScanWorkspace {
    uint32_t* partial = nullptr;
    int capacity = 0;
};

hipError_t scan_workspace_create(ScanWorkspace* ws, int capacity);
void scan_workspace_destroy(ScanWorkspace* ws);

// Items per thread for a tile size (tile = 256 * items, items in {2, 4, 8, 16}); 0 if unsupported.
int scan_items_for_tile(int tile);

// Exclusive prefix sum of n uint32 values; d_in may equal d_out; the total goes to *d_total
// (device pointer, may be null). n > capacity -> hipErrorInvalidValue with no side effects.
hipError_t scan_exclusive_u32(const uint32_t* d_in, uint32_t* d_out, int n, uint32_t* d_total, ScanWorkspace& ws,
                              hipStream_t stream, int tile = kDefaultScanTile);

// Stable compaction: d_out[k] = d_values[i] for the k-th i with d_flags[i] != 0; *d_count gets
// the number selected. n > capacity -> hipErrorInvalidValue with no side effects.
hipError_t compact_u32(const uint32_t* d_values, const uint32_t* d_flags, uint32_t* d_out, uint32_t* d_count, int n,
                       ScanWorkspace& ws, hipStream_t stream, int tile = kDefaultScanTile);

// Allocation-free compaction: as compact_u32, with the per-tile partial sums in a static device
// array of kNoAllocMaxN / kMinScanTile + 1 words instead of a workspace. n > kNoAllocMaxN ->
// hipErrorInvalidValue with no side effects. No hipMalloc, hipFree or allocating library call is
// made; calls that overlap on one device share the array and race.
constexpr int kNoAllocMaxN = 1 << 26;
hipError_t compact_u32_noalloc(const uint32_t* d_values, const uint32_t* d_flags, uint32_t* d_out, uint32_t* d_count,
                               int n, hipStream_t stream, int tile = kDefaultScanTile);

// Test hook: lanes i < n with d_flags[i] != 0 push i through wave_push.
hipError_t test_queue_push(const uint32_t* d_flags, int n, uint32_t* d_items, uint32_t capacity, uint32_t* d_counter,
                           uint32_t* d_overflow, hipStream_t stream);

}  // namespace wf::gpu
