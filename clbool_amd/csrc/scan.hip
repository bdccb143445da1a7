// scan.hip -- device-wide exclusive scans.
//
// Replaces prefix_sum (src/common/cl_operations.cpp:4-88; kernels
// src/cl/prefix_sum.cl:11-83).  The reference runs a Blelloch scan per
// 512-element tile, one host-waited launch per level, and reads the total
// back to the host every time.  Here a level is reduce-then-scan over
// 4096-element tiles (256 threads x 16 items, wave64 shuffle scans), levels
// recurse on the device without host waits, and the total stays in device
// memory unless the caller asks for it.
#include "clb_internal.hpp"
#include "device_utils.hpp"

namespace clb {

constexpr int SCAN_WG = 256;
constexpr int SCAN_ITEMS = 16;
constexpr u64 SCAN_TILE = (u64)SCAN_WG * SCAN_ITEMS;

template <typename In, typename T>
__global__ __launch_bounds__(SCAN_WG) void k_scan_reduce(const In *__restrict__ in, u64 n, T *__restrict__ partial) {
    __shared__ T lds[SCAN_WG / dev::WAVE + 1];
    const u64 base = (u64)blockIdx.x * SCAN_TILE + (u64)threadIdx.x * SCAN_ITEMS;
    T s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        const u64 idx = base + i;
        if (idx < n) s += static_cast<T>(in[idx]);
    }
    T t = dev::block_sum<T, SCAN_WG>(s, lds);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// Scans each tile and adds the tile's offset (offsets == nullptr: zero).
template <typename In, typename T>
__global__ __launch_bounds__(SCAN_WG) void k_scan_apply(const In *in, T *out, u64 n, const T *__restrict__ offsets,
                                                        T *total) {
    __shared__ T lds[SCAN_WG / dev::WAVE + 1];
    const u64 base = (u64)blockIdx.x * SCAN_TILE + (u64)threadIdx.x * SCAN_ITEMS;
    T v[SCAN_ITEMS];
    T s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        const u64 idx = base + i;
        v[i] = idx < n ? static_cast<T>(in[idx]) : T(0);
        s += v[i];
    }
    T tile_total;
    T p = dev::block_excl_scan<T, SCAN_WG>(s, lds, tile_total);
    const T off = offsets ? offsets[blockIdx.x] : T(0);
    p += off;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        const u64 idx = base + i;
        if (idx < n) out[idx] = p;
        p += v[i];
    }
    if (total && blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) *total = off + tile_total;
}

template <typename In, typename T>
static void scan_impl(clb_controls_s *ctl, const In *in, T *out, u64 n, T *total) {
    if (n == 0) {
        if (total) {
            if (sizeof(T) == 8) fill_u64(ctl, reinterpret_cast<u64 *>(total), 1, 0);
            else fill_u32(ctl, reinterpret_cast<u32 *>(total), 1, 0);
        }
        return;
    }
    const u64 tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
    if (tiles == 1) {
        k_scan_apply<In, T><<<1, SCAN_WG, 0, ctl->stream>>>(in, out, n, nullptr, total);
        CLB_LAUNCHED("scan_apply");
        return;
    }
    DevArray<T> partial(ctl, tiles);
    k_scan_reduce<In, T><<<static_cast<u32>(tiles), SCAN_WG, 0, ctl->stream>>>(in, n, partial.ptr);
    CLB_LAUNCHED("scan_reduce");
    DevArray<T> tot;
    T *total_dst = total;
    // partial is scanned in place; its total is the grand total
    scan_impl<T, T>(ctl, partial.ptr, partial.ptr, tiles, total_dst);
    k_scan_apply<In, T><<<static_cast<u32>(tiles), SCAN_WG, 0, ctl->stream>>>(in, out, n, partial.ptr, nullptr);
    CLB_LAUNCHED("scan_apply");
}

void exclusive_scan_u32(clb_controls_s *ctl, const u32 *in, u32 *out, u64 n, u32 *total) {
    scan_impl<u32, u32>(ctl, in, out, n, total);
}

void exclusive_scan_u64(clb_controls_s *ctl, const u64 *in, u64 *out, u64 n, u64 *total) {
    scan_impl<u64, u64>(ctl, in, out, n, total);
}

void exclusive_scan_u32_to_u64(clb_controls_s *ctl, const u32 *in, u64 *out, u64 n, u64 *total) {
    scan_impl<u32, u64>(ctl, in, out, n, total);
}

}  // namespace clb
