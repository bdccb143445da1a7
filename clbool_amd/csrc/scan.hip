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

#include <map>
#include <mutex>

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

// ---------------------------------------------------------------------------
// The prefix sum without allocation (clb_prefix_sum_noalloc): the tile sums
// of every level live in one file-scope device array instead of pool blocks.
// Scans of at most 2^28 elements need 2^16 level-1 tile sums, 16 at level 2
// and the total: NOALLOC_WORDS (65,600 words, 256 KiB) per GPU, held until
// the process ends.  Every call on a GPU shares that GPU's copy; calls that
// overlap on one GPU (from several streams or threads) race.  Nothing in it
// has to start at zero: every word a scan reads was written earlier in the
// same call.
constexpr u64 NOALLOC_MAX_N = 1ull << 28;
constexpr u64 NOALLOC_WORDS = NOALLOC_MAX_N / SCAN_TILE + 64;
static __device__ __align__(256) u32 scratch[NOALLOC_WORDS];

// Device address of `scratch` on the controls' device, looked up once per
// device with that device current (each GPU holds its own copy).
static u32 *scratch_on(clb_controls_s *ctl) {
    static std::mutex mu;
    static std::map<int, u32 *> addr;
    std::lock_guard<std::mutex> lk(mu);
    auto it = addr.find(ctl->device);
    if (it != addr.end()) return it->second;
    int prev = 0;
    CLB_HIP(hipGetDevice(&prev), RUN_KERNEL_FAILURE);
    if (prev != ctl->device) CLB_HIP(hipSetDevice(ctl->device), RUN_KERNEL_FAILURE);
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (prev != ctl->device) (void)hipSetDevice(prev);
    if (e != hipSuccess || !p) {
        (void)hipGetLastError();
        raise(RUN_KERNEL_FAILURE, std::string("hipGetSymbolAddress of the scan scratch failed: ") +
                                      hipGetErrorString(e));
    }
    addr[ctl->device] = static_cast<u32 *>(p);
    return static_cast<u32 *>(p);
}

// scan_impl with the tile sums of each level at ws (the levels' tile counts
// added up must fit the workspace).
static void scan_ws(clb_controls_s *ctl, const u32 *in, u32 *out, u64 n, u32 *total, u32 *ws) {
    if (n == 0) {
        if (total) fill_u32(ctl, total, 1, 0);
        return;
    }
    const u64 tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
    if (tiles == 1) {
        k_scan_apply<u32, u32><<<1, SCAN_WG, 0, ctl->stream>>>(in, out, n, nullptr, total);
        CLB_LAUNCHED("scan_apply");
        return;
    }
    u32 *partial = ws;
    k_scan_reduce<u32, u32><<<static_cast<u32>(tiles), SCAN_WG, 0, ctl->stream>>>(in, n, partial);
    CLB_LAUNCHED("scan_reduce");
    scan_ws(ctl, partial, partial, tiles, total, ws + tiles);
    k_scan_apply<u32, u32><<<static_cast<u32>(tiles), SCAN_WG, 0, ctl->stream>>>(in, out, n, partial, nullptr);
    CLB_LAUNCHED("scan_apply");
}

}  // namespace clb

using namespace clb;

extern "C" {

clb_status clb_prefix_sum_noalloc(clb_controls ctl, clb_buffer buf, uint32_t n, uint32_t *total) {
    return run_guarded([&] {
        if (!ctl || !total) raise(INVALID_ARGUMENT, "prefix_sum_noalloc: null argument");
        *total = 0;
        if ((u64)n > NOALLOC_MAX_N)
            raise(INVALID_ARGUMENT, "prefix_sum_noalloc: at most 2^28 elements (its fixed scratch holds the tile "
                                    "sums of that many); use clb_prefix_sum");
        u32 *ws = scratch_on(ctl);
        if (!n) return;
        if (!buf || 4ull * n > buf->bytes) raise(INVALID_ARGUMENT, "prefix_sum_noalloc: buffer too small");
        u32 *tot = ws + NOALLOC_WORDS - 1;  // past every level's tile sums
        scan_ws(ctl, u32p(buf), u32p(buf), n, tot, ws);
        readback(ctl, total, tot, 4);  // through the controls' pinned staging
    });
}

}  // extern "C"
