// clb_internal.hpp -- shared internals of the MI355X boolean sparse library.
//
// The runtime replaces clBool's OpenCL layer (src/core/controls.hpp,
// kernel.hpp, kernel_cache.hpp, error.hpp): one HIP stream per Controls, a
// per-device caching allocator instead of per-call cl::Buffer allocation,
// ahead-of-time compiled kernels instead of a JIT program cache, and C++
// exceptions carrying clBool's Status codes that the C-ABI turns into return
// codes.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/clbool_c.h"
#include "block_heap.hpp"

namespace clb {

using u32 = uint32_t;
using u64 = uint64_t;
using i64 = int64_t;

// Status codes of clBool (src/core/error.hpp:130-151), same order.
enum Status : int32_t {
    NO_SUCH_PROGRAM = 0,
    CREATE_KERNEL_ERROR,
    BUILD_PROGRAM_ERROR,
    INCOMPLETE_KERNEL,
    RUN_KERNEL_FAILURE,
    EVENT_WAITING_ERROR,
    CREATE_BUFFER_ERROR,
    COPY_BUFFER_ERROR,
    READ_BUFFER_ERROR,
    WRITE_BUFFER_ERROR,
    CREATE_CONTROLS_ERROR,
    INVALID_ARGUMENT,
    INVALID_VALUE,
    INITIALIZATION_ERROR
};

struct Error : std::runtime_error {
    Status status;
    Error(Status s, const std::string &m) : std::runtime_error(m), status(s) {}
};

[[noreturn]] void raise(Status s, const std::string &msg);
void check_hip(hipError_t e, Status s, const char *what);
#define CLB_HIP(expr, status) ::clb::check_hip((expr), (status), #expr)
// After a kernel launch: report launch-configuration errors immediately.
#define CLB_LAUNCHED(name) ::clb::check_hip(hipGetLastError(), ::clb::RUN_KERNEL_FAILURE, name)

// ---------------------------------------------------------------------------
// Per-device caching allocator.  Memory comes from HIP in segments (at least
// SEG_MIN bytes for large requests, SMALL_SEG for requests up to SMALL_MAX)
// that the pool keeps; a request takes the smallest free block that holds it
// and splits off the rest, and a released block merges with free neighbours
// in its segment.  Segments go back to HIP only when they are entirely free
// and an allocation fails (or on trim()), so after the first steps of a
// workload the pool serves every size from memory it already holds: recycled
// device memory makes hipMalloc slow (hundreds of milliseconds per call once
// the device has freed much memory), whatever the request size.
class Pool {
  public:
    static constexpr u64 SMALL_MAX = 1ull << 20;  // requests up to 1 MiB: 256-byte units
    static constexpr u64 SMALL_SEG = 2ull << 20;  // segments of the small requests
    static constexpr u64 SEG_MIN = 1ull << 30;    // smallest segment of the large requests
    explicit Pool(int device) : device_(device) {}
    ~Pool();
    void *alloc(u64 bytes, u64 *cap_out);
    void release(void *p, u64 cap);
    void trim();
    u64 cached_bytes() const {
        std::lock_guard<std::mutex> lk(mu_);
        return small_.free_bytes() + large_.free_bytes();
    }
    int device() const { return device_; }
    // [0] free bytes inside the pool's segments, [1] hipMalloc calls, [2]
    // bytes they allocated, [3] times free segments were handed back after a
    // failed hipMalloc
    void stats(u64 out[4]) const {
        std::lock_guard<std::mutex> lk(mu_);
        out[0] = small_.free_bytes() + large_.free_bytes();
        out[1] = mallocs_;
        out[2] = malloc_bytes_;
        out[3] = drops_;
    }

  private:
    // hands entirely free segments back to HIP (counted in drops_ after a
    // failed hipMalloc)
    void drop_free_segments(bool after_failure);
    int device_;
    mutable std::mutex mu_;
    BlockHeap small_{256}, large_{2ull << 20};
    u64 mallocs_ = 0, malloc_bytes_ = 0, drops_ = 0;
};
Pool &pool_for(int device);

}  // namespace clb

// Opaque handles of the C-ABI.
struct clb_buffer_s {
    void *ptr = nullptr;
    clb::u64 bytes = 0;  // requested size
    clb::u64 cap = 0;    // allocated capacity
    int device = 0;
    std::atomic<int> refs{1};
};

struct clb_controls_s {
    int device = 0;
    hipStream_t stream = nullptr;
    clb::u32 cu_count = 0;
    clb::u32 max_wg_size = 1024;
    clb::u32 lds_bytes = 0;
    clb::u64 *pinned = nullptr;   // small pinned staging area for readbacks
    clb_tuning tuning{};          // bin thresholds and kernel shapes
    clb_spgemm_stats last{};      // statistics of the last SpGEMM
    void *learned = nullptr;      // learned row-size predictor state (learned.hip)
    static constexpr int NSIDE = 3;
    hipStream_t side[NSIDE] = {};  // default-priority side streams for per-bin launches
    hipEvent_t ev_fork = nullptr;
    hipEvent_t ev_join[NSIDE] = {};
    // default-priority streams for the last phase of a one-pass block
    // (compaction, dense-row emission); a pipelined plan lets them run under
    // the next block's passes
    static constexpr int NTAIL = 2;
    hipStream_t tail[NTAIL] = {};
    std::mutex mu;
};

namespace clb {

clb_buffer_s *buffer_new(clb_controls_s *ctl, u64 bytes);
void buffer_retain(clb_buffer_s *b);
void buffer_release(clb_buffer_s *b);
inline u32 *u32p(clb_buffer_s *b) { return b ? static_cast<u32 *>(b->ptr) : nullptr; }
inline const u32 *u32p(const clb_buffer_s *b) { return b ? static_cast<const u32 *>(b->ptr) : nullptr; }

// RAII temporary device array from the pool.
template <typename T>
struct DevArray {
    T *ptr = nullptr;
    u64 n = 0;
    u64 cap = 0;
    int device = 0;
    DevArray() = default;
    DevArray(clb_controls_s *ctl, u64 count) { reset(ctl, count); }
    void reset(clb_controls_s *ctl, u64 count) {
        free_();
        n = count;
        device = ctl->device;
        if (count) ptr = static_cast<T *>(pool_for(device).alloc(count * sizeof(T), &cap));
    }
    ~DevArray() { free_(); }
    DevArray(const DevArray &) = delete;
    DevArray &operator=(const DevArray &) = delete;
    DevArray(DevArray &&o) noexcept : ptr(o.ptr), n(o.n), cap(o.cap), device(o.device) { o.ptr = nullptr; }
    void free_() {
        if (ptr) pool_for(device).release(ptr, cap);
        ptr = nullptr;
    }
};

// Fork the side streams off the main stream / join them back.
void fork_side(clb_controls_s *ctl);
void join_side(clb_controls_s *ctl);
inline hipStream_t side_stream(clb_controls_s *ctl, int k) { return ctl->side[k % clb_controls_s::NSIDE]; }

// Small synchronous readback through the pinned staging area.
void readback(clb_controls_s *ctl, void *host, const void *dev, u64 bytes);
void sync(clb_controls_s *ctl);

// Launch-size helpers.
inline u32 div_up(u64 a, u64 b) { return static_cast<u32>((a + b - 1) / b); }
constexpr u32 EMPTY = 0xFFFFFFFFu;

// Fill kernels (used instead of hipMemsetAsync so paths stay graph-safe).
void fill_u32(clb_controls_s *ctl, u32 *p, u64 n, u32 v);
void fill_u32_on(hipStream_t st, u32 *p, u64 n, u32 v);  // on a given stream
void fill_u64(clb_controls_s *ctl, u64 *p, u64 n, u64 v);

// Device-wide exclusive scans (scan.hip).  out may alias in.  If total is
// non-null the grand total is written there (device memory).
void exclusive_scan_u32(clb_controls_s *ctl, const u32 *in, u32 *out, u64 n, u32 *total);
void exclusive_scan_u64(clb_controls_s *ctl, const u64 *in, u64 *out, u64 n, u64 *total);
// Exclusive scan of u32 inputs widened to u64 outputs.
void exclusive_scan_u32_to_u64(clb_controls_s *ctl, const u32 *in, u64 *out, u64 n, u64 *total);

// Runs body, mapping exceptions to C-ABI status codes (runtime.hip).
clb_status run_guarded(const std::function<void()> &body);

// Frees the learned predictor's device state of a Controls (learned.hip).
void learned_destroy_state(clb_controls_s *ctl);

// Kernel registry for the kernel<> compatibility shim.
bool kernel_registered(const std::string &program, const std::string &kernel);

}  // namespace clb
