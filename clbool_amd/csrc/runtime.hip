// runtime.hip -- Controls, device buffers, caching allocator, errors.
//
// Replaces clBool's OpenCL runtime layer: Controls (src/core/controls.hpp:12-28,
// one context and two queues) becomes one HIP stream on one device;
// create_controls (src/common/env.cpp:7-36) validates platform/device ids the
// same way; cl::Buffer becomes a reference-counted pool allocation; the
// exception model of src/core/error.hpp:130-271 becomes C++ exceptions that
// the C-ABI maps to return codes plus a thread-local message.
#include "clb_internal.hpp"

#include <cstdio>
#include <sstream>

namespace clb {

static thread_local std::string g_last_error;

void raise(Status s, const std::string &msg) { throw Error(s, msg); }

void check_hip(hipError_t e, Status s, const char *what) {
    if (e != hipSuccess) {
        std::ostringstream os;
        os << what << ": " << hipGetErrorString(e) << " (" << static_cast<int>(e) << ")";
        raise(s, os.str());
    }
}

// ---------------------------------------------------------------------------
Pool::~Pool() { trim(); }

static u64 round_capacity(u64 bytes) {
    // 256-byte units for small blocks, 2 MiB for large ones
    if (bytes <= Pool::SMALL_MAX) return (bytes + 255) & ~255ull;
    const u64 g = 2ull << 20;
    return (bytes + g - 1) / g * g;
}

void *Pool::alloc(u64 bytes, u64 *cap_out) {
    const u64 cap = round_capacity(bytes ? bytes : 1);
    const bool small = cap <= SMALL_MAX;
    BlockHeap &heap = small ? small_ : large_;
    {
        std::lock_guard<std::mutex> lk(mu_);
        if (BlockHeap::Block *b = heap.take(cap)) {
            *cap_out = b->size;
            return b->ptr;
        }
    }
    // a new segment: SEG_MIN or the request, whichever is larger (only the
    // request itself when that much is not available)
    const u64 seg = small ? SMALL_SEG : (cap > SEG_MIN ? cap : SEG_MIN);
    int prev = 0;
    CLB_HIP(hipGetDevice(&prev), CREATE_BUFFER_ERROR);
    if (prev != device_) CLB_HIP(hipSetDevice(device_), CREATE_BUFFER_ERROR);
    void *p = nullptr;
    u64 got = seg;
    auto try_malloc = [&]() {
        got = seg;
        hipError_t e = hipMalloc(&p, got);
        if (e != hipSuccess && seg > cap) {
            (void)hipGetLastError();
            got = cap;
            e = hipMalloc(&p, got);
        }
        return e;
    };
    hipError_t e = try_malloc();
    if (e != hipSuccess) {
        (void)hipGetLastError();
        drop_free_segments(true);
        e = try_malloc();
    }
    if (prev != device_) (void)hipSetDevice(prev);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        std::ostringstream os;
        os << "device allocation of " << cap << " bytes failed: " << hipGetErrorString(e);
        raise(CREATE_BUFFER_ERROR, os.str());
    }
    std::lock_guard<std::mutex> lk(mu_);
    ++mallocs_;
    malloc_bytes_ += got;
    heap.add_segment(static_cast<char *>(p), got);
    BlockHeap::Block *b = heap.take(cap);  // fits: the new segment holds cap bytes
    *cap_out = b->size;
    return b->ptr;
}

void Pool::release(void *p, u64) {
    if (!p) return;
    std::lock_guard<std::mutex> lk(mu_);
    if (!small_.give_back(static_cast<char *>(p))) (void)large_.give_back(static_cast<char *>(p));
}

void Pool::drop_free_segments(bool after_failure) {
    std::vector<std::pair<char *, u64>> drop;
    {
        std::lock_guard<std::mutex> lk(mu_);
        small_.collect_free_segments(drop);
        large_.collect_free_segments(drop);
        if (after_failure) ++drops_;
    }
    if (drop.empty()) return;
    int prev = 0;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(device_);
    (void)hipDeviceSynchronize();
    for (auto &kv : drop) (void)hipFree(kv.first);
    (void)hipSetDevice(prev);
}

void Pool::trim() { drop_free_segments(false); }

Pool &pool_for(int device) {
    static std::mutex mu;
    static std::map<int, Pool *> pools;  // intentionally leaked: lives to process end
    std::lock_guard<std::mutex> lk(mu);
    auto it = pools.find(device);
    if (it != pools.end()) return *it->second;
    Pool *p = new Pool(device);
    pools[device] = p;
    return *p;
}

// ---------------------------------------------------------------------------
clb_buffer_s *buffer_new(clb_controls_s *ctl, u64 bytes) {
    if (bytes == 0) raise(CREATE_BUFFER_ERROR, "cannot create a buffer of size 0");
    auto *b = new clb_buffer_s();
    b->device = ctl->device;
    b->bytes = bytes;
    try {
        b->ptr = pool_for(ctl->device).alloc(bytes, &b->cap);
    } catch (...) {
        delete b;
        throw;
    }
    return b;
}

void buffer_retain(clb_buffer_s *b) {
    if (b) b->refs.fetch_add(1, std::memory_order_relaxed);
}

void buffer_release(clb_buffer_s *b) {
    if (!b) return;
    if (b->refs.fetch_sub(1, std::memory_order_acq_rel) == 1) {
        pool_for(b->device).release(b->ptr, b->cap);
        delete b;
    }
}

void sync(clb_controls_s *ctl) { CLB_HIP(hipStreamSynchronize(ctl->stream), EVENT_WAITING_ERROR); }

void fork_side(clb_controls_s *ctl) {
    CLB_HIP(hipEventRecord(ctl->ev_fork, ctl->stream), EVENT_WAITING_ERROR);
    for (int k = 0; k < clb_controls_s::NSIDE; ++k)
        CLB_HIP(hipStreamWaitEvent(ctl->side[k], ctl->ev_fork, 0), EVENT_WAITING_ERROR);
}

void join_side(clb_controls_s *ctl) {
    for (int k = 0; k < clb_controls_s::NSIDE; ++k) {
        CLB_HIP(hipEventRecord(ctl->ev_join[k], ctl->side[k]), EVENT_WAITING_ERROR);
        CLB_HIP(hipStreamWaitEvent(ctl->stream, ctl->ev_join[k], 0), EVENT_WAITING_ERROR);
    }
}

void readback(clb_controls_s *ctl, void *host, const void *dev, u64 bytes) {
    if (bytes <= 4096) {
        CLB_HIP(hipMemcpyAsync(ctl->pinned, dev, bytes, hipMemcpyDeviceToHost, ctl->stream), READ_BUFFER_ERROR);
        sync(ctl);
        std::memcpy(host, ctl->pinned, bytes);
    } else {
        CLB_HIP(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, ctl->stream), READ_BUFFER_ERROR);
        sync(ctl);
    }
}

// ---------------------------------------------------------------------------
__global__ void k_fill_u32(u32 *__restrict__ p, u64 n, u32 v) {
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

__global__ void k_fill_u64(u64 *__restrict__ p, u64 n, u64 v) {
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

static u32 fill_grid(clb_controls_s *ctl, u64 n) {
    u64 g = (n + 255) / 256;
    u64 cap = (u64)ctl->cu_count * 16;
    if (g > cap) g = cap;
    return static_cast<u32>(g ? g : 1);
}

void fill_u32(clb_controls_s *ctl, u32 *p, u64 n, u32 v) {
    if (!n) return;
    k_fill_u32<<<fill_grid(ctl, n), 256, 0, ctl->stream>>>(p, n, v);
    CLB_LAUNCHED("fill_with");
}

void fill_u32_on(hipStream_t st, u32 *p, u64 n, u32 v) {
    if (!n) return;
    k_fill_u32<<<(u32)((n + 255) / 256 < 1024 ? (n + 255) / 256 : 1024), 256, 0, st>>>(p, n, v);
    CLB_LAUNCHED("fill_with");
}

void fill_u64(clb_controls_s *ctl, u64 *p, u64 n, u64 v) {
    if (!n) return;
    k_fill_u64<<<fill_grid(ctl, n), 256, 0, ctl->stream>>>(p, n, v);
    CLB_LAUNCHED("fill_with_u64");
}

// ---------------------------------------------------------------------------
// Names of the compiled programs and kernels, for the kernel<> shim.  The
// program names are the reference's .cl file stems (src/cl/**), mapped to the
// HIP kernels that replace them.
struct KName { const char *program, *kernel; };
static const KName kRegistry[] = {
    {"prefix_sum", "scan_blelloch"}, {"prefix_sum", "update_pref_sum"},
    {"initialization", "fill_with"},
    {"count_workload", "count_workload"},
    {"prepare_positions", "prepare_for_shift_empty_rows"},
    {"prepare_positions", "prepare_array_for_rows_positions"},
    {"set_positions", "set_positions_pointers_and_rows"},
    {"set_positions", "set_positions_rows"},
    {"set_positions", "set_positions_pointers_and_rows_csr"},
    {"hash/hash_pwarp", "hash_symbolic_pwarp"}, {"hash/hash_pwarp", "hash_numeric_pwarp"},
    {"hash/hash_tb", "hash_symbolic_tb"}, {"hash/hash_tb", "hash_numeric_tb"},
    {"hash/hash_global", "hash_symbolic_global"}, {"hash/hash_global", "hash_numeric_global"},
    {"merge_path", "merge"},
    {"coo_reduce_duplicates", "init_duplicates"}, {"coo_reduce_duplicates", "reduce_duplicates_tb"},
    {"coo_reduce_duplicates", "shift_tb"},
    {"csr_addition", "addition_symbolic"}, {"csr_addition", "addition_numeric"},
    {"dcsr_kronecker", "count_nnz_per_row"}, {"dcsr_kronecker", "calculate_kronecker_product"},
    {"dscr_to_coo", "dscr_to_coo"},
    {"conversions", "dcsr_to_csr_set_size"},
    {"coo_bitonic_sort", "local_bitonic_begin"}, {"coo_bitonic_sort", "bitonic_global_step"},
    {"coo_bitonic_sort", "bitonic_local_endings"},
};

bool kernel_registered(const std::string &program, const std::string &kernel) {
    for (const auto &k : kRegistry)
        if (program == k.program && kernel == k.kernel) return true;
    return false;
}

static bool program_registered(const std::string &program) {
    for (const auto &k : kRegistry)
        if (program == k.program) return true;
    return false;
}

static void default_tuning(clb_tuning *t) {
    std::memset(t, 0, sizeof(*t));
    t->hash_max_slots = 16384;
    t->bitmap_min_products = 8192;
    t->bitmap_slot_ratio = 1;
    t->tiny_max_products = 32;
    t->learned_margin_pct = 25;
    // the one pass sizes bitmap rows by the prediction too (R-MAT 20 A * A:
    // 3 row blocks instead of 5, 62.4 ms per step instead of 64.8)
    t->reserved[CLB_TUNE_BITMAP_PRED_CAPS] = 1;
}

void set_last_error(const std::string &m) { g_last_error = m; }
const char *last_error() { return g_last_error.c_str(); }

}  // namespace clb

using namespace clb;

// Runs `body`, mapping exceptions to status codes.
template <typename F>
static clb_status guarded(F &&body) {
    try {
        body();
        return CLB_SUCCESS;
    } catch (const Error &e) {
        set_last_error(e.what());
        return static_cast<clb_status>(e.status) + 1;
    } catch (const std::bad_alloc &) {
        set_last_error("host allocation failed");
        return static_cast<clb_status>(CREATE_BUFFER_ERROR) + 1;
    } catch (const std::exception &e) {
        set_last_error(e.what());
        return static_cast<clb_status>(INVALID_VALUE) + 1;
    }
}

extern "C" {

const char *clb_last_error(void) { return clb::last_error(); }

clb_status clb_device_count(uint32_t *count) {
    return guarded([&] {
        int n = 0;
        hipError_t e = hipGetDeviceCount(&n);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            n = 0;
        }
        *count = static_cast<uint32_t>(n);
    });
}

clb_status clb_device_info_get(uint32_t device_id, clb_device_info *info) {
    return guarded([&] {
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || device_id >= static_cast<uint32_t>(n)) {
            (void)hipGetLastError();
            raise(INITIALIZATION_ERROR, "No such device: " + std::to_string(device_id));
        }
        hipDeviceProp_t p;
        CLB_HIP(hipGetDeviceProperties(&p, device_id), INITIALIZATION_ERROR);
        std::memset(info, 0, sizeof(*info));
        std::snprintf(info->name, sizeof(info->name), "%s", p.name);
        std::snprintf(info->arch, sizeof(info->arch), "%s", p.gcnArchName);
        info->compute_units = p.multiProcessorCount;
        info->max_wg_size = p.maxThreadsPerBlock;
        info->lds_bytes = static_cast<uint32_t>(p.sharedMemPerBlock);
        info->global_mem_bytes = p.totalGlobalMem;
        info->clock_khz = p.clockRate;
        info->mem_clock_khz = p.memoryClockRate;
        info->mem_bus_width = p.memoryBusWidth;
    });
}

clb_status clb_controls_create(uint32_t platform_id, uint32_t device_id, clb_controls *out) {
    return guarded([&] {
        *out = nullptr;
        // env.cpp:15-20: an unknown platform id raises INITIALIZATION_ERROR.
        if (platform_id != 0)
            raise(INITIALIZATION_ERROR, "No such platform: " + std::to_string(platform_id) +
                                            ". Run show_devices() to enumerate available platforms and devices.");
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess) {
            (void)hipGetLastError();
            n = 0;
        }
        // env.cpp:22-27: an unknown device id raises INITIALIZATION_ERROR.
        if (device_id >= static_cast<uint32_t>(n))
            raise(INITIALIZATION_ERROR, "No such device: " + std::to_string(device_id) +
                                            ". Run show_devices() to enumerate available platforms and devices.");
        auto *c = new clb_controls_s();
        c->device = static_cast<int>(device_id);
        try {
            CLB_HIP(hipSetDevice(c->device), CREATE_CONTROLS_ERROR);
            hipDeviceProp_t p;
            CLB_HIP(hipGetDeviceProperties(&p, c->device), CREATE_CONTROLS_ERROR);
            c->cu_count = p.multiProcessorCount;
            c->max_wg_size = p.maxThreadsPerBlock;
            c->lds_bytes = static_cast<u32>(p.sharedMemPerBlock);
            // default priority, non-blocking with respect to the null stream
            CLB_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), CREATE_CONTROLS_ERROR);
            for (int k = 0; k < clb_controls_s::NSIDE; ++k) {
                CLB_HIP(hipStreamCreateWithFlags(&c->side[k], hipStreamNonBlocking), CREATE_CONTROLS_ERROR);
                CLB_HIP(hipEventCreateWithFlags(&c->ev_join[k], hipEventDisableTiming), CREATE_CONTROLS_ERROR);
            }
            CLB_HIP(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming), CREATE_CONTROLS_ERROR);
            for (int k = 0; k < clb_controls_s::NTAIL; ++k)
                CLB_HIP(hipStreamCreateWithFlags(&c->tail[k], hipStreamNonBlocking), CREATE_CONTROLS_ERROR);
            CLB_HIP(hipHostMalloc(reinterpret_cast<void **>(&c->pinned), 4096, hipHostMallocDefault),
                    CREATE_CONTROLS_ERROR);
            default_tuning(&c->tuning);
        } catch (...) {
            for (int k = 0; k < clb_controls_s::NSIDE; ++k) {
                if (c->side[k]) (void)hipStreamDestroy(c->side[k]);
                if (c->ev_join[k]) (void)hipEventDestroy(c->ev_join[k]);
            }
            if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
            for (int k = 0; k < clb_controls_s::NTAIL; ++k)
                if (c->tail[k]) (void)hipStreamDestroy(c->tail[k]);
            if (c->stream) (void)hipStreamDestroy(c->stream);
            delete c;
            throw;
        }
        *out = c;
    });
}

clb_status clb_controls_destroy(clb_controls ctl) {
    return guarded([&] {
        if (!ctl) return;
        (void)hipStreamSynchronize(ctl->stream);
        clb::learned_destroy_state(ctl);
        (void)hipHostFree(ctl->pinned);
        for (int k = 0; k < clb_controls_s::NSIDE; ++k) {
            (void)hipStreamSynchronize(ctl->side[k]);
            (void)hipStreamDestroy(ctl->side[k]);
            (void)hipEventDestroy(ctl->ev_join[k]);
        }
        (void)hipEventDestroy(ctl->ev_fork);
        for (int k = 0; k < clb_controls_s::NTAIL; ++k) {
            (void)hipStreamSynchronize(ctl->tail[k]);
            (void)hipStreamDestroy(ctl->tail[k]);
        }
        (void)hipStreamDestroy(ctl->stream);
        delete ctl;
    });
}

clb_status clb_controls_sync(clb_controls ctl) { return guarded([&] { sync(ctl); }); }

clb_status clb_controls_max_wg_size(clb_controls ctl, uint32_t *out) {
    return guarded([&] { *out = ctl->max_wg_size; });
}

clb_status clb_get_tuning(clb_controls ctl, clb_tuning *out) { return guarded([&] { *out = ctl->tuning; }); }

clb_status clb_set_tuning(clb_controls ctl, const clb_tuning *t) {
    return guarded([&] {
        const u32 s = t->hash_max_slots;
        if (s < 256 || s > 16384 || (s & (s - 1)))
            raise(INVALID_ARGUMENT, "hash_max_slots must be a power of two in [256, 16384]");
        if (t->bitmap_slot_ratio == 0) raise(INVALID_ARGUMENT, "bitmap_slot_ratio must be positive");
        if (t->tiny_max_products > 64) raise(INVALID_ARGUMENT, "tiny_max_products must be <= 64");
        ctl->tuning = *t;
    });
}

clb_status clb_last_stats(clb_controls ctl, clb_spgemm_stats *out) {
    return guarded([&] { *out = ctl->last; });
}

clb_status clb_pool_trim(clb_controls ctl) { return guarded([&] { pool_for(ctl->device).trim(); }); }

clb_status clb_pool_stats(clb_controls ctl, uint64_t out[6]) {
    return guarded([&] {
        if (!ctl || !out) raise(INVALID_ARGUMENT, "null argument");
        pool_for(ctl->device).stats(out);
        size_t free_b = 0, total_b = 0;
        int prev = 0;
        CLB_HIP(hipGetDevice(&prev), INVALID_ARGUMENT);
        if (prev != ctl->device) CLB_HIP(hipSetDevice(ctl->device), INVALID_ARGUMENT);
        const hipError_t e = hipMemGetInfo(&free_b, &total_b);
        if (prev != ctl->device) (void)hipSetDevice(prev);
        CLB_HIP(e, INVALID_ARGUMENT);
        out[4] = free_b;
        out[5] = total_b;
    });
}

clb_status clb_buffer_create(clb_controls ctl, uint64_t bytes, clb_buffer *out) {
    return guarded([&] { *out = buffer_new(ctl, bytes); });
}

clb_status clb_buffer_from_host(clb_controls ctl, const void *src, uint64_t bytes, clb_buffer *out) {
    return guarded([&] {
        clb_buffer_s *b = buffer_new(ctl, bytes);
        hipError_t e = hipMemcpyAsync(b->ptr, src, bytes, hipMemcpyHostToDevice, ctl->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctl->stream);
        if (e != hipSuccess) {
            buffer_release(b);
            check_hip(e, WRITE_BUFFER_ERROR, "hipMemcpyAsync(H2D)");
        }
        *out = b;
    });
}

clb_status clb_buffer_read(clb_controls ctl, clb_buffer buf, uint64_t offset, uint64_t bytes, void *dst) {
    return guarded([&] {
        if (!bytes) return;
        if (!buf || offset + bytes > buf->bytes) raise(READ_BUFFER_ERROR, "read out of buffer bounds");
        CLB_HIP(hipMemcpyAsync(dst, static_cast<char *>(buf->ptr) + offset, bytes, hipMemcpyDeviceToHost,
                               ctl->stream),
                READ_BUFFER_ERROR);
        sync(ctl);
    });
}

clb_status clb_buffer_write(clb_controls ctl, clb_buffer buf, uint64_t offset, uint64_t bytes, const void *src) {
    return guarded([&] {
        if (!bytes) return;
        if (!buf || offset + bytes > buf->bytes) raise(WRITE_BUFFER_ERROR, "write out of buffer bounds");
        CLB_HIP(hipMemcpyAsync(static_cast<char *>(buf->ptr) + offset, src, bytes, hipMemcpyHostToDevice,
                               ctl->stream),
                WRITE_BUFFER_ERROR);
        sync(ctl);
    });
}

clb_status clb_buffer_copy(clb_controls ctl, clb_buffer src, uint64_t so, clb_buffer dst, uint64_t dof,
                           uint64_t bytes) {
    return guarded([&] {
        if (!bytes) return;
        if (!src || !dst || so + bytes > src->bytes || dof + bytes > dst->bytes)
            raise(COPY_BUFFER_ERROR, "copy out of buffer bounds");
        CLB_HIP(hipMemcpyAsync(static_cast<char *>(dst->ptr) + dof, static_cast<char *>(src->ptr) + so, bytes,
                               hipMemcpyDeviceToDevice, ctl->stream),
                COPY_BUFFER_ERROR);
        sync(ctl);
    });
}

clb_status clb_buffer_retain(clb_buffer buf) { return guarded([&] { buffer_retain(buf); }); }
clb_status clb_buffer_release(clb_buffer buf) { return guarded([&] { buffer_release(buf); }); }
void *clb_buffer_device_ptr(clb_buffer buf) { return buf ? buf->ptr : nullptr; }
uint64_t clb_buffer_size(clb_buffer buf) { return buf ? buf->bytes : 0; }

clb_status clb_dcsr_release(clb_dcsr *m) {
    return guarded([&] {
        buffer_release(m->rpt);
        buffer_release(m->rows);
        buffer_release(m->cols);
        m->rpt = m->rows = m->cols = nullptr;
        m->nnz = m->nzr = 0;
    });
}

clb_status clb_coo_release(clb_coo *m) {
    return guarded([&] {
        buffer_release(m->rows);
        buffer_release(m->cols);
        m->rows = m->cols = nullptr;
        m->nnz = 0;
    });
}

clb_status clb_csr_release(clb_csr *m) {
    return guarded([&] {
        buffer_release(m->rpt);
        buffer_release(m->cols);
        m->rpt = m->cols = nullptr;
        m->nnz = 0;
    });
}

clb_status clb_dcsr_from_host(clb_controls ctl, const uint32_t *rpt, const uint32_t *rows, const uint32_t *cols,
                              uint32_t nzr, uint32_t nnz, uint32_t nrows, uint32_t ncols, clb_dcsr *out) {
    return guarded([&] {
        clb_dcsr m{};
        m.nrows = nrows;
        m.ncols = ncols;
        // matrices_conversions.cpp:94-96: no columns -> empty matrix of that shape
        if (nnz != 0) {
            if (nzr == 0) raise(INVALID_ARGUMENT, "DCSR with nonzeros must have stored rows");
            m.rpt = buffer_new(ctl, 4ull * (nzr + 1));
            m.rows = buffer_new(ctl, 4ull * nzr);
            m.cols = buffer_new(ctl, 4ull * nnz);
            hipError_t e = hipMemcpyAsync(m.rpt->ptr, rpt, 4ull * (nzr + 1), hipMemcpyHostToDevice, ctl->stream);
            if (e == hipSuccess)
                e = hipMemcpyAsync(m.rows->ptr, rows, 4ull * nzr, hipMemcpyHostToDevice, ctl->stream);
            if (e == hipSuccess)
                e = hipMemcpyAsync(m.cols->ptr, cols, 4ull * nnz, hipMemcpyHostToDevice, ctl->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(ctl->stream);
            if (e != hipSuccess) {
                clb_dcsr_release(&m);
                check_hip(e, WRITE_BUFFER_ERROR, "upload DCSR");
            }
            m.nnz = nnz;
            m.nzr = nzr;
        }
        *out = m;
    });
}

clb_status clb_csr_from_host(clb_controls ctl, const uint32_t *rpt, const uint32_t *cols, uint32_t nrows,
                             uint32_t ncols, uint32_t nnz, clb_csr *out) {
    return guarded([&] {
        clb_csr m{};
        m.nrows = nrows;
        m.ncols = ncols;
        if (nnz != 0) {  // matrices_conversions.cpp:200-202
            m.rpt = buffer_new(ctl, 4ull * (nrows + 1ull));
            m.cols = buffer_new(ctl, 4ull * nnz);
            hipError_t e =
                hipMemcpyAsync(m.rpt->ptr, rpt, 4ull * (nrows + 1ull), hipMemcpyHostToDevice, ctl->stream);
            if (e == hipSuccess)
                e = hipMemcpyAsync(m.cols->ptr, cols, 4ull * nnz, hipMemcpyHostToDevice, ctl->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(ctl->stream);
            if (e != hipSuccess) {
                clb_csr_release(&m);
                check_hip(e, WRITE_BUFFER_ERROR, "upload CSR");
            }
            m.nnz = nnz;
        }
        *out = m;
    });
}

clb_status clb_fill_with(clb_controls ctl, clb_buffer buf, uint32_t n, uint32_t value) {
    return guarded([&] {
        if (!n) return;
        if (!buf || 4ull * n > buf->bytes) raise(INVALID_ARGUMENT, "fill_with: buffer too small");
        fill_u32(ctl, u32p(buf), n, value);
        sync(ctl);
    });
}

clb_status clb_prefix_sum(clb_controls ctl, clb_buffer buf, uint32_t n, uint32_t *total) {
    return guarded([&] {
        *total = 0;
        if (!n) return;
        if (!buf || 4ull * n > buf->bytes) raise(INVALID_ARGUMENT, "prefix_sum: buffer too small");
        DevArray<u32> tot(ctl, 1);
        exclusive_scan_u32(ctl, u32p(buf), u32p(buf), n, tot.ptr);
        readback(ctl, total, tot.ptr, 4);
    });
}

clb_status clb_kernel_lookup(const char *program, const char *kernel) {
    return guarded([&] {
        std::string p = program ? program : "", k = kernel ? kernel : "";
        // kernel_cache.hpp:31-32: unknown program -> CLBOOL_NO_SUCH_PROGRAM
        if (!program_registered(p)) raise(NO_SUCH_PROGRAM, "Cannot find " + p);
        // kernel_cache.hpp:53: unknown kernel in a program -> CREATE_KERNEL_ERROR
        if (!kernel_registered(p, k)) raise(CREATE_KERNEL_ERROR, "no kernel " + k + " in program " + p);
    });
}

}  // extern "C"

// Shared by the other translation units: run a body under the same mapping.
namespace clb {
clb_status run_guarded(const std::function<void()> &body) { return guarded(body); }
}  // namespace clb
