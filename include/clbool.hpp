// clbool.hpp -- drop-in C++ API of clBool (include/clbool.hpp of the
// reference) over the MI355X C-ABI of clbool_c.h.
//
// Code written against the reference's public headers compiles against this
// one: the same namespaces, class names, constructors, getters and free
// functions (src/core/controls.hpp, matrix_{base,dcsr,coo,csr}.hpp,
// cpu_matrices.hpp, src/common/env.hpp, matrices_conversions.hpp,
// cl_operations.hpp, src/dcsr/dcsr.hpp, src/coo/coo.hpp, src/csr/csr.hpp,
// src/core/kernel.hpp, src/core/error.hpp).  Differences: device buffers are
// clbool::buffer (reference-counted HIP allocations) instead of cl::Buffer,
// Controls wraps a HIP device and stream instead of an OpenCL context and
// queues, and kernel<> only checks names (kernels are compiled ahead of time).
// Provided beside the SpGEMM path: dcsr::matrix_multiplication (the
// expand-sort-compress SpGEMM), dcsr::transpose, submatrix and reduce.  Not
// provided: coo::kronecker_product.  Link with -lclbool_amd.
#pragma once

#include <algorithm>
#include <cstdint>
#include <exception>
#include <iostream>
#include <memory>
#include <sstream>
#include <string>
#include <utility>
#include <vector>

#include "clbool_c.h"

namespace clbool {

// ---- errors: src/core/error.hpp:130-213 ------------------------------------
enum Status {
    CLBOOL_NO_SUCH_PROGRAM,
    CLBOOL_CREATE_KERNEL_ERROR,
    CLBOOL_BUILD_PROGRAM_ERROR,
    CLBOOL_INCOMPLETE_KERNEL,
    CLBOOL_RUN_KERNEL_FAILURE,
    CLBOOL_EVENT_WAITING_ERROR,
    CLBOOL_CREATE_BUFFER_ERROR,
    CLBOOL_COPY_BUFFER_ERROR,
    CLBOOL_READ_BUFFER_ERROR,
    CLBOOL_WRITE_BUFFER_ERROR,
    CLBOOL_CREATE_CONTROLS_ERROR,
    CLBOOL_INVALID_ARGUMENT,
    CLBOOL_INVALID_VALUE,
    CLBOOL_INITIALIZATION_ERROR
};

inline const char *status_name(Status s) {
    static const char *names[] = {"CLBOOL_NO_SUCH_PROGRAM",      "CLBOOL_CREATE_KERNEL_ERROR",
                                  "CLBOOL_BUILD_PROGRAM_ERROR",  "CLBOOL_INCOMPLETE_KERNEL",
                                  "CLBOOL_RUN_KERNEL_FAILURE",   "CLBOOL_EVENT_WAITING_ERROR",
                                  "CLBOOL_CREATE_BUFFER_ERROR",  "CLBOOL_COPY_BUFFER_ERROR",
                                  "CLBOOL_READ_BUFFER_ERROR",    "CLBOOL_WRITE_BUFFER_ERROR",
                                  "CLBOOL_CREATE_CONTROLS_ERROR", "CLBOOL_INVALID_ARGUMENT",
                                  "CLBOOL_INVALID_VALUE",        "CLBOOL_INITIALIZATION_ERROR"};
    return (s >= 0 && s <= CLBOOL_INITIALIZATION_ERROR) ? names[s] : "CLBOOL_UNKNOWN";
}

inline std::ostream &operator<<(std::ostream &out, const Status s) { return out << status_name(s); }

class Exception : public std::exception {
    std::string message_;
    Status status_;
    std::string file_, function_;
    size_t line_;
    mutable std::string what_;

  public:
    Exception(std::string message, Status status, std::string file = "", std::string function = "", size_t line = 0)
        : message_(std::move(message)), status_(status), file_(std::move(file)), function_(std::move(function)),
          line_(line) {}
    const char *what() const noexcept override {
        std::stringstream s;
        s << "[" << status_ << "] \"" << message_ << "\"" << std::endl
          << file_ << ": line: " << line_ << " function: " << function_ << std::endl;
        what_ = s.str();
        return what_.c_str();
    }
    std::string get_message() const { return message_; }
    Status status() const { return status_; }
};

#define CLB_RAISE(msg, status)                                                            \
    do {                                                                                  \
        throw ::clbool::Exception(msg, status, __FILE__, __FUNCTION__, __LINE__);         \
    } while (0)
#define CLB_CHECK(cond, msg, status)                                                      \
    do {                                                                                  \
        if (!(cond)) throw ::clbool::Exception(msg, status, __FILE__, __FUNCTION__, __LINE__); \
    } while (0)
// error.hpp:257-265: a failure inside `run` is reported as RUN_KERNEL_FAILURE
#define CLB_RUN(run)                                                                      \
    try {                                                                                 \
        run;                                                                              \
    } catch (const ::clbool::Exception &e) {                                              \
        throw ::clbool::Exception(e.get_message(), ::clbool::CLBOOL_RUN_KERNEL_FAILURE, __FILE__, __FUNCTION__, \
                                  __LINE__);                                              \
    }

namespace details {
inline void check(clb_status rc, const char *where) {
    if (rc != CLB_SUCCESS) throw Exception(clb_last_error(), static_cast<Status>(rc - 1), "libclbool_amd", where, 0);
}
}  // namespace details
#define CLBOOL_CALL(expr) ::clbool::details::check((expr), #expr)

// ---- device buffer (replaces cl::Buffer) ------------------------------------
class buffer {
    clb_buffer h_ = nullptr;

  public:
    buffer() = default;
    explicit buffer(clb_buffer adopt) : h_(adopt) {}
    buffer(const buffer &o) : h_(o.h_) {
        if (h_) clb_buffer_retain(h_);
    }
    buffer(buffer &&o) noexcept : h_(o.h_) { o.h_ = nullptr; }
    buffer &operator=(buffer o) noexcept {
        std::swap(h_, o.h_);
        return *this;
    }
    ~buffer() {
        if (h_) clb_buffer_release(h_);
    }
    clb_buffer get() const { return h_; }
    clb_buffer release_handle() {
        clb_buffer h = h_;
        h_ = nullptr;
        return h;
    }
    size_t size() const { return h_ ? static_cast<size_t>(clb_buffer_size(h_)) : 0; }
    void *device_ptr() const { return h_ ? clb_buffer_device_ptr(h_) : nullptr; }
    explicit operator bool() const { return h_ != nullptr; }
};

// ---- controls: src/core/controls.hpp:12-28, src/common/env.hpp ---------------
struct Controls {
    std::shared_ptr<clb_controls_s> handle;
    uint32_t max_wg_size = 1024;
    const uint32_t block_size = uint32_t(256);

    Controls(uint32_t platform_id, uint32_t device_id) {
        clb_controls c = nullptr;
        CLBOOL_CALL(clb_controls_create(platform_id, device_id, &c));
        handle = std::shared_ptr<clb_controls_s>(c, [](clb_controls p) { clb_controls_destroy(p); });
        CLBOOL_CALL(clb_controls_max_wg_size(c, &max_wg_size));
    }
    Controls(const Controls &o) : handle(o.handle), max_wg_size(o.max_wg_size) {}
    clb_controls get() const { return handle.get(); }
    void sync() const { CLBOOL_CALL(clb_controls_sync(handle.get())); }
};

inline Controls create_controls(uint32_t platform_id = 0, uint32_t device_id = 0) {
    return Controls(platform_id, device_id);
}

inline void show_devices() {
    uint32_t n = 0;
    CLBOOL_CALL(clb_device_count(&n));
    std::cout << "platform id: 0 (HIP)\n";
    for (uint32_t i = 0; i < n; ++i) {
        clb_device_info info;
        CLBOOL_CALL(clb_device_info_get(i, &info));
        std::cout << "  device " << i << ": " << info.arch << ", " << info.compute_units << " CUs, "
                  << info.global_mem_bytes / (1ull << 30) << " GiB\n";
    }
}

// ---- host matrices: src/core/cpu_matrices.hpp --------------------------------
using coordinates = std::pair<uint32_t, uint32_t>;
using matrix_coo_cpu_pairs = std::vector<coordinates>;
using cpu_buffer = std::vector<uint32_t>;

class matrix_dcsr_cpu {
    cpu_buffer _rpt, _rows, _cols;

  public:
    matrix_dcsr_cpu(cpu_buffer rpt, cpu_buffer rows, cpu_buffer cols)
        : _rpt(std::move(rpt)), _rows(std::move(rows)), _cols(std::move(cols)) {}
    matrix_dcsr_cpu() = default;
    cpu_buffer &rpt() { return _rpt; }
    cpu_buffer &rows() { return _rows; }
    cpu_buffer &cols() { return _cols; }
    const cpu_buffer &rpt() const { return _rpt; }
    const cpu_buffer &rows() const { return _rows; }
    const cpu_buffer &cols() const { return _cols; }
};

class matrix_coo_cpu {
    cpu_buffer _rows, _cols;

  public:
    matrix_coo_cpu(cpu_buffer rows, cpu_buffer cols) : _rows(std::move(rows)), _cols(std::move(cols)) {}
    matrix_coo_cpu() = default;
    cpu_buffer &rows() { return _rows; }
    cpu_buffer &cols() { return _cols; }
    const cpu_buffer &rows() const { return _rows; }
    const cpu_buffer &cols() const { return _cols; }
};

class matrix_csr_cpu {
    cpu_buffer _rpt, _cols;
    uint32_t _nrows = 0, _ncols = 0;

  public:
    matrix_csr_cpu(cpu_buffer rpt, cpu_buffer cols, uint32_t nrows, uint32_t ncols)
        : _rpt(std::move(rpt)), _cols(std::move(cols)), _nrows(nrows), _ncols(ncols) {}
    matrix_csr_cpu(uint32_t nrows, uint32_t ncols) : _nrows(nrows), _ncols(ncols) {}
    matrix_csr_cpu() = default;
    cpu_buffer &rpt() { return _rpt; }
    cpu_buffer &cols() { return _cols; }
    const cpu_buffer &rpt() const { return _rpt; }
    const cpu_buffer &cols() const { return _cols; }
    uint32_t nrows() const { return _nrows; }
    uint32_t ncols() const { return _ncols; }
};

// ---- device matrices: src/core/matrix_*.hpp ---------------------------------
namespace details {
class matrix_base {
  public:
    using index_type = uint32_t;

  protected:
    index_type _nrows = 0, _ncols = 0, _nnz = 0;

  public:
    matrix_base() = default;
    matrix_base(index_type n_rows, index_type n_cols, index_type n_entities)
        : _nrows(n_rows), _ncols(n_cols), _nnz(n_entities) {}
    index_type nrows() const { return _nrows; }
    index_type ncols() const { return _ncols; }
    index_type nnz() const { return _nnz; }
    bool empty() const { return _nnz == 0; }
};
}  // namespace details

class matrix_dcsr : public details::matrix_base {
    buffer _rpt_gpu, _rows_gpu, _cols_gpu;
    uint32_t _nzr = 0;

  public:
    matrix_dcsr() = default;
    matrix_dcsr(buffer rpt_gpu, buffer rows_gpu, buffer cols_gpu, uint32_t n_rows, uint32_t n_cols, uint32_t nnz,
                uint32_t nzr)
        : details::matrix_base(n_rows, n_cols, nnz), _rpt_gpu(std::move(rpt_gpu)), _rows_gpu(std::move(rows_gpu)),
          _cols_gpu(std::move(cols_gpu)), _nzr(nzr) {}
    matrix_dcsr(uint32_t n_rows, uint32_t n_cols) : details::matrix_base(n_rows, n_cols, 0) {}
    const buffer &rpt_gpu() const { return _rpt_gpu; }
    const buffer &rows_gpu() const { return _rows_gpu; }
    const buffer &cols_gpu() const { return _cols_gpu; }
    buffer &rpt_gpu() { return _rpt_gpu; }
    buffer &rows_gpu() { return _rows_gpu; }
    buffer &cols_gpu() { return _cols_gpu; }
    const uint32_t &nzr() const { return _nzr; }

    clb_dcsr c() const {
        return clb_dcsr{_rpt_gpu.get(), _rows_gpu.get(), _cols_gpu.get(), _nrows, _ncols, _nnz, _nzr};
    }
    static matrix_dcsr adopt(const clb_dcsr &m) {
        return matrix_dcsr(buffer(m.rpt), buffer(m.rows), buffer(m.cols), m.nrows, m.ncols, m.nnz, m.nzr);
    }
};

class matrix_csr : public details::matrix_base {
    buffer _rpt_gpu, _cols_gpu;

  public:
    matrix_csr() = default;
    matrix_csr(Controls &, uint32_t nrows, uint32_t ncols, uint32_t nvals) : matrix_base(nrows, ncols, nvals) {}
    matrix_csr(buffer rpt_gpu, buffer cols_gpu, uint32_t nrows, uint32_t ncols, uint32_t nvals)
        : matrix_base(nrows, ncols, nvals), _rpt_gpu(std::move(rpt_gpu)), _cols_gpu(std::move(cols_gpu)) {}
    matrix_csr(uint32_t nrows, uint32_t ncols) : details::matrix_base(nrows, ncols, 0) {}
    const buffer &rpt_gpu() const { return _rpt_gpu; }
    const buffer &cols_gpu() const { return _cols_gpu; }

    clb_csr c() const { return clb_csr{_rpt_gpu.get(), _cols_gpu.get(), _nrows, _ncols, _nnz}; }
    static matrix_csr adopt(const clb_csr &m) {
        return matrix_csr(buffer(m.rpt), buffer(m.cols), m.nrows, m.ncols, m.nnz);
    }
};

class matrix_coo : public details::matrix_base {
    buffer _rows, _cols;

  public:
    matrix_coo() = default;
    matrix_coo(index_type nrows, index_type ncols) : matrix_base(nrows, ncols, 0) {}
    matrix_coo(buffer &rows, buffer &cols, index_type nrows, index_type ncols, index_type nnz)
        : matrix_base(nrows, ncols, nnz), _rows(rows), _cols(cols) {}
    // matrix_coo.cpp:17-47
    matrix_coo(Controls &controls, const index_type *rows_indices, const index_type *cols_indices, index_type nrows,
               index_type ncols, index_type nnz, bool sorted = true, bool noDuplicates = true) {
        clb_coo m{};
        CLBOOL_CALL(clb_coo_from_host(controls.get(), rows_indices, cols_indices, nrows, ncols, nnz, sorted,
                                      noDuplicates, &m));
        *this = adopt(m);
    }
    // matrix_coo.cpp:50-68
    matrix_coo(Controls &controls, index_type nrows, index_type ncols, index_type nnz, buffer &rows, buffer &cols,
               bool sorted = true, bool noDuplicates = true) {
        clb_coo m{};
        CLBOOL_CALL(clb_coo_from_buffers(controls.get(), nrows, ncols, nnz, rows.get(), cols.get(), sorted,
                                         noDuplicates, &m));
        *this = adopt(m);
    }
    const buffer &rows_gpu() const { return _rows; }
    const buffer &cols_gpu() const { return _cols; }
    buffer &rows_gpu() { return _rows; }
    buffer &cols_gpu() { return _cols; }

    clb_coo c() const { return clb_coo{_rows.get(), _cols.get(), _nrows, _ncols, _nnz}; }
    static matrix_coo adopt(const clb_coo &m) {
        matrix_coo r(m.nrows, m.ncols);
        r._rows = buffer(m.rows);
        r._cols = buffer(m.cols);
        r._nnz = m.nnz;
        return r;
    }
};

// ---- primitives: src/common/cl_operations.hpp --------------------------------
inline void prefix_sum(Controls &controls, buffer &array, uint32_t &total_sum, uint32_t array_size) {
    CLBOOL_CALL(clb_prefix_sum(controls.get(), array.get(), array_size, &total_sum));
}
inline void fill_with(Controls &controls, buffer &array, uint32_t size, uint32_t value) {
    CLBOOL_CALL(clb_fill_with(controls.get(), array.get(), size, value));
}
inline void fill_with_zeroes(Controls &controls, buffer &array, uint32_t size) { fill_with(controls, array, size, 0); }

// ---- kernel<>: src/core/kernel.hpp (names only; compiled ahead of time) ------
template <typename... Args>
class kernel {
    std::string _program_name, _kernel_name;
    uint32_t _block_size = 0, _req_work_size = 0;
    bool _async = false;

  public:
    kernel(std::string program_name, std::string kernel_name) noexcept
        : _program_name(std::move(program_name)), _kernel_name(std::move(kernel_name)) {}
    kernel &set_kernel_name(std::string k) {
        _kernel_name = std::move(k);
        return *this;
    }
    kernel &set_block_size(uint32_t b) {
        _block_size = b;
        return *this;
    }
    kernel &set_work_size(uint32_t w) {
        _req_work_size = w;
        return *this;
    }
    template <typename OptionType>
    kernel &add_option(std::string, const OptionType &) {
        return *this;
    }
    kernel &set_async(bool a) {
        _async = a;
        return *this;
    }
    // Looks the kernel up; unknown names throw as in kernel_cache.hpp:31-53.
    // The library's kernels are launched by its operations, not through here.
    void run(Controls &, Args...) {
        CLBOOL_CALL(clb_kernel_lookup(_program_name.c_str(), _kernel_name.c_str()));
        CLB_RAISE("kernel " + _kernel_name + " of " + _program_name +
                      " is compiled ahead of time and launched by the library's operations",
                  CLBOOL_INVALID_VALUE);
    }
};

// ---- conversions: src/common/matrices_conversions.hpp -----------------------
inline matrix_coo dcsr_to_coo_shallow(Controls &controls, matrix_dcsr &a) {
    clb_dcsr in = a.c();
    clb_coo m{};
    CLBOOL_CALL(clb_dcsr_to_coo_shallow(controls.get(), &in, &m));
    return matrix_coo::adopt(m);
}
inline matrix_coo dcsr_to_coo_deep(Controls &controls, const matrix_dcsr &a) {
    clb_dcsr in = a.c();
    clb_coo m{};
    CLBOOL_CALL(clb_dcsr_to_coo_deep(controls.get(), &in, &m));
    return matrix_coo::adopt(m);
}
inline matrix_dcsr coo_to_dcsr_shallow(Controls &controls, const matrix_coo &a) {
    clb_coo in = a.c();
    clb_dcsr m{};
    CLBOOL_CALL(clb_coo_to_dcsr_shallow(controls.get(), &in, &m));
    return matrix_dcsr::adopt(m);
}
inline matrix_dcsr matrix_dcsr_from_cpu(Controls &controls, const matrix_dcsr_cpu &m, uint32_t size) {
    clb_dcsr out{};
    CLBOOL_CALL(clb_dcsr_from_host(controls.get(), m.rpt().data(), m.rows().data(), m.cols().data(),
                                   static_cast<uint32_t>(m.rows().size()), static_cast<uint32_t>(m.cols().size()),
                                   size, size, &out));
    return matrix_dcsr::adopt(out);
}
inline matrix_dcsr_cpu matrix_dcsr_from_gpu(Controls &controls, matrix_dcsr &m) {
    if (m.empty()) return matrix_dcsr_cpu();
    cpu_buffer rpt(m.nzr() + 1), rows(m.nzr()), cols(m.nnz());
    CLBOOL_CALL(clb_buffer_read(controls.get(), m.rpt_gpu().get(), 0, 4ull * rpt.size(), rpt.data()));
    CLBOOL_CALL(clb_buffer_read(controls.get(), m.rows_gpu().get(), 0, 4ull * rows.size(), rows.data()));
    CLBOOL_CALL(clb_buffer_read(controls.get(), m.cols_gpu().get(), 0, 4ull * cols.size(), cols.data()));
    return matrix_dcsr_cpu(rpt, rows, cols);
}
inline matrix_coo_cpu matrix_coo_from_gpu(Controls &controls, matrix_coo &m) {
    if (m.empty()) return matrix_coo_cpu();
    cpu_buffer rows(m.nnz()), cols(m.nnz());
    CLBOOL_CALL(clb_buffer_read(controls.get(), m.rows_gpu().get(), 0, 4ull * rows.size(), rows.data()));
    CLBOOL_CALL(clb_buffer_read(controls.get(), m.cols_gpu().get(), 0, 4ull * cols.size(), cols.data()));
    return matrix_coo_cpu(rows, cols);
}
inline matrix_csr_cpu matrix_csr_from_gpu(Controls &controls, const matrix_csr &m) {
    if (m.empty()) return matrix_csr_cpu(m.nrows(), m.ncols());
    cpu_buffer rpt(m.nrows() + 1), cols(m.nnz());
    CLBOOL_CALL(clb_buffer_read(controls.get(), m.rpt_gpu().get(), 0, 4ull * rpt.size(), rpt.data()));
    CLBOOL_CALL(clb_buffer_read(controls.get(), m.cols_gpu().get(), 0, 4ull * cols.size(), cols.data()));
    return matrix_csr_cpu(rpt, cols, m.nrows(), m.ncols());
}
// matrices_conversions.cpp:158-177
inline matrix_csr_cpu csr_cpu_from_pairs(const matrix_coo_cpu_pairs &mat, uint32_t m, uint32_t n) {
    cpu_buffer rpt(m + 1), cols(mat.size());
    rpt[m] = static_cast<uint32_t>(cols.size());
    size_t ptr = 0, j = 0;
    for (uint32_t i = 0; i < m; ++i) {
        rpt[i] = static_cast<uint32_t>(ptr);
        for (; j < mat.size() && mat[j].first == i; ++j, ++ptr) cols[j] = mat[j].second;
    }
    return matrix_csr_cpu(std::move(rpt), std::move(cols), m, n);
}
inline matrix_csr_cpu csr_cpu_from_coo_cpu(const matrix_coo_cpu &mat, uint32_t m, uint32_t n) {
    matrix_coo_cpu_pairs p(mat.rows().size());
    for (size_t i = 0; i < p.size(); ++i) p[i] = {mat.rows()[i], mat.cols()[i]};
    return csr_cpu_from_pairs(p, m, n);
}
inline matrix_csr csr_from_cpu(Controls &controls, const matrix_csr_cpu &m) {
    clb_csr out{};
    CLBOOL_CALL(clb_csr_from_host(controls.get(), m.rpt().data(), m.cols().data(), m.nrows(), m.ncols(),
                                  static_cast<uint32_t>(m.cols().size()), &out));
    return matrix_csr::adopt(out);
}
inline matrix_dcsr csr_to_dcsr(Controls &controls, const matrix_csr &m) {
    clb_csr in = m.c();
    clb_dcsr out{};
    CLBOOL_CALL(clb_csr_to_dcsr(controls.get(), &in, &out));
    return matrix_dcsr::adopt(out);
}
inline matrix_csr dcsr_to_csr(Controls &controls, const matrix_dcsr &m) {
    clb_dcsr in = m.c();
    clb_csr out{};
    CLBOOL_CALL(clb_dcsr_to_csr(controls.get(), &in, &out));
    return matrix_csr::adopt(out);
}

// ---- operations: src/dcsr/dcsr.hpp, src/coo/coo.hpp, src/csr/csr.hpp -------
namespace dcsr {
inline void matrix_multiplication_hash(Controls &controls, matrix_dcsr &matrix_out, const matrix_dcsr &a,
                                       const matrix_dcsr &b) {
    clb_dcsr ca = a.c(), cbm = b.c(), out{};
    CLBOOL_CALL(clb_dcsr_spgemm_hash(controls.get(), &ca, &cbm, &out));
    matrix_out = matrix_dcsr::adopt(out);  // the output may alias an input
}
// dcsr::matrix_multiplication (src/dcsr/dcsr.hpp:28-31): the same product
// by expand-sort-compress (bit-identical output).
inline void matrix_multiplication(Controls &controls, matrix_dcsr &matrix_out, const matrix_dcsr &a,
                                  const matrix_dcsr &b) {
    clb_dcsr ca = a.c(), cbm = b.c(), out{};
    CLBOOL_CALL(clb_dcsr_spgemm_esc(controls.get(), &ca, &cbm, &out));
    matrix_out = matrix_dcsr::adopt(out);
}
// The same product through the learned one-pass mode (bit-identical output).
inline void matrix_multiplication_learned(Controls &controls, matrix_dcsr &matrix_out, const matrix_dcsr &a,
                                          const matrix_dcsr &b) {
    clb_dcsr ca = a.c(), cbm = b.c(), out{};
    clb_spgemm_opts o{CLB_SPGEMM_LEARNED, 0, 0, 0};
    CLBOOL_CALL(clb_dcsr_spgemm(controls.get(), &ca, &cbm, &o, &out, nullptr));
    matrix_out = matrix_dcsr::adopt(out);
}
inline void kronecker_product(Controls &controls, matrix_dcsr &matrix_c, const matrix_dcsr &matrix_a,
                              const matrix_dcsr &matrix_b) {
    clb_dcsr ca = matrix_a.c(), cbm = matrix_b.c(), out{};
    CLBOOL_CALL(clb_dcsr_kronecker(controls.get(), &ca, &cbm, &out));
    matrix_c = matrix_dcsr::adopt(out);
}
inline void matrix_addition(Controls &controls, matrix_dcsr &matrix_out, const matrix_dcsr &a, const matrix_dcsr &b) {
    clb_dcsr ca = a.c(), cbm = b.c(), out{};
    CLBOOL_CALL(clb_dcsr_addition(controls.get(), &ca, &cbm, &out));
    matrix_out = matrix_dcsr::adopt(out);
}
inline void transpose(Controls &controls, matrix_dcsr &matrix_out, const matrix_dcsr &matrix_in) {
    clb_dcsr ci = matrix_in.c(), out{};
    CLBOOL_CALL(clb_dcsr_transpose(controls.get(), &ci, &out));
    matrix_out = matrix_dcsr::adopt(out);
}
inline void submatrix(Controls &controls, matrix_dcsr &matrix_out, const matrix_dcsr &matrix_in, uint32_t i,
                      uint32_t j, uint32_t nrows, uint32_t ncols) {
    clb_dcsr ci = matrix_in.c(), out{};
    CLBOOL_CALL(clb_dcsr_submatrix(controls.get(), &ci, i, j, nrows, ncols, &out));
    matrix_out = matrix_dcsr::adopt(out);
}
inline void reduce(Controls &controls, matrix_dcsr &matrix_out, const matrix_dcsr &matrix_in) {
    clb_dcsr ci = matrix_in.c(), out{};
    CLBOOL_CALL(clb_dcsr_reduce(controls.get(), &ci, &out));
    matrix_out = matrix_dcsr::adopt(out);
}
}  // namespace dcsr

namespace coo {
inline void matrix_addition(Controls &controls, matrix_coo &matrix_out, const matrix_coo &a, const matrix_coo &b) {
    clb_coo ca = a.c(), cbm = b.c(), out{};
    CLBOOL_CALL(clb_coo_addition(controls.get(), &ca, &cbm, &out));
    matrix_out = matrix_coo::adopt(out);
}
inline void sort_arrays(Controls &controls, buffer &rows_gpu, buffer &cols_gpu, uint32_t n) {
    CLBOOL_CALL(clb_coo_sort_arrays(controls.get(), rows_gpu.get(), cols_gpu.get(), n));
}
}  // namespace coo

namespace csr {
inline void matrix_addition(Controls &controls, matrix_csr &c, const matrix_csr &a, const matrix_csr &b) {
    clb_csr ca = a.c(), cbm = b.c(), out{};
    CLBOOL_CALL(clb_csr_addition(controls.get(), &ca, &cbm, &out));
    c = matrix_csr::adopt(out);
}
}  // namespace csr

}  // namespace clbool
