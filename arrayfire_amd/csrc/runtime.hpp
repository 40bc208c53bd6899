// Host runtime of libafhip: the af_array handle, the device memory pool, the
// library stream, and the error machinery that maps C++ exceptions onto
// af_err codes at the C ABI (no exception crosses it).
//
// Reference counterparts: handle layer src/api/c/handle.hpp:46-90, error
// macros src/backend/common/err_common.hpp:149-200 and their mapping
// src/backend/common/err_common.cpp:125-175, memory manager
// src/backend/common/DefaultMemoryManager.cpp.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/af/defines.h"
#include "../../include/af/hip.h"

namespace afhip {

// ---------------------------------------------------------------- errors
class AfError : public std::exception {
   public:
    AfError(af_err code, std::string msg) : code_(code), msg_(std::move(msg)) {}
    af_err code() const { return code_; }
    const char *what() const noexcept override { return msg_.c_str(); }

   private:
    af_err code_;
    std::string msg_;
};

void set_last_error(const std::string &msg);
af_err process_exception();  // call inside catch(...)

#define AFHIP_THROW(code, msg) throw ::afhip::AfError((code), (msg))
#define ARG_ASSERT(INDEX, COND)                                                   \
    do {                                                                          \
        if (!(COND))                                                              \
            AFHIP_THROW(AF_ERR_ARG, std::string("Invalid argument at index ") +   \
                                        std::to_string(INDEX) + ": " #COND);      \
    } while (0)
#define DIM_ASSERT(INDEX, COND)                                                   \
    do {                                                                          \
        if (!(COND))                                                              \
            AFHIP_THROW(AF_ERR_SIZE, std::string("Invalid dimension at index ") + \
                                         std::to_string(INDEX) + ": " #COND);     \
    } while (0)
#define TYPE_ERROR(INDEX, T)                                                      \
    AFHIP_THROW(AF_ERR_TYPE, std::string("Invalid type at index ") +              \
                                 std::to_string(INDEX) + ": " + std::to_string((int)(T)))
#define CATCHALL                                   \
    catch (...) {                                  \
        return ::afhip::process_exception();       \
    }

// Launch-side check of a kernel-ABI return code.
void check_k(int rc, const char *what);
void check_hip(hipError_t e, const char *what);

// ---------------------------------------------------------------- dims
struct dim4 {
    dim_t d[4];
    dim4() : d{0, 0, 0, 0} {}
    dim4(dim_t a, dim_t b = 1, dim_t c = 1, dim_t e = 1) : d{a, b, c, e} {}
    // dim4(n, ptr): missing entries are 1 (src/backend/common/dim4.cpp:35-37)
    dim4(unsigned n, const dim_t *p) {
        for (unsigned i = 0; i < 4; ++i) d[i] = n > i ? p[i] : 1;
    }
    dim_t &operator[](int i) { return d[i]; }
    dim_t operator[](int i) const { return d[i]; }
    dim_t elements() const { return d[0] * d[1] * d[2] * d[3]; }
    // src/backend/common/dim4.cpp:48-58
    dim_t ndims() const {
        const dim_t n = elements();
        if (n == 0) return 0;
        if (n == 1) return 1;
        if (d[3] != 1) return 4;
        if (d[2] != 1) return 3;
        if (d[1] != 1) return 2;
        return 1;
    }
    bool operator==(const dim4 &o) const {
        return d[0] == o.d[0] && d[1] == o.d[1] && d[2] == o.d[2] && d[3] == o.d[3];
    }
    bool operator!=(const dim4 &o) const { return !(*this == o); }
};

size_t dtype_size(af_dtype t);  // throws AF_ERR_TYPE for unknown types
bool dtype_known(int t);

// ---------------------------------------------------------------- device
hipStream_t stream();            // the library stream (created on first use)
void sync_stream();
void ensure_device();

// Stream-ordered caching pool. Every kernel runs on the one library stream,
// so a buffer returned to the pool may be handed to the next allocation at
// once: later work on the same stream is ordered after earlier readers.
void *pool_alloc(size_t bytes);
void pool_free(void *p);
void pool_gc();
// While a graph is captured on the library stream, buffers returned to the
// pool by this thread are deferred into *keep instead (the graph still uses
// them); pool_keep(nullptr) ends that. The caller returns them later with
// pool_free.
void pool_keep(std::vector<void *> *keep);
void pool_info(size_t *alloc_bytes, size_t *alloc_buffers, size_t *lock_bytes,
               size_t *lock_buffers);

struct Buffer {
    void *ptr = nullptr;
    size_t bytes = 0;
    bool user_owned = false;  // af_device_array: hipFree on release
    bool locked = false;      // af_lock_array: never freed or pooled by the library
    ~Buffer();
};

// ---------------------------------------------------------------- arrays
struct ArrayImpl {
    af_dtype type;
    dim4 dims;
    std::shared_ptr<Buffer> buf;
    void *data() const { return buf ? buf->ptr : nullptr; }
    dim_t elements() const { return dims.elements(); }
};

ArrayImpl &get(const af_array a);  // validates the handle
af_array make_handle(af_dtype type, const dim4 &dims);  // allocates storage
af_array make_handle_shared(const ArrayImpl &src);     // shares storage
ArrayImpl *impl_of(af_array a);

// Scratch device memory owned by a shared_ptr (returned to the pool).
std::shared_ptr<Buffer> scratch(size_t bytes);

// Releases a handle made by make_handle; unknown handles are ignored.
void destroy_handle(af_array a) noexcept;

// Owns a handle made by make_handle until release() hands it to the caller:
// an exception thrown before then (a failed launch, a type error) releases
// the handle and its storage instead of leaking them.
class NewHandle {
   public:
    explicit NewHandle(af_array h) : h_(h) {}
    NewHandle(const NewHandle &) = delete;
    NewHandle &operator=(const NewHandle &) = delete;
    ~NewHandle() {
        if (h_) destroy_handle(h_);
    }
    af_array get() const { return h_; }
    void *data() const;
    af_array release() {
        af_array h = h_;
        h_ = nullptr;
        return h;
    }

   private:
    af_array h_;
};

}  // namespace afhip
