// Host runtime of libafhip: device/stream set-up, caching memory pool,
// handle registry and error state.
#include "runtime.hpp"

#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <unordered_map>
#include <unordered_set>

#include "../../include/af/array.h"
#include "../../include/af/device.h"

namespace afhip {

// ---------------------------------------------------------------- errors
static thread_local std::string g_last_error;

void set_last_error(const std::string &msg) { g_last_error = msg; }

// src/backend/common/err_common.cpp:125-175: AfError subclasses keep their
// code, anything else becomes AF_ERR_INTERNAL/AF_ERR_UNKNOWN.
af_err process_exception() {
    try {
        throw;
    } catch (const AfError &e) {
        set_last_error(e.what());
        if (std::getenv("AF_PRINT_ERRORS")) fprintf(stderr, "afhip: %s\n", e.what());
        return e.code();
    } catch (const std::bad_alloc &) {
        set_last_error("host allocation failed");
        return AF_ERR_NO_MEM;
    } catch (const std::exception &e) {
        set_last_error(e.what());
        return AF_ERR_INTERNAL;
    } catch (...) {
        set_last_error("unknown error");
        return AF_ERR_UNKNOWN;
    }
}

void check_hip(hipError_t e, const char *what) {
    if (e == hipSuccess) return;
    const af_err code = (e == hipErrorOutOfMemory) ? AF_ERR_NO_MEM : AF_ERR_RUNTIME;
    AFHIP_THROW(code, std::string(what) + ": " + hipGetErrorString(e));
}

void check_k(int rc, const char *what) {
    if (rc == 0) return;
    std::string msg = std::string(what) + " failed with code " + std::to_string(rc);
    if (rc == AF_ERR_RUNTIME) msg += std::string(" (") + hipGetErrorString(hipGetLastError()) + ")";
    AFHIP_THROW((af_err)rc, msg);
}

// ---------------------------------------------------------------- dtypes
bool dtype_known(int t) { return (t >= 0 && t <= 13) || t == (int)afhip_bf16; }

size_t dtype_size(af_dtype t) {
    switch ((int)t) {
        case f32: return 4;
        case c32: return 8;
        case f64: return 8;
        case c64: return 16;
        case b8: return 1;
        case s32: return 4;
        case u32: return 4;
        case u8: return 1;
        case s64: return 8;
        case u64: return 8;
        case s16: return 2;
        case u16: return 2;
        case f16: return 2;
        case s8: return 1;
        case (int)afhip_bf16: return 2;
        default: TYPE_ERROR(0, t);
    }
}

// ---------------------------------------------------------------- device
static std::once_flag g_dev_once;
static hipStream_t g_stream = nullptr;
static int g_dev_status = 0;

void ensure_device() {
    std::call_once(g_dev_once, [] {
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n < 1) {
            g_dev_status = 1;
            return;
        }
        if (hipSetDevice(0) != hipSuccess ||
            hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking) != hipSuccess) {
            g_dev_status = 2;
        }
    });
    if (g_dev_status == 1) AFHIP_THROW(AF_ERR_DEVICE, "no HIP device available");
    if (g_dev_status == 2) AFHIP_THROW(AF_ERR_DRIVER, "HIP device/stream set-up failed");
}

hipStream_t stream() {
    ensure_device();
    return g_stream;
}

void sync_stream() { check_hip(hipStreamSynchronize(stream()), "hipStreamSynchronize"); }

// ---------------------------------------------------------------- pool
namespace {
struct Pool {
    std::mutex m;
    std::multimap<size_t, void *> free_list;  // size -> ptr
    std::unordered_map<void *, size_t> sizes;  // every live or cached buffer
    size_t bytes_total = 0, bytes_locked = 0, bufs_locked = 0;
};
Pool &pool() {
    static Pool *p = new Pool();  // never destroyed: outlives static handles
    return *p;
}
size_t round_size(size_t b) {
    if (b == 0) b = 1;
    const size_t step = 1024;
    return (b + step - 1) / step * step;
}
}  // namespace

void *pool_alloc(size_t bytes) {
    ensure_device();
    const size_t sz = round_size(bytes);
    Pool &P = pool();
    {
        std::lock_guard<std::mutex> g(P.m);
        auto it = P.free_list.lower_bound(sz);
        // reuse a cached buffer no more than twice as large as needed
        if (it != P.free_list.end() && it->first <= 2 * sz) {
            void *p = it->second;
            P.bytes_locked += it->first;
            P.bufs_locked += 1;
            P.free_list.erase(it);
            return p;
        }
    }
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, sz);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        pool_gc();
        e = hipMalloc(&p, sz);
    }
    check_hip(e, "hipMalloc");
    std::lock_guard<std::mutex> g(P.m);
    P.sizes[p] = sz;
    P.bytes_total += sz;
    P.bytes_locked += sz;
    P.bufs_locked += 1;
    return p;
}

namespace {
thread_local std::vector<void *> *t_keep = nullptr;
}

void pool_keep(std::vector<void *> *keep) { t_keep = keep; }

void pool_free(void *p) {
    if (!p) return;
    if (t_keep) {
        t_keep->push_back(p);
        return;
    }
    Pool &P = pool();
    std::lock_guard<std::mutex> g(P.m);
    auto it = P.sizes.find(p);
    if (it == P.sizes.end()) return;
    P.free_list.emplace(it->second, p);
    P.bytes_locked -= it->second;
    P.bufs_locked -= 1;
}

void pool_gc() {
    Pool &P = pool();
    std::lock_guard<std::mutex> g(P.m);
    if (P.free_list.empty()) return;
    // Cached buffers may still be read by queued kernels: drain first.
    if (g_stream) (void)hipStreamSynchronize(g_stream);
    for (auto &kv : P.free_list) {
        (void)hipFree(kv.second);
        P.bytes_total -= kv.first;
        P.sizes.erase(kv.second);
    }
    P.free_list.clear();
}

void pool_info(size_t *alloc_bytes, size_t *alloc_buffers, size_t *lock_bytes,
               size_t *lock_buffers) {
    Pool &P = pool();
    std::lock_guard<std::mutex> g(P.m);
    if (alloc_bytes) *alloc_bytes = P.bytes_total;
    if (alloc_buffers) *alloc_buffers = P.sizes.size();
    if (lock_bytes) *lock_bytes = P.bytes_locked;
    if (lock_buffers) *lock_buffers = P.bufs_locked;
}

Buffer::~Buffer() {
    if (!ptr || locked) return;  // a locked buffer belongs to the caller
    if (user_owned) {
        if (g_stream) (void)hipStreamSynchronize(g_stream);
        (void)hipFree(ptr);
    } else {
        pool_free(ptr);
    }
}

std::shared_ptr<Buffer> scratch(size_t bytes) {
    auto b = std::make_shared<Buffer>();
    b->ptr = pool_alloc(bytes);
    b->bytes = bytes;
    return b;
}

// ---------------------------------------------------------------- handles
namespace {
std::mutex g_h_mutex;
std::unordered_set<ArrayImpl *> &live() {
    static auto *s = new std::unordered_set<ArrayImpl *>();
    return *s;
}
}  // namespace

ArrayImpl *impl_of(af_array a) {
    std::lock_guard<std::mutex> g(g_h_mutex);
    auto *p = static_cast<ArrayImpl *>(a);
    return live().count(p) ? p : nullptr;
}

ArrayImpl &get(const af_array a) {
    ArrayImpl *p = impl_of(a);
    if (!p) AFHIP_THROW(AF_ERR_ARG, "Input Array not created on current device");
    return *p;
}

static af_array register_impl(ArrayImpl *p) {
    std::lock_guard<std::mutex> g(g_h_mutex);
    live().insert(p);
    return static_cast<af_array>(p);
}

af_array make_handle(af_dtype type, const dim4 &dims) {
    auto *p = new ArrayImpl();
    p->type = type;
    p->dims = dims;
    const size_t bytes = (size_t)dims.elements() * dtype_size(type);
    if (bytes > 0) {
        p->buf = std::make_shared<Buffer>();
        p->buf->ptr = pool_alloc(bytes);
        p->buf->bytes = bytes;
    }
    return register_impl(p);
}

af_array make_handle_shared(const ArrayImpl &src) {
    auto *p = new ArrayImpl(src);
    return register_impl(p);
}

void destroy_handle(af_array a) noexcept {
    ArrayImpl *p = nullptr;
    {
        std::lock_guard<std::mutex> g(g_h_mutex);
        auto *q = static_cast<ArrayImpl *>(a);
        if (live().erase(q)) p = q;
    }
    delete p;
}

void *NewHandle::data() const { return ::afhip::get(h_).data(); }

static void release_handle(af_array a) {
    ArrayImpl *p = nullptr;
    {
        std::lock_guard<std::mutex> g(g_h_mutex);
        auto *q = static_cast<ArrayImpl *>(a);
        if (live().erase(q)) p = q;
    }
    if (!p) AFHIP_THROW(AF_ERR_ARG, "release of an unknown af_array");
    delete p;
}

}  // namespace afhip

using namespace afhip;

// ================================================================ C API
extern "C" {

AFAPI void af_get_last_error(char **msg, dim_t *len) {
    // include/af/exception.h:65, src/api/c/error.cpp:19-40
    std::string &s = g_last_error;
    const size_t n = std::min<size_t>(s.size(), 1024);
    if (len && n == 0) {
        *len = 0;
        *msg = nullptr;
        return;
    }
    char *buf = static_cast<char *>(std::malloc(n + 1));
    if (!buf) {
        if (len) *len = 0;
        *msg = nullptr;
        return;
    }
    std::memcpy(buf, s.data(), n);
    buf[n] = '\0';
    *msg = buf;
    s.clear();
    if (len) *len = (dim_t)n;
}

AFAPI af_err af_alloc_host(void **ptr, const dim_t bytes) {
    try {
        ARG_ASSERT(0, ptr != nullptr);
        *ptr = std::malloc(bytes > 0 ? (size_t)bytes : 1);
        if (!*ptr) AFHIP_THROW(AF_ERR_NO_MEM, "af_alloc_host");
    }
    CATCHALL;
    return AF_SUCCESS;
}

AFAPI af_err af_free_host(void *ptr) {
    std::free(ptr);
    return AF_SUCCESS;
}

AFAPI af_err af_init(void) {
    try {
        ensure_device();
        (void)stream();
    }
    CATCHALL;
    return AF_SUCCESS;
}

AFAPI af_err af_get_device_count(int *num) {
    try {
        ARG_ASSERT(0, num != nullptr);
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
        *num = n > 0 ? 1 : 0;
    }
    CATCHALL;
    return AF_SUCCESS;
}

AFAPI af_err af_set_device(const int device) {
    try {
        ARG_ASSERT(0, device == 0);
        ensure_device();
    }
    CATCHALL;
    return AF_SUCCESS;
}

AFAPI af_err af_get_device(int *device) {
    try {
        ARG_ASSERT(0, device != nullptr);
        *device = 0;
    }
    CATCHALL;
    return AF_SUCCESS;
}

AFAPI af_err af_sync(const int device) {
    try {
        ARG_ASSERT(0, device == -1 || device == 0);
        sync_stream();
    }
    CATCHALL;
    return AF_SUCCESS;
}

AFAPI af_err afhip_get_stream(void **out, int id) {
    try {
        ARG_ASSERT(0, out != nullptr);
        ARG_ASSERT(1, id == 0);
        *out = (void *)stream();
    }
    CATCHALL;
    return AF_SUCCESS;
}

AFAPI af_err af_device_gc(void) {
    try {
        pool_gc();
    }
    CATCHALL;
    return AF_SUCCESS;
}

AFAPI af_err af_device_mem_info(size_t *alloc_bytes, size_t *alloc_buffers,
                                size_t *lock_bytes, size_t *lock_buffers) {
    pool_info(alloc_bytes, alloc_buffers, lock_bytes, lock_buffers);
    return AF_SUCCESS;
}

AFAPI af_err af_create_array(af_array *arr, const void *const data, const unsigned ndims,
                             const dim_t *const dims, const af_dtype type) {
    try {
        ARG_ASSERT(0, arr != nullptr);
        ARG_ASSERT(2, ndims >= 1 && ndims <= 4);
        ARG_ASSERT(3, dims != nullptr);
        if (!dtype_known(type)) TYPE_ERROR(4, type);
        dim4 d(ndims, dims);
        for (int i = 0; i < 4; ++i) DIM_ASSERT(3, d[i] >= 0);
        af_array h = make_handle(type, d);
        ArrayImpl &a = get(h);
        const size_t bytes = (size_t)d.elements() * dtype_size(type);
        if (bytes > 0) {
            ARG_ASSERT(1, data != nullptr);
            hipError_t e = hipMemcpyAsync(a.data(), data, bytes, hipMemcpyHostToDevice, stream());
            if (e == hipSuccess) e = hipStreamSynchronize(stream());
            if (e != hipSuccess) {
                release_handle(h);
                check_hip(e, "af_create_array copy");
            }
        }
        *arr = h;
    }
    CATCHALL;
    return AF_SUCCESS;
}

AFAPI af_err af_device_array(af_array *arr, void *data, const unsigned ndims,
                             const dim_t *const dims, const af_dtype type) {
    try {
        ARG_ASSERT(0, arr != nullptr);
        ARG_ASSERT(1, data != nullptr);
        ARG_ASSERT(2, ndims >= 1 && ndims <= 4);
        if (!dtype_known(type)) TYPE_ERROR(4, type);
        ensure_device();
        auto *p = new ArrayImpl();
        p->type = type;
        p->dims = dim4(ndims, dims);
        p->buf = std::make_shared<Buffer>();
        p->buf->ptr = data;
        p->buf->bytes = (size_t)p->dims.elements() * dtype_size(type);
        p->buf->user_owned = true;
        *arr = register_impl(p);
    }
    CATCHALL;
    return AF_SUCCESS;
}

// include/af/device.h:543-565: a locked buffer is left to the caller (never
// freed or returned to the pool); unlocking hands it back to the library.
AFAPI af_err af_lock_array(const af_array arr) {
    try {
        ArrayImpl &a = get(arr);
        if (a.buf) a.buf->locked = true;
    }
    CATCHALL;
    return AF_SUCCESS;
}

AFAPI af_err af_unlock_array(const af_array arr) {
    try {
        ArrayImpl &a = get(arr);
        if (a.buf) a.buf->locked = false;
    }
    CATCHALL;
    return AF_SUCCESS;
}

AFAPI af_err af_is_locked_array(bool *res, const af_array arr) {
    try {
        ARG_ASSERT(0, res != nullptr);
        ArrayImpl &a = get(arr);
        *res = a.buf && a.buf->locked;
    }
    CATCHALL;
    return AF_SUCCESS;
}

AFAPI af_err af_get_device_ptr(void **ptr, const af_array arr) {
    try {
        ARG_ASSERT(0, ptr != nullptr);
        *ptr = get(arr).data();
    }
    CATCHALL;
    return AF_SUCCESS;
}

AFAPI af_err af_get_data_ptr(void *data, const af_array arr) {
    try {
        ArrayImpl &a = get(arr);
        const size_t bytes = (size_t)a.elements() * dtype_size(a.type);
        if (bytes == 0) return AF_SUCCESS;
        ARG_ASSERT(0, data != nullptr);
        check_hip(hipMemcpyAsync(data, a.data(), bytes, hipMemcpyDeviceToHost, stream()),
                  "af_get_data_ptr copy");
        sync_stream();
    }
    CATCHALL;
    return AF_SUCCESS;
}

AFAPI af_err af_release_array(af_array arr) {
    try {
        if (arr == nullptr) return AF_SUCCESS;
        release_handle(arr);
    }
    CATCHALL;
    return AF_SUCCESS;
}

AFAPI af_err af_retain_array(af_array *out, const af_array in) {
    try {
        ARG_ASSERT(0, out != nullptr);
        *out = make_handle_shared(get(in));
    }
    CATCHALL;
    return AF_SUCCESS;
}

AFAPI af_err af_eval(af_array in) {
    try {
        (void)get(in);
    }
    CATCHALL;
    return AF_SUCCESS;
}

AFAPI af_err af_get_elements(dim_t *elems, const af_array arr) {
    try {
        ARG_ASSERT(0, elems != nullptr);
        *elems = get(arr).elements();
    }
    CATCHALL;
    return AF_SUCCESS;
}

AFAPI af_err af_get_type(af_dtype *type, const af_array arr) {
    try {
        ARG_ASSERT(0, type != nullptr);
        *type = get(arr).type;
    }
    CATCHALL;
    return AF_SUCCESS;
}

AFAPI af_err af_get_dims(dim_t *d0, dim_t *d1, dim_t *d2, dim_t *d3, const af_array arr) {
    try {
        const ArrayImpl &a = get(arr);
        if (d0) *d0 = a.dims[0];
        if (d1) *d1 = a.dims[1];
        if (d2) *d2 = a.dims[2];
        if (d3) *d3 = a.dims[3];
    }
    CATCHALL;
    return AF_SUCCESS;
}

AFAPI af_err af_get_numdims(unsigned *result, const af_array arr) {
    try {
        ARG_ASSERT(0, result != nullptr);
        *result = (unsigned)get(arr).dims.ndims();
    }
    CATCHALL;
    return AF_SUCCESS;
}

}  // extern "C"
