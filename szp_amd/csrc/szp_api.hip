// C-ABI of libszp_amd.so: the reference's fastmode1 float32 entry points
// (include/szp.h) on host pointers, and the device-resident API
// (include/szp_amd.h).  The host-pointer entry points copy through cached
// device buffers and run the same kernels as the device API.
#include <hip/hip_runtime.h>

#include <dlfcn.h>
#include <sched.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>

#include "szp.h"
#include "szp_amd.h"
#include "szp_rw.h"
#include "szp_internal.h"

using namespace szp_amd;

extern "C" {
// SZp/src/szp.cc:17-20.  The stream's eb is always written big-endian, as
// the reference does on little-endian hosts.
int szp_versionNumber[4] = {1, 0, 0, 0};
int szp_dataEndianType = LITTLE_ENDIAN_DATA;
int szp_sysEndianType = LITTLE_ENDIAN_SYSTEM;
}

namespace {

thread_local std::string t_last_error;
int g_threads_override = 0;

void set_error(const char *msg, int code = 0)
{
    t_last_error = msg;
    if (code) {
        t_last_error += ": ";
        t_last_error += szp_amd_status_string(code);
    }
}

int map_device_errors(uint32_t bits)
{
    if (bits & kErrCapacity) return SZP_AMD_ERR_CAPACITY;
    if (bits & kErrStream) return SZP_AMD_ERR_STREAM;
    if (bits & kErrTimeout) return SZP_AMD_ERR_TIMEOUT;
    return SZP_AMD_OK;
}

// Device buffers reused by the host-pointer entry points of one device.
struct Session {
    std::mutex mu;
    hipStream_t stream = nullptr;
    void *buf[4] = {nullptr, nullptr, nullptr, nullptr};
    size_t cap[4] = {0, 0, 0, 0};

    void *ensure(int slot, size_t bytes)
    {
        if (bytes == 0) bytes = 256;
        if (cap[slot] >= bytes) return buf[slot];
        if (buf[slot]) (void)hipFree(buf[slot]);
        buf[slot] = nullptr;
        cap[slot] = 0;
        if (hipMalloc(&buf[slot], bytes) != hipSuccess) return nullptr;
        cap[slot] = bytes;
        return buf[slot];
    }
};

Session *session_for_current_device()
{
    static Session sessions[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nullptr;
    Session *s = &sessions[dev];
    if (!s->stream) {
        std::lock_guard<std::mutex> lk(s->mu);
        if (!s->stream && hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess)
            return nullptr;
    }
    return s;
}

uint64_t read_u64(const unsigned char *p)
{
    uint64_t v;
    memcpy(&v, p, 8);
    return v;
}

// Length of the part of a stream after the eb (4 bytes for float, 8 for
// double), from the chunk offsets and a walk over the last chunk's width
// bytes; 0 when malformed.  fastMode 2 walks the last chunk's units, each a
// 4-byte head and a block of len - 1 values (SZp/src/szpd_float.cc:470-520).
// The reference trusts the header; here every position read is first checked
// against the longest body a valid stream can have (compress_bound less the
// eb), so a damaged last-chunk offset is refused instead of read through.
uint64_t body_length(const unsigned char *body, uint64_t n, int B, int T, int fast = 1,
                     int dtype = 0)
{
    Geom g;
    if (!make_geom(n, B, T, &g, 4, (uint32_t)fast)) return 0;
    const uint64_t hdr = 8ull * T;
    if (n == 0) return hdr;
    const uint64_t lim = compress_bound(n, B, T, dtype, fast) - (dtype == SZ_DOUBLE ? 8 : 4);
    const uint64_t last = read_u64(body + hdr - 8);
    if (last > lim - hdr) return 0;
    const uint64_t nl = n - (uint64_t)(T - 1) * g.tbs;
    if (fast == 2) {
        uint64_t pos = hdr + last;
        for (uint64_t k = 0; k < g.nblast; ++k) {
            const uint64_t rem = nl - k * (uint64_t)B;
            const uint64_t L = (rem < (uint64_t)B ? rem : (uint64_t)B) - 1;
            pos += 4;
            if (pos >= lim) return 0;
            const uint32_t b = body[pos];
            if (L && b > 32) return 0;
            pos += L ? block_bytes(b, (uint32_t)L) : 1;
        }
        return pos <= lim ? pos : 0;
    }
    uint64_t pos = hdr + last + 4;
    for (uint64_t k = 0; k < g.nblast; ++k) {
        if (pos >= lim) return 0;
        const uint32_t b = body[pos];
        if (b > 32) return 0;
        const uint64_t L = (k + 1 == g.nblast) ? nl - 1 - k * (uint64_t)B : (uint64_t)B;
        pos += block_bytes(b, (uint32_t)L);
    }
    return pos <= lim ? pos : 0;
}

// Memory-access count returned by the reference's instrumented single-thread
// forms, from the block widths of a T = 1 stream body (after the eb):
//   compress   SZp/src/szp_float.cc:607-750: 8 for the offset; 4 + 8 for the
//              head; per block of L values 4L (input) + 5L (temporaries) + 1
//              (width), and for a width b > 0 also L + 4L (reading the
//              temporaries) + the block's sign, plane and residual bytes;
//              then 8 + 8 + 24 + 16 of bookkeeping and twice the chunk bytes;
//   decompress SZp/src/szpd_float.cc:216-317: 8 for the offset; 8 + 8 for the
//              head; per block 1 (width), then 8L for b = 0, else its sign,
//              plane and residual bytes + L + 4L + 13L.
// The body is a stream this library wrote or accepted (widths <= 32).
size_t single_thread_record(const unsigned char *body, uint64_t n, int B, bool decompress)
{
    uint64_t comp = 8, dec = 8;
    if (n == 0) return decompress ? dec : comp + 56;
    comp += 12;
    dec += 16;
    uint64_t pos = 8 + read_u64(body) + 4, chunk = 4;
    for (uint64_t i = 1; i < n; i += (uint64_t)B) {
        const uint64_t L = std::min<uint64_t>((uint64_t)B, n - i);
        const uint32_t b = body[pos];
        const uint64_t bb = block_bytes(b, (uint32_t)L);
        pos += bb;
        chunk += bb;
        comp += 9 * L + 1 + (b ? 5 * L + bb - 1 : 0);
        dec += 1 + (b ? bb - 1 + 18 * L : 8 * L);
    }
    return decompress ? dec : comp + 56 + 2 * chunk;
}

// Bytes per value of a dataType (SZ_FLOAT 4, SZ_DOUBLE 8).
uint64_t value_bytes(int dtype) { return dtype == SZ_DOUBLE ? 8 : 4; }

// Compress host data of type dtype through the device with fastMode `fast`;
// returns 0 and fills *stream (malloc'd unless `into` is given) or a status
// code.
int host_compress(int fast, int dtype, const void *data, uint64_t n, int mode, double abs_eb,
                  double rel, int B, int T, unsigned char *into, unsigned char **stream,
                  size_t *out_size)
{
    Session *s = session_for_current_device();
    if (!s) return SZP_AMD_ERR_HIP;
    std::lock_guard<std::mutex> lk(s->mu);
    const uint64_t vb = value_bytes(dtype);
    const uint64_t cap = compress_bound(n, B, T, dtype, fast);
    const uint64_t wsb = compress_workspace_bytes(n, B, T, fast);
    if (!cap || !wsb) return SZP_AMD_ERR_ARG;
    void *d_in = s->ensure(0, n * vb);
    unsigned char *d_out = (unsigned char *)s->ensure(1, cap);
    void *d_ws = s->ensure(2, wsb);
    uint64_t *d_size = (uint64_t *)s->ensure(3, 256);
    if (!d_in || !d_out || !d_ws || !d_size) return SZP_AMD_ERR_NOMEM;
    if (n && hipMemcpyAsync(d_in, data, n * vb, hipMemcpyHostToDevice, s->stream) != hipSuccess)
        return SZP_AMD_ERR_HIP;
    int rc = launch_compress(dtype, d_in, n, mode, abs_eb, rel, B, T, d_out, cap, d_size, d_ws,
                             wsb, s->stream, nullptr, 0, fast);
    if (rc) return rc;
    rc = szp_amd_check_status(d_ws, s->stream);
    if (rc) return rc;
    uint64_t size = 0;
    if (hipMemcpy(&size, d_size, 8, hipMemcpyDeviceToHost) != hipSuccess) return SZP_AMD_ERR_HIP;
    unsigned char *dst = into ? into : (unsigned char *)malloc(size ? size : 1);
    if (!dst) return SZP_AMD_ERR_NOMEM;
    if (hipMemcpy(dst, d_out, size, hipMemcpyDeviceToHost) != hipSuccess) {
        if (!into) free(dst);
        return SZP_AMD_ERR_HIP;
    }
    *stream = dst;
    *out_size = size;
    return SZP_AMD_OK;
}

// Decompress a host stream of type dtype and fastMode `fast` given as (eb
// bytes, big-endian, value_bytes(dtype) of them; body) into `into` or a new
// malloc'd array.
int host_decompress(int fast, int dtype, const unsigned char *eb_be, const unsigned char *body,
                    uint64_t n, int B, int T, void *into, void **result)
{
    const uint64_t vb = value_bytes(dtype);
    const uint64_t blen = body_length(body, n, B, T, fast, dtype);
    if (!blen) return SZP_AMD_ERR_STREAM;
    const uint64_t len = vb + blen;
    Session *s = session_for_current_device();
    if (!s) return SZP_AMD_ERR_HIP;
    std::lock_guard<std::mutex> lk(s->mu);
    const uint64_t wsb = decompress_workspace_bytes(len, n, B, T, fast);
    if (!wsb) return SZP_AMD_ERR_ARG;
    unsigned char *d_st = (unsigned char *)s->ensure(1, len);
    void *d_out = s->ensure(0, n * vb);
    void *d_ws = s->ensure(2, wsb);
    if (!d_st || !d_out || !d_ws) return SZP_AMD_ERR_NOMEM;
    if (hipMemcpyAsync(d_st, eb_be, vb, hipMemcpyHostToDevice, s->stream) != hipSuccess ||
        hipMemcpyAsync(d_st + vb, body, blen, hipMemcpyHostToDevice, s->stream) != hipSuccess)
        return SZP_AMD_ERR_HIP;
    int rc = launch_decompress(dtype, d_st, len, n, B, T, d_out, d_ws, wsb, s->stream, fast);
    if (rc) return rc;
    rc = szp_amd_check_status(d_ws, s->stream);
    if (rc) return rc;
    void *dst = into ? into : malloc(n ? n * vb : vb);
    if (!dst) return SZP_AMD_ERR_NOMEM;
    if (n && hipMemcpy(dst, d_out, n * vb, hipMemcpyDeviceToHost) != hipSuccess) {
        if (!into) free(dst);
        return SZP_AMD_ERR_HIP;
    }
    *result = dst;
    return SZP_AMD_OK;
}

// floatToBytes / doubleToBytes (SZp/include/szp_CompressionToolkit.h:520-546):
// the value's bytes, most significant first.
void float_to_be(float v, unsigned char b[4])
{
    uint32_t u;
    memcpy(&u, &v, 4);
    for (int i = 0; i < 4; ++i) b[i] = (unsigned char)(u >> (24 - 8 * i));
}

void double_to_be(double v, unsigned char b[8])
{
    uint64_t u;
    memcpy(&u, &v, 8);
    for (int i = 0; i < 8; ++i) b[i] = (unsigned char)(u >> (56 - 8 * i));
}

// Shared body of the lower-level compress entry points (szp_float.h /
// szp_double.h forms): ABS bound, T from szp_amd_default_threads(); writes
// into `into` when given (the _arg forms), else returns a malloc'd stream.
unsigned char *lower_compress(const char *who, int fast, int dtype, const void *data,
                              size_t *outSize, double eb, size_t nbEle, int blockSize,
                              unsigned char *into, int T = 0)
{
    if (!outSize || (!data && nbEle)) {
        set_error(who, SZP_AMD_ERR_ARG);
        return NULL;
    }
    unsigned char *stream = NULL;
    size_t size = 0;
    const int rc = host_compress(fast, dtype, data, nbEle, ABS, eb, 0.0, blockSize,
                                 T > 0 ? T : szp_amd_default_threads(), into, &stream, &size);
    if (rc) {
        set_error(who, rc);
        *outSize = 0;
        return NULL;
    }
    *outSize = size;
    return stream;
}

// Shared body of the lower-level decompress entry points (szpd_float.h /
// szpd_double.h forms): cmpBytes points just after the eb, which the caller
// passes as a value; writes into `into` when given (the _arg forms), else
// returns a malloc'd array.
void *lower_decompress(const char *who, int fast, int dtype, double eb, size_t nbEle,
                       int blockSize, unsigned char *cmpBytes, void *into, int T = 0)
{
    if (!cmpBytes) {
        set_error(who, SZP_AMD_ERR_ARG);
        return NULL;
    }
    unsigned char ebb[8];
    if (dtype == SZ_DOUBLE)
        double_to_be(eb, ebb);
    else
        float_to_be((float)eb, ebb);
    void *out = NULL;
    const int rc = host_decompress(fast, dtype, ebb, cmpBytes, nbEle, blockSize,
                                   T > 0 ? T : szp_amd_default_threads(), into, &out);
    if (rc) {
        set_error(who, rc);
        return NULL;
    }
    return out;
}

}  // namespace

extern "C" {

const char *szp_amd_status_string(int status)
{
    switch (status) {
    case SZP_AMD_OK: return "ok";
    case SZP_AMD_ERR_ARG: return "invalid argument";
    case SZP_AMD_ERR_CAPACITY: return "output or workspace capacity too small";
    case SZP_AMD_ERR_HIP: return "HIP runtime error";
    case SZP_AMD_ERR_STREAM: return "malformed compressed stream";
    case SZP_AMD_ERR_TIMEOUT: return "device wait timed out";
    case SZP_AMD_ERR_UNSUPPORTED: return "only fastMode 1 or 2 with SZ_FLOAT or SZ_DOUBLE is supported";
    case SZP_AMD_ERR_NOMEM: return "allocation failed";
    default: return "unknown status";
    }
}

const char *szp_amd_last_error(void) { return t_last_error.c_str(); }

int szp_amd_default_threads(void)
{
    if (g_threads_override > 0) return g_threads_override;
    const char *e = getenv("SZP_AMD_NUM_THREADS");
    if (e && atoi(e) > 0) return atoi(e);
    typedef int (*omp_fn)(void);
    omp_fn f = (omp_fn)dlsym(RTLD_DEFAULT, "omp_get_max_threads");
    if (f) {
        int v = f();
        if (v > 0) return v;
    }
    e = getenv("OMP_NUM_THREADS");
    if (e && atoi(e) > 0) return atoi(e);
    cpu_set_t set;
    if (sched_getaffinity(0, sizeof(set), &set) == 0 && CPU_COUNT(&set) > 0) return CPU_COUNT(&set);
    return 1;
}

void szp_amd_set_threads(int nbThreads) { g_threads_override = nbThreads > 0 ? nbThreads : 0; }

int szp_amd_bit_width(uint32_t maxMagnitude) { return bit_width_host(maxMagnitude); }

void szp_amd_profile_enable(int enable) { profile_enable(enable); }

void szp_amd_profile_enable_only(const char *kernel) { profile_enable(1, kernel); }

int szp_amd_profile_count(void) { return profile_count(); }

int szp_amd_profile_entry(int index, char *name, int nameCapacity, uint64_t *launches,
                          double *totalMs)
{
    return profile_entry(index, name, nameCapacity, launches, totalMs);
}

uint64_t szp_amd_compress_bound(uint64_t nbEle, int blockSize, int nbThreads)
{
    return compress_bound(nbEle, blockSize, nbThreads, SZ_FLOAT);
}

static bool typed_ok(int fastMode, int dataType)
{
    return (fastMode == SZP_NONRANDOMACCESS || fastMode == SZP_RANDOMACCESS) &&
           (dataType == SZ_FLOAT || dataType == SZ_DOUBLE);
}

uint64_t szp_amd_compress_bound_typed(int fastMode, int dataType, uint64_t nbEle, int blockSize,
                                      int nbThreads)
{
    if (!typed_ok(fastMode, dataType)) return 0;
    return compress_bound(nbEle, blockSize, nbThreads, dataType, fastMode);
}

uint64_t szp_amd_compress_workspace_size_typed(int fastMode, uint64_t nbEle, int blockSize,
                                               int nbThreads)
{
    if (!typed_ok(fastMode, SZ_FLOAT)) return 0;
    return compress_workspace_bytes(nbEle, blockSize, nbThreads, fastMode);
}

uint64_t szp_amd_decompress_workspace_size_typed(int fastMode, uint64_t streamBytes,
                                                 uint64_t nbEle, int blockSize, int nbThreads)
{
    if (!typed_ok(fastMode, SZ_FLOAT)) return 0;
    return decompress_workspace_bytes(streamBytes, nbEle, blockSize, nbThreads, fastMode);
}

uint64_t szp_amd_index_size_typed(int fastMode, uint64_t nbEle, int blockSize, int nbThreads)
{
    if (!typed_ok(fastMode, SZ_FLOAT)) return 0;
    return index_bytes(nbEle, blockSize, nbThreads, fastMode);
}

uint64_t szp_amd_compress_workspace_size(uint64_t nbEle, int blockSize, int nbThreads)
{
    return compress_workspace_bytes(nbEle, blockSize, nbThreads);
}

uint64_t szp_amd_decompress_workspace_size(uint64_t streamBytes, uint64_t nbEle, int blockSize,
                                           int nbThreads)
{
    return decompress_workspace_bytes(streamBytes, nbEle, blockSize, nbThreads);
}

int szp_amd_compress_dev(const float *d_data, uint64_t nbEle, int errBoundMode,
                         double absErrBound, double relBoundRatio, int blockSize, int nbThreads,
                         unsigned char *d_out, uint64_t outCapacity, uint64_t *d_outSize,
                         void *d_work, uint64_t workBytes, void *stream)
{
    return launch_compress(SZ_FLOAT, d_data, nbEle, errBoundMode, absErrBound, relBoundRatio,
                           blockSize, nbThreads, d_out, outCapacity, d_outSize, d_work, workBytes,
                           (hipStream_t)stream);
}

int szp_amd_compress_dev_noalloc(const float *d_data, uint64_t nbEle, int errBoundMode,
                                 double absErrBound, double relBoundRatio, int blockSize,
                                 int nbThreads, unsigned char *d_out, uint64_t outCapacity,
                                 uint64_t *d_outSize, void *stream)
{
    const int rc = launch_compress_noalloc(d_data, nbEle, errBoundMode, absErrBound,
                                           relBoundRatio, blockSize, nbThreads, d_out,
                                           outCapacity, d_outSize, (hipStream_t)stream);
    if (rc == SZP_AMD_ERR_CAPACITY)
        set_error("szp_amd_compress_dev_noalloc: the field's workspace exceeds the "
                  "8 MiB static device scratch (SZP_AMD_NOALLOC_SCRATCH_BYTES)");
    else if (rc)
        set_error("szp_amd_compress_dev_noalloc", rc);
    return rc;
}

int szp_amd_decompress_dev(const unsigned char *d_stream, uint64_t streamBytes, uint64_t nbEle,
                           int blockSize, int nbThreads, float *d_out, void *d_work,
                           uint64_t workBytes, void *stream)
{
    return launch_decompress(SZ_FLOAT, d_stream, streamBytes, nbEle, blockSize, nbThreads, d_out,
                             d_work, workBytes, (hipStream_t)stream);
}

uint64_t szp_amd_index_size(uint64_t nbEle, int blockSize, int nbThreads)
{
    return index_bytes(nbEle, blockSize, nbThreads);
}

int szp_amd_compress_dev_indexed(const float *d_data, uint64_t nbEle, int errBoundMode,
                                 double absErrBound, double relBoundRatio, int blockSize,
                                 int nbThreads, unsigned char *d_out, uint64_t outCapacity,
                                 uint64_t *d_outSize, void *d_index, uint64_t indexBytes,
                                 void *d_work, uint64_t workBytes, void *stream)
{
    if (!d_index) return SZP_AMD_ERR_ARG;
    return launch_compress(SZ_FLOAT, d_data, nbEle, errBoundMode, absErrBound, relBoundRatio,
                           blockSize, nbThreads, d_out, outCapacity, d_outSize, d_work, workBytes,
                           (hipStream_t)stream, d_index, indexBytes);
}

int szp_amd_decompress_dev_indexed(const unsigned char *d_stream, uint64_t streamBytes,
                                   const void *d_index, uint64_t indexBytes, uint64_t nbEle,
                                   int blockSize, int nbThreads, float *d_out, void *d_work,
                                   uint64_t workBytes, void *stream)
{
    return launch_decompress_indexed(SZ_FLOAT, d_stream, streamBytes, d_index, indexBytes, nbEle,
                                     blockSize, nbThreads, d_out, d_work, workBytes,
                                     (hipStream_t)stream);
}

int szp_amd_compress_dev_typed(int fastMode, int dataType, const void *d_data, uint64_t nbEle,
                               int errBoundMode, double absErrBound, double relBoundRatio,
                               int blockSize, int nbThreads, unsigned char *d_out,
                               uint64_t outCapacity, uint64_t *d_outSize, void *d_index,
                               uint64_t indexBytes, void *d_work, uint64_t workBytes, void *stream)
{
    if (!typed_ok(fastMode, dataType)) return SZP_AMD_ERR_UNSUPPORTED;
    return launch_compress(dataType, d_data, nbEle, errBoundMode, absErrBound, relBoundRatio,
                           blockSize, nbThreads, d_out, outCapacity, d_outSize, d_work, workBytes,
                           (hipStream_t)stream, d_index, indexBytes, fastMode);
}

int szp_amd_decompress_dev_typed(int fastMode, int dataType, const unsigned char *d_stream,
                                 uint64_t streamBytes, const void *d_index, uint64_t indexBytes,
                                 uint64_t nbEle, int blockSize, int nbThreads, void *d_out,
                                 void *d_work, uint64_t workBytes, void *stream)
{
    if (!typed_ok(fastMode, dataType)) return SZP_AMD_ERR_UNSUPPORTED;
    if (d_index)
        return launch_decompress_indexed(dataType, d_stream, streamBytes, d_index, indexBytes,
                                         nbEle, blockSize, nbThreads, d_out, d_work, workBytes,
                                         (hipStream_t)stream, fastMode);
    return launch_decompress(dataType, d_stream, streamBytes, nbEle, blockSize, nbThreads, d_out,
                             d_work, workBytes, (hipStream_t)stream, fastMode);
}

static_assert(sizeof(szp_amd::BatchPlan) <= sizeof(szp_amd_batch), "batch plan size");
static_assert(SZP_AMD_NOALLOC_SCRATCH_BYTES == szp_amd::kNoallocScratchBytes, "scratch size");
static_assert(SZP_AMD_BATCH_GRID_PER_TILE == szp_amd::kBatchGridPerTile, "per-tile grid value");
static_assert(sizeof(szp_amd_field) == sizeof(szp_amd::BatchField), "field layout");

uint64_t szp_amd_batch_workspace_size(const szp_amd_field *fields, int numFields, int blockSize,
                                      int nbThreads, int decompress)
{
    return batch_workspace_bytes(reinterpret_cast<const szp_amd::BatchField *>(fields), numFields,
                                 blockSize, nbThreads, decompress);
}

int szp_amd_batch_prepare(const szp_amd_field *fields, int numFields, int blockSize,
                          int nbThreads, int decompress, uint64_t *d_sizes, void *d_work,
                          uint64_t workBytes, szp_amd_batch *batch)
{
    if (!batch) return SZP_AMD_ERR_ARG;
    szp_amd::BatchPlan plan;
    memset(&plan, 0, sizeof(plan));
    const int rc = batch_prepare(reinterpret_cast<const szp_amd::BatchField *>(fields), numFields,
                                 blockSize, nbThreads, decompress, d_sizes, d_work, workBytes,
                                 &plan);
    memset(batch, 0, sizeof(*batch));
    if (rc == 0) memcpy(batch, &plan, sizeof(plan));
    return rc;
}

int szp_amd_batch_launch(const szp_amd_batch *batch, uint32_t ticket, uint32_t grid, void *stream)
{
    if (!batch) return SZP_AMD_ERR_ARG;
    szp_amd::BatchPlan plan;
    memcpy(&plan, batch, sizeof(plan));
    return batch_launch(plan, ticket, grid, (hipStream_t)stream);
}

uint32_t szp_amd_batch_resident_grid(const szp_amd_batch *batch)
{
    if (!batch) return 0;
    szp_amd::BatchPlan plan;
    memcpy(&plan, batch, sizeof(plan));
    return batch_resident_grid(plan);
}

int szp_amd_check_status(void *d_work, void *stream)
{
    if (!d_work) return SZP_AMD_ERR_ARG;
    if (hipStreamSynchronize((hipStream_t)stream) != hipSuccess) return SZP_AMD_ERR_HIP;
    uint32_t bits = 0;
    if (hipMemcpy(&bits, workspace_error_word(d_work), 4, hipMemcpyDeviceToHost) != hipSuccess)
        return SZP_AMD_ERR_HIP;
    return map_device_errors(bits);
}

uint64_t szp_amd_stream_length(const unsigned char *bytes, uint64_t nbEle, int blockSize,
                               int nbThreads)
{
    if (!bytes) return 0;
    const uint64_t b = body_length(bytes + 4, nbEle, blockSize, nbThreads);
    return b ? 4 + b : 0;
}

uint64_t szp_amd_stream_length_typed(int fastMode, int dataType, const unsigned char *bytes,
                                     uint64_t nbEle, int blockSize, int nbThreads)
{
    if (!bytes || !typed_ok(fastMode, dataType)) return 0;
    const uint64_t vb = value_bytes(dataType);
    const uint64_t b = body_length(bytes + vb, nbEle, blockSize, nbThreads, fastMode, dataType);
    return b ? vb + b : 0;
}

unsigned char *szp_compress(int fastMode, int dataType, void *data, size_t *outSize,
                            int errBoundMode, double absErrBound, double relBoundRatio,
                            size_t nbEle, int blockSize)
{
    // SZp/src/szp.cc:23-115: fastMode 1 or 2, SZ_FLOAT or SZ_DOUBLE; any
    // other fastMode returns NULL.
    if ((fastMode != SZP_NONRANDOMACCESS && fastMode != SZP_RANDOMACCESS) ||
        (dataType != SZ_FLOAT && dataType != SZ_DOUBLE)) {
        set_error("szp_compress", SZP_AMD_ERR_UNSUPPORTED);
        return NULL;
    }
    if (!outSize || (!data && nbEle)) {
        set_error("szp_compress", SZP_AMD_ERR_ARG);
        return NULL;
    }
    unsigned char *stream = NULL;
    size_t size = 0;
    const int rc = host_compress(fastMode, dataType, data, nbEle, errBoundMode, absErrBound,
                                 relBoundRatio, blockSize, szp_amd_default_threads(), NULL,
                                 &stream, &size);
    if (rc) {
        set_error("szp_compress", rc);
        return NULL;
    }
    *outSize = size;
    return stream;
}

void *szp_decompress(int fastMode, int dataType, unsigned char *bytes, size_t byteLength,
                     size_t nbEle, int blockSize)
{
    (void)byteLength;  // unused by the reference too (SZp/src/szp.cc:118-153)
    if (dataType != SZ_FLOAT && dataType != SZ_DOUBLE) {
        set_error("szp_decompress", SZP_AMD_ERR_UNSUPPORTED);
        return NULL;
    }
    if (!bytes) {
        set_error("szp_decompress", SZP_AMD_ERR_ARG);
        return NULL;
    }
    szp_sysEndianType = LITTLE_ENDIAN_SYSTEM;  // SZp/src/szp.cc:120-125 on x86-64
    // Every fastMode other than 1 decodes as random access (SZp/src/szp.cc:135-151).
    const int fast = fastMode == SZP_NONRANDOMACCESS ? 1 : 2;
    void *out = NULL;
    const uint64_t vb = value_bytes(dataType);
    const int rc = host_decompress(fast, dataType, bytes, bytes + vb, nbEle, blockSize,
                                   szp_amd_default_threads(), NULL, &out);
    if (rc) {
        set_error("szp_decompress", rc);
        return NULL;
    }
    return out;
}

// ---- lower-level entry points (szp_float.h, szpd_float.h, szp_double.h,
// szpd_double.h): ABS bound given as the value type, T from
// szp_amd_default_threads().

unsigned char *szp_float_openmp_threadblock(float *oriData, size_t *outSize, float absErrBound,
                                            size_t nbEle, int blockSize)
{
    return lower_compress("szp_float_openmp_threadblock", 1, SZ_FLOAT, oriData, outSize,
                          (double)absErrBound, nbEle, blockSize, NULL);
}

void szp_float_openmp_threadblock_arg(unsigned char *output, float *oriData, size_t *outSize,
                                      float absErrBound, size_t nbEle, int blockSize)
{
    if (!output) {
        set_error("szp_float_openmp_threadblock_arg", SZP_AMD_ERR_ARG);
        return;
    }
    lower_compress("szp_float_openmp_threadblock_arg", 1, SZ_FLOAT, oriData, outSize,
                   (double)absErrBound, nbEle, blockSize, output);
}

float *szp_float_decompress_openmp_threadblock(size_t nbEle, float absErrBound, int blockSize,
                                               unsigned char *cmpBytes)
{
    return (float *)lower_decompress("szp_float_decompress_openmp_threadblock", 1, SZ_FLOAT,
                                     absErrBound, nbEle, blockSize, cmpBytes, NULL);
}

void szp_float_decompress_openmp_threadblock_arg(float *newData, size_t nbEle, float absErrBound,
                                                 int blockSize, unsigned char *cmpBytes)
{
    if (!newData && nbEle) {
        set_error("szp_float_decompress_openmp_threadblock_arg", SZP_AMD_ERR_ARG);
        return;
    }
    lower_decompress("szp_float_decompress_openmp_threadblock_arg", 1, SZ_FLOAT, absErrBound,
                     nbEle, blockSize, cmpBytes, newData);
}

void szp_float_single_thread_arg(unsigned char *output, float *oriData, size_t *outSize,
                                 float absErrBound, size_t nbEle, int blockSize)
{
    if (!output) {
        set_error("szp_float_single_thread_arg", SZP_AMD_ERR_ARG);
        return;
    }
    lower_compress("szp_float_single_thread_arg", 1, SZ_FLOAT, oriData, outSize,
                   (double)absErrBound, nbEle, blockSize, output, 1);
}

size_t szp_float_single_thread_arg_record(unsigned char *output, float *oriData, size_t *outSize,
                                          float absErrBound, size_t nbEle, int blockSize)
{
    if (!output) {
        set_error("szp_float_single_thread_arg_record", SZP_AMD_ERR_ARG);
        return 0;
    }
    if (!lower_compress("szp_float_single_thread_arg_record", 1, SZ_FLOAT, oriData, outSize,
                        (double)absErrBound, nbEle, blockSize, output, 1))
        return 0;
    return single_thread_record(output + 4, nbEle, blockSize, false);
}

void szp_float_decompress_single_thread_arg(float *newData, size_t nbEle, float absErrBound,
                                            int blockSize, unsigned char *cmpBytes)
{
    if (!newData && nbEle) {
        set_error("szp_float_decompress_single_thread_arg", SZP_AMD_ERR_ARG);
        return;
    }
    lower_decompress("szp_float_decompress_single_thread_arg", 1, SZ_FLOAT, absErrBound, nbEle,
                     blockSize, cmpBytes, newData, 1);
}

size_t szp_float_decompress_single_thread_arg_record(float *newData, size_t nbEle,
                                                     float absErrBound, int blockSize,
                                                     unsigned char *cmpBytes)
{
    if (!newData && nbEle) {
        set_error("szp_float_decompress_single_thread_arg_record", SZP_AMD_ERR_ARG);
        return 0;
    }
    if (!lower_decompress("szp_float_decompress_single_thread_arg_record", 1, SZ_FLOAT,
                          absErrBound, nbEle, blockSize, cmpBytes, newData, 1))
        return 0;
    return single_thread_record(cmpBytes, nbEle, blockSize, true);
}

unsigned char *szp_float_openmp_threadblock_randomaccess(float *oriData, size_t *outSize,
                                                         float absErrBound, size_t nbEle,
                                                         int blockSize)
{
    return lower_compress("szp_float_openmp_threadblock_randomaccess", 2, SZ_FLOAT, oriData,
                          outSize, (double)absErrBound, nbEle, blockSize, NULL);
}

void szp_float_openmp_threadblock_randomaccess_arg(unsigned char *output, float *oriData,
                                                   size_t *outSize, float absErrBound,
                                                   size_t nbEle, int blockSize)
{
    if (!output) {
        set_error("szp_float_openmp_threadblock_randomaccess_arg", SZP_AMD_ERR_ARG);
        return;
    }
    lower_compress("szp_float_openmp_threadblock_randomaccess_arg", 2, SZ_FLOAT, oriData, outSize,
                   (double)absErrBound, nbEle, blockSize, output);
}

float *szp_float_decompress_openmp_threadblock_randomaccess(size_t nbEle, float absErrBound,
                                                            int blockSize, unsigned char *cmpBytes)
{
    return (float *)lower_decompress("szp_float_decompress_openmp_threadblock_randomaccess", 2,
                                     SZ_FLOAT, absErrBound, nbEle, blockSize, cmpBytes, NULL);
}

void szp_float_decompress_openmp_threadblock_randomaccess_arg(float *newData, size_t nbEle,
                                                              float absErrBound, int blockSize,
                                                              unsigned char *cmpBytes)
{
    if (!newData && nbEle) {
        set_error("szp_float_decompress_openmp_threadblock_randomaccess_arg", SZP_AMD_ERR_ARG);
        return;
    }
    lower_decompress("szp_float_decompress_openmp_threadblock_randomaccess_arg", 2, SZ_FLOAT,
                     absErrBound, nbEle, blockSize, cmpBytes, newData);
}

unsigned char *szp_double_openmp_threadblock(double *oriData, size_t *outSize, double absErrBound,
                                             size_t nbEle, int blockSize)
{
    return lower_compress("szp_double_openmp_threadblock", 1, SZ_DOUBLE, oriData, outSize,
                          absErrBound, nbEle, blockSize, NULL);
}

void szp_double_openmp_threadblock_arg(unsigned char *output, double *oriData, size_t *outSize,
                                       double absErrBound, size_t nbEle, int blockSize)
{
    if (!output) {
        set_error("szp_double_openmp_threadblock_arg", SZP_AMD_ERR_ARG);
        return;
    }
    lower_compress("szp_double_openmp_threadblock_arg", 1, SZ_DOUBLE, oriData, outSize,
                   absErrBound, nbEle, blockSize, output);
}

double *szp_double_decompress_openmp_threadblock(size_t nbEle, double absErrBound, int blockSize,
                                                 unsigned char *cmpBytes)
{
    return (double *)lower_decompress("szp_double_decompress_openmp_threadblock", 1, SZ_DOUBLE,
                                      absErrBound, nbEle, blockSize, cmpBytes, NULL);
}

void szp_double_decompress_openmp_threadblock_arg(double *newData, size_t nbEle,
                                                  double absErrBound, int blockSize,
                                                  unsigned char *cmpBytes)
{
    if (!newData && nbEle) {
        set_error("szp_double_decompress_openmp_threadblock_arg", SZP_AMD_ERR_ARG);
        return;
    }
    lower_decompress("szp_double_decompress_openmp_threadblock_arg", 1, SZ_DOUBLE, absErrBound,
                     nbEle, blockSize, cmpBytes, newData);
}

unsigned char *szp_double_openmp_threadblock_randomaccess(double *oriData, size_t *outSize,
                                                          double absErrBound, size_t nbEle,
                                                          int blockSize)
{
    return lower_compress("szp_double_openmp_threadblock_randomaccess", 2, SZ_DOUBLE, oriData,
                          outSize, absErrBound, nbEle, blockSize, NULL);
}

void szp_double_openmp_threadblock_randomaccess_arg(unsigned char *output, double *oriData,
                                                    size_t *outSize, double absErrBound,
                                                    size_t nbEle, int blockSize)
{
    if (!output) {
        set_error("szp_double_openmp_threadblock_randomaccess_arg", SZP_AMD_ERR_ARG);
        return;
    }
    lower_compress("szp_double_openmp_threadblock_randomaccess_arg", 2, SZ_DOUBLE, oriData,
                   outSize, absErrBound, nbEle, blockSize, output);
}

double *szp_double_decompress_openmp_threadblock_randomaccess(size_t nbEle, double absErrBound,
                                                              int blockSize,
                                                              unsigned char *cmpBytes)
{
    return (double *)lower_decompress("szp_double_decompress_openmp_threadblock_randomaccess", 2,
                                      SZ_DOUBLE, absErrBound, nbEle, blockSize, cmpBytes, NULL);
}

void szp_double_decompress_openmp_threadblock_randomaccess_arg(double *newData, size_t nbEle,
                                                               double absErrBound, int blockSize,
                                                               unsigned char *cmpBytes)
{
    if (!newData && nbEle) {
        set_error("szp_double_decompress_openmp_threadblock_randomaccess_arg", SZP_AMD_ERR_ARG);
        return;
    }
    lower_decompress("szp_double_decompress_openmp_threadblock_randomaccess_arg", 2, SZ_DOUBLE,
                     absErrBound, nbEle, blockSize, cmpBytes, newData);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Raw-file helpers (include/szp_rw.h)
// ---------------------------------------------------------------------------

namespace {

// Reads a whole file; nullptr (status SZ_FERR) when it cannot be opened or
// read.
unsigned char *read_file(const char *path, size_t *length, int *status)
{
    *status = SZ_FERR;
    FILE *f = fopen(path, "rb");
    if (!f) return nullptr;
    if (fseek(f, 0, SEEK_END) != 0) {
        fclose(f);
        return nullptr;
    }
    const long end = ftell(f);
    if (end < 0 || fseek(f, 0, SEEK_SET) != 0) {
        fclose(f);
        return nullptr;
    }
    const size_t n = (size_t)end;
    unsigned char *buf = (unsigned char *)malloc(n ? n : 1);
    if (!buf || fread(buf, 1, n, f) != n) {
        free(buf);
        fclose(f);
        return nullptr;
    }
    fclose(f);
    *length = n;
    *status = SZ_SCES;
    return buf;
}

// The values of a data file as host-order values of `width` bytes, swapping
// each value's bytes when the data's byte order differs from the host's
// (SZp/src/szp_rw.cc:424-459).
void *read_values(const char *path, size_t width, size_t *nbEle, int *status)
{
    size_t len = 0;
    unsigned char *buf = read_file(path, &len, status);
    if (!buf) return nullptr;
    *nbEle = len / width;
    if (szp_dataEndianType != szp_sysEndianType) {
        for (size_t i = 0; i < *nbEle; ++i)
            std::reverse(buf + i * width, buf + (i + 1) * width);
    }
    return buf;
}

void write_file(const void *data, size_t length, const char *path, int *status)
{
    *status = SZ_FERR;
    FILE *f = fopen(path, "wb");
    if (!f) return;
    const bool ok = length == 0 || fwrite(data, 1, length, f) == length;
    if (fclose(f) == 0 && ok) *status = SZ_SCES;
}

}  // namespace

extern "C" {

size_t szp_checkFileSize(char *srcFilePath, int *status)
{
    *status = SZ_FERR;
    FILE *f = fopen(srcFilePath, "rb");
    if (!f) return 0;
    size_t n = 0;
    if (fseek(f, 0, SEEK_END) == 0) {
        const long end = ftell(f);
        if (end >= 0) {
            n = (size_t)end;
            *status = SZ_SCES;
        }
    }
    fclose(f);
    return n;
}

unsigned char *szp_readByteData(char *srcFilePath, size_t *byteLength, int *status)
{
    return read_file(srcFilePath, byteLength, status);
}

float *szp_readFloatData(char *srcFilePath, size_t *nbEle, int *status)
{
    return (float *)read_values(srcFilePath, sizeof(float), nbEle, status);
}

double *szp_readDoubleData(char *srcFilePath, size_t *nbEle, int *status)
{
    return (double *)read_values(srcFilePath, sizeof(double), nbEle, status);
}

void szp_writeByteData(unsigned char *bytes, size_t byteLength, char *tgtFilePath, int *status)
{
    write_file(bytes, byteLength, tgtFilePath, status);
}

void szp_writeFloatData_inBytes(float *data, size_t nbEle, char *tgtFilePath, int *status)
{
    write_file(data, nbEle * sizeof(float), tgtFilePath, status);
}

void szp_writeDoubleData_inBytes(double *data, size_t nbEle, char *tgtFilePath, int *status)
{
    write_file(data, nbEle * sizeof(double), tgtFilePath, status);
}

}  // extern "C"
