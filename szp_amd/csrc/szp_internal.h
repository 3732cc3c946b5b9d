// Internal declarations shared by the HIP kernels and the C-ABI layer.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "szp_geom.h"

namespace szp_amd {

// dtype: 0 float (SZ_FLOAT), 1 double (SZ_DOUBLE); mode: the reference's
// fastMode, 1 (SZP_NONRANDOMACCESS) or 2 (SZP_RANDOMACCESS).
uint64_t compress_bound(uint64_t n, int B, int T, int dtype = 0, int mode = 1);
uint64_t compress_workspace_bytes(uint64_t n, int B, int T, int mode = 1);
uint64_t decompress_workspace_bytes(uint64_t stream_bytes, uint64_t n, int B, int T,
                                    int mode = 1);

// Error bits accumulated in the workspace control word.
enum : uint32_t {
    kErrCapacity = 1u,   // output capacity too small
    kErrStream = 2u,     // malformed compressed stream
    kErrTimeout = 4u,    // a look-back wait exceeded its bound
};

// Asynchronous launches on `stream`; all buffers are device memory.
// Return 0 or a positive szp_amd_status code for argument errors.
// d_index (optional, 256-byte aligned, index_bytes() long) receives the side
// index: each block's stream position and each tile's carry.
int launch_compress(int dtype, const void *d_x, uint64_t n, int err_mode, double abs_eb,
                    double rel, int B, int T, uint8_t *d_out, uint64_t cap, uint64_t *d_size,
                    void *d_work, uint64_t work_bytes, hipStream_t stream,
                    void *d_index = nullptr, uint64_t idx_bytes = 0, int mode = 1);

uint64_t index_bytes(uint64_t n, int B, int T, int mode = 1);

// launch_compress (float, fastMode 1) with its workspace in a static device
// array of kNoallocScratchBytes per device: allocates nothing.  Returns kCap
// (2) when the field's workspace exceeds the array.  A device-side error
// makes *d_size read ~0.  Calls that overlap on one device race.
constexpr uint64_t kNoallocScratchBytes = 8ull << 20;
// batch_launch's grid value for one compress workgroup per tile.
constexpr uint32_t kBatchGridPerTile = 0xFFFFFFFFu;
int launch_compress_noalloc(const float *d_x, uint64_t n, int err_mode, double abs_eb,
                            double rel, int B, int T, uint8_t *d_out, uint64_t cap,
                            uint64_t *d_size, hipStream_t stream);

int launch_decompress_indexed(int dtype, const uint8_t *d_stream, uint64_t stream_bytes,
                              const void *d_index, uint64_t idx_bytes, uint64_t n, int B, int T,
                              void *d_out, void *d_work, uint64_t work_bytes, hipStream_t stream,
                              int mode = 1);

int launch_decompress(int dtype, const uint8_t *d_stream, uint64_t stream_bytes, uint64_t n, int B,
                      int T, void *d_out, void *d_work, uint64_t work_bytes, hipStream_t stream,
                      int mode = 1);

// ---- batched forms: many fields in one persistent launch ----
// One field of a batch.  Compress: in = float[n], out = stream buffer of
// out_cap bytes.  Decompress: in = stream of in_bytes bytes, out = float[n],
// index = the side index written when it was compressed (required).
struct BatchField {
    const void *in;
    void *out;
    uint64_t n;
    uint64_t in_bytes;
    uint64_t out_cap;
    void *index;
    int err_mode;
    double abs_eb;
    double rel;
};

// Host copy of a prepared batch: where its tables live in the workspace and
// the sizes its launches need.
struct BatchPlan {
    void *work;
    uint64_t jobs, qmap, smap, parts, status, nstatus;  // byte offsets / u64 count
    uint64_t tiles, segs;
    uint32_t F, B, T, spb, index, decode;
};

uint64_t batch_workspace_bytes(const BatchField *f, int F, int B, int T, int decode);
// Builds the batch tables on the host and copies them into d_work (returns
// after the copy).  d_sizes: device u64[F] for compressed sizes (compress).
int batch_prepare(const BatchField *f, int F, int B, int T, int decode, uint64_t *d_sizes,
                  void *d_work, uint64_t work_bytes, BatchPlan *plan);
// Enqueues the batch's kernels only (capturable in a graph).  ticket: queue
// entries a workgroup takes at a time; grid: persistent workgroups (0 = one
// per resident slot).
int batch_launch(const BatchPlan &plan, uint32_t ticket, uint32_t grid, hipStream_t stream);
// Workgroups of the batch's persistent kernel that fit the device at once.
uint32_t batch_resident_grid(const BatchPlan &plan);

// Device address of the workspace's error word (u32).
uint32_t *workspace_error_word(void *d_work);

// Per-kernel event timing (off by default).
void profile_enable(int enable, const char *only = nullptr);
int profile_count();
int profile_entry(int i, char *name, int cap, uint64_t *launches, double *total_ms);

// Bit width of a block maximum exactly as the reference computes it
// ((int)log2f((float)max) + 1), via clz and a threshold table (host copy).
int bit_width_host(uint32_t m);

}  // namespace szp_amd
