// Internal declarations shared by the HIP kernels and the C-ABI layer.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace szp_amd {

constexpr int kWG = 256;                 // threads per workgroup (4 wave64)
constexpr int kMaxBlock = 2048;          // largest SZp block size handled on the GPU
constexpr int kOutBufBytes = 8704 + 64;  // LDS bytes for one round's encoded blocks
constexpr int kMaxRounds = 8;            // rounds of K items per tile
constexpr int kOutCap = 32768;           // LDS bytes buffering a compress tile before placement
constexpr int kDecTileBytes = 24576;     // LDS bytes holding a decode tile's stream bytes

// Geometry of the item sequence of one field.  The stream body is the
// concatenation, in element order, of "items": the raw head value of each
// non-empty chunk (4 bytes) followed by that chunk's SZp blocks.  A tile is K
// consecutive items; every item owns spb = ceil(B/8) thread slots ("octets").
struct Geom {
    uint64_t n;        // elements
    uint64_t tbs;      // floor(n / T), the size of chunks 0..T-2
    uint64_t c0;       // items in chunks 0..T-2 (0 when tbs == 0)
    uint64_t clast;    // items in chunk T-1
    uint64_t nitems;   // (T-1)*c0 + clast
    uint64_t nb0;      // SZp blocks in chunks 0..T-2
    uint64_t nblast;   // SZp blocks in chunk T-1
    uint64_t ntiles;   // ceil(nitems / K)
    uint64_t body;     // 4 + 8T: offset of chunk 0 in the stream
    uint32_t T;        // chunks (the reference's OpenMP team size)
    uint32_t B;        // SZp block size
    uint32_t spb;      // slots per item
    uint32_t K;        // items per tile
    uint32_t spb_magic;  // slot / spb == (slot * spb_magic) >> 16 for slot < 256
    uint32_t R;          // rounds per tile (1 when chunks are smaller than a tile)
};

// Host-side construction; returns false when (n, B, T) is outside what the
// kernels handle (B > kMaxBlock, T < 1).
bool make_geom(uint64_t n, int B, int T, Geom *g);

uint64_t compress_bound(uint64_t n, int B, int T);
uint64_t compress_workspace_bytes(uint64_t n, int B, int T);
uint64_t decompress_workspace_bytes(uint64_t stream_bytes, uint64_t n, int B, int T);

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
int launch_compress(const float *d_x, uint64_t n, int err_mode, double abs_eb, double rel,
                    int B, int T, uint8_t *d_out, uint64_t cap, uint64_t *d_size,
                    void *d_work, uint64_t work_bytes, hipStream_t stream,
                    void *d_index = nullptr, uint64_t idx_bytes = 0);

uint64_t index_bytes(uint64_t n, int B, int T);

int launch_decompress_indexed(const uint8_t *d_stream, uint64_t stream_bytes, const void *d_index,
                              uint64_t idx_bytes, uint64_t n, int B, int T, float *d_out,
                              void *d_work, uint64_t work_bytes, hipStream_t stream);

int launch_decompress(const uint8_t *d_stream, uint64_t stream_bytes, uint64_t n, int B, int T,
                      float *d_out, void *d_work, uint64_t work_bytes, hipStream_t stream);

// Device address of the workspace's error word (u32).
uint32_t *workspace_error_word(void *d_work);

// Per-kernel event timing (off by default).
void profile_enable(int enable);
int profile_count();
int profile_entry(int i, char *name, int cap, uint64_t *launches, double *total_ms);

// Bit width of a block maximum exactly as the reference computes it
// ((int)log2f((float)max) + 1), via clz and a threshold table (host copy).
int bit_width_host(uint32_t m);

}  // namespace szp_amd
