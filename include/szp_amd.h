/*
 * szp_amd.h -- device-resident API of libszp_amd.so (SZp compress/decompress
 * on MI355X: fastMode 1 and 2, float32 and float64).  Plain pointers and sizes; `stream` is a hipStream_t passed as
 * void* (NULL = the default stream).
 *
 * The *_dev entry points enqueue kernels and return at once: no host
 * synchronisation and no allocation.  The caller provides device buffers:
 * input/output arrays, the output stream (16-byte aligned) and a workspace
 * (256-byte aligned) of the size the *_workspace_size queries return.  Calls
 * that share a workspace must be ordered (same stream).  Errors found on the
 * device (capacity, malformed stream) accumulate in the workspace and are
 * read by szp_amd_check_status().
 *
 * Streams are byte-identical to the reference's szp_compress(SZP_NONRANDOMACCESS,
 * SZ_FLOAT, ...) run with an OpenMP team of nbThreads threads
 * (SZp/src/szp_float.cc:148-309); decompressed values are bit-identical to
 * SZp/src/szpd_float.cc:23-128.
 */
#ifndef SZP_AMD_API_H
#define SZP_AMD_API_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum szp_amd_status {
    SZP_AMD_OK = 0,
    SZP_AMD_ERR_ARG = 1,         /* invalid argument (sizes, modes, alignment) */
    SZP_AMD_ERR_CAPACITY = 2,    /* output or workspace too small */
    SZP_AMD_ERR_HIP = 3,         /* HIP runtime error */
    SZP_AMD_ERR_STREAM = 4,      /* malformed compressed stream */
    SZP_AMD_ERR_TIMEOUT = 5,     /* a device-side wait exceeded its bound */
    SZP_AMD_ERR_UNSUPPORTED = 6, /* fastMode or dataType the library does not serve */
    SZP_AMD_ERR_NOMEM = 7        /* host or device allocation failed */
} szp_amd_status;

/* No reference counterpart (the reference mallocs 4N+4 and can overflow,
 * SZp/src/szp_float.cc:156): the largest stream for (nbEle, blockSize, T). */
uint64_t szp_amd_compress_bound(uint64_t nbEle, int blockSize, int nbThreads);

uint64_t szp_amd_compress_workspace_size(uint64_t nbEle, int blockSize, int nbThreads);
uint64_t szp_amd_decompress_workspace_size(uint64_t streamBytes, uint64_t nbEle, int blockSize,
                                           int nbThreads);

/* Device form of szp_compress(SZP_NONRANDOMACCESS, SZ_FLOAT, ...)
 * (SZp/src/szp.cc:23-58): errBoundMode ABS (0) or REL (1).  Writes the
 * stream to d_out and its length (u64) to *d_outSize. */
int szp_amd_compress_dev(const float *d_data, uint64_t nbEle, int errBoundMode,
                         double absErrBound, double relBoundRatio, int blockSize, int nbThreads,
                         unsigned char *d_out, uint64_t outCapacity, uint64_t *d_outSize,
                         void *d_work, uint64_t workBytes, void *stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): szp_amd_compress_dev with its workspace taken from a
 * static device array of SZP_AMD_NOALLOC_SCRATCH_BYTES per device, held until
 * the process ends.  Allocates nothing (the array's address is looked up on a
 * device's first call).  A field whose compress workspace
 * (szp_amd_compress_workspace_size) exceeds the array is refused with
 * SZP_AMD_ERR_CAPACITY before any launch (about 8.5e9 elements when blockSize
 * is a multiple of 8).
 * Every call on a device shares that device's array: calls that overlap on
 * one device race, so order them (one stream).  Device-side errors (output
 * capacity, wait timeout) make *d_outSize read UINT64_MAX. */
#define SZP_AMD_NOALLOC_SCRATCH_BYTES (8u << 20)
int szp_amd_compress_dev_noalloc(const float *d_data, uint64_t nbEle, int errBoundMode,
                                 double absErrBound, double relBoundRatio, int blockSize,
                                 int nbThreads, unsigned char *d_out, uint64_t outCapacity,
                                 uint64_t *d_outSize, void *stream);

/* Device form of szp_decompress(SZP_NONRANDOMACCESS, SZ_FLOAT, ...)
 * (SZp/src/szp.cc:118-139).  streamBytes is the stream length or any upper
 * bound within the d_stream allocation. */
int szp_amd_decompress_dev(const unsigned char *d_stream, uint64_t streamBytes, uint64_t nbEle,
                           int blockSize, int nbThreads, float *d_out, void *d_work,
                           uint64_t workBytes, void *stream);

/* Side index (no reference counterpart; the stream bytes are unchanged):
 * the stream position of every block and, per tile of the codec, the
 * quantized value preceding the tile.  The reference's format has no block
 * index, so a decoder must otherwise recover block boundaries by parsing
 * (szp_amd_decompress_dev does).  The *_indexed pair writes and consumes it;
 * the index is only valid for the stream it was written with. */
uint64_t szp_amd_index_size(uint64_t nbEle, int blockSize, int nbThreads);

int szp_amd_compress_dev_indexed(const float *d_data, uint64_t nbEle, int errBoundMode,
                                 double absErrBound, double relBoundRatio, int blockSize,
                                 int nbThreads, unsigned char *d_out, uint64_t outCapacity,
                                 uint64_t *d_outSize, void *d_index, uint64_t indexBytes,
                                 void *d_work, uint64_t workBytes, void *stream);

int szp_amd_decompress_dev_indexed(const unsigned char *d_stream, uint64_t streamBytes,
                                   const void *d_index, uint64_t indexBytes, uint64_t nbEle,
                                   int blockSize, int nbThreads, float *d_out, void *d_work,
                                   uint64_t workBytes, void *stream);

/* Batched forms (no reference counterpart: the reference compresses one
 * field per call, SZp/src/szp.cc:23-115).  Many fields of any sizes share
 * one persistent launch: workgroups take tickets of consecutive tiles from a
 * device-side queue that lists every tile of every field.  All fields of a
 * batch use the same block size B and team size T; each field's stream is
 * byte-identical to a single-field call's.
 *
 * Compress: in = float[n] (device), out = stream buffer of out_cap bytes
 * (16-byte aligned, szp_amd_compress_bound() is always enough), err_mode /
 * abs_eb / rel as for szp_amd_compress_dev, index = optional side index
 * (szp_amd_index_size() bytes, 256-byte aligned; all fields or none).
 * Decompress: in = stream (16-byte aligned) of in_bytes bytes, out =
 * float[n], index = the side index written when it was compressed
 * (required).  n >= 1 for every field. */
typedef struct szp_amd_field {
    const void *in;
    void *out;
    uint64_t n;
    uint64_t in_bytes;
    uint64_t out_cap;
    void *index;
    int err_mode;
    double abs_eb;
    double rel;
} szp_amd_field;

/* A prepared batch (opaque). */
typedef struct szp_amd_batch {
    uint64_t opaque[16];
} szp_amd_batch;

uint64_t szp_amd_batch_workspace_size(const szp_amd_field *fields, int numFields, int blockSize,
                                      int nbThreads, int decompress);

/* Builds the batch's tables and copies them into d_work (256-byte aligned,
 * workspace_size bytes); returns after the copy.  d_sizes: device uint64[F]
 * receiving the compressed sizes (compress; may be NULL for decompress).
 * The fields' buffers must stay valid while the batch is launched. */
int szp_amd_batch_prepare(const szp_amd_field *fields, int numFields, int blockSize,
                          int nbThreads, int decompress, uint64_t *d_sizes, void *d_work,
                          uint64_t workBytes, szp_amd_batch *batch);

/* Enqueues the batch on `stream` (kernels only, so a graph may capture it
 * and replay it).  ticket: tiles a workgroup takes at a time (>= 1;
 * decompress only, compress takes one tile at a time because a tile waits for
 * the tiles before it to start); grid: persistent workgroups (0 = one per
 * resident slot), or, for compress only, SZP_AMD_BATCH_GRID_PER_TILE: one
 * workgroup per tile, each taking one ticket.  Status as for the
 * single-field calls: szp_amd_check_status(d_work, stream). */
#define SZP_AMD_BATCH_GRID_PER_TILE 0xFFFFFFFFu
int szp_amd_batch_launch(const szp_amd_batch *batch, uint32_t ticket, uint32_t grid,
                         void *stream);

/* Workgroups of the batch's persistent kernel that fit on the current device
 * at once (CUs x workgroups per CU by occupancy): the resident grid that
 * grid = 0 launches, and the base of the autotuner's grid candidates. */
uint32_t szp_amd_batch_resident_grid(const szp_amd_batch *batch);

/* Synchronises `stream` and returns the status accumulated in d_work. */
int szp_amd_check_status(void *d_work, void *stream);

/* Exact length of a host-resident stream, from its header and block widths. */
uint64_t szp_amd_stream_length(const unsigned char *bytes, uint64_t nbEle, int blockSize,
                               int nbThreads);

/* Typed forms: fastMode SZP_NONRANDOMACCESS (1) or SZP_RANDOMACCESS (2),
 * dataType SZ_FLOAT (0) or SZ_DOUBLE (1).  No reference counterpart as
 * device calls; each stream is byte-identical to the reference's
 * szp_compress(fastMode, dataType, ...) with T = nbThreads:
 *   fastMode 1, float:  SZp/src/szp_float.cc:148-309 / SZp/src/szpd_float.cc:23-128
 *   fastMode 1, double: SZp/src/szp_double.cc:148-318 / SZp/src/szpd_double.cc:23-106
 *     (an 8-byte big-endian eb; the quantizer (int)(x * (1 / eb)) in double;
 *     REL bound from the double range, SZp/src/szp.cc:60-84)
 *   fastMode 2, float:  SZp/src/szp_float.cc:752-870 / SZp/src/szpd_float.cc:425-530
 *   fastMode 2, double: SZp/src/szp_double.cc:759-920 / SZp/src/szpd_double.cc:431-534
 *     (each chunk cut into units of blockSize values from its start; a unit
 *     is the raw quantized first value and one block of the others).
 * d_data / d_out are float[nbEle] or double[nbEle].  d_index is optional on
 * both sides: compress writes the side index when given; decompress uses it
 * when given, else parses the stream.  Workspace and index sizes depend on
 * fastMode, not on dataType. */
uint64_t szp_amd_compress_bound_typed(int fastMode, int dataType, uint64_t nbEle, int blockSize,
                                      int nbThreads);
uint64_t szp_amd_compress_workspace_size_typed(int fastMode, uint64_t nbEle, int blockSize,
                                               int nbThreads);
uint64_t szp_amd_decompress_workspace_size_typed(int fastMode, uint64_t streamBytes,
                                                 uint64_t nbEle, int blockSize, int nbThreads);
uint64_t szp_amd_index_size_typed(int fastMode, uint64_t nbEle, int blockSize, int nbThreads);

int szp_amd_compress_dev_typed(int fastMode, int dataType, const void *d_data, uint64_t nbEle,
                               int errBoundMode, double absErrBound, double relBoundRatio,
                               int blockSize, int nbThreads, unsigned char *d_out,
                               uint64_t outCapacity, uint64_t *d_outSize, void *d_index,
                               uint64_t indexBytes, void *d_work, uint64_t workBytes,
                               void *stream);

int szp_amd_decompress_dev_typed(int fastMode, int dataType, const unsigned char *d_stream,
                                 uint64_t streamBytes, const void *d_index, uint64_t indexBytes,
                                 uint64_t nbEle, int blockSize, int nbThreads, void *d_out,
                                 void *d_work, uint64_t workBytes, void *stream);

uint64_t szp_amd_stream_length_typed(int fastMode, int dataType, const unsigned char *bytes,
                                     uint64_t nbEle, int blockSize, int nbThreads);

/* T used by the drop-in entry points of szp.h; set_threads(0) restores auto. */
int szp_amd_default_threads(void);
void szp_amd_set_threads(int nbThreads);

/* Width of a block maximum as the reference computes it: 0 for 0, else
 * (int)log2f((float)max) + 1 (SZp/src/szp_float.cc:253-262). */
int szp_amd_bit_width(uint32_t maxMagnitude);

/* Per-kernel device timing, the device counterpart of the reference's
 * memory-access-count instrumentation (SZp/src/szp_float.cc:607-750).  While
 * enabled, each kernel launch is bracketed by HIP events on its stream;
 * reading syncs those events and returns the per-kernel launch count and
 * summed milliseconds.  enable(1) and enable(0) both clear the totals. */
void szp_amd_profile_enable(int enable);
/* As szp_amd_profile_enable(1), timing only the launches of one kernel
 * (e.g. "k_compress"), so that the events disturb the others least. */
void szp_amd_profile_enable_only(const char *kernel);
int szp_amd_profile_count(void);
int szp_amd_profile_entry(int index, char *name, int nameCapacity, uint64_t *launches,
                          double *totalMs);

const char *szp_amd_status_string(int status);
/* Message of the last failure of a szp.h entry point in this thread. */
const char *szp_amd_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* SZP_AMD_API_H */
