/*
 * szp_oracle.h -- CPU restatement of SZp fastmode1 compress/decompress
 * (float32 and float64).
 *
 * TEST INFRASTRUCTURE, NOT PRODUCT CODE.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load this library, and only as the
 * checker or as the timed CPU baseline.  The product (szp_amd/) never links,
 * loads or falls back to it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the SZp repository root).
 */
#ifndef SZP_ORACLE_H
#define SZP_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* b = 0 if max == 0 else (int)log2f((float)max) + 1 (SZp/src/szp_float.cc:253-262). */
int szpo_bit_width(uint32_t max_mag);

/* Same value through clz and the glibc-log2f threshold table used by the GPU kernels. */
int szpo_bit_width_table(uint32_t max_mag);

/* Exhaustively compares szpo_bit_width and szpo_bit_width_table over [lo, hi].
 * Returns the number of mismatches; *first_bad receives the first one. */
uint64_t szpo_check_bit_width_range(uint64_t lo, uint64_t hi, uint32_t *first_bad);

/* fl32(1 / eb): the FP32 division of SZp/src/szp_float.cc:185. */
float szpo_inverse_bound(float eb);

/* q = (int)((double)x * (double)inv) with the x86 cvttsd2si rule: NaN and
 * out-of-range products give INT32_MIN (SZp/src/szp_float.cc:209,228). */
int32_t szpo_quantize(float x, float inv);

/* Chunk t of T: [lo, hi) (SZp/src/szp_float.cc:186,194-199). */
void szpo_chunk(uint64_t n, int T, int t, uint64_t *lo, uint64_t *hi);

/* Bytes of one encoded block of L values with bit width b
 * (SZp/src/szp_float.cc:253-273, SZp/src/szp_TypeManager.cc:20-31). */
uint64_t szpo_block_bytes(int b, uint64_t L);

/* Encodes L int32 deltas as one block; returns bytes written. */
uint64_t szpo_encode_block(const int32_t *delta, uint64_t L, uint8_t *out);

/* Decodes one block of L deltas; returns bytes consumed, 0 on a bad width byte. */
uint64_t szpo_decode_block(const uint8_t *in, uint64_t L, int32_t *delta);

/* REL bound: fl32(fl32(max - min) * fl32(rel)) with the serial min/max scan of
 * SZp/src/szp.cc:36-47. */
float szpo_rel_bound(const float *x, uint64_t n, double rel);

/* Largest stream szpo_compress can emit for (n, B, T). */
uint64_t szpo_compress_bound(uint64_t n, int B, int T);

/* fastmode1 compress with explicit team size T (SZp/src/szp_float.cc:148-309).
 * workers = OpenMP threads used to run the T chunks (0 = T).  Returns 0 on
 * success, -1 on bad arguments or too small a buffer. */
int szpo_compress(const float *x, uint64_t n, float eb, int B, int T, int workers,
                  uint8_t *out, uint64_t cap, uint64_t *out_size);

/* fastmode1 decompress of a whole stream (SZp/src/szp.cc:118-139,
 * SZp/src/szpd_float.cc:23-128).  Returns 0 on success, -1 on a malformed
 * stream (out of range offsets, widths > 32, truncation). */
int szpo_decompress(const uint8_t *stream, uint64_t len, uint64_t n, int B, int T, int workers,
                    float *out);

/* szp_compress(fastMode=1, SZ_FLOAT, ...) semantics (SZp/src/szp.cc:23-58). */
int szpo_compress_api(const float *x, uint64_t n, int err_mode, double abs_eb, double rel,
                      int B, int T, int workers, uint8_t *out, uint64_t cap, uint64_t *out_size);

/* ---- float64 (SZp/src/szp_double.cc:148-318, SZp/src/szpd_double.cc:23-106):
 * the same chunks and blocks, an 8-byte big-endian eb (header es + 8T with
 * es = 8) and the quantizer (int)(x * (1 / eb)) evaluated in double. */
int32_t szpo_quantize_f64(double x, double inv);
double szpo_rel_bound_f64(const double *x, uint64_t n, double rel);
int szpo_compress_f64(const double *x, uint64_t n, double eb, int B, int T, int workers,
                      uint8_t *out, uint64_t cap, uint64_t *out_size);
int szpo_decompress_f64(const uint8_t *stream, uint64_t len, uint64_t n, int B, int T,
                        int workers, double *out);
int szpo_compress_api_f64(const double *x, uint64_t n, int err_mode, double abs_eb, double rel,
                          int B, int T, int workers, uint8_t *out, uint64_t cap,
                          uint64_t *out_size);

/* ---- any fastMode (1, or 2 = random access: SZp/src/szp_float.cc:752-870,
 * SZp/src/szpd_float.cc:425-530; each chunk cut into units of B values, a
 * unit = [q(first) i32][block of the other len - 1 deltas]) and dataType
 * (0 float, 1 double).  Decompress treats any fast != 1 as random access,
 * as szp_decompress does (SZp/src/szp.cc:135-151). */
int szpo_compress_mode(int fast, int dtype, const void *x, uint64_t n, int err_mode,
                       double abs_eb, double rel, int B, int T, int workers, uint8_t *out,
                       uint64_t cap, uint64_t *out_size);
int szpo_decompress_mode(int fast, int dtype, const uint8_t *stream, uint64_t len, uint64_t n,
                         int B, int T, int workers, void *out);
uint64_t szpo_compress_bound_mode(int fast, int dtype, uint64_t n, int B, int T);

/* Memory-access count returned by szp_float_single_thread_arg_record
 * (decompress = 0, SZp/src/szp_float.cc:607-750) or
 * szp_float_decompress_single_thread_arg_record (decompress = 1,
 * SZp/src/szpd_float.cc:216-317) for the T = 1 stream whose body (the bytes
 * after the 4-byte eb) is given. */
uint64_t szpo_record_count(const uint8_t *body, uint64_t n, int B, int decompress);

#ifdef __cplusplus
}
#endif
#endif
