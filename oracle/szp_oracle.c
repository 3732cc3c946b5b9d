/*
 * szp_oracle.c -- CPU restatement of SZp fastmode1 compress/decompress
 * (float32, and float64 through the same chunk/block code).
 *
 * TEST INFRASTRUCTURE, NOT PRODUCT CODE.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load this library, and only as the
 * checker or as the timed CPU baseline; the product never routes through it.
 *
 * Written from the reference's observable behaviour (paths relative to the SZp
 * repository root):
 *   SZp/src/szp.cc:23-58,118-139        API, eb derivation, header eb
 *   SZp/src/szp_float.cc:148-309        fastmode1 compress (chunks, blocks, concat)
 *   SZp/src/szpd_float.cc:23-128        fastmode1 decompress
 *   SZp/src/szp_TypeManager.cc:13-94    byte planes + residual bits (pack)
 *   SZp/src/szp_TypeManager.cc:436-517  (unpack)
 *   SZp/src/szp_TypeManager.cc:820-844,902-944  sign bitmap (pack / unpack)
 *   SZp/include/szp_CompressionToolkit.h:511-546 big-endian eb bytes
 *   SZp/src/szp.cc:60-84, SZp/src/szp_double.cc:148-318,
 *   SZp/src/szpd_double.cc:23-106       the double path (8-byte eb, FP64 quantizer)
 *   SZp/src/szp_float.cc:752-870,
 *   SZp/src/szpd_float.cc:425-530       fastMode 2 (random access) chunks of units
 *
 * Stream format (all multi-byte integers little-endian unless noted):
 *   [eb: f32 big-endian][T x u64 chunk offsets, exclusive scan of chunk sizes]
 *   [chunk 0]...[chunk T-1]
 *   chunk t = [] if empty, else [q(lo_t): i32][block 0][block 1]...
 *   block  = [b] if b == 0, else
 *            [b][ceil(L/8) sign bytes, element i -> bit (7 - i%8) of byte i/8]
 *            [(b/8)*L bytes: element-major little-endian bytes of (mag >> b%8)]
 *            [ceil((b%8)*L/8) bytes: MSB-first stream of the low b%8 bits]
 */
#include "szp_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* SZp/src/szp_float.cc:253-262: bit_count = (int)(log2f(max)) + 1, 0 for max == 0. */
int szpo_bit_width(uint32_t max_mag)
{
    if (max_mag == 0) return 0;
    return (int)log2f((float)max_mag) + 1;
}

/* glibc log2f rounds log2 of values just below 2^k up to k for k >= 21, so the
 * width is one more than the number of significant bits for max in
 * [kWidthBump[k], 2^k).  Pinned exhaustively against szpo_bit_width. */
static const uint32_t kWidthBump[33] = {
    0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0,
    2097151u, 4194302u, 8388603u, 16777205u, 33554410u, 67108819u,
    134217637u, 268435273u, 536870545u, 1073741089u, 2147482177u, 4294964353u};

int szpo_bit_width_table(uint32_t m)
{
    if (m == 0) return 0;
    int k = 32 - __builtin_clz(m);
    return k + ((k >= 21 && m >= kWidthBump[k]) ? 1 : 0);
}

uint64_t szpo_check_bit_width_range(uint64_t lo, uint64_t hi, uint32_t *first_bad)
{
    uint64_t bad = 0;
    uint64_t first = UINT64_MAX;
#pragma omp parallel for reduction(+ : bad) reduction(min : first) schedule(static)
    for (uint64_t v = lo; v <= hi; ++v) {
        if (szpo_bit_width((uint32_t)v) != szpo_bit_width_table((uint32_t)v)) {
            ++bad;
            if (v < first) first = v;
        }
    }
    if (first_bad) *first_bad = (first == UINT64_MAX) ? 0u : (uint32_t)first;
    return bad;
}

/* SZp/src/szp_float.cc:171,185: inver_bound = 1 / absErrBound is an int/float
 * division, evaluated in single precision and then widened to double. */
float szpo_inverse_bound(float eb)
{
    volatile float one = 1.0f;
    return one / eb;
}

/* SZp/src/szp_float.cc:209,228: (int)(float * double).  The product of two
 * floats is exact in double; x86 cvttsd2si truncates toward zero and returns
 * 0x80000000 for NaN and for values outside the int32 range. */
int32_t szpo_quantize(float x, float inv)
{
    double p = (double)x * (double)inv;
    if (!(p > -2147483649.0 && p < 2147483648.0)) return INT32_MIN;
    return (int32_t)p;
}

/* SZp/src/szp_double.cc:185,209,228: inver_bound = 1 / absErrBound in double,
 * q = (int)(x * inver_bound) with the double product rounded, then the same
 * x86 truncation rule. */
int32_t szpo_quantize_f64(double x, double inv)
{
    volatile double p = x * inv;
    if (!(p > -2147483649.0 && p < 2147483648.0)) return INT32_MIN;
    return (int32_t)p;
}

/* SZp/src/szp_float.cc:186,194-199 (compressor, 32-bit) and
 * SZp/src/szpd_float.cc:40-47 (decompressor, 64-bit): both give
 * lo = t * floor(n / T), hi = lo + floor(n / T), the last chunk ends at n, for
 * every n <= 2^32. */
void szpo_chunk(uint64_t n, int T, int t, uint64_t *lo, uint64_t *hi)
{
    uint64_t tbs = n / (uint64_t)T;
    *lo = (uint64_t)t * tbs;
    *hi = (t == T - 1) ? n : (uint64_t)(t + 1) * tbs;
}

/* SZp/src/szp_float.cc:253-273 and SZp/src/szp_TypeManager.cc:20-31. */
uint64_t szpo_block_bytes(int b, uint64_t L)
{
    if (b == 0) return 1;
    return 1 + (L + 7) / 8 + (uint64_t)(b / 8) * L + ((uint64_t)(b % 8) * L + 7) / 8;
}

/* Magnitude and sign of one Lorenzo delta (SZp/src/szp_float.cc:231-251):
 * sign = delta < 0, magnitude = |delta| as an unsigned 32-bit value. */
static inline uint32_t mag_of(int32_t d) { return d < 0 ? 0u - (uint32_t)d : (uint32_t)d; }

uint64_t szpo_encode_block(const int32_t *delta, uint64_t L, uint8_t *out)
{
    uint32_t maxm = 0;
    for (uint64_t i = 0; i < L; ++i) {
        uint32_t m = mag_of(delta[i]);
        if (m > maxm) maxm = m;
    }
    int b = szpo_bit_width(maxm);
    out[0] = (uint8_t)b;
    if (b == 0) return 1;

    /* Sign bitmap, MSB first (SZp/src/szp_TypeManager.cc:820-844). */
    uint64_t S = (L + 7) / 8;
    uint8_t *p = out + 1;
    memset(p, 0, S);
    for (uint64_t i = 0; i < L; ++i)
        if (delta[i] < 0) p[i >> 3] |= (uint8_t)(0x80u >> (i & 7));
    p += S;

    /* Whole bytes of mag >> r, element-major little-endian
     * (SZp/src/szp_TypeManager.cc:36-52). */
    int nb = b / 8, r = b % 8;
    for (uint64_t i = 0; i < L; ++i) {
        uint32_t v = mag_of(delta[i]) >> r;
        for (int j = 0; j < nb; ++j) *p++ = (uint8_t)(v >> (8 * j));
    }

    /* Low r bits, one MSB-first bit stream (SZp/src/szp_TypeManager.cc:53-91 and
     * the 1..7-bit packers at :122-434). */
    uint64_t tail = ((uint64_t)r * L + 7) / 8;
    if (r) {
        memset(p, 0, tail);
        uint64_t bit = 0;
        for (uint64_t i = 0; i < L; ++i) {
            uint32_t v = mag_of(delta[i]) & ((1u << r) - 1u);
            for (int k = r - 1; k >= 0; --k, ++bit)
                if ((v >> k) & 1u) p[bit >> 3] |= (uint8_t)(0x80u >> (bit & 7));
        }
    }
    return 1 + S + (uint64_t)nb * L + tail;
}

uint64_t szpo_decode_block(const uint8_t *in, uint64_t L, int32_t *delta)
{
    int b = in[0];
    if (b > 32) return 0;
    if (b == 0) {
        for (uint64_t i = 0; i < L; ++i) delta[i] = 0;
        return 1;
    }
    uint64_t S = (L + 7) / 8;
    const uint8_t *sg = in + 1;
    const uint8_t *hi = sg + S;
    int nb = b / 8, r = b % 8;
    const uint8_t *lo = hi + (uint64_t)nb * L;
    for (uint64_t i = 0; i < L; ++i) {
        uint32_t v = 0;
        for (int j = 0; j < nb; ++j) v |= (uint32_t)hi[i * nb + j] << (8 * j);
        uint32_t low = 0;
        for (int k = 0; k < r; ++k) {
            uint64_t bit = i * (uint64_t)r + k;
            low = (low << 1) | ((lo[bit >> 3] >> (7 - (bit & 7))) & 1u);
        }
        uint32_t m = (r ? (v << r) : v) | low;
        int neg = (sg[i >> 3] >> (7 - (i & 7))) & 1;
        /* SZp/src/szpd_float.cc:101-110: diff = sign ? 0 - mag : mag. */
        delta[i] = (int32_t)(neg ? 0u - m : m);
    }
    return 1 + S + (uint64_t)nb * L + ((uint64_t)r * L + 7) / 8;
}

/* SZp/src/szp.cc:36-47. */
float szpo_rel_bound(const float *x, uint64_t n, double rel)
{
    float mn = x[0], mx = x[0];
    for (uint64_t i = 0; i < n; ++i) {
        float v = x[i];
        if (mn > v) mn = v;
        else if (mx < v) mx = v;
    }
    float range = mx - mn;
    return range * (float)rel;
}

uint64_t szpo_compress_bound(uint64_t n, int B, int T)
{
    uint64_t total = 4 + 8 * (uint64_t)T;
    for (int t = 0; t < T; ++t) {
        uint64_t lo, hi;
        szpo_chunk(n, T, t, &lo, &hi);
        uint64_t m = hi - lo;
        if (m == 0) continue;
        uint64_t nblk = (m - 1 + (uint64_t)B - 1) / (uint64_t)B;
        total += 4 + nblk * (1 + ((uint64_t)B + 7) / 8) + 4 * (m - 1);
    }
    return total;
}

/* The values of one field: float (f, inv_f) or double (d, inv_d); fast is
 * the reference's fastMode (1, or 2 for random access). */
typedef struct {
    const float *f;
    const double *d;
    float inv_f;
    double inv_d;
    int fast;
} Field;

static inline int32_t quant_at(const Field *v, uint64_t i)
{
    return v->f ? szpo_quantize(v->f[i], v->inv_f) : szpo_quantize_f64(v->d[i], v->inv_d);
}

/* One chunk of SZp/src/szp_float.cc:201-275 (SZp/src/szp_double.cc:201-275)
 * into dst; returns bytes written. */
static uint64_t compress_chunk(const Field *x, uint64_t lo, uint64_t hi, int B, uint8_t *dst)
{
    if (lo >= hi) return 0;
    uint8_t *p = dst;
    int32_t prior = quant_at(x, lo);
    memcpy(p, &prior, 4);
    p += 4;
    int32_t *d = (int32_t *)malloc(sizeof(int32_t) * (size_t)B);
    for (uint64_t i = lo + 1; i < hi; i += (uint64_t)B) {
        uint64_t L = (i + (uint64_t)B > hi) ? hi - i : (uint64_t)B;
        for (uint64_t j = 0; j < L; ++j) {
            int32_t cur = quant_at(x, i + j);
            d[j] = (int32_t)((uint32_t)cur - (uint32_t)prior);
            prior = cur;
        }
        p += szpo_encode_block(d, L, p);
    }
    free(d);
    return (uint64_t)(p - dst);
}

/* One chunk of fastMode 2 (SZp/src/szp_float.cc:817-870): units of B values
 * from lo, each [q(first) i32][block of the other len - 1 deltas]; a unit of
 * one value still writes its zero width byte (:857-862). */
static uint64_t compress_chunk_ra(const Field *x, uint64_t lo, uint64_t hi, int B, uint8_t *dst)
{
    uint8_t *p = dst;
    int32_t *d = (int32_t *)malloc(sizeof(int32_t) * (size_t)B);
    for (uint64_t i = lo; i < hi; i += (uint64_t)B) {
        uint64_t len = (i + (uint64_t)B > hi) ? hi - i : (uint64_t)B;
        int32_t prior = quant_at(x, i);
        memcpy(p, &prior, 4);
        p += 4;
        for (uint64_t j = 1; j < len; ++j) {
            int32_t cur = quant_at(x, i + j);
            d[j - 1] = (int32_t)((uint32_t)cur - (uint32_t)prior);
            prior = cur;
        }
        p += szpo_encode_block(d, len - 1, p);
    }
    free(d);
    return (uint64_t)(p - dst);
}

/* Largest chunk of m values: fastMode 1 = head + blocks of width 32;
 * fastMode 2 = per unit 4 + 1 + ceil((B-1)/8) bytes plus 4 per other value. */
static uint64_t chunk_bound(int fast, uint64_t m, int B)
{
    if (!m) return 0;
    if (fast == 2) {
        uint64_t units = (m + (uint64_t)B - 1) / (uint64_t)B;
        return units * (5 + ((uint64_t)B - 1 + 7) / 8) + 4 * (m - units);
    }
    return 4 + ((m - 1 + B - 1) / B) * (1 + (B + 7) / 8) + 4 * (m - 1);
}

static void put_be_f32(uint8_t *b, float v)
{
    uint32_t u;
    memcpy(&u, &v, 4);
    b[0] = (uint8_t)(u >> 24); b[1] = (uint8_t)(u >> 16); b[2] = (uint8_t)(u >> 8); b[3] = (uint8_t)u;
}

static float get_be_f32(const uint8_t *b)
{
    uint32_t u = ((uint32_t)b[0] << 24) | ((uint32_t)b[1] << 16) | ((uint32_t)b[2] << 8) | b[3];
    float v;
    memcpy(&v, &u, 4);
    return v;
}

static void put_be_f64(uint8_t *b, double v)
{
    uint64_t u;
    memcpy(&u, &v, 8);
    for (int i = 0; i < 8; ++i) b[i] = (uint8_t)(u >> (56 - 8 * i));
}

static double get_be_f64(const uint8_t *b)
{
    uint64_t u = 0;
    for (int i = 0; i < 8; ++i) u = (u << 8) | b[i];
    double v;
    memcpy(&v, &u, 8);
    return v;
}

/* Chunks, header offsets and concatenation (SZp/src/szp_float.cc:175-303);
 * es = 4 (float) or 8 (double) bytes of eb already written at out[0]. */
static int compress_body(const Field *x, uint64_t n, uint64_t es, int B, int T, int workers,
                         uint8_t *out, uint64_t cap, uint64_t *out_size);

int szpo_compress(const float *x, uint64_t n, float eb, int B, int T, int workers,
                  uint8_t *out, uint64_t cap, uint64_t *out_size)
{
    if (B < 1 || T < 1 || !out || !out_size) return -1;
    if (cap < 4 + 8 * (uint64_t)T) return -1;
    Field f = {x, NULL, szpo_inverse_bound(eb), 0.0, 1};
    put_be_f32(out, eb);
    return compress_body(&f, n, 4, B, T, workers, out, cap, out_size);
}

int szpo_compress_f64(const double *x, uint64_t n, double eb, int B, int T, int workers,
                      uint8_t *out, uint64_t cap, uint64_t *out_size)
{
    if (B < 1 || T < 1 || !out || !out_size) return -1;
    if (cap < 8 + 8 * (uint64_t)T) return -1;
    volatile double one = 1.0;
    Field f = {NULL, x, 0.0f, one / eb, 1};
    put_be_f64(out, eb);
    return compress_body(&f, n, 8, B, T, workers, out, cap, out_size);
}

static int compress_body(const Field *x, uint64_t n, uint64_t es, int B, int T, int workers,
                         uint8_t *out, uint64_t cap, uint64_t *out_size)
{
    uint64_t hdr = es + 8 * (uint64_t)T;
    if (workers <= 0) workers = T;

    uint8_t **part = (uint8_t **)calloc((size_t)T, sizeof(uint8_t *));
    uint64_t *size = (uint64_t *)calloc((size_t)T, sizeof(uint64_t));
    int fail = 0;
#pragma omp parallel for num_threads(workers) schedule(static, 1) reduction(| : fail)
    for (int t = 0; t < T; ++t) {
        uint64_t lo, hi;
        szpo_chunk(n, T, t, &lo, &hi);
        uint64_t m = hi - lo;
        uint64_t bound = chunk_bound(x->fast, m, B);
        part[t] = (uint8_t *)malloc(bound ? bound : 1);
        if (!part[t]) { fail = 1; continue; }
        size[t] = x->fast == 2 ? compress_chunk_ra(x, lo, hi, B, part[t])
                               : compress_chunk(x, lo, hi, B, part[t]);
    }
    uint64_t off = 0;
    for (int t = 0; t < T; ++t) {
        memcpy(out + es + 8 * (uint64_t)t, &off, 8);
        off += size[t];
    }
    if (!fail && hdr + off > cap) fail = 1;
    if (!fail) {
        uint64_t pos = hdr;
        for (int t = 0; t < T; ++t) {
            memcpy(out + pos, part[t], size[t]);
            pos += size[t];
        }
        *out_size = hdr + off;
    }
    for (int t = 0; t < T; ++t) free(part[t]);
    free(part);
    free(size);
    return fail ? -1 : 0;
}

/* x-hat = (V)q * eb in the value type: fl32((float)q * eb) for float
 * (SZp/src/szpd_float.cc:63-70,101-116), (double)q * eb for double
 * (SZp/src/szpd_double.cc:45,90). */
typedef struct {
    float *f;
    double *d;
    float eb_f;
    double eb_d;
    int fast;
} Out;

static inline void put_val(const Out *o, uint64_t i, int32_t q)
{
    if (o->f) o->f[i] = (float)q * o->eb_f;
    else o->d[i] = (double)q * o->eb_d;
}

/* One chunk of SZp/src/szpd_float.cc:58-117; returns 0 or -1 when the chunk
 * would read past `end`. */
static int decompress_chunk(const uint8_t *p, const uint8_t *end, uint64_t lo, uint64_t hi,
                            const Out *out, int B)
{
    if (lo >= hi) return 0;
    if (p + 4 > end) return -1;
    int32_t prior;
    memcpy(&prior, p, 4);
    p += 4;
    put_val(out, lo, prior);
    int32_t *d = (int32_t *)malloc(sizeof(int32_t) * (size_t)B);
    int rc = 0;
    for (uint64_t i = lo + 1; i < hi; i += (uint64_t)B) {
        uint64_t L = (i + (uint64_t)B > hi) ? hi - i : (uint64_t)B;
        if (p >= end || p[0] > 32 || p + szpo_block_bytes(p[0], L) > end) { rc = -1; break; }
        p += szpo_decode_block(p, L, d);
        for (uint64_t j = 0; j < L; ++j) {
            prior = (int32_t)((uint32_t)prior + (uint32_t)d[j]);
            put_val(out, i + j, prior);
        }
    }
    free(d);
    return rc;
}

/* One chunk of fastMode 2 (SZp/src/szpd_float.cc:470-520): a unit of one
 * value has no width byte to read (the compressor's zero byte is skipped
 * only because nothing follows it in the chunk). */
static int decompress_chunk_ra(const uint8_t *p, const uint8_t *end, uint64_t lo, uint64_t hi,
                               const Out *out, int B)
{
    int32_t *d = (int32_t *)malloc(sizeof(int32_t) * (size_t)B);
    int rc = 0;
    for (uint64_t i = lo; i < hi; i += (uint64_t)B) {
        uint64_t len = (i + (uint64_t)B > hi) ? hi - i : (uint64_t)B;
        if (p + 4 > end) { rc = -1; break; }
        int32_t prior;
        memcpy(&prior, p, 4);
        p += 4;
        put_val(out, i, prior);
        if (len == 1) continue;
        uint64_t L = len - 1;
        if (p >= end || p[0] > 32 || p + szpo_block_bytes(p[0], L) > end) { rc = -1; break; }
        p += szpo_decode_block(p, L, d);
        for (uint64_t j = 0; j < L; ++j) {
            prior = (int32_t)((uint32_t)prior + (uint32_t)d[j]);
            put_val(out, i + 1 + j, prior);
        }
    }
    free(d);
    return rc;
}

static int decompress_body(const uint8_t *stream, uint64_t len, uint64_t es, uint64_t n, int B,
                           int T, int workers, const Out *out)
{
    if (B < 1 || T < 1) return -1;
    uint64_t hdr = es + 8 * (uint64_t)T;
    if (len < hdr) return -1;
    if (workers <= 0) workers = T;
    const uint8_t *body = stream + hdr;
    const uint8_t *end = stream + len;
    int fail = 0;
#pragma omp parallel for num_threads(workers) schedule(static, 1) reduction(| : fail)
    for (int t = 0; t < T; ++t) {
        uint64_t lo, hi, off;
        szpo_chunk(n, T, t, &lo, &hi);
        memcpy(&off, stream + es + 8 * (uint64_t)t, 8);
        if (off > len - hdr) { fail = 1; continue; }
        int rc = out->fast == 2 ? decompress_chunk_ra(body + off, end, lo, hi, out, B)
                                : decompress_chunk(body + off, end, lo, hi, out, B);
        if (rc) fail = 1;
    }
    return fail ? -1 : 0;
}

int szpo_decompress(const uint8_t *stream, uint64_t len, uint64_t n, int B, int T, int workers,
                    float *out)
{
    if (len < 4) return -1;
    Out o = {out, NULL, get_be_f32(stream), 0.0, 1};
    return decompress_body(stream, len, 4, n, B, T, workers, &o);
}

int szpo_decompress_f64(const uint8_t *stream, uint64_t len, uint64_t n, int B, int T,
                        int workers, double *out)
{
    if (len < 8) return -1;
    Out o = {NULL, out, 0.0f, get_be_f64(stream), 1};
    return decompress_body(stream, len, 8, n, B, T, workers, &o);
}

/* SZp/src/szp.cc:29-58 (float, fastMode 1). */
int szpo_compress_api(const float *x, uint64_t n, int err_mode, double abs_eb, double rel,
                      int B, int T, int workers, uint8_t *out, uint64_t cap, uint64_t *out_size)
{
    float eb = (float)abs_eb;
    if (err_mode == 1) {
        if (n == 0) return -1;
        eb = szpo_rel_bound(x, n, rel);
    }
    return szpo_compress(x, n, eb, B, T, workers, out, cap, out_size);
}

/* SZp/src/szp.cc:60-84: REL range in double, eb = range * rel. */
double szpo_rel_bound_f64(const double *x, uint64_t n, double rel)
{
    double mn = x[0], mx = x[0];
    for (uint64_t i = 0; i < n; ++i) {
        double v = x[i];
        if (mn > v) mn = v;
        else if (mx < v) mx = v;
    }
    double range = mx - mn;
    return range * rel;
}

/* szp_compress(fastMode=1, SZ_DOUBLE, ...) semantics (SZp/src/szp.cc:60-98). */
int szpo_compress_api_f64(const double *x, uint64_t n, int err_mode, double abs_eb, double rel,
                          int B, int T, int workers, uint8_t *out, uint64_t cap,
                          uint64_t *out_size)
{
    double eb = abs_eb;
    if (err_mode == 1) {
        if (n == 0) return -1;
        eb = szpo_rel_bound_f64(x, n, rel);
    }
    return szpo_compress_f64(x, n, eb, B, T, workers, out, cap, out_size);
}

/* szp_compress(fast, dtype, ...) for fastMode 1 or 2 and SZ_FLOAT (0) or
 * SZ_DOUBLE (1) (SZp/src/szp.cc:23-115). */
int szpo_compress_mode(int fast, int dtype, const void *x, uint64_t n, int err_mode,
                       double abs_eb, double rel, int B, int T, int workers, uint8_t *out,
                       uint64_t cap, uint64_t *out_size)
{
    if ((fast != 1 && fast != 2) || (dtype != 0 && dtype != 1)) return -1;
    if (B < 1 || T < 1 || !out || !out_size) return -1;
    if (err_mode == 1 && n == 0) return -1;
    uint64_t es = dtype ? 8 : 4;
    if (cap < es + 8 * (uint64_t)T) return -1;
    Field f = {NULL, NULL, 0.0f, 0.0, fast};
    if (dtype == 0) {
        float eb = err_mode == 1 ? szpo_rel_bound((const float *)x, n, rel) : (float)abs_eb;
        f.f = (const float *)x;
        f.inv_f = szpo_inverse_bound(eb);
        put_be_f32(out, eb);
    } else {
        double eb = err_mode == 1 ? szpo_rel_bound_f64((const double *)x, n, rel) : abs_eb;
        volatile double one = 1.0;
        f.d = (const double *)x;
        f.inv_d = one / eb;
        put_be_f64(out, eb);
    }
    return compress_body(&f, n, es, B, T, workers, out, cap, out_size);
}

/* szp_decompress(fast, dtype, ...) (SZp/src/szp.cc:118-179): fastMode 1
 * decodes SZP_NONRANDOMACCESS, any other value random access. */
int szpo_decompress_mode(int fast, int dtype, const uint8_t *stream, uint64_t len, uint64_t n,
                         int B, int T, int workers, void *out)
{
    if (dtype != 0 && dtype != 1) return -1;
    uint64_t es = dtype ? 8 : 4;
    if (len < es) return -1;
    Out o = {NULL, NULL, 0.0f, 0.0, fast == 1 ? 1 : 2};
    if (dtype == 0) {
        o.f = (float *)out;
        o.eb_f = get_be_f32(stream);
    } else {
        o.d = (double *)out;
        o.eb_d = get_be_f64(stream);
    }
    return decompress_body(stream, len, es, n, B, T, workers, &o);
}

uint64_t szpo_compress_bound_mode(int fast, int dtype, uint64_t n, int B, int T)
{
    uint64_t total = (dtype ? 8 : 4) + 8 * (uint64_t)T;
    for (int t = 0; t < T; ++t) {
        uint64_t lo, hi;
        szpo_chunk(n, T, t, &lo, &hi);
        total += chunk_bound(fast, hi - lo, B);
    }
    return total;
}

/* The memory-access tally of the reference's instrumented single-thread
 * forms, replayed event by event over a T = 1 fastMode-1 float stream body
 * (the bytes after the eb): compress SZp/src/szp_float.cc:607-750,
 * decompress SZp/src/szpd_float.cc:216-317.  Sizes follow the reference's
 * types: size_t 8, float 4, int 4, unsigned char 1. */
uint64_t szpo_record_count(const uint8_t *body, uint64_t n, int B, int decompress)
{
    uint64_t c = 8;                         /* offsets[0] */
    uint64_t chunk = 0;                     /* outSize_perthread */
    uint64_t pos;
    memcpy(&pos, body, 8);
    pos += 8;
    if (n > 0) {
        c += decompress ? 4 * 2 + 4 * 2     /* head read + prior; value + output */
                        : 4 + 4 * 2;        /* input read; head memcpy */
        pos += 4;
        chunk += 4;
    }
    for (uint64_t i = 1; i < n; i += (uint64_t)B) {
        const uint64_t L = (i + (uint64_t)B > n) ? n - i : (uint64_t)B;
        const int b = body[pos];
        const uint64_t bytes = szpo_block_bytes(b, L);
        const uint64_t signs = (L - 1) / 8 + 1, packed = bytes - 1 - (b ? signs : 0);
        c += 1;                             /* the width byte */
        if (!decompress) {
            for (uint64_t j = 0; j < L; ++j) c += 4 + 1 + 4;  /* input, temp sign, temp mag */
            if (b) c += L + signs + 4 * L + packed;
        } else if (!b) {
            for (uint64_t j = 0; j < L; ++j) c += 4 * 2;
        } else {
            c += signs + L + packed + 4 * L;
            for (uint64_t j = 0; j < L; ++j) c += 1 + 4 + 4 * 2;
        }
        pos += bytes;
        chunk += bytes;
    }
    if (!decompress) c += 8 + 8 + 8 * 3 + 8 * 2 + chunk * 2;
    return c;
}
