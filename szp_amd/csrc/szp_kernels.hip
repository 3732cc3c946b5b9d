// HIP kernels for SZp fastmode1 float32 compress / decompress on CDNA4 (gfx950).
//
// Format and arithmetic follow the reference exactly (paths relative to the
// SZp repository root):
//   quantizer       SZp/src/szp_float.cc:185,209,228
//   Lorenzo delta   SZp/src/szp_float.cc:201-251 (prior carried across blocks)
//   bit width       SZp/src/szp_float.cc:253-262 (glibc log2f, see c_width_bump)
//   block layout    SZp/src/szp_TypeManager.cc:13-94,820-844
//   stream layout   SZp/src/szp_float.cc:156-165,180-213,277-303
//   decode          SZp/src/szpd_float.cc:23-128, SZp/src/szp_TypeManager.cc:436-944
//
// The design is GPU-first: the byte stream is cut into "items" (chunk heads
// and SZp blocks) in element order, a workgroup encodes K items with one
// thread per 8 elements, and a decoupled look-back over tile byte counts
// places every tile's output without a second pass over the input.  Decoding
// recovers block boundaries (the stream has no index) with a speculative
// parse whose tiles only wait for their predecessor when they cannot
// synchronise by themselves, then decodes tiles of K blocks with a segmented
// look-back scan of the Lorenzo deltas.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>

#include "szp_internal.h"

namespace szp_amd {

// ---------------------------------------------------------------------------
// Constants and small helpers
// ---------------------------------------------------------------------------

constexpr uint64_t kFlagAgg = 1ull << 62;
constexpr uint64_t kFlagIncl = 2ull << 62;
constexpr uint64_t kValMask = (1ull << 62) - 1;
constexpr uint64_t kHeadBit = 1ull << 32;
constexpr uint32_t kSpinLimit = 1u << 21;
constexpr uint16_t kDead = 0xFFFF;

// Values m in [c_width_bump[k], 2^k) with k significant bits get bit width
// k + 1, because glibc's log2f rounds log2(m) up to k (the reference calls
// (int)log2f((float)m) + 1, SZp/src/szp_float.cc:261).  Checked against log2f
// over all 2^32 inputs by the test suite.
__constant__ uint32_t c_width_bump[33] = {
    0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu,
    0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu,
    0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu,
    2097151u,    4194302u,    8388603u,    16777205u,   33554410u,   67108819u,   134217637u,
    268435273u,  536870545u,  1073741089u, 2147482177u, 4294964353u};

static const uint32_t h_width_bump[33] = {
    0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu,
    0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu,
    0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu,
    2097151u,    4194302u,    8388603u,    16777205u,   33554410u,   67108819u,   134217637u,
    268435273u,  536870545u,  1073741089u, 2147482177u, 4294964353u};

int bit_width_host(uint32_t m)
{
    if (!m) return 0;
    int k = 32 - __builtin_clz(m);
    return k + (m >= h_width_bump[k] ? 1 : 0);
}

__device__ __forceinline__ uint32_t bit_width(uint32_t m)
{
    if (!m) return 0;
    uint32_t k = 32u - __clz(m);
    return k + (m >= c_width_bump[k] ? 1u : 0u);
}

// Bytes of one block of L values at width b (SZp/src/szp_TypeManager.cc:20-31).
__host__ __device__ __forceinline__ uint32_t block_bytes(uint32_t b, uint32_t L)
{
    if (!b) return 1;
    return 1 + ((L + 7) >> 3) + (b >> 3) * L + (((b & 7) * L + 7) >> 3);
}

// (int)((double)x * (double)inv) with x86 cvttsd2si semantics: NaN and
// out-of-range products give INT32_MIN (SZp/src/szp_float.cc:209,228).
__device__ __forceinline__ int32_t quantize(float x, float inv)
{
    double p = (double)x * (double)inv;
    return (p > -2147483649.0 && p < 2147483648.0) ? (int32_t)p : INT32_MIN;
}

__device__ __forceinline__ uint64_t ld_status(const uint64_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void st_status(uint64_t *p, uint64_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint32_t ld_ctl(const uint32_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void raise_error(uint32_t *ctl, uint32_t bits)
{
    __hip_atomic_fetch_or(&ctl[1], bits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Waits until the status word at p carries a flag.  A bounded wait: on
// timeout (or when another workgroup has timed out) it records kErrTimeout
// and returns an inclusive zero so that every loop above it terminates.
__device__ uint64_t wait_status(const uint64_t *p, uint32_t *ctl)
{
    uint64_t s = ld_status(p);
    uint32_t spins = 0;
    while ((s >> 62) == 0) {
        __builtin_amdgcn_s_sleep(2);
        s = ld_status(p);
        if (++spins >= kSpinLimit) {
            raise_error(ctl, kErrTimeout);
            return kFlagIncl;
        }
        if ((spins & 255u) == 0 && (ld_ctl(&ctl[1]) & kErrTimeout)) return kFlagIncl;
    }
    return s;
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v)
{
#pragma unroll
    for (int o = 32; o; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v)
{
#pragma unroll
    for (int o = 32; o; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__device__ __forceinline__ uint32_t wave_incl_scan_u32(uint32_t v, int lane)
{
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t t = __shfl_up(v, d);
        if (lane >= d) v += t;
    }
    return v;
}

// Decoupled look-back for a plain sum (one wave calls it).  Tiles
// [first, tile) form the scan domain; returns the exclusive prefix of `tile`
// and publishes its inclusive value.
__device__ uint64_t lookback_sum(uint64_t *st, uint64_t tile, uint64_t first, uint64_t agg,
                                 uint32_t *ctl)
{
    const int lane = __lane_id();
    if (tile == first) {
        if (lane == 0) st_status(&st[tile], kFlagIncl | agg);
        return 0;
    }
    if (lane == 0) st_status(&st[tile], kFlagAgg | agg);
    uint64_t excl = 0;
    int64_t hi = (int64_t)tile - 1;
    while (true) {
        int64_t idx = hi - lane;
        uint64_t s = (idx >= (int64_t)first) ? wait_status(&st[idx], ctl) : kFlagIncl;
        uint64_t stop = __ballot((s >> 62) == 2);
        if (stop) {
            int f = __ffsll((long long)stop) - 1;
            excl += wave_sum_u64(lane <= f ? (s & kValMask) : 0);
            break;
        }
        excl += wave_sum_u64(s & kValMask);
        hi -= 64;
    }
    if (lane == 0) st_status(&st[tile], kFlagIncl | (excl + agg));
    return excl;
}

// ---------------------------------------------------------------------------
// Item geometry
// ---------------------------------------------------------------------------

__device__ __forceinline__ void locate(const Geom &g, uint64_t item, uint32_t &t, uint64_t &j)
{
    uint64_t lead = (uint64_t)(g.T - 1) * g.c0;
    if (item < lead) {
        t = (uint32_t)(item / g.c0);
        j = item - (uint64_t)t * g.c0;
    } else {
        t = g.T - 1;
        j = item - lead;
    }
}

__device__ __forceinline__ void advance(const Geom &g, uint32_t &t, uint64_t &j, uint32_t d)
{
    j += d;
    while (t < g.T - 1 && j >= g.c0) {
        j -= g.c0;
        ++t;
    }
}

__host__ __device__ __forceinline__ uint64_t chunk_lo(const Geom &g, uint32_t t)
{
    return (uint64_t)t * g.tbs;
}

__host__ __device__ __forceinline__ uint64_t chunk_len(const Geom &g, uint32_t t)
{
    return t == g.T - 1 ? g.n - chunk_lo(g, t) : g.tbs;
}

__device__ __forceinline__ uint64_t item_elem(const Geom &g, uint32_t t, uint64_t j)
{
    return chunk_lo(g, t) + (j ? 1 + (j - 1) * (uint64_t)g.B : 0);
}

__device__ __forceinline__ uint32_t item_len(const Geom &g, uint32_t t, uint64_t j)
{
    if (!j) return 1;
    uint64_t rem = chunk_len(g, t) - 1 - (j - 1) * (uint64_t)g.B;
    return rem < g.B ? (uint32_t)rem : g.B;
}

// First element of item `item` (n for item == nitems).
__device__ __forceinline__ uint64_t item_start(const Geom &g, uint64_t item)
{
    if (item >= g.nitems) return g.n;
    uint32_t t;
    uint64_t j;
    locate(g, item, t, j);
    return item_elem(g, t, j);
}

// Copies tot bytes from LDS to out[G0, G0 + tot): 16-byte stores for the
// aligned middle, byte stores for the ragged ends shared with neighbours.
__device__ void copy_out_bytes(uint8_t *__restrict__ out, uint64_t G0, uint32_t tot,
                               const uint8_t *ob, int tid)
{
    const uint64_t G1 = G0 + tot;
    const uint64_t a0 = (G0 + 15) & ~15ull, a1 = G1 & ~15ull;
    if (a0 >= a1) {
        for (uint32_t i = tid; i < tot; i += kWG) out[G0 + i] = ob[i];
        return;
    }
    const uint32_t head = (uint32_t)(a0 - G0), tail = (uint32_t)(G1 - a1);
    if ((uint32_t)tid < head) out[G0 + tid] = ob[tid];
    if ((uint32_t)tid < tail) out[a1 + tid] = ob[(uint32_t)(a1 - G0) + tid];
    const uint32_t *ob32 = reinterpret_cast<const uint32_t *>(ob);
    const uint32_t nv = (uint32_t)((a1 - a0) >> 4);
    const uint32_t sh = head & 3;
    for (uint32_t v = tid; v < nv; v += kWG) {
        const uint32_t p = head + 16 * v;
        const uint32_t w = p >> 2;
        uint32_t d0 = ob32[w], d1 = ob32[w + 1], d2 = ob32[w + 2], d3 = ob32[w + 3];
        uint4 val;
        if (sh) {
            uint32_t d4 = ob32[w + 4];
            val.x = __builtin_amdgcn_alignbyte(d1, d0, sh);
            val.y = __builtin_amdgcn_alignbyte(d2, d1, sh);
            val.z = __builtin_amdgcn_alignbyte(d3, d2, sh);
            val.w = __builtin_amdgcn_alignbyte(d4, d3, sh);
        } else {
            val = make_uint4(d0, d1, d2, d3);
        }
        *reinterpret_cast<uint4 *>(out + a0 + 16ull * v) = val;
    }
}

// ---------------------------------------------------------------------------
// Compress
// ---------------------------------------------------------------------------

// Clears the tile counter, error word and look-back status words of one call.
__global__ void k_clear(uint32_t *ctl, uint64_t *status, uint64_t nstatus)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 8) ctl[i] = 0;
    for (uint64_t k = i; k < nstatus; k += (uint64_t)gridDim.x * blockDim.x) status[k] = 0;
}

struct RangePart {
    float mn;
    float mx;
    uint64_t imn;
    uint64_t imx;
    uint64_t pad;
};

// (v, i) beats (w, j) as a minimum: NaN never wins, the earlier index wins
// ties.  This reproduces the serial scan "if (min > v) min = v" of
// SZp/src/szp.cc:36-45 (first occurrence kept; only +0/-0 can tie with
// different bits).
__device__ __forceinline__ bool min_beats(float v, uint64_t i, float w, uint64_t j)
{
    return !isnan(v) && (isnan(w) || v < w || (v == w && i < j));
}

__device__ __forceinline__ bool max_beats(float v, uint64_t i, float w, uint64_t j)
{
    return !isnan(v) && (isnan(w) || v > w || (v == w && i < j));
}

__device__ void range_reduce_wg(float &mn, uint64_t &imn, float &mx, uint64_t &imx)
{
    __shared__ float smn[kWG / 64], smx[kWG / 64];
    __shared__ uint64_t simn[kWG / 64], simx[kWG / 64];
    const int lane = __lane_id(), wv = threadIdx.x >> 6;
#pragma unroll
    for (int o = 32; o; o >>= 1) {
        float v = __shfl_xor(mn, o);
        uint64_t i = __shfl_xor(imn, o);
        if (min_beats(v, i, mn, imn)) { mn = v; imn = i; }
        v = __shfl_xor(mx, o);
        i = __shfl_xor(imx, o);
        if (max_beats(v, i, mx, imx)) { mx = v; imx = i; }
    }
    if (lane == 0) { smn[wv] = mn; simn[wv] = imn; smx[wv] = mx; simx[wv] = imx; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < kWG / 64; ++k) {
            if (min_beats(smn[k], simn[k], mn, imn)) { mn = smn[k]; imn = simn[k]; }
            if (max_beats(smx[k], simx[k], mx, imx)) { mx = smx[k]; imx = simx[k]; }
        }
    }
}

template <bool VEC>
__global__ __launch_bounds__(kWG) void k_range_partial(const float *__restrict__ x, uint64_t n,
                                                       uint64_t seg, RangePart *part)
{
    const uint64_t s = (uint64_t)blockIdx.x * seg;
    const uint64_t e = min(s + seg, n);
    float mn = __builtin_nanf(""), mx = __builtin_nanf("");
    uint64_t imn = ~0ull, imx = ~0ull;
    if (VEC) {
        // seg is a multiple of 4 and x is 16-byte aligned.
        for (uint64_t v = s + 4ull * threadIdx.x; v < e; v += 4ull * kWG) {
            float f[4];
            if (v + 4 <= e) {
                float4 q = *reinterpret_cast<const float4 *>(x + v);
                f[0] = q.x; f[1] = q.y; f[2] = q.z; f[3] = q.w;
            } else {
                for (int k = 0; k < 4; ++k) f[k] = (v + k < e) ? x[v + k] : __builtin_nanf("");
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                if (min_beats(f[k], v + k, mn, imn)) { mn = f[k]; imn = v + k; }
                if (max_beats(f[k], v + k, mx, imx)) { mx = f[k]; imx = v + k; }
            }
        }
    } else {
        for (uint64_t i = s + threadIdx.x; i < e; i += kWG) {
            float f = x[i];
            if (min_beats(f, i, mn, imn)) { mn = f; imn = i; }
            if (max_beats(f, i, mx, imx)) { mx = f; imx = i; }
        }
    }
    range_reduce_wg(mn, imn, mx, imx);
    if (threadIdx.x == 0) part[blockIdx.x] = RangePart{mn, mx, imn, imx, 0};
}

// eb = fl32(fl32(max - min) * fl32(rel)), NaN when x[0] is NaN
// (SZp/src/szp.cc:36-47).
__global__ __launch_bounds__(kWG) void k_range_final(const float *__restrict__ x,
                                                     const RangePart *part, int nparts,
                                                     float relf, float *d_eb)
{
    float mn = __builtin_nanf(""), mx = __builtin_nanf("");
    uint64_t imn = ~0ull, imx = ~0ull;
    for (int k = threadIdx.x; k < nparts; k += kWG) {
        RangePart p = part[k];
        if (min_beats(p.mn, p.imn, mn, imn)) { mn = p.mn; imn = p.imn; }
        if (max_beats(p.mx, p.imx, mx, imx)) { mx = p.mx; imx = p.imx; }
    }
    range_reduce_wg(mn, imn, mx, imx);
    if (threadIdx.x == 0) {
        float eb;
        if (isnan(x[0])) {
            eb = x[0];
        } else {
            float range = mx - mn;
            eb = range * relf;
        }
        *d_eb = eb;
    }
}

// Stream of an empty field: eb and T zero chunk offsets.
__global__ void k_empty_stream(uint8_t *out, const float *d_eb, float eb_host, uint32_t T,
                               uint64_t *d_size)
{
    const float eb = d_eb ? *d_eb : eb_host;
    const uint32_t u = __float_as_uint(eb);
    for (uint32_t i = threadIdx.x; i < 4 + 8 * T; i += blockDim.x) {
        uint8_t v = i < 4 ? (uint8_t)(u >> (24 - 8 * i)) : 0;
        out[i] = v;
    }
    if (threadIdx.x == 0) *d_size = 4 + 8ull * T;
}

template <bool VEC>
__global__ __launch_bounds__(kWG) void k_compress(const float *__restrict__ x,
                                                  const float *__restrict__ d_eb, float eb_host,
                                                  Geom g, uint8_t *__restrict__ out, uint64_t cap,
                                                  uint64_t *__restrict__ status,
                                                  uint32_t *__restrict__ ctl,
                                                  uint64_t *__restrict__ d_size)
{
    constexpr int kQ = 2052;  // one tile: <= 2048 elements + the previous one + alignment
    __shared__ int32_t qs[kQ + kQ / 8 + 8];
    __shared__ __align__(16) uint8_t ob[kOutBufBytes];
    __shared__ uint32_t s_max[kWG];
    __shared__ uint32_t s_off[kWG];
    __shared__ uint64_t s_tile, s_prefix;
    __shared__ uint32_t s_total;

    const int tid = threadIdx.x;
    if (tid == 0)
        s_tile = __hip_atomic_fetch_add(&ctl[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    s_max[tid] = 0;
    s_off[tid] = 0;
    __syncthreads();
    const uint64_t tile = s_tile;
    if (tile >= g.ntiles) return;

    const float eb = d_eb ? *d_eb : eb_host;
    const float inv = 1.0f / eb;  // FP32 division, SZp/src/szp_float.cc:185
    const uint64_t item0 = tile * g.K;
    const uint64_t itemN = min(item0 + g.K, g.nitems);
    uint32_t t0;
    uint64_t j0;
    locate(g, item0, t0, j0);
    const uint64_t E0 = item_elem(g, t0, j0);
    const uint64_t E1 = item_start(g, itemN);
    const uint64_t base = E0 ? E0 - 1 : 0;

    // Stage q = quantize(x) for elements [base, E1) in LDS (padded index).
    for (uint64_t v = (base & ~3ull) + 4ull * tid; v < E1; v += 4ull * kWG) {
        float f[4];
        if (VEC && v + 4 <= g.n) {
            float4 q = *reinterpret_cast<const float4 *>(x + v);
            f[0] = q.x; f[1] = q.y; f[2] = q.z; f[3] = q.w;
        } else {
#pragma unroll
            for (int k = 0; k < 4; ++k) f[k] = (v + k < g.n) ? x[v + k] : 0.0f;
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint64_t e = v + k;
            if (e >= base && e < E1) {
                const uint32_t r = (uint32_t)(e - base);
                qs[r + (r >> 3)] = quantize(f[k], inv);
            }
        }
    }

    // Slot -> (item, octet).
    const uint32_t d = ((uint32_t)tid * g.spb_magic) >> 16;
    const uint32_t o = (uint32_t)tid - d * g.spb;
    const bool active = d < g.K && item0 + d < itemN;
    uint32_t t = t0;
    uint64_t j = j0;
    if (active) advance(g, t, j, d);
    const bool head = active && j == 0;
    const uint32_t L = active ? item_len(g, t, j) : 0;
    const uint64_t es = active ? item_elem(g, t, j) : 0;
    __syncthreads();

    // Lorenzo deltas, signs, magnitudes and the block maximum.
    uint32_t mg[8];
    uint32_t sgn = 0, m = 0;
    int32_t qhead = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) mg[i] = 0;
    if (active && !head && 8 * o < L) {
        m = min(8u, L - 8 * o);
        const uint32_t r0 = (uint32_t)(es + 8 * o - base);
        int32_t prev = qs[(r0 - 1) + ((r0 - 1) >> 3)];
        uint32_t omax = 0;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            if ((uint32_t)i < m) {
                const uint32_t r = r0 + i;
                const int32_t c = qs[r + (r >> 3)];
                const uint32_t dd = (uint32_t)c - (uint32_t)prev;
                prev = c;
                const bool neg = (int32_t)dd < 0;
                mg[i] = neg ? 0u - dd : dd;
                sgn |= (uint32_t)neg << (7 - i);
                omax = max(omax, mg[i]);
            }
        }
        if (omax) atomicMax(&s_max[d], omax);
    }
    if (head && o == 0) {
        const uint32_t r = (uint32_t)(es - base);
        qhead = qs[r + (r >> 3)];
    }
    __syncthreads();

    uint32_t b = 0;
    if (active) {
        uint32_t len = 4;
        if (!head) {
            b = bit_width(s_max[d]);
            len = block_bytes(b, L);
        }
        if (o == 0) s_off[d] = len;
    }
    __syncthreads();

    // Exclusive scan of item byte lengths inside the tile.
    if (tid < 64) {
        const uint32_t a0 = s_off[4 * tid], a1 = s_off[4 * tid + 1];
        const uint32_t a2 = s_off[4 * tid + 2], a3 = s_off[4 * tid + 3];
        const uint32_t sum = a0 + a1 + a2 + a3;
        const uint32_t incl = wave_incl_scan_u32(sum, tid);
        const uint32_t ex = incl - sum;
        s_off[4 * tid] = ex;
        s_off[4 * tid + 1] = ex + a0;
        s_off[4 * tid + 2] = ex + a0 + a1;
        s_off[4 * tid + 3] = ex + a0 + a1 + a2;
        if (tid == 63) s_total = incl;
    }
    __syncthreads();
    if (tid < 64) {
        const uint64_t p = lookback_sum(status, tile, 0, s_total, ctl);
        if (tid == 0) s_prefix = p;
    }
    __syncthreads();
    const uint64_t P = s_prefix;
    const uint32_t tot = s_total;
    const uint64_t G0 = g.body + P;
    if (itemN == g.nitems && tid == 0) *d_size = G0 + tot;
    if (G0 + tot > cap || (ld_ctl(&ctl[1]) & kErrTimeout)) {
        if (tid == 0 && G0 + tot > cap) raise_error(ctl, kErrCapacity);
        return;
    }

    // Encode into LDS; every output byte of the tile is written exactly once.
    if (active) {
        const uint32_t ib = s_off[d];
        if (head) {
            if (o == 0) {
                ob[ib] = (uint8_t)qhead;
                ob[ib + 1] = (uint8_t)((uint32_t)qhead >> 8);
                ob[ib + 2] = (uint8_t)((uint32_t)qhead >> 16);
                ob[ib + 3] = (uint8_t)((uint32_t)qhead >> 24);
                // Chunk offset relative to the body (SZp/src/szp_float.cc:280-288).
                const uint64_t off = P + ib;
                uint32_t *h = reinterpret_cast<uint32_t *>(out + 4 + 8ull * t);
                h[0] = (uint32_t)off;
                h[1] = (uint32_t)(off >> 32);
            }
        } else {
            if (o == 0) ob[ib] = (uint8_t)b;
            if (b && m) {
                const uint32_t S8 = (L + 7) >> 3;
                ob[ib + 1 + o] = (uint8_t)sgn;
                const uint32_t nb = b >> 3, r = b & 7;
                uint32_t p1 = ib + 1 + S8 + 8 * o * nb;
                if (nb) {
#pragma unroll
                    for (int i = 0; i < 8; ++i) {
                        if ((uint32_t)i < m) {
                            const uint32_t v = mg[i] >> r;
                            for (uint32_t k = 0; k < nb; ++k) ob[p1++] = (uint8_t)(v >> (8 * k));
                        }
                    }
                }
                if (r) {
                    const uint32_t mask = (1u << r) - 1;
                    uint64_t acc = 0;
#pragma unroll
                    for (int i = 0; i < 8; ++i) acc = (acc << r) | (mg[i] & mask);
                    const uint32_t nbytes = (m * r + 7) >> 3;
                    const uint32_t p2 = ib + 1 + S8 + nb * L + o * r;
                    for (uint32_t k = 0; k < nbytes; ++k)
                        ob[p2 + k] = (uint8_t)(acc >> (8 * (r - 1 - k)));
                }
            }
        }
    }
    if (tile == 0) {
        if (tid == 0) {
            const uint32_t u = __float_as_uint(eb);
            out[0] = (uint8_t)(u >> 24);
            out[1] = (uint8_t)(u >> 16);
            out[2] = (uint8_t)(u >> 8);
            out[3] = (uint8_t)u;
        }
        // Empty chunks (n < T) all start at offset 0.
        if (g.tbs == 0) {
            for (uint32_t k = tid; k + 1 < g.T; k += kWG) {
                uint32_t *h = reinterpret_cast<uint32_t *>(out + 4 + 8ull * k);
                h[0] = 0;
                h[1] = 0;
            }
        }
    }
    __syncthreads();
    copy_out_bytes(out, G0, tot, ob, tid);
}

// ---------------------------------------------------------------------------
// Decompress
// ---------------------------------------------------------------------------

__device__ __forceinline__ uint64_t read_u64_le4(const uint8_t *p)
{
    const uint32_t *w = reinterpret_cast<const uint32_t *>(p);
    return (uint64_t)w[0] | ((uint64_t)w[1] << 32);
}

// Reads the chunk offsets, validates them and lays out the parse tiles:
// ctab = [cstart[T] | aend[T] | ttab[T+1]] where chunk t's blocks occupy
// [cstart[t] + 4, aend[t]) and its parse tiles are [ttab[t], ttab[t+1]).
// All workgroups also clear the look-back status words of this call.
__global__ void k_dsetup(const uint8_t *__restrict__ st, uint64_t len, Geom g, uint32_t S,
                         uint64_t *__restrict__ ctab, uint64_t *__restrict__ clr, uint64_t nclr,
                         uint32_t *__restrict__ ctl)
{
    const uint64_t gi = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (uint64_t k = gi; k < nclr; k += (uint64_t)gridDim.x * blockDim.x) clr[k] = 0;
    if (gi != 0) return;
    for (int k = 0; k < 8; ++k) ctl[k] = 0;
    uint64_t *cstart = ctab, *aend = ctab + g.T, *ttab = ctab + 2 * g.T;
    uint64_t tiles = 0;
    bool bad = len < g.body;
    for (uint32_t t = 0; t < g.T && !bad; ++t) {
        const uint64_t nt = chunk_len(g, t);
        const uint64_t off = read_u64_le4(st + 4 + 8ull * t);
        const uint64_t cs = g.body + off;
        ttab[t] = tiles;
        cstart[t] = cs;
        aend[t] = cs;
        if (nt == 0) {
            if (off != 0) bad = true;
            continue;
        }
        const uint64_t ce = (t + 1 < g.T) ? g.body + read_u64_le4(st + 4 + 8ull * (t + 1)) : len;
        if (off > len || cs + 4 > ce || ce > len) {
            bad = true;
            break;
        }
        aend[t] = ce;
        const uint64_t nb = (t == g.T - 1) ? g.nblast : g.nb0;
        const uint64_t alen = ce - (cs + 4);
        if (nb && !alen) bad = true;
        if (nb) tiles += (alen + S - 1) / S;
    }
    if (bad) {
        raise_error(ctl, kErrStream);
        tiles = 0;
        for (uint32_t t = 0; t <= g.T; ++t) ttab[t] = 0;
    }
    ttab[g.T] = tiles;
}

// Recovers the start of every SZp block.  Parse tile i of chunk t covers
// bytes [s, te) of the chunk's block area.  The first block starting at or
// after s is the "entry"; the first at or after te is the "exit" that the
// next tile takes as its entry.  A tile walks one speculative chain, then
// checks that every chain starting in [s, s + W) (W = longest block) either
// joins it, reaches an invalid width byte, or leaves the tile at the same
// exit.  If so the exit does not depend on the entry and is published before
// the predecessor finishes; otherwise the tile waits for its predecessor's
// exit.  Block counts are then scanned with a decoupled look-back and every
// block's stream position is written to blkpos[].
__global__ __launch_bounds__(kWG) void k_parse(const uint8_t *__restrict__ st, uint64_t len,
                                               Geom g, uint32_t S, uint32_t W,
                                               const uint64_t *__restrict__ ctab,
                                               uint64_t *__restrict__ exit_st,
                                               uint64_t *__restrict__ cnt_st,
                                               uint64_t *__restrict__ blkpos,
                                               uint32_t *__restrict__ ctl)
{
    extern __shared__ __align__(16) uint8_t dyn[];
    uint8_t *buf = dyn;
    uint16_t *rank = reinterpret_cast<uint16_t *>(dyn + ((S + 32 + 15) & ~15u));
    uint16_t *lst = rank + S;
    __shared__ uint32_t s_lenB[33];
    __shared__ uint64_t s_tile, s_spec_exit, s_prefix, s_lastpos;
    __shared__ uint32_t s_spec_cnt, s_npre, s_r0, s_cnt;
    __shared__ int s_spec_ok, s_nonsync, s_broken, s_merged;

    const int tid = threadIdx.x;
    if (tid == 0)
        s_tile = __hip_atomic_fetch_add(&ctl[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (tid < 33) s_lenB[tid] = block_bytes(tid, g.B);
    if (tid == 0) { s_nonsync = 0; s_spec_ok = 0; s_spec_cnt = 0; s_spec_exit = 0; }
    __syncthreads();
    const uint64_t *cstart = ctab, *aend = ctab + g.T, *ttab = ctab + 2 * g.T;
    const uint64_t gt = s_tile;
    if (gt >= ttab[g.T]) return;
    uint32_t lo = 0, hi = g.T - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (ttab[mid] <= gt) lo = mid; else hi = mid - 1;
    }
    const uint32_t t = lo;
    const uint64_t first = ttab[t];
    const uint64_t i = gt - first;
    const uint64_t as = cstart[t] + 4, ae = aend[t];
    const uint64_t s = as + i * S;
    const uint64_t te = min(s + S, ae);
    const uint32_t span = (uint32_t)(te - s);
    const uint64_t s0 = s & ~15ull;
    const uint32_t sh = (uint32_t)(s - s0);

    for (uint64_t a = s0 + 16ull * tid; a < te; a += 16ull * kWG) {
        if (a + 16 <= len) {
            *reinterpret_cast<uint4 *>(buf + (a - s0)) = *reinterpret_cast<const uint4 *>(st + a);
        } else {
            for (int k = 0; k < 16; ++k)
                if (a + k < len) buf[a - s0 + k] = st[a + k];
        }
    }
    for (uint32_t k = tid; k < span; k += kWG) rank[k] = 0;
    __syncthreads();

    if (i > 0) {
        // One speculative chain: the first candidate in [s, s + W) whose walk
        // reaches te.  Positions of failed walks are marked dead, so the
        // search is linear in the tile size.
        if (tid == 0) {
            const uint32_t cmax = min(W, span);
            int ok = 0;
            uint32_t cnt = 0, p = 0;
            for (uint32_t c = 0; c < cmax && !ok; ++c) {
                if (rank[c] == kDead) continue;
                p = c;
                cnt = 0;
                while (p < span) {
                    const uint32_t b = buf[p + sh];
                    if (b > 32 || rank[p] == kDead) break;
                    lst[cnt++] = (uint16_t)p;
                    p += s_lenB[b];
                }
                if (p >= span) ok = 1;
                else for (uint32_t k = 0; k < cnt; ++k) rank[lst[k]] = kDead;
            }
            s_spec_ok = ok;
            s_spec_cnt = ok ? cnt : 0;
            s_spec_exit = s + p;
            s_nonsync = !ok;
        }
        __syncthreads();
        for (uint32_t k = tid; k < s_spec_cnt; k += kWG) rank[lst[k]] = (uint16_t)(k + 1);
        __syncthreads();
        if (s_spec_ok) {
            const uint32_t cmax = min(W, span);
            const uint64_t sx = s_spec_exit;
            for (uint32_t c = tid; c < cmax; c += kWG) {
                uint32_t p = c;
                while (true) {
                    const uint16_t rk = rank[p];
                    if (rk) break;  // joins the chain, or is known dead
                    const uint32_t b = buf[p + sh];
                    if (b > 32) break;
                    p += s_lenB[b];
                    if (p >= span) {
                        if (s + p != sx) s_nonsync = 1;
                        break;
                    }
                }
            }
        }
        __syncthreads();
        if (tid == 0 && !s_nonsync) st_status(&exit_st[gt], kFlagIncl | s_spec_exit);
    }

    // The true chain from the entry.
    if (tid == 0) {
        uint64_t e = as;
        if (i > 0) e = wait_status(&exit_st[gt - 1], ctl) & kValMask;
        uint32_t npre = 0, r0 = 0, cnt = 0;
        int broken = 0, merged = 0;
        uint64_t ex = te;
        if (e < s) {
            broken = 1;
        } else {
            uint64_t p = e;
            // Dead marks only steer the speculation: past the true end of the
            // last chunk the true chain may run into them, so they are
            // stepped over here and only an invalid width byte stops it.
            while (p < te) {
                const uint16_t rk = rank[p - s];
                if (rk && rk != kDead) { merged = 1; break; }
                const uint32_t b = buf[p - s0];
                if (b > 32) { broken = 1; break; }
                lst[npre++] = (uint16_t)(p - s);
                p += s_lenB[b];
            }
            if (broken) {
                cnt = npre;
                ex = te;
            } else if (merged) {
                r0 = rank[p - s] - 1;
                cnt = npre + s_spec_cnt - r0;
                ex = s_spec_exit;
            } else {
                cnt = npre;
                ex = p;
            }
        }
        if (i == 0 || s_nonsync) st_status(&exit_st[gt], kFlagIncl | ex);
        s_npre = npre;
        s_r0 = r0;
        s_cnt = cnt;
        s_broken = broken;
        s_merged = merged;
        s_lastpos = ~0ull;
    }
    __syncthreads();
    if (tid < 64) {
        const uint64_t pre = lookback_sum(cnt_st, gt, first, s_cnt, ctl);
        if (tid == 0) s_prefix = pre;
    }
    __syncthreads();

    const uint64_t pre = s_prefix;
    const uint64_t nbt = (t == g.T - 1) ? g.nblast : g.nb0;
    const uint64_t gbase = (uint64_t)t * g.nb0;
    const uint32_t npre = s_npre, r0 = s_r0;
    for (uint32_t k = tid; k < npre; k += kWG) {
        const uint64_t bi = pre + k;
        if (bi < nbt) {
            blkpos[gbase + bi] = s + lst[k];
            if (bi == nbt - 1) s_lastpos = s + lst[k];
        }
    }
    if (s_merged) {
        for (uint32_t q = tid; q < span; q += kWG) {
            const uint16_t rk = rank[q];
            if (rk && rk != kDead && (uint32_t)(rk - 1) >= r0) {
                const uint64_t bi = pre + npre + (rk - 1 - r0);
                if (bi < nbt) {
                    blkpos[gbase + bi] = s + q;
                    if (bi == nbt - 1) s_lastpos = s + q;
                }
            }
        }
    }
    __syncthreads();
    if (tid == 0) {
        bool bad = false;
        if (s_broken && pre + npre < nbt) bad = true;
        if (gt + 1 == ttab[t + 1] && pre + s_cnt < nbt) bad = true;  // truncated chunk
        if (s_lastpos != ~0ull) {
            // The last block is shorter when (n_t - 1) % B != 0; its end must
            // be the chunk end (exact for all but the last chunk).
            const uint64_t pl = s_lastpos;
            const uint32_t b = st[pl];
            const uint64_t Ll = chunk_len(g, t) - 1 - (nbt - 1) * (uint64_t)g.B;
            const uint64_t end = pl + block_bytes(b, (uint32_t)Ll);
            if (b > 32 || (t + 1 < g.T ? end != ae : end > len)) bad = true;
            if (t + 1 == g.T) ctl[4] = (uint32_t)end, ctl[5] = (uint32_t)(end >> 32);
        }
        if (bad) raise_error(ctl, kErrStream);
    }
}

template <bool VEC>
__global__ __launch_bounds__(kWG) void k_decode(const uint8_t *__restrict__ st, uint64_t len,
                                                Geom g, const uint64_t *__restrict__ ctab,
                                                const uint64_t *__restrict__ blkpos,
                                                float *__restrict__ out,
                                                uint64_t *__restrict__ status,
                                                uint32_t *__restrict__ ctl)
{
    __shared__ __align__(16) uint8_t buf[kOutBufBytes + 32];
    __shared__ __align__(16) float fs[2048 + 8];
    __shared__ uint64_t s_ipos[kWG];
    __shared__ uint32_t s_wh[kWG / 64], s_wv[kWG / 64];
    __shared__ uint64_t s_tile, s_P0, s_P1;
    __shared__ uint32_t s_carry, s_bad;

    const int tid = threadIdx.x, lane = __lane_id(), wv = tid >> 6;
    if (tid == 0) {
        s_tile = __hip_atomic_fetch_add(&ctl[3], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s_bad = 0;
    }
    __syncthreads();
    const uint64_t tile = s_tile;
    if (tile >= g.ntiles) return;
    if (ld_ctl(&ctl[1]) & (kErrStream | kErrTimeout)) {
        // Successors may already wait on this tile: publish before leaving.
        if (tid == 0) st_status(&status[tile], kFlagIncl);
        return;
    }

    const uint32_t ebu = ((uint32_t)st[0] << 24) | ((uint32_t)st[1] << 16) |
                         ((uint32_t)st[2] << 8) | (uint32_t)st[3];
    const float eb = __uint_as_float(ebu);
    const uint64_t item0 = tile * g.K;
    const uint64_t itemN = min(item0 + g.K, g.nitems);
    const uint32_t nit = (uint32_t)(itemN - item0);
    uint32_t t0;
    uint64_t j0;
    locate(g, item0, t0, j0);
    const uint64_t E0 = item_elem(g, t0, j0);
    const uint64_t E1 = item_start(g, itemN);
    const uint64_t Eb = E0 & ~3ull;

    const uint32_t d = ((uint32_t)tid * g.spb_magic) >> 16;
    const uint32_t o = (uint32_t)tid - d * g.spb;
    const bool active = d < g.K && d < nit;
    uint32_t t = t0;
    uint64_t j = j0;
    if (active) advance(g, t, j, d);
    const bool head = active && j == 0;
    const uint32_t L = active ? item_len(g, t, j) : 0;
    const uint64_t es = active ? item_elem(g, t, j) : 0;
    if (active && o == 0) s_ipos[d] = head ? ctab[t] : blkpos[(uint64_t)t * g.nb0 + (j - 1)];
    __syncthreads();
    if (tid == 0) {
        const uint64_t P0 = s_ipos[0], Pl = s_ipos[nit - 1];
        uint32_t tl;
        uint64_t jl;
        locate(g, itemN - 1, tl, jl);
        uint64_t P1;
        if (jl == 0) {
            P1 = Pl + 4;
        } else {
            const uint32_t b = Pl < len ? st[Pl] : 255u;
            P1 = b > 32 ? ~0ull : Pl + block_bytes(b, item_len(g, tl, jl));
        }
        if (P1 > len || P0 > Pl || P1 - P0 > kOutBufBytes) {
            s_bad = 1;
            raise_error(ctl, kErrStream);
        }
        s_P0 = P0;
        s_P1 = P1;
    }
    __syncthreads();
    // A malformed stream stops here; later tiles may wait on this one, so
    // publish an inclusive zero first.
    if (s_bad) {
        if (tid == 0) st_status(&status[tile], kFlagIncl);
        return;
    }
    const uint64_t P0 = s_P0, P1 = s_P1;
    const uint64_t pb = P0 & ~15ull;
    for (uint64_t a = pb + 16ull * tid; a < P1; a += 16ull * kWG) {
        if (a + 16 <= len) {
            *reinterpret_cast<uint4 *>(buf + (a - pb)) = *reinterpret_cast<const uint4 *>(st + a);
        } else {
            for (int k = 0; k < 16; ++k)
                if (a + k < len) buf[a - pb + k] = st[a + k];
        }
    }
    __syncthreads();

    // Unpack this slot's (up to) 8 deltas (SZp/src/szpd_float.cc:81-110).
    int32_t dl[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) dl[i] = 0;
    uint32_t m = 0;
    uint32_t hflag = 0, agg = 0;
    if (active) {
        const uint32_t lp = (uint32_t)(s_ipos[d] - pb);
        if (head) {
            if (o == 0) {
                m = 1;
                hflag = 1;
                agg = (uint32_t)buf[lp] | ((uint32_t)buf[lp + 1] << 8) |
                      ((uint32_t)buf[lp + 2] << 16) | ((uint32_t)buf[lp + 3] << 24);
            }
        } else if (8 * o < L) {
            m = min(8u, L - 8 * o);
            const uint32_t b = buf[lp];
            if (b > 32 || s_ipos[d] + block_bytes(b, L) > P1) {
                s_bad = 1;
            } else if (b) {
                const uint32_t S8 = (L + 7) >> 3;
                const uint32_t sg = buf[lp + 1 + o];
                const uint32_t nb = b >> 3, r = b & 7;
                const uint32_t p1 = lp + 1 + S8 + 8 * o * nb;
                uint64_t acc = 0;
                if (r) {
                    const uint32_t nbytes = (m * r + 7) >> 3;
                    const uint32_t p2 = lp + 1 + S8 + nb * L + o * r;
                    for (uint32_t k = 0; k < nbytes; ++k) acc = (acc << 8) | buf[p2 + k];
                    acc <<= 8 * (r - nbytes);
                }
                const uint32_t mask = (1u << r) - 1;
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    if ((uint32_t)i < m) {
                        uint32_t v = 0;
                        for (uint32_t k = 0; k < nb; ++k) v |= (uint32_t)buf[p1 + i * nb + k] << (8 * k);
                        const uint32_t low = r ? (uint32_t)(acc >> (r * (7 - i))) & mask : 0u;
                        const uint32_t mag = (r ? (v << r) : v) | low;
                        const bool neg = (sg >> (7 - i)) & 1u;
                        dl[i] = (int32_t)(neg ? 0u - mag : mag);
                    }
                }
            }
        }
    }
    // Inclusive prefix inside the slot, then a segmented scan over slots:
    // a head slot restarts the running value with its absolute q.
    uint32_t run = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        run += (uint32_t)dl[i];
        dl[i] = (int32_t)run;
    }
    if (!hflag) agg = run;
    uint32_t h = hflag, v = agg;
#pragma unroll
    for (int dd = 1; dd < 64; dd <<= 1) {
        const uint32_t ph = __shfl_up(h, dd), pv = __shfl_up(v, dd);
        if (lane >= dd) {
            v = h ? v : pv + v;
            h |= ph;
        }
    }
    if (lane == 63) { s_wh[wv] = h; s_wv[wv] = v; }
    __syncthreads();
    // Exclusive (h, v) of this slot within the tile.
    uint32_t eh = __shfl_up(h, 1), ev = __shfl_up(v, 1);
    if (lane == 0) { eh = 0; ev = 0; }
    {
        uint32_t wh = 0, wvv = 0;
        for (int k = 0; k < wv; ++k) {
            wvv = s_wh[k] ? s_wv[k] : wvv + s_wv[k];
            wh |= s_wh[k];
        }
        ev = eh ? ev : wvv + ev;
        eh |= wh;
    }
    uint32_t th = 0, tv = 0;
    for (int k = 0; k < kWG / 64; ++k) {
        tv = s_wh[k] ? s_wv[k] : tv + s_wv[k];
        th |= s_wh[k];
    }
    // Look-back: a tile whose aggregate contains a head is absolute by itself.
    if (wv == 0) {
        if (th && lane == 0) st_status(&status[tile], kFlagIncl | tv);
        uint32_t carry = 0;
        if (tile > 0 && !(j0 == 0)) {
            if (!th && lane == 0) st_status(&status[tile], kFlagAgg | tv);
            int64_t hi = (int64_t)tile - 1;
            while (true) {
                const int64_t idx = hi - lane;
                const uint64_t sx = idx >= 0 ? wait_status(&status[idx], ctl) : kFlagIncl;
                const uint64_t stop = __ballot((sx >> 62) == 2 || (sx & kHeadBit));
                if (stop) {
                    const int f = __ffsll((long long)stop) - 1;
                    carry += wave_sum_u32(lane <= f ? (uint32_t)sx : 0u);
                    break;
                }
                carry += wave_sum_u32((uint32_t)sx);
                hi -= 64;
            }
            if (!th && lane == 0) st_status(&status[tile], kFlagIncl | (uint32_t)(carry + tv));
        } else if (!th && lane == 0) {
            st_status(&status[tile], kFlagIncl | tv);
        }
        if (lane == 0) s_carry = carry;
    }
    // (th is always set for the tile holding item 0; the AGG flag of a tile
    // whose aggregate contains a head is never needed since it publishes INCL.)
    __syncthreads();
    const uint32_t carry = s_carry;
    if (m) {
        const uint32_t before = eh ? ev : carry + ev;
        const uint32_t base_e = (uint32_t)(es + (head ? 0 : 8 * o) - Eb);
        if (head) {
            fs[base_e] = (float)(int32_t)agg * eb;
        } else {
#pragma unroll
            for (int i = 0; i < 8; ++i)
                if ((uint32_t)i < m) fs[base_e + i] = (float)(int32_t)(before + (uint32_t)dl[i]) * eb;
        }
    }
    __syncthreads();
    if (s_bad) {
        if (tid == 0) raise_error(ctl, kErrStream);
        return;
    }
    // Store [E0, E1): aligned float4 middle, scalar ends.
    const uint64_t a0 = (E0 + 3) & ~3ull, a1 = E1 & ~3ull;
    if (!VEC || a0 >= a1) {
        for (uint64_t e = E0 + tid; e < E1; e += kWG) out[e] = fs[e - Eb];
        return;
    }
    if (E0 + tid < a0) out[E0 + tid] = fs[E0 + tid - Eb];
    if (a1 + tid < E1) out[a1 + tid] = fs[a1 + tid - Eb];
    for (uint64_t e = a0 + 4ull * tid; e < a1; e += 4ull * kWG)
        *reinterpret_cast<float4 *>(out + e) = *reinterpret_cast<const float4 *>(&fs[e - Eb]);
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

bool make_geom(uint64_t n, int B, int T, Geom *g)
{
    if (B < 1 || B > kMaxBlock || T < 1) return false;
    const uint64_t b = (uint64_t)B;
    auto items_of = [&](uint64_t m) -> uint64_t { return m ? 1 + (m - 1 + b - 1) / b : 0; };
    auto blocks_of = [&](uint64_t m) -> uint64_t { return m ? (m - 1 + b - 1) / b : 0; };
    g->n = n;
    g->T = (uint32_t)T;
    g->B = (uint32_t)B;
    g->tbs = n / (uint64_t)T;
    const uint64_t nlast = n - (uint64_t)(T - 1) * g->tbs;
    g->c0 = items_of(g->tbs);
    g->clast = items_of(nlast);
    g->nb0 = blocks_of(g->tbs);
    g->nblast = blocks_of(nlast);
    g->nitems = (uint64_t)(T - 1) * g->c0 + g->clast;
    g->spb = (uint32_t)((B + 7) / 8);
    g->K = (uint32_t)kWG / g->spb;
    g->spb_magic = (65536u + g->spb - 1) / g->spb;
    g->ntiles = (g->nitems + g->K - 1) / g->K;
    g->body = 4 + 8ull * (uint64_t)T;
    g->pad = 0;
    return true;
}

uint64_t compress_bound(uint64_t n, int B, int T)
{
    Geom g;
    if (!make_geom(n, B, T, &g)) return 0;
    const uint64_t per_block = 1 + ((uint64_t)B + 7) / 8;
    uint64_t total = g.body;
    const uint64_t nl = n - (uint64_t)(T - 1) * g.tbs;
    if (g.tbs) total += (uint64_t)(T - 1) * (4 + g.nb0 * per_block + 4 * (g.tbs - 1));
    if (nl) total += 4 + g.nblast * per_block + 4 * (nl - 1);
    return total;
}

namespace {

constexpr uint64_t kAlign = 256;
constexpr int kRangeParts = 2048;

uint64_t align_up(uint64_t v) { return (v + kAlign - 1) & ~(kAlign - 1); }

struct CompressWs {
    uint32_t *ctl;
    float *eb;
    RangePart *parts;
    uint64_t *status;
};

uint64_t compress_ws_layout(const Geom &g, uint8_t *base, CompressWs *w)
{
    uint64_t off = 0;
    if (w) w->ctl = reinterpret_cast<uint32_t *>(base + off);
    off += align_up(64);
    if (w) w->eb = reinterpret_cast<float *>(base + off);
    off += align_up(16);
    if (w) w->parts = reinterpret_cast<RangePart *>(base + off);
    off += align_up(sizeof(RangePart) * kRangeParts);
    if (w) w->status = reinterpret_cast<uint64_t *>(base + off);
    off += align_up(8 * (g.ntiles ? g.ntiles : 1));
    return off;
}

struct DecompressWs {
    uint32_t *ctl;
    uint64_t *ctab;
    uint64_t *exit_st;
    uint64_t *cnt_st;
    uint64_t *dec_st;
    uint64_t *blkpos;
    uint64_t nclr;
};

uint32_t parse_tile_bytes(int B)
{
    const uint32_t W = block_bytes(32, (uint32_t)B);
    uint32_t S = W > 4096 ? W : 4096;
    return (S + 255) & ~255u;
}

uint64_t decompress_ws_layout(const Geom &g, uint64_t len, uint8_t *base, DecompressWs *w)
{
    const uint32_t S = parse_tile_bytes((int)g.B);
    const uint64_t nt1 = len / S + 1 + g.T;
    const uint64_t nblocks = (uint64_t)(g.T - 1) * g.nb0 + g.nblast;
    uint64_t off = 0;
    if (w) w->ctl = reinterpret_cast<uint32_t *>(base + off);
    off += align_up(64);
    if (w) w->ctab = reinterpret_cast<uint64_t *>(base + off);
    off += align_up(8 * (3ull * g.T + 1));
    const uint64_t clr0 = off;
    if (w) w->exit_st = reinterpret_cast<uint64_t *>(base + off);
    off += 8 * nt1;
    if (w) w->cnt_st = reinterpret_cast<uint64_t *>(base + off);
    off += 8 * nt1;
    if (w) w->dec_st = reinterpret_cast<uint64_t *>(base + off);
    off += 8 * (g.ntiles ? g.ntiles : 1);
    if (w) w->nclr = (off - clr0) / 8;
    off = align_up(off);
    if (w) w->blkpos = reinterpret_cast<uint64_t *>(base + off);
    off += align_up(8 * (nblocks ? nblocks : 1));
    return off;
}

}  // namespace

uint64_t compress_workspace_bytes(uint64_t n, int B, int T)
{
    Geom g;
    if (!make_geom(n, B, T, &g)) return 0;
    return compress_ws_layout(g, nullptr, nullptr);
}

uint64_t decompress_workspace_bytes(uint64_t stream_bytes, uint64_t n, int B, int T)
{
    Geom g;
    if (!make_geom(n, B, T, &g)) return 0;
    return decompress_ws_layout(g, stream_bytes, nullptr, nullptr);
}

uint32_t *workspace_error_word(void *d_work)
{
    return reinterpret_cast<uint32_t *>(d_work) + 1;
}

// Status codes shared with the public header (szp_amd_status).
enum { kOk = 0, kArg = 1, kCap = 2, kHip = 3 };

int launch_compress(const float *d_x, uint64_t n, int err_mode, double abs_eb, double rel, int B,
                    int T, uint8_t *d_out, uint64_t cap, uint64_t *d_size, void *d_work,
                    uint64_t work_bytes, hipStream_t stream)
{
    Geom g;
    if (!make_geom(n, B, T, &g) || !d_out || !d_size || !d_work) return kArg;
    if (err_mode != 0 && err_mode != 1) return kArg;
    if (err_mode == 1 && n == 0) return kArg;
    if (n && !d_x) return kArg;
    if ((reinterpret_cast<uintptr_t>(d_out) & 15) || (reinterpret_cast<uintptr_t>(d_work) & 255))
        return kArg;
    if (cap < g.body) return kCap;
    CompressWs w;
    if (compress_ws_layout(g, reinterpret_cast<uint8_t *>(d_work), &w) > work_bytes) return kCap;

    const uint64_t nst = g.ntiles ? g.ntiles : 1;
    const uint32_t cgrid = (uint32_t)std::min<uint64_t>((nst + kWG - 1) / kWG, 4096);
    hipLaunchKernelGGL(k_clear, dim3(cgrid), dim3(kWG), 0, stream, w.ctl, w.status, nst);
    const float *d_eb = nullptr;
    const float eb_host = (float)abs_eb;  // SZp/src/szp.cc:32
    const bool vec = (reinterpret_cast<uintptr_t>(d_x) & 15) == 0;
    if (err_mode == 1) {
        uint64_t seg = (n + kRangeParts - 1) / kRangeParts;
        seg = (seg + 4095) & ~4095ull;
        const int np = (int)((n + seg - 1) / seg);
        if (vec)
            hipLaunchKernelGGL(k_range_partial<true>, dim3(np), dim3(kWG), 0, stream, d_x, n, seg, w.parts);
        else
            hipLaunchKernelGGL(k_range_partial<false>, dim3(np), dim3(kWG), 0, stream, d_x, n, seg, w.parts);
        hipLaunchKernelGGL(k_range_final, dim3(1), dim3(kWG), 0, stream, d_x, w.parts, np,
                           (float)rel, w.eb);
        d_eb = w.eb;
    }
    if (n == 0) {
        hipLaunchKernelGGL(k_empty_stream, dim3(1), dim3(kWG), 0, stream, d_out, d_eb, eb_host,
                           g.T, d_size);
    } else if (vec) {
        hipLaunchKernelGGL(k_compress<true>, dim3((uint32_t)g.ntiles), dim3(kWG), 0, stream, d_x,
                           d_eb, eb_host, g, d_out, cap, w.status, w.ctl, d_size);
    } else {
        hipLaunchKernelGGL(k_compress<false>, dim3((uint32_t)g.ntiles), dim3(kWG), 0, stream, d_x,
                           d_eb, eb_host, g, d_out, cap, w.status, w.ctl, d_size);
    }
    return hipGetLastError() == hipSuccess ? kOk : kHip;
}

int launch_decompress(const uint8_t *d_stream, uint64_t len, uint64_t n, int B, int T,
                      float *d_out, void *d_work, uint64_t work_bytes, hipStream_t stream)
{
    Geom g;
    if (!make_geom(n, B, T, &g) || !d_stream || !d_work) return kArg;
    if (n && !d_out) return kArg;
    if ((reinterpret_cast<uintptr_t>(d_stream) & 15) || (reinterpret_cast<uintptr_t>(d_work) & 255))
        return kArg;
    DecompressWs w;
    if (decompress_ws_layout(g, len, reinterpret_cast<uint8_t *>(d_work), &w) > work_bytes)
        return kCap;
    const uint32_t S = parse_tile_bytes(B);
    const uint32_t W = block_bytes(32, (uint32_t)B);
    const uint32_t cgrid = (uint32_t)std::min<uint64_t>((w.nclr + kWG - 1) / kWG, 4096);
    hipLaunchKernelGGL(k_dsetup, dim3(cgrid ? cgrid : 1), dim3(kWG), 0, stream, d_stream, len, g,
                       S, w.ctab, w.exit_st, w.nclr, w.ctl);
    if (n == 0) return hipGetLastError() == hipSuccess ? kOk : kHip;
    const uint64_t nt1 = len / S + 1 + g.T;
    const size_t lds = ((S + 32 + 15) & ~15u) + 4ull * S;
    hipLaunchKernelGGL(k_parse, dim3((uint32_t)nt1), dim3(kWG), lds, stream, d_stream, len, g, S, W,
                       w.ctab, w.exit_st, w.cnt_st, w.blkpos, w.ctl);
    const bool vec = (reinterpret_cast<uintptr_t>(d_out) & 15) == 0;
    if (vec)
        hipLaunchKernelGGL(k_decode<true>, dim3((uint32_t)g.ntiles), dim3(kWG), 0, stream, d_stream,
                           len, g, w.ctab, w.blkpos, d_out, w.dec_st, w.ctl);
    else
        hipLaunchKernelGGL(k_decode<false>, dim3((uint32_t)g.ntiles), dim3(kWG), 0, stream, d_stream,
                           len, g, w.ctab, w.blkpos, d_out, w.dec_st, w.ctl);
    return hipGetLastError() == hipSuccess ? kOk : kHip;
}

}  // namespace szp_amd
