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
#include <stdio.h>
#include <stdlib.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

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
    const uint32_t k = 32u - __clz(m);
    if (__builtin_expect(k < 21, 1)) return k;
    return k + (m >= c_width_bump[k] ? 1u : 0u);
}

// (int)((double)x * (double)inv) with x86 cvttsd2si semantics: NaN and
// out-of-range products give INT32_MIN (SZp/src/szp_float.cc:209,228).  The
// product of two floats is exact in double.  |p| < 2^31 admits every product
// whose truncation fits int32 except (-2^31 - 1, -2^31], whose truncation is
// INT32_MIN anyway; NaN fails the compare.
__device__ __forceinline__ int32_t quantize(float x, float inv)
{
    const double p = (double)x * (double)inv;
    return __builtin_fabs(p) < 2147483648.0 ? (int32_t)p : INT32_MIN;
}

// (int)(x * inv) for double fields (SZp/src/szp_double.cc:209,228): the
// product is rounded to double, then truncated with the same x86 rule.
__device__ __forceinline__ int32_t quantize(double x, double inv)
{
    const double p = x * inv;
    return __builtin_fabs(p) < 2147483648.0 ? (int32_t)p : INT32_MIN;
}

// min/max that return the non-NaN operand, per value type.
__device__ __forceinline__ float vmin(float a, float b) { return fminf(a, b); }
__device__ __forceinline__ float vmax(float a, float b) { return fmaxf(a, b); }
__device__ __forceinline__ double vmin(double a, double b) { return fmin(a, b); }
__device__ __forceinline__ double vmax(double a, double b) { return fmax(a, b); }

// The error bound as the stream's first sizeof(V) bytes, big-endian
// (floatToBytes / doubleToBytes, SZp/include/szp_CompressionToolkit.h:520-546).
__device__ __forceinline__ void put_eb(uint8_t *out, float eb)
{
    const uint32_t u = __float_as_uint(eb);
#pragma unroll
    for (int i = 0; i < 4; ++i) out[i] = (uint8_t)(u >> (24 - 8 * i));
}
__device__ __forceinline__ void put_eb(uint8_t *out, double eb)
{
    const uint64_t u = (uint64_t)__double_as_longlong(eb);
#pragma unroll
    for (int i = 0; i < 8; ++i) out[i] = (uint8_t)(u >> (56 - 8 * i));
}
template <typename V>
__device__ __forceinline__ V get_eb(const uint8_t *st);
template <>
__device__ __forceinline__ float get_eb<float>(const uint8_t *st)
{
    return __uint_as_float(((uint32_t)st[0] << 24) | ((uint32_t)st[1] << 16) |
                           ((uint32_t)st[2] << 8) | (uint32_t)st[3]);
}
template <>
__device__ __forceinline__ double get_eb<double>(const uint8_t *st)
{
    uint64_t u = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) u = (u << 8) | st[i];
    return __longlong_as_double((long long)u);
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

// Inclusive wave64 prefix sum in six DPP steps: row_shr 1, 2, 4, 8 scan each
// row of 16 lanes (a lane whose source lies before its row keeps the 0 it
// was given), then row_bcast:15 adds row 0's total to row 1 and row 2's to
// row 3, and row_bcast:31 adds the total of rows 0-1 (lane 31) to rows 2-3.
__device__ __forceinline__ uint32_t wave_incl_scan_u32(uint32_t v, int /*lane*/)
{
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xF, 0xF, false);  // row_shr:1
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xF, 0xF, false);  // row_shr:2
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xF, 0xF, false);  // row_shr:4
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xF, 0xF, false);  // row_shr:8
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xA, 0xF, false);  // row_bcast:15
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xC, 0xF, false);  // row_bcast:31
    return v;
}

// Publishes a tile's aggregate (its inclusive value for the first tile of
// the scan domain) so that successors can look back past it.
__device__ __forceinline__ void lookback_publish(uint64_t *st, uint64_t tile, uint64_t first,
                                                 uint64_t agg)
{
    if (__lane_id() == 0) st_status(&st[tile], (tile == first ? kFlagIncl : kFlagAgg) | agg);
}

// Decoupled look-back for a plain sum (one wave calls it).  Tiles
// [first, tile) form the scan domain; returns the exclusive prefix of `tile`
// and publishes its inclusive value (and, unless `published`, first its
// aggregate).
__device__ uint64_t lookback_sum(uint64_t *st, uint64_t tile, uint64_t first, uint64_t agg,
                                 uint32_t *ctl, bool published = false)
{
    const int lane = __lane_id();
    if (!published) lookback_publish(st, tile, first, agg);
    if (tile == first) return 0;
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

// Decoupled look-back with the whole workgroup: every step reads the status
// of kWG predecessors at once (all threads call it; s_lb is 8 u64 of LDS).
// Returns the exclusive prefix of `tile` in every thread and publishes its
// inclusive value (and, unless `published`, first its aggregate).  WRAP32:
// the values are sums mod 2^32 (Lorenzo carries).
template <bool WRAP32 = false>
__device__ uint64_t lookback_sum_wg(uint64_t *st, uint64_t tile, uint64_t first, uint64_t agg,
                                    uint32_t *ctl, bool published, uint64_t *s_lb)
{
    const int tid = threadIdx.x, lane = __lane_id(), wv = tid >> 6;
    if (!published && tid < 64) lookback_publish(st, tile, first, agg);
    if (tile == first) return 0;
    uint64_t excl = 0;
    int64_t hi = (int64_t)tile - 1;
    for (uint32_t step = 0;; ++step) {
        const int64_t idx = hi - tid;
        const uint64_t s = (idx >= (int64_t)first) ? wait_status(&st[idx], ctl) : kFlagIncl;
        const uint64_t incl = __ballot((s >> 62) == 2);
        const int f = incl ? __ffsll((long long)incl) - 1 : 64;
        // This wave's statuses up to (and including) its nearest inclusive one.
        const uint64_t upto = wave_sum_u64(lane <= f ? (s & kValMask) : 0);
        uint64_t *slot = s_lb + (step & 1) * 4;  // alternating: one barrier per step
        if (lane == 0) slot[wv] = (upto << 1) | (f < 64 ? 1u : 0u);
        __syncthreads();
        bool found = false;
#pragma unroll
        for (int w = 0; w < kWG / 64; ++w) {
            const uint64_t e = slot[w];
            if (!found) {
                excl += e >> 1;
                found = e & 1;
            }
        }
        if (found) break;
        hi -= kWG;
    }
    if (WRAP32) excl = (uint32_t)excl;
    if (tid == 0) st_status(&st[tile], kFlagIncl | (WRAP32 ? (uint32_t)(excl + agg) : excl + agg));
    return excl;
}

// ---- compress look-back over tiles and groups of tiles ----
//
// With 1 500+ tiles in flight the nearest tile whose inclusive prefix is
// published lies several hundred tiles back (measured: ~460 at the median on
// a 512^3 field), and a look-back over single tiles then needs several
// dependent steps of kWG statuses.  The compress kernels therefore also keep
// one status word per group of kLbGroup consecutive tiles.  A tile status
// keeps the tile's own byte count in its low bits in both states.  A group's
// last tile publishes the group's total (flag kFlagAgg) right after its own
// byte count, from its predecessors' statuses, and its inclusive prefix as
// the group's (flag kFlagIncl) at the end of its look-back.  A tile reads the
// earlier tiles of its group and of the group before (waves 0-1) and 128
// earlier groups (waves 2-3) in one step: group totals do not wait for any
// look-back, so that step almost always reaches an inclusive group.  Each
// memory round trip costs ~2 us while the kernel streams, so one step
// matters.  Group words sit one 128-byte line apart, spreading the polling
// of the few open groups over memory channels.
//
// Tile status: [flag 2][inclusive prefix 40][tile bytes 22]; group status:
// [flag 2][value 62].
constexpr uint32_t kLbGroup = 64;       // tiles per look-back group
constexpr int kLbRawWaves = 2;          // look-back waves that read tile statuses
constexpr uint32_t kGroupStride = 16;   // u64 words between group statuses
constexpr uint32_t kAggBits = 22;
constexpr uint64_t kAggMask = (1ull << kAggBits) - 1;
constexpr uint64_t kInclMask = (1ull << (62 - kAggBits)) - 1;
// A tile holds at most kMaxRounds * 2048 values; each costs at most 4 head
// bytes or 1/8 + 4 block bytes plus a width byte per block.
static_assert((uint64_t)kMaxRounds * 2048 * 10 <= kAggMask, "a tile's byte count fits its field");
constexpr uint64_t kMaxCompressStream = kInclMask;  // streams beyond 1 TiB are refused

SZP_HD uint64_t lb_groups(uint64_t ntiles) { return (ntiles + kLbGroup - 1) / kLbGroup; }

// Publishes a compress tile's byte count (tile 0: also its inclusive prefix);
// wave 0 calls it.  A group's last tile then also publishes the group's total
// from its predecessors' statuses (published earlier, as they started
// earlier): the one wait here is a single round trip for one tile in
// kLbGroup, and it makes the group's total available a whole encode pass
// before the tile's own look-back.
__device__ __forceinline__ void lookback_publish_c(uint64_t *st, uint64_t *gst, uint64_t tile,
                                                   uint64_t agg, uint32_t *ctl)
{
    const int lane = __lane_id();
    if (lane == 0)
        st_status(&st[tile], tile == 0 ? (kFlagIncl | (agg << kAggBits) | agg) : (kFlagAgg | agg));
    if (tile % kLbGroup == kLbGroup - 1) {
        const uint64_t w = lane < (int)kLbGroup - 1 ? wait_status(&st[tile - 1 - lane], ctl) : 0;
        const uint64_t tot = wave_sum_u64(w & kAggMask) + agg;
        if (lane == 0) st_status(&gst[(tile / kLbGroup) * kGroupStride], kFlagAgg | tot);
    }
}

// Sum of a wave's values up to and including its nearest inclusive one
// (lanes ordered nearest first; v: each lane's value, incl: whether it is
// inclusive), and whether it met one.
__device__ __forceinline__ uint64_t wave_lookback_part(uint64_t v, bool incl, bool &found)
{
    const uint64_t m = __ballot(incl);
    const int f = m ? __ffsll((long long)m) - 1 : 64;
    found = m != 0;
    return wave_sum_u64(__lane_id() <= f ? v : 0);
}

// Look-back of compress tile `tile` after lookback_publish_c (all threads
// call it): returns its exclusive prefix in every thread and publishes its
// inclusive prefix, and for a group's last tile the group's.  gst: group statuses (kGroupStride apart); s_lb: 8 u64
// of LDS.
__device__ uint64_t lookback_compress(uint64_t *st, uint64_t *gst, uint64_t tile, uint64_t agg,
                                      uint32_t *ctl, uint64_t *s_lb)
{
    const int tid = threadIdx.x, wv = tid >> 6, lane = __lane_id();
    const uint64_t g = tile / kLbGroup;
    const uint32_t j = (uint32_t)(tile % kLbGroup);
    const bool last = j + 1 == kLbGroup;
    uint64_t excl = 0;
    if (tile != 0) {
        uint64_t v;
        bool incl;
        if (wv < kLbRawWaves) {
            // The earlier tiles of this group and of the kLbRawWaves - 1
            // groups before, nearest first (a group's total comes one round
            // trip after its last tile's byte count, and after the slowest
            // of its tiles).
            const uint32_t back = (uint32_t)min<uint64_t>(g, kLbRawWaves - 1) * kLbGroup;
            const bool in = (uint32_t)tid < j + back;
            const uint64_t w = in ? wait_status(&st[tile - 1 - tid], ctl) : 0;
            incl = in && (w >> 62) == 2;
            v = incl ? ((w >> kAggBits) & kInclMask) : (w & kAggMask);
        } else {
            // The groups before those, nearest first.
            const int64_t gi = (int64_t)g - kLbRawWaves - (int64_t)(tid - 64 * kLbRawWaves);
            const uint64_t w = gi >= 0 ? wait_status(&gst[(uint64_t)gi * kGroupStride], ctl) : kFlagIncl;
            incl = (w >> 62) == 2;
            v = w & kValMask;
        }
        bool found;
        const uint64_t part = wave_lookback_part(v, incl, found);
        if (lane == 0) s_lb[wv] = (part << 1) | (found ? 1u : 0u);
        __syncthreads();
        bool done = false;
#pragma unroll
        for (int w = 0; w < kWG / 64; ++w) {
            const uint64_t e = s_lb[w];
            if (!done) {
                excl += e >> 1;
                done = e & 1;
            }
        }
        // Further back, every thread reads a group (only when no group
        // read above has its inclusive prefix yet).
        int64_t hi = (int64_t)g - kLbRawWaves - (kWG - 64 * kLbRawWaves);
        while (!done) {  // uniform: every thread combines the same slots
            const int64_t gi = hi - tid;
            const uint64_t w = gi >= 0 ? wait_status(&gst[(uint64_t)gi * kGroupStride], ctl) : kFlagIncl;
            bool fw;
            const uint64_t pw = wave_lookback_part(w & kValMask, (w >> 62) == 2, fw);
            __syncthreads();  // every thread has read the previous slots
            if (lane == 0) s_lb[4 + wv] = (pw << 1) | (fw ? 1u : 0u);
            __syncthreads();
#pragma unroll
            for (int w2 = 0; w2 < kWG / 64; ++w2) {
                const uint64_t e = s_lb[4 + w2];
                if (!done) {
                    excl += e >> 1;
                    done = e & 1;
                }
            }
            hi -= kWG;
        }
    }
    if (tid == 0) {
        const uint64_t inc = excl + agg;
        st_status(&st[tile], kFlagIncl | ((inc & kInclMask) << kAggBits) | agg);
        if (last) st_status(&gst[g * kGroupStride], kFlagIncl | inc);
    }
    return excl;
}

// ---- foreign-stream decode look-back over tiles and groups of tiles ----
//
// The same grouping for the delta sums (mod 2^32) of the look-back decode.
// Its status words live at `st` (one per tile, the value in the low 32
// bits), then a tile's own first word (never overwritten: the group's last
// tile sums them even after a tile's status has turned inclusive), then the
// group statuses, kGroupStride words apart.  A tile that holds a chunk head
// is absolute: its first word is already inclusive, and a group's word is
// inclusive from the group's last absolute tile on.  Only fastMode-1 simple
// tiles read group words, so fastMode 2 publishes none.
__device__ __forceinline__ uint64_t *dec_own(uint64_t *st, const Geom &g) { return st + g.ntiles; }
__device__ __forceinline__ uint64_t *dec_groups(uint64_t *st, const Geom &g) { return st + 2 * g.ntiles; }

// Publishes a decode tile's first word (kFlagAgg | sum, or kFlagIncl | value
// for an absolute tile); wave 0 calls it (all 64 lanes).  A group's last
// tile then publishes the group's word from the group's 64 first words.
__device__ void lookback_publish_d(uint64_t *st, const Geom &g, uint64_t tile, uint64_t word,
                                   uint32_t *ctl)
{
    const int lane = __lane_id();
    if (lane == 0) {
        st_status(&dec_own(st, g)[tile], word);
        st_status(&st[tile], word);
    }
    if (g.mode != 1 || tile % kLbGroup != kLbGroup - 1) return;
    const uint64_t w = lane == 63 ? word : wait_status(&dec_own(st, g)[tile - 63 + lane], ctl);
    const uint64_t inc = __ballot((w >> 62) == 2);
    const int h = inc ? 63 - __builtin_clzll(inc) : 0;  // the last absolute tile
    const uint32_t sum = wave_sum_u32(lane >= h ? (uint32_t)w : 0u);
    if (lane == 0)
        st_status(&dec_groups(st, g)[(tile / kLbGroup) * kGroupStride], (inc ? kFlagIncl : kFlagAgg) | sum);
}

// Publishes a decode tile's inclusive value, and its group's for a group's
// last tile (lane 0 of the caller).
__device__ __forceinline__ void lookback_finish_d(uint64_t *st, const Geom &g, uint64_t tile,
                                                  uint32_t inc)
{
    st_status(&st[tile], kFlagIncl | inc);
    if (g.mode == 1 && tile % kLbGroup == kLbGroup - 1)
        st_status(&dec_groups(st, g)[(tile / kLbGroup) * kGroupStride], kFlagIncl | inc);
}

// A tile that stops on an error still publishes (inclusive, value 0) its
// status, its first word and its group's word, so that no tile waits for it.
__device__ __forceinline__ void publish_abort_d(uint64_t *st, const Geom &g, uint64_t tile)
{
    st_status(&dec_own(st, g)[tile], kFlagIncl);
    lookback_finish_d(st, g, tile, 0);
}

// Look-back of a decode simple tile (never absolute, never tile 0) after
// lookback_publish_d, in the shape of lookback_compress: returns the
// exclusive prefix mod 2^32 in every thread and publishes the inclusive one.
__device__ uint32_t lookback_decode(uint64_t *st, const Geom &g, uint64_t tile, uint32_t agg,
                                    uint32_t *ctl, uint64_t *s_lb)
{
    const int tid = threadIdx.x, wv = tid >> 6, lane = __lane_id();
    uint64_t *gst = dec_groups(st, g);
    const uint64_t gr = tile / kLbGroup;
    const uint32_t j = (uint32_t)(tile % kLbGroup);
    uint64_t v;
    bool incl;
    if (wv < kLbRawWaves) {
        const uint32_t back = (uint32_t)min<uint64_t>(gr, kLbRawWaves - 1) * kLbGroup;
        const bool in = (uint32_t)tid < j + back;
        const uint64_t w = in ? wait_status(&st[tile - 1 - tid], ctl) : 0;
        incl = in && (w >> 62) == 2;
        v = (uint32_t)w;
    } else {
        const int64_t gi = (int64_t)gr - kLbRawWaves - (int64_t)(tid - 64 * kLbRawWaves);
        const uint64_t w = gi >= 0 ? wait_status(&gst[(uint64_t)gi * kGroupStride], ctl) : kFlagIncl;
        incl = (w >> 62) == 2;
        v = (uint32_t)w;
    }
    bool found;
    const uint64_t part = wave_lookback_part(v, incl, found);
    if (lane == 0) s_lb[wv] = (part << 1) | (found ? 1u : 0u);
    __syncthreads();
    uint64_t excl = 0;
    bool done = false;
#pragma unroll
    for (int w = 0; w < kWG / 64; ++w) {
        const uint64_t e = s_lb[w];
        if (!done) {
            excl += e >> 1;
            done = e & 1;
        }
    }
    int64_t hi = (int64_t)gr - kLbRawWaves - (kWG - 64 * kLbRawWaves);
    while (!done) {  // uniform
        const int64_t gi = hi - tid;
        const uint64_t w = gi >= 0 ? wait_status(&gst[(uint64_t)gi * kGroupStride], ctl) : kFlagIncl;
        bool fw;
        const uint64_t pw = wave_lookback_part((uint32_t)w, (w >> 62) == 2, fw);
        __syncthreads();
        if (lane == 0) s_lb[4 + wv] = (pw << 1) | (fw ? 1u : 0u);
        __syncthreads();
#pragma unroll
        for (int w2 = 0; w2 < kWG / 64; ++w2) {
            const uint64_t e = s_lb[4 + w2];
            if (!done) {
                excl += e >> 1;
                done = e & 1;
            }
        }
        hi -= kWG;
    }
    if (tid == 0) lookback_finish_d(st, g, tile, (uint32_t)excl + agg);
    return (uint32_t)excl;
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
    const uint32_t nv = (uint32_t)((a1 - a0) >> 4);
    for (uint32_t v = tid; v < nv; v += kWG) {
        // One 16-byte LDS read at any byte offset (unaligned LDS access).
        uint4 val;
        __builtin_memcpy(&val, ob + head + 16 * v, 16);
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

template <typename V>
struct RangePart {
    V mn;
    V mx;
};

// The serial scan of SZp/src/szp.cc:36-45 keeps the first occurrence of the
// minimum and maximum, skips NaN (every comparison with it is false) and
// returns x[0] for both when all values are equal.  Its result enters the
// stream only through fl32(max - min), and that difference does not depend on
// which of two equal values was kept: with max != min one side is nonzero and
// the other's zero sign cannot change the exact difference, and with max ==
// min the scan subtracts x[0] from itself, i.e. max - max.  So the GPU keeps
// no indices: fminf/fmaxf (which return the non-NaN operand) in any order.
template <typename V>
__device__ __forceinline__ void range_reduce_wg(V &mn, V &mx)
{
    __shared__ V smn[kWG / 64], smx[kWG / 64];
    const int lane = __lane_id(), wv = threadIdx.x >> 6;
#pragma unroll
    for (int o = 32; o; o >>= 1) {
        mn = vmin(mn, __shfl_xor(mn, o));
        mx = vmax(mx, __shfl_xor(mx, o));
    }
    if (lane == 0) {
        smn[wv] = mn;
        smx[wv] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < kWG / 64; ++k) {
            mn = vmin(mn, smn[k]);
            mx = vmax(mx, smx[k]);
        }
    }
}

// Min and max of one 16-byte vector of values.
__device__ __forceinline__ void vec_minmax(const uint8_t *p, float &mn, float &mx)
{
    float4 q;
    __builtin_memcpy(&q, p, 16);
    mn = fminf(mn, fminf(fminf(q.x, q.y), fminf(q.z, q.w)));
    mx = fmaxf(mx, fmaxf(fmaxf(q.x, q.y), fmaxf(q.z, q.w)));
}
__device__ __forceinline__ void vec_minmax(const uint8_t *p, double &mn, double &mx)
{
    double2 q;
    __builtin_memcpy(&q, p, 16);
    mn = fmin(mn, fmin(q.x, q.y));
    mx = fmax(mx, fmax(q.x, q.y));
}

// Min and max of x[s, s + seg) per workgroup.  VEC: x is 16-byte aligned and
// seg a multiple of 4 * kWG * kRangeUnroll elements (only the last segment
// is ragged).
constexpr int kRangeUnroll = 4;
constexpr int kRangeParts = 2048;  // segments (workgroups) of the pass at most

// Segments are taken from the field's end backwards (workgroup b reads
// segment gridDim.x - 1 - b), with non-temporal loads.  Measured on a 512^3
// float field: with ordinary loads the pass ran at 88 us after a compress
// (the field's head, read last here and first by the compress kernel, was
// still in the 256 MB memory-side cache) but at 134 us after a decode (the
// cache then held the decode's dirty output, which had to be written back
// first); with non-temporal loads it runs at 82 us after either, and the
// compress kernel's loss of those cache hits (233 -> 255 us) is smaller than
// the gain (compress + decompress step 0.530 -> 0.503 ms).
//
// The pass also clears the compress call's control and look-back words (the
// compress kernel follows in stream order), which saves a launch.
template <typename V, bool VEC>
__global__ __launch_bounds__(kWG) void k_range_partial(const V *__restrict__ x, uint64_t n,
                                                       uint64_t seg, RangePart<V> *part,
                                                       uint32_t *ctl, uint64_t *status,
                                                       uint64_t nstatus)
{
    {
        const uint64_t per = (nstatus + gridDim.x - 1) / gridDim.x;
        const uint64_t b0 = (uint64_t)blockIdx.x * per, b1 = min(b0 + per, nstatus);
        for (uint64_t k = b0 + threadIdx.x; k < b1; k += kWG) status[k] = 0;
        if (blockIdx.x == 0 && threadIdx.x < 8) ctl[threadIdx.x] = 0;
    }
    constexpr uint32_t kPer = 16 / sizeof(V);  // values per 16-byte load
    const uint32_t sg = gridDim.x - 1 - blockIdx.x;
    const uint64_t s = (uint64_t)sg * seg;
    const uint32_t len = (uint32_t)(min(s + seg, n) - s);
    const V *xs = x + s;
    V mn = (V)__builtin_nan(""), mx = (V)__builtin_nan("");
    if (VEC) {
        constexpr uint32_t kStep = kPer * kWG * kRangeUnroll;
        const uint32_t full = len / kStep * kStep;
        for (uint32_t v = kPer * threadIdx.x; v < full; v += kStep) {
            typedef uint32_t u4v __attribute__((ext_vector_type(4)));
            u4v q[kRangeUnroll];
#pragma unroll
            for (int u = 0; u < kRangeUnroll; ++u)
                q[u] = __builtin_nontemporal_load(reinterpret_cast<const u4v *>(xs + v + kPer * kWG * u));
#pragma unroll
            for (int u = 0; u < kRangeUnroll; ++u)
                vec_minmax(reinterpret_cast<const uint8_t *>(&q[u]), mn, mx);
        }
        for (uint32_t i = full + threadIdx.x; i < len; i += kWG) {
            mn = vmin(mn, xs[i]);
            mx = vmax(mx, xs[i]);
        }
    } else {
        for (uint32_t i = threadIdx.x; i < len; i += kWG) {
            mn = vmin(mn, xs[i]);
            mx = vmax(mx, xs[i]);
        }
    }
    range_reduce_wg(mn, mx);
    if (threadIdx.x == 0) part[sg] = RangePart<V>{mn, mx};
}

// eb = (max - min) * rel in the value type (float: fl32(fl32(max - min) *
// fl32(rel)), SZp/src/szp.cc:36-47; double: SZp/src/szp.cc:70-84), NaN when
// x[0] is NaN.
template <typename V>
__global__ __launch_bounds__(kWG) void k_range_final(const V *__restrict__ x,
                                                     const RangePart<V> *part, int nparts,
                                                     V relf, V *d_eb)
{
    // Every part is loaded before any is combined (one memory round trip).
    constexpr int kPer = kRangeParts / kWG;
    RangePart<V> p[kPer];
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
        const int k = (int)threadIdx.x + u * kWG;
        p[u] = k < nparts ? part[k] : RangePart<V>{(V)__builtin_nan(""), (V)__builtin_nan("")};
    }
    V mn = (V)__builtin_nan(""), mx = (V)__builtin_nan("");
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
        mn = vmin(mn, p[u].mn);
        mx = vmax(mx, p[u].mx);
    }
    range_reduce_wg(mn, mx);
    if (threadIdx.x == 0) {
        V eb;
        if (isnan(x[0])) {
            eb = x[0];
        } else {
            const V range = (mx == mn) ? mx - mx : mx - mn;
            eb = range * relf;
        }
        *d_eb = eb;
    }
}

// Stream of an empty field: eb and T zero chunk offsets.
template <typename V>
__global__ void k_empty_stream(uint8_t *out, const V *d_eb, V eb_host, uint32_t T,
                               uint64_t *d_size)
{
    const V eb = d_eb ? *d_eb : eb_host;
    constexpr uint32_t es = sizeof(V);
    for (uint32_t i = threadIdx.x; i < es + 8 * T; i += blockDim.x)
        if (i >= es) out[i] = 0;
    if (threadIdx.x == 0) {
        put_eb(out, eb);
        *d_size = es + 8ull * T;
    }
}

// Unaligned LDS accesses (gfx950 runs with unaligned LDS access enabled; the
// compiler emits ds_read/ds_write_b32/b64/b128 for these copies).
__device__ __forceinline__ void lds_st32(uint8_t *p, uint32_t v) { __builtin_memcpy(p, &v, 4); }
__device__ __forceinline__ void lds_st64(uint8_t *p, uint32_t lo, uint32_t hi)
{
    const uint2 v = make_uint2(lo, hi);
    __builtin_memcpy(p, &v, 8);
}
__device__ __forceinline__ void lds_st128(uint8_t *p, uint32_t a, uint32_t b, uint32_t c, uint32_t d)
{
    const uint4 v = make_uint4(a, b, c, d);
    __builtin_memcpy(p, &v, 16);
}
__device__ __forceinline__ uint2 lds_ld64(const uint8_t *p)
{
    uint2 v;
    __builtin_memcpy(&v, p, 8);
    return v;
}
__device__ __forceinline__ uint4 lds_ld128(const uint8_t *p)
{
    uint4 v;
    __builtin_memcpy(&v, p, 16);
    return v;
}
// Writes the n (<= 8) low-order bytes of w to p, little-endian.
__device__ __forceinline__ void lds_st_bytes(uint8_t *p, uint64_t w, uint32_t n)
{
    if (n & 4) {
        lds_st32(p, (uint32_t)w);
        p += 4;
        w >>= 32;
    }
    if (n & 2) {
        const uint16_t v = (uint16_t)w;
        __builtin_memcpy(p, &v, 2);
        p += 2;
        w >>= 16;
    }
    if (n & 1) *p = (uint8_t)w;
}

// v[0]: the element before the octet (blocks only); v[1..8]: the octet.
template <typename V>
struct OctIn {
    V v[9];
};

template <typename V>
__device__ __forceinline__ void octet_load(const V *__restrict__ x, const Oct &s, OctIn<V> &in)
{
    if (s.m == 8) {
        __builtin_memcpy(&in.v[1], x + s.e0, 8 * sizeof(V));
        in.v[0] = x[s.e0 - 1];
    } else {
        in.v[0] = (s.m && !s.head) ? x[s.e0 - 1] : (V)0;
#pragma unroll
        for (int i = 0; i < 8; ++i) in.v[i + 1] = (uint32_t)i < s.m ? x[s.e0 + i] : (V)0;
    }
}

// Lorenzo deltas of the octet: magnitudes, MSB-first sign byte, maximum
// (SZp/src/szp_float.cc:220-251); the head's raw q for a head slot.
template <typename V>
__device__ __forceinline__ uint32_t octet_deltas(const Oct &s, const OctIn<V> &in, V inv,
                                                 uint32_t mg[8], uint32_t &sgn, int32_t &qhead)
{
    sgn = 0;
    qhead = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) mg[i] = 0;
    if (!s.m) return 0;
    if (s.head) {
        qhead = quantize(in.v[1], inv);
        return 0;
    }
    int32_t prev = quantize(in.v[0], inv);
    uint32_t omax = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const int32_t c = quantize(in.v[i + 1], inv);
        const uint32_t dd = (uint32_t)c - (uint32_t)prev;
        prev = c;
        const bool keep = (uint32_t)i < s.m;
        const bool neg = keep && (int32_t)dd < 0;
        const uint32_t mag = keep ? (neg ? 0u - dd : dd) : 0u;
        mg[i] = mag;
        sgn |= (uint32_t)neg << (7 - i);
        omax = max(omax, mag);
    }
    return omax;
}

// Writes the octet's bytes of its block at ob[ib]: the width byte (octet 0),
// its sign byte, its whole bytes of mag >> r (element-major little-endian)
// and its r-bit residuals (MSB-first), SZp/src/szp_TypeManager.cc:13-94,820-844.
__device__ __forceinline__ void encode_octet(uint8_t *ob, uint32_t ib, const Oct &s, uint32_t o,
                                             uint32_t b, const uint32_t mg[8], uint32_t sgn,
                                             int32_t qhead)
{
    if (!s.m) {
        // A fastMode-2 unit of one value still writes its block's zero width
        // byte (SZp/src/szp_float.cc:857-862).
        if (s.live && !s.head && s.L == 0 && o == 0) ob[ib] = 0;
        return;
    }
    if (s.head) {
        lds_st32(ob + ib, (uint32_t)qhead);
        return;
    }
    if (o == 0) ob[ib] = (uint8_t)b;
    if (!b) return;
    const uint32_t S8 = (s.L + 7) >> 3;
    ob[ib + 1 + o] = (uint8_t)sgn;
    const uint32_t nb = b >> 3, r = b & 7;
    uint8_t *p1 = ob + ib + 1 + S8 + 8 * o * nb;
    uint32_t v[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) v[i] = mg[i] >> r;  // < 2^(8 nb)
    if (nb) {
        if (s.m == 8) {
            if (nb == 1) {
                lds_st64(p1, v[0] | (v[1] << 8) | (v[2] << 16) | (v[3] << 24),
                         v[4] | (v[5] << 8) | (v[6] << 16) | (v[7] << 24));
            } else if (nb == 2) {
                lds_st128(p1, v[0] | (v[1] << 16), v[2] | (v[3] << 16), v[4] | (v[5] << 16),
                          v[6] | (v[7] << 16));
            } else if (nb == 3) {
                lds_st128(p1, v[0] | (v[1] << 24), (v[1] >> 8) | (v[2] << 16),
                          (v[2] >> 16) | (v[3] << 8), v[4] | (v[5] << 24));
                lds_st64(p1 + 16, (v[5] >> 8) | (v[6] << 16), (v[6] >> 16) | (v[7] << 8));
            } else {
                lds_st128(p1, v[0], v[1], v[2], v[3]);
                lds_st128(p1 + 16, v[4], v[5], v[6], v[7]);
            }
        } else {
            for (uint32_t i = 0; i < s.m; ++i)
                for (uint32_t k = 0; k < nb; ++k) p1[i * nb + k] = (uint8_t)(v[i] >> (8 * k));
        }
    }
    if (r) {
        const uint32_t mask = (1u << r) - 1;
        uint64_t acc = 0;
#pragma unroll
        for (int i = 0; i < 8; ++i) acc = (acc << r) | (mg[i] & mask);
        const uint32_t nbytes = s.m == 8 ? r : (s.m * r + 7) >> 3;
        // Big-endian bytes of the 8r-bit stream, first byte lowest in w.
        const uint64_t w = __builtin_bswap64(acc << (64 - 8 * r));
        lds_st_bytes(ob + ib + 1 + S8 + nb * s.L + o * r, w, nbytes);
    }
}

// Full-octet deltas for the simple-tile path: p = element before the octet,
// a, b = the octet.  mag = max(d, -d) is exact for every int32 d, INT32_MIN
// included (it maps to 0x80000000 as the reference's unsigned negation does).
template <typename V>
__device__ __forceinline__ uint32_t octet8_deltas(V p, const V v[8], V inv, uint32_t mg[8],
                                                  uint32_t &sgn)
{
    int32_t prev = quantize(p, inv);
    uint32_t omax = 0, sg = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const int32_t c = quantize(v[i], inv);
        const int32_t dd = (int32_t)((uint32_t)c - (uint32_t)prev);
        prev = c;
        const uint32_t mag = (uint32_t)max(dd, (int32_t)(0u - (uint32_t)dd));
        mg[i] = mag;
        sg = (sg << 1) | ((uint32_t)dd >> 31);
        omax = max(omax, mag);
    }
    sgn = sg;
    return omax;
}

// Plane bytes of the first 7 values of an octet (the last octet of a
// fastMode-2 unit's block, which holds 8 * SPB - 1 values): 7 * nb bytes,
// element-major little-endian, nothing written past them.
__device__ __forceinline__ void store_planes7(uint8_t *p1, uint32_t nb, const uint32_t v[8])
{
    if (nb == 1) {
        const uint64_t w = (uint64_t)(v[0] | (v[1] << 8) | (v[2] << 16) | (v[3] << 24)) |
                           ((uint64_t)(v[4] | (v[5] << 8) | (v[6] << 16)) << 32);
        lds_st_bytes(p1, w, 7);
    } else if (nb == 2) {
        lds_st64(p1, v[0] | (v[1] << 16), v[2] | (v[3] << 16));
        lds_st_bytes(p1 + 8, (uint64_t)(v[4] | (v[5] << 16)) | ((uint64_t)v[6] << 32), 6);
    } else if (nb == 3) {
        lds_st128(p1, v[0] | (v[1] << 24), (v[1] >> 8) | (v[2] << 16), (v[2] >> 16) | (v[3] << 8),
                  v[4] | (v[5] << 24));
        lds_st_bytes(p1 + 16, (uint64_t)((v[5] >> 8) | (v[6] << 16)) | ((uint64_t)(v[6] >> 16) << 32),
                     5);
    } else if (nb == 4) {
        lds_st128(p1, v[0], v[1], v[2], v[3]);
        lds_st64(p1 + 16, v[4], v[5]);
        lds_st32(p1 + 24, v[6]);
    }
}

// Encodes a full octet (8 values) of a block of L values at width b > 0;
// SEVEN: the block's last octet, of 7 values (mg[7] == 0, sign bit 0).
template <bool SEVEN = false>
__device__ __forceinline__ void encode_octet8(uint8_t *ob, uint32_t ib, uint32_t o, uint32_t L,
                                              uint32_t b, const uint32_t mg[8], uint32_t sgn)
{
    const uint32_t S8 = (L + 7) >> 3;
    ob[ib + 1 + o] = (uint8_t)sgn;
    const uint32_t nb = b >> 3, r = b & 7;
    uint8_t *p1 = ob + ib + 1 + S8 + 8 * o * nb;
    if (SEVEN) {
        uint32_t v[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) v[i] = mg[i] >> r;
        store_planes7(p1, nb, v);
    } else if (nb == 1) {
        uint32_t v[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) v[i] = mg[i] >> r;
        lds_st64(p1, v[0] | (v[1] << 8) | (v[2] << 16) | (v[3] << 24),
                 v[4] | (v[5] << 8) | (v[6] << 16) | (v[7] << 24));
    } else if (nb == 2) {
        uint32_t v[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) v[i] = mg[i] >> r;
        lds_st128(p1, v[0] | (v[1] << 16), v[2] | (v[3] << 16), v[4] | (v[5] << 16),
                  v[6] | (v[7] << 16));
    } else if (nb == 3) {
        uint32_t v[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) v[i] = mg[i] >> r;
        lds_st128(p1, v[0] | (v[1] << 24), (v[1] >> 8) | (v[2] << 16), (v[2] >> 16) | (v[3] << 8),
                  v[4] | (v[5] << 24));
        lds_st64(p1 + 16, (v[5] >> 8) | (v[6] << 16), (v[6] >> 16) | (v[7] << 8));
    } else if (nb == 4) {
        lds_st128(p1, mg[0], mg[1], mg[2], mg[3]);
        lds_st128(p1 + 16, mg[4], mg[5], mg[6], mg[7]);
    }
    if (r) {
        // 4r <= 28 bits per half: pack the residual stream in two 32-bit words.
        const uint32_t mask = (1u << r) - 1;
        uint32_t h = 0, l = 0;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            h = (h << r) | (mg[i] & mask);
            l = (l << r) | (mg[i + 4] & mask);
        }
        const uint64_t A = ((uint64_t)h << (4 * r)) | l;  // 8r bits, first value on top
        const uint64_t w = __builtin_bswap64(A << (64 - 8 * r));
        lds_st_bytes(ob + ib + 1 + S8 + nb * L + o * r, w, r);
    }
}

// Byte lengths of a round's items and their exclusive offsets (s_off),
// after every slot has folded its octet maximum into s_max.  Returns the
// round's bytes.  Called by all threads; three barriers.
__device__ __forceinline__ uint32_t round_layout(bool live, bool head, uint32_t L, uint32_t d,
                                                 uint32_t o, uint32_t K, uint32_t *s_max,
                                                 uint32_t *s_off, uint32_t *s_tot, uint32_t &b)
{
    const int tid = threadIdx.x;
    // The scan below reads all kWG entries; items past the round add 0.  The
    // previous round's readers of s_off finished before its closing barrier.
    s_off[tid] = 0;
    __syncthreads();
    uint32_t len = 0;
    b = 0;
    if (live) {
        if (head) {
            len = 4;
        } else {
            b = bit_width(s_max[d]);
            len = block_bytes(b, L);
        }
    }
    if (o == 0 && d < K) s_off[d] = len;
    __syncthreads();
    s_max[tid] = 0;
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
        if (tid == 63) *s_tot = incl;
    }
    __syncthreads();
    return *s_tot;
}

// Maximum of v over the SPB consecutive lanes of an item (SPB a power of two
// up to 32; items start at multiples of SPB).  DPP: quad_perm [1,0,3,2] and
// [2,3,0,1], then the mirror of each 8-lane half-row and of each 16-lane row;
// the 32-lane step swaps the two rows of each half-wave.
template <int SPB>
__device__ __forceinline__ uint32_t item_max(uint32_t v)
{
    if (SPB >= 2) v = max(v, (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, 0xB1, 0xF, 0xF, false));
    if (SPB >= 4) v = max(v, (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x4E, 0xF, 0xF, false));
    if (SPB >= 8) v = max(v, (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x141, 0xF, 0xF, false));
    if (SPB >= 16) v = max(v, (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x140, 0xF, 0xF, false));
    if (SPB >= 32) v = max(v, (uint32_t)__shfl_xor((int)v, 16));
    return v;
}

// v of lane l - 1 (lane 0 gets `first`): DPP wave_shr:1.
__device__ __forceinline__ float lane_before(float v, float first)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(first), __float_as_int(v),
                                                      0x138, 0xF, 0xF, false));
}
__device__ __forceinline__ double lane_before(double v, double first)
{
    const uint64_t u = (uint64_t)__double_as_longlong(v), f = (uint64_t)__double_as_longlong(first);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp((int)(uint32_t)f, (int)(uint32_t)u,
                                                              0x138, 0xF, 0xF, false);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp((int)(uint32_t)(f >> 32),
                                                              (int)(uint32_t)(u >> 32), 0x138, 0xF,
                                                              0xF, false);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// Copies tot bytes from src (LDS) to out[G, G + tot) with the 64 lanes of one
// wave: 16-byte stores for the aligned middle, byte stores for the ragged ends.
__device__ __forceinline__ void copy_out_wave(uint8_t *__restrict__ out, uint64_t G, uint32_t tot,
                                              const uint8_t *src, uint32_t lane)
{
    const uint64_t G1 = G + tot;
    const uint64_t a0 = (G + 15) & ~15ull, a1 = G1 & ~15ull;
    if (a0 >= a1) {
        for (uint32_t i = lane; i < tot; i += 64) out[G + i] = src[i];
        return;
    }
    const uint32_t head = (uint32_t)(a0 - G), tail = (uint32_t)(G1 - a1);
    if (lane < head) out[G + lane] = src[lane];
    if (lane < tail) out[a1 + lane] = src[(uint32_t)(a1 - G) + lane];
    const uint32_t nv = (uint32_t)((a1 - a0) >> 4);
    // Non-temporal: the stream is not read again by this kernel (k_compress
    // 203 -> 199 us on the 512^3 field, 0.598 -> 0.606 of peak end to end).
    typedef uint32_t u4 __attribute__((ext_vector_type(4)));
    for (uint32_t v = lane; v < nv; v += 64) {
        u4 val;
        __builtin_memcpy(&val, src + head + 16 * v, 16);  // any LDS byte offset
        __builtin_nontemporal_store(val, reinterpret_cast<u4 *>(out + a0 + 16ull * v));
    }
}

// LDS bytes of one wave's region in the simple path (a quarter of ob).
constexpr uint32_t kWaveCap = (uint32_t)kOutCap / 4;

// Lorenzo deltas of a quantized octet q against prev (the q of the element
// before it): magnitudes, MSB-first sign byte, maximum (SZp/src/szp_float.cc:220-251).
// mag = max(d, -d) is exact for every int32 d, INT32_MIN included (it maps to
// 0x80000000 as the reference's unsigned negation does).
__device__ __forceinline__ uint32_t octet_mags(const int32_t q[8], int32_t prev, uint32_t mg[8],
                                               uint32_t &sgn)
{
    uint32_t omax = 0, sg = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const int32_t dd = (int32_t)((uint32_t)q[i] - (uint32_t)prev);
        prev = q[i];
        const uint32_t mag = (uint32_t)max(dd, (int32_t)(0u - (uint32_t)dd));
        mg[i] = mag;
        sg = (sg << 1) | ((uint32_t)dd >> 31);
        omax = max(omax, mag);
    }
    sgn = sg;
    return omax;
}

// The largest magnitude of an octet's deltas (octet_mags' return value) from
// the deltas' maximum and minimum: |d| peaks at the largest positive or the
// most negative delta (INT32_MIN gives 0x80000000, as octet_mags does).
__device__ __forceinline__ uint32_t octet_max_mag(const int32_t q[8], int32_t prev)
{
    int32_t hi = 0, lo = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const int32_t dd = (int32_t)((uint32_t)q[i] - (uint32_t)prev);
        prev = q[i];
        hi = max(hi, dd);
        lo = min(lo, dd);
    }
    return max((uint32_t)hi, 0u - (uint32_t)lo);
}

// The q of the element before the lane's octet: the previous lane's last q
// (DPP wave_shr:1); lane 0 keeps qp, the q of its own preceding element.
__device__ __forceinline__ int32_t q_before(const int32_t q[8], int32_t qp)
{
    return __builtin_amdgcn_update_dpp(qp, q[7], 0x138, 0xF, 0xF, false);
}

// Fast path of k_compress for a simple tile (only full blocks of one chunk)
// when SPB = B / 8 is a power of two up to 32.  Wave w encodes the tile's
// w-th quarter, elements [2048 w, 2048 w + 2048) = blocks [w K, (w + 1) K)
// (K = 256 / SPB), in four steps of 512 elements: lane l holds elements
// 2048 w + 512 s + 8 l + [0, 8) of step s.  All element loads are issued
// before the first is used (idle lanes, past the tile's blocks, read the
// tile's first octet; their values are never used; Et >= 1, since the tile's
// first item is a block).
//
// Pass 1 quantizes every step (the q replace the values in registers), takes
// each block's maximum over its SPB lanes with DPP, and gives each block its
// byte offset in the quarter with one wave scan per step; one barrier then
// gives the tile's total, which is published for the look-back before any
// byte is encoded, and each quarter's offset in the tile.  Pass 2 takes the
// deltas again from the q and encodes each wave's quarter into its own LDS
// region (kWaveCap bytes); the quarters are consecutive in the stream.
//
// The wave's results: qbase (its quarter's offset in the tile), sbase[s]
// (step s's offset in the quarter; sbase[4] = the quarter's bytes), nfit and
// fitbytes (the steps encoded in the region and their bytes: a step that
// does not fit, and every later one, is encoded again after the look-back,
// one step at a time) and boff[s], the offset in the quarter of the lane's
// block in step s.
template <typename V, int SPB, bool INDEX>
__device__ __forceinline__ void compress_simple_waves(
    const V *__restrict__ x, V inv, uint64_t Et, uint32_t nti, uint8_t *ob, uint32_t *s_wtot,
    uint64_t *status, uint64_t *gst, uint32_t *ctl, uint64_t tile, uint32_t *idx_carry,
    uint32_t &run, uint32_t &qbase,
    uint32_t (&sbase)[5], uint32_t &nfit, uint32_t &fitbytes, uint32_t (&boff)[4])
{
    constexpr uint32_t K = 256 / SPB, KS = 64 / SPB, B = 8 * SPB;
    const int tid = threadIdx.x, lane = __lane_id(), wv = tid >> 6;
    const uint32_t o = (uint32_t)lane % SPB;
    // The lane's block in step s is b0 + s KS (simple_wave_block), its
    // elements Et + simple_wave_off(wv, s, lane) + [0, 8).
    const uint32_t b0 = (uint32_t)wv * K + (uint32_t)lane / SPB;
    int32_t q[4][8], qp[4];
    {
        V va[4][8];
        V p0[4];
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const uint64_t e = Et + ((b0 + (uint32_t)s * KS < nti)
                                         ? simple_wave_off((uint32_t)wv, (uint32_t)s, (uint32_t)lane)
                                         : 0u);
            __builtin_memcpy(&va[s][0], x + e, 8 * sizeof(V));
            p0[s] = x[e - 1];
        }
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            qp[s] = quantize(p0[s], inv);
#pragma unroll
            for (int i = 0; i < 8; ++i) q[s][i] = quantize(va[s][i], inv);
        }
    }
    // Side index: the q of the element before the tile (lane 0 of wave 0).
    if (INDEX && tid == 0) idx_carry[tile] = (uint32_t)qp[0];
    uint32_t bw[4];
    uint32_t base = 0;
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const bool live = b0 + (uint32_t)s * KS < nti;
        const uint32_t omax = item_max<SPB>(octet_max_mag(q[s], q_before(q[s], qp[s])));
        bw[s] = live ? bit_width(omax) : 0u;
        const uint32_t sz = live ? (bw[s] ? 1u + SPB * (bw[s] + 1u) : 1u) : 0u;  // block_bytes(bw, B)
        const uint32_t v = o == 0 ? sz : 0u;
        const uint32_t incl = wave_incl_scan_u32(v, lane);
        // sz is the same in every lane of a block, so incl - sz is the
        // block's start in each of them.
        boff[s] = base + incl - sz;
        sbase[s] = base;
        base += __builtin_amdgcn_readlane(incl, 63);
    }
    sbase[4] = base;
    if (lane == 0) s_wtot[wv] = base;
    __syncthreads();
    run = 0;
    qbase = 0;
#pragma unroll
    for (int w = 0; w < kWG / 64; ++w) {
        const uint32_t t = s_wtot[w];
        qbase += w < wv ? t : 0u;
        run += t;
    }
    if (tid < 64) lookback_publish_c(status, gst, tile, run, ctl);
    __builtin_amdgcn_s_setprio(0);
    uint8_t *reg = ob + (uint32_t)wv * kWaveCap;
    nfit = 0;
    fitbytes = 0;
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        if (sbase[s + 1] > kWaveCap) break;  // uniform in the wave
        // The deltas are taken again: an empty asm statement that "changes"
        // the q keeps the compiler from holding pass 1's magnitudes (32
        // registers) across the barrier instead.
#pragma unroll
        for (int i = 0; i < 8; ++i) asm volatile("" : "+v"(q[s][i]));
        asm volatile("" : "+v"(qp[s]));
        const int32_t pv = q_before(q[s], qp[s]);  // DPP with every lane active
        if (b0 + (uint32_t)s * KS < nti) {
            uint32_t mg[8], sg;
            octet_mags(q[s], pv, mg, sg);
            if (o == 0) reg[boff[s]] = (uint8_t)bw[s];
            if (bw[s]) encode_octet8(reg, boff[s], o, B, bw[s], mg, sg);
        }
        nfit = s + 1;
        fitbytes = sbase[s + 1];
    }
}

// Step s of a simple tile's quarter again, after the look-back: loads,
// deltas and widths as in compress_simple_waves, encoded at the start of the
// wave's region (the step's bytes always fit: at most KS blocks of
// block_bytes(32, B) bytes).
template <typename V, int SPB>
__device__ __forceinline__ void simple_step_again(const V *__restrict__ x, V inv, uint64_t Et,
                                                  uint32_t nti, uint32_t s, uint8_t *reg,
                                                  uint32_t off)
{
    constexpr uint32_t K = 256 / SPB, KS = 64 / SPB, B = 8 * SPB;
    const int lane = __lane_id(), wv = threadIdx.x >> 6;
    const uint32_t o = (uint32_t)lane % SPB;
    const uint32_t b = (uint32_t)wv * K + s * KS + (uint32_t)lane / SPB;
    const bool live = b < nti;
    const uint64_t e = Et + (live ? simple_wave_off((uint32_t)wv, s, (uint32_t)lane) : 0u);
    V v[8];
    __builtin_memcpy(&v[0], x + e, 8 * sizeof(V));
    const int32_t qp = quantize(x[e - 1], inv);
    int32_t q[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) q[i] = quantize(v[i], inv);
    uint32_t mg[8], sg;
    const uint32_t omax = item_max<SPB>(octet_mags(q, q_before(q, qp), mg, sg));
    const uint32_t bw = live ? bit_width(omax) : 0u;
    if (live) {
        if (o == 0) reg[off] = (uint8_t)bw;
        if (bw) encode_octet8(reg, off, o, B, bw, mg, sg);
    }
}

// Fast path of k_compress for a fastMode-2 tile of whole units (unit_tile)
// when SPB = B / 8 is a power of two up to 32.  A unit is B values: its
// first value's q (the head) and a block of the other B - 1 deltas.  Fast
// round r is K = 256 / SPB units (2048 values); thread tid owns values
// r * 2048 + 8 * tid + [0, 8), octet o = tid % SPB of unit tid / SPB.  For
// the block's layout the thread encodes the deltas of unit values 8o + 1 ..
// 8o + 8 (block indices 8o .. 8o + 7), taking value 8o + 8 from the next
// lane (DPP wave_shl:1; the unit's last lane has 7).  A tile holds at most
// 4096 values, whose stream bytes always fit in ob, so every round is
// encoded at once and the tile's total is published before any byte is.
// uoff[r]: in the lane that starts a unit (o == 0), the unit's offset in the
// tile.
template <typename V, int SPB>
__device__ __forceinline__ void compress_units_fast(
    const V *__restrict__ x, V inv, uint64_t Et, uint32_t nti, uint8_t *ob, uint32_t *s_wtot,
    uint64_t *status, uint64_t *gst, uint32_t *ctl, uint64_t tile, uint32_t &run,
    uint32_t (&uoff)[kMaxRounds / 2])
{
    constexpr uint32_t K = 256 / SPB, B = 8 * SPB;
    constexpr int kRounds = kMaxRounds / 2;
    static_assert(kRounds % 2 == 0, "the size scan packs rounds in pairs");
    // A unit's bytes: 4 + block_bytes(b, B - 1) <= 1 + 33 SPB.
    static_assert(2u * K * (1u + 33u * SPB) <= (uint32_t)kOutCap, "a unit tile's bytes fit in ob");
    const int tid = threadIdx.x, lane = __lane_id(), wv = tid >> 6;
    const uint32_t d = (uint32_t)tid / SPB, o = (uint32_t)tid % SPB;
    const uint32_t nu = nti / 2;  // units in the tile
    const bool last = o + 1 == SPB;

    // Unconditional loads (idle slots read the tile's first octet; their
    // units are not live), so that both rounds' loads go out back to back.
    V va[kRounds][8];
#pragma unroll
    for (int r = 0; r < kRounds; ++r)
        __builtin_memcpy(&va[r][0],
                         x + Et + ((r * K + d < nu) ? (uint64_t)r * 2048u + 8u * (uint32_t)tid : 0u),
                         8 * sizeof(V));
    uint32_t mg[kRounds][8], sg[kRounds], bw[kRounds], sz[kRounds];
    int32_t hq[kRounds];
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
        int32_t q[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) q[i] = quantize(va[r][i], inv);
        const int32_t q8 = __builtin_amdgcn_update_dpp(0, q[0], 0x130, 0xF, 0xF, false);
        hq[r] = q[0];
        uint32_t omax = 0, sgn = 0;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const bool keep = k < 7 || !last;
            const int32_t dd = (int32_t)((uint32_t)(k < 7 ? q[k + 1] : q8) - (uint32_t)q[k]);
            const uint32_t mag = keep ? (uint32_t)max(dd, (int32_t)(0u - (uint32_t)dd)) : 0u;
            mg[r][k] = mag;
            sgn = (sgn << 1) | (keep ? (uint32_t)dd >> 31 : 0u);
            omax = max(omax, mag);
        }
        sg[r] = sgn;
        const bool live = r * K + d < nu;
        omax = item_max<SPB>(omax);
        bw[r] = live ? bit_width(omax) : 0u;
        // 4 head bytes + block_bytes(b, 8 * SPB - 1) = 1 + SPB * (b + 1) - b / 8.
        sz[r] = live ? 4u + (bw[r] ? 1u + SPB * (bw[r] + 1u) - (bw[r] >> 3) : 1u) : 0u;
    }
    uint32_t ex[kRounds];
#pragma unroll
    for (int r = 0; r < kRounds; r += 2) {
        const uint32_t v = o == 0 ? (sz[r] | (sz[r + 1] << 16)) : 0u;
        const uint32_t incl = wave_incl_scan_u32(v, lane);
        ex[r] = (incl & 0xFFFFu) - sz[r];
        ex[r + 1] = (incl >> 16) - sz[r + 1];
        const uint32_t wt = __builtin_amdgcn_readlane(incl, 63);
        if (lane == 0) {
            s_wtot[r * 4 + wv] = wt & 0xFFFFu;
            s_wtot[(r + 1) * 4 + wv] = wt >> 16;
        }
    }
    __syncthreads();
    uint32_t acc = 0;
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
        uint32_t before = 0, rt = 0;
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            const uint32_t t = s_wtot[r * 4 + w];
            before += w < wv ? t : 0u;
            rt += t;
        }
        uoff[r] = acc + before + ex[r];
        acc += rt;
    }
    run = acc;
    if (tid < 64) lookback_publish_c(status, gst, tile, acc, ctl);
    __builtin_amdgcn_s_setprio(0);
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
        if (r * K + d >= nu) continue;
        if (o == 0) {
            lds_st32(ob + uoff[r], (uint32_t)hq[r]);
            ob[uoff[r] + 4] = (uint8_t)bw[r];
        }
        if (bw[r]) {
            if (last)
                encode_octet8<true>(ob, uoff[r] + 4, o, B - 1, bw[r], mg[r], sg[r]);
            else
                encode_octet8(ob, uoff[r] + 4, o, B - 1, bw[r], mg[r], sg[r]);
        }
    }
}

// A tile is g.R rounds of K items.  Simple and unit tiles take the fast
// paths above; every other tile (chunk heads, short blocks, B not a multiple
// of 8) is encoded one round at a time, each thread owning one octet per
// round, loaded one round ahead of its use.  Rounds are encoded into the LDS
// buffer ob (kOutCap bytes) as long as they fit; the first round that does
// not fit and all later ones are encoded again after the tile's offset is
// known.  One decoupled look-back per tile places the tile; tiles are ordered
// by blockIdx.x (each XCD dispatches its workgroups in index order, and every
// wait is bounded).
//
// With INDEX the kernel also writes a side index beside the stream (the
// stream itself is unchanged): the stream position of every block
// (idx_pos[global block]) and, per tile, the q of the element before the
// tile's first element (idx_carry[tile]), so that a decoder needs neither a
// parse nor a scan across tiles.
// LDS of one compress workgroup: the tile's encoded bytes (four wave regions
// in the simple path) and the per-item scan arrays of the general path.
struct CompressLds {
    __align__(16) uint8_t ob[kOutCap + 64];
    uint32_t s_max[kWG];
    uint32_t s_off[kWG];
    uint32_t s_tot;
    uint32_t s_rbase[kMaxRounds + 1];
    uint32_t s_wtot[kMaxRounds * 4];
    uint64_t s_lb[8];
    int s_stop;
    uint64_t s_prefix;
};

// One round of a general tile: deltas of the thread's octet `cur` (loaded in
// `in`), the block widths and item offsets of the round (s_off), and the
// octet's bytes when the round fits in ob at `run`.  Returns the round's
// bytes; ib = the thread's item's offset in the round.  Barriers inside.
template <typename V>
__device__ __forceinline__ uint32_t general_round(const Oct &cur, const OctIn<V> &in, V inv,
                                                  uint32_t d, uint32_t o, uint32_t K,
                                                  CompressLds &L, uint32_t run, bool encode,
                                                  uint32_t &ib)
{
    const int tid = threadIdx.x;
    uint32_t mg[8], sgn;
    int32_t qh;
    L.s_off[tid] = 0;  // the scan reads all kWG entries; items >= K contribute 0
    const uint32_t omax = octet_deltas(cur, in, inv, mg, sgn, qh);
    if (omax) atomicMax(&L.s_max[d], omax);
    __syncthreads();
    uint32_t b = 0, len = 0;
    if (cur.live) {
        if (cur.head) {
            len = 4;
        } else {
            b = bit_width(L.s_max[d]);
            len = block_bytes(b, cur.L);
        }
    }
    if (o == 0 && d < K) L.s_off[d] = len;
    __syncthreads();
    L.s_max[tid] = 0;
    if (tid < 64) {
        const uint32_t a0 = L.s_off[4 * tid], a1 = L.s_off[4 * tid + 1];
        const uint32_t a2 = L.s_off[4 * tid + 2], a3 = L.s_off[4 * tid + 3];
        const uint32_t sum = a0 + a1 + a2 + a3;
        const uint32_t incl = wave_incl_scan_u32(sum, tid);
        const uint32_t ex = incl - sum;
        L.s_off[4 * tid] = ex;
        L.s_off[4 * tid + 1] = ex + a0;
        L.s_off[4 * tid + 2] = ex + a0 + a1;
        L.s_off[4 * tid + 3] = ex + a0 + a1 + a2;
        if (tid == 63) L.s_tot = incl;
    }
    __syncthreads();
    const uint32_t rt = L.s_tot;
    ib = cur.live ? L.s_off[d] : 0u;
    if (encode && run + rt <= (uint32_t)kOutCap) encode_octet(L.ob, run + ib, cur, o, b, mg, sgn, qh);
    __syncthreads();  // s_off and s_tot are read before the next round writes them
    return rt;
}

// One compress tile (all threads of the workgroup call it; every exit is
// uniform, so a persistent workgroup can go on to its next tile).
// EARLY_ERR: read the error word before the look-back (the single-field
// kernel; the batched kernel reads it after).
template <typename V, bool INDEX, int SPB, bool EARLY_ERR = true>
__device__ void compress_tile(const V *__restrict__ x, const V *__restrict__ d_eb,
                              V eb_host, const Geom &g, uint8_t *__restrict__ out,
                              uint64_t cap, uint64_t *__restrict__ status,
                              uint32_t *__restrict__ ctl, uint64_t *__restrict__ d_size,
                              uint64_t *__restrict__ idx_pos, uint32_t *__restrict__ idx_carry,
                              uint64_t tile, CompressLds &L)
{
    // Until the tile's byte count is published its waves issue first: every
    // later tile's look-back waits for the slowest recent predecessor to
    // publish, so the work before a publish is on everyone's critical path,
    // the work after it only on the tile's own (k_compress 214 -> 204 us on
    // a 512^3 field).
    __builtin_amdgcn_s_setprio(2);
    const int tid = threadIdx.x, lane = __lane_id();
    const V eb = d_eb ? *d_eb : eb_host;
    // 1 / eb in the value type: FP32 division for float (SZp/src/szp_float.cc:185),
    // FP64 for double (SZp/src/szp_double.cc:185).
    const V inv = (V)1 / eb;
    const uint64_t tI0 = tile * (uint64_t)g.K * g.R;
    const uint64_t tIN = min(tI0 + (uint64_t)g.K * g.R, g.nitems);
    const uint32_t d = ((uint32_t)tid * g.spb_magic) >> 16;
    const uint32_t o = (uint32_t)tid - d * g.spb;
    uint32_t ct;
    uint64_t cj;
    locate(g, tI0, ct, cj);
    const uint32_t nti = (uint32_t)(tIN - tI0);
    const bool simple = SPB && simple_tile(g, ct, cj, nti);
    const bool units = SPB && !simple && unit_tile(g, ct, cj, nti);

    uint32_t run = 0;       // the tile's bytes
    uint64_t *gst = status + g.ntiles;  // look-back group statuses
    // Simple path (per wave).
    uint32_t qbase = 0, nfit = 0, fitbytes = 0, sbase[5] = {0, 0, 0, 0, 0},
             boff[4] = {0, 0, 0, 0};
    // Unit path: unit offsets in the tile (lanes o == 0).
    uint32_t uoff[kMaxRounds / 2] = {0, 0};
    // General path: rounds held in ob, and each round's item offset in the
    // tile for the thread that starts the item (o == 0).
    uint32_t nbuf = 0;
    uint32_t ioff[kMaxRounds] = {0, 0, 0, 0};
    if (simple) {
        compress_simple_waves<V, (SPB ? SPB : 1), INDEX>(x, inv, item_elem(g, ct, cj), nti, L.ob,
                                                         L.s_wtot, status, gst, ctl, tile,
                                                         idx_carry, run,
                                                         qbase, sbase, nfit, fitbytes, boff);
    } else if (units) {
        if (INDEX && tid == 0)
            idx_carry[tile] = cj ? (uint32_t)quantize(x[item_elem(g, ct, cj) - 1], inv) : 0u;
        compress_units_fast<V, (SPB ? SPB : 1)>(x, inv, item_elem(g, ct, cj), nti, L.ob, L.s_wtot,
                                                status, gst, ctl, tile, run, uoff);
        nbuf = g.R;  // the whole tile is in ob
    } else {
        if (INDEX && tid == 0)
            idx_carry[tile] = cj ? (uint32_t)quantize(x[item_elem(g, ct, cj) - 1], inv) : 0u;
        L.s_max[tid] = 0;
        __syncthreads();
        // Round r + 1's octet is loaded while round r is encoded.
        uint32_t t = ct;
        uint64_t j = cj;
        Oct cur{0, 0, 0, 0, false, false, false};
        OctIn<V> cin;
        {
            const uint32_t nr = (uint32_t)min<uint64_t>(g.K, tIN - tI0);
            cur = octet_of(g, t, j, nr, d, o);
            octet_load(x, cur, cin);
            advance(g, t, j, nr);
        }
        bool fits = true;
#pragma unroll
        for (uint32_t r = 0; r < (uint32_t)kMaxRounds; ++r) {
            const uint64_t ri0 = tI0 + (uint64_t)r * g.K;
            if (r >= g.R || ri0 >= tIN) break;
            Oct nxt{0, 0, 0, 0, false, false, false};
            OctIn<V> nin;
            const uint64_t rn0 = ri0 + g.K;
            if (r + 1 < g.R && rn0 < tIN) {
                const uint32_t nr = (uint32_t)min<uint64_t>(g.K, tIN - rn0);
                nxt = octet_of(g, t, j, nr, d, o);
                octet_load(x, nxt, nin);
                advance(g, t, j, nr);
            }
            if (tid == 0) L.s_rbase[r] = run;
            uint32_t ib;
            const uint32_t rt = general_round<V>(cur, cin, inv, d, o, g.K, L, run, fits, ib);
            if (o == 0 && cur.live) ioff[r] = run + ib;
            fits = fits && run + rt <= (uint32_t)kOutCap;
            if (fits) nbuf = r + 1;
            run += rt;
            cur = nxt;
            cin = nin;
        }
    }

    // The error word is read before the look-back, so that its latency
    // overlaps the wait instead of following it (a timeout raised during
    // this tile's own look-back still stops later tiles, and this tile's
    // writes stay within the capacity check below).
    const uint32_t err0 = (EARLY_ERR && tid == 0) ? ld_ctl(&ctl[1]) : 0u;
    {
        if (!(simple || units) && tid < 64) lookback_publish_c(status, gst, tile, run, ctl);
        __builtin_amdgcn_s_setprio(0);
        const uint64_t p = lookback_compress(status, gst, tile, run, ctl, L.s_lb);
        if (tid == 0) L.s_prefix = p;
    }
    __syncthreads();
    const uint64_t prefix = L.s_prefix;
    const uint64_t G0 = g.body + prefix;
    if (tIN == g.nitems && tid == 0) *d_size = G0 + run;
    if (tid == 0) {
        L.s_stop = G0 + run > cap || ((EARLY_ERR ? err0 : ld_ctl(&ctl[1])) & kErrTimeout);
        if (G0 + run > cap) raise_error(ctl, kErrCapacity);
    }
    __syncthreads();
    if (L.s_stop) return;
    if (tile == 0) {
        if (tid == 0) put_eb(out, eb);
        // Empty chunks (n < T) all start at offset 0.
        if (g.tbs == 0) {
            for (uint32_t k = tid; k + 1 < g.T; k += kWG) {
                uint32_t *h = reinterpret_cast<uint32_t *>(out + g.es + 8ull * k);
                h[0] = 0;
                h[1] = 0;
            }
        }
    }

    if (simple) {
        constexpr uint32_t SP = SPB ? SPB : 1;
        constexpr uint32_t K = 256 / SP, KS = 64 / SP;
        const uint32_t wv = (uint32_t)tid >> 6;
        const uint32_t ow = (uint32_t)lane % SP, b0 = wv * K + (uint32_t)lane / SP;
        const uint64_t Gq = G0 + qbase;
        if (INDEX && ow == 0) {
            // Blocks cj .. cj + nti - 1 of chunk ct: consecutive index entries.
            uint64_t *ip = idx_pos + (uint64_t)ct * g.nb0 + (cj - 1);
#pragma unroll
            for (int s = 0; s < 4; ++s)
                if (b0 + (uint32_t)s * KS < nti) ip[b0 + s * KS] = Gq + boff[s];
        }
        uint8_t *reg = L.ob + wv * kWaveCap;
        copy_out_wave(out, Gq, fitbytes, reg, (uint32_t)lane);
        if (nfit < 4) {
            const uint64_t Et = item_elem(g, ct, cj);
#pragma unroll
            for (uint32_t s = 0; s < 4; ++s) {
                if (s < nfit) continue;  // uniform in the wave
                simple_step_again<V, SP>(x, inv, Et, nti, s, reg, boff[s] - sbase[s]);
                copy_out_wave(out, Gq + sbase[s], sbase[s + 1] - sbase[s], reg, (uint32_t)lane);
            }
        }
        return;
    }
    if (units) {
        constexpr uint32_t SP = SPB ? SPB : 1;
        constexpr uint32_t K = 256 / SP;
        if (o == 0) {
#pragma unroll
            for (int r = 0; r < kMaxRounds / 2; ++r) {
                if (r * K + d >= nti / 2) continue;
                if (INDEX) idx_pos[(uint64_t)ct * g.nb0 + cj / 2 + r * K + d] = G0 + uoff[r];
            }
        }
        // The chunk's first unit: its offset goes in the header
        // (SZp/src/szp_float.cc:280-288).
        if (cj == 0 && tid == 0) {
            uint32_t *h = reinterpret_cast<uint32_t *>(out + g.es + 8ull * ct);
            h[0] = (uint32_t)prefix;
            h[1] = (uint32_t)(prefix >> 32);
        }
        copy_out_bytes(out, G0, run, L.ob, tid);
        return;
    }
    // General tile: index entries of its blocks (units in fastMode 2) and the
    // header offsets of the chunks it starts (SZp/src/szp_float.cc:280-288).
    if (o == 0) {
#pragma unroll
        for (uint32_t r = 0; r < (uint32_t)kMaxRounds; ++r) {
            if (r >= g.R || (uint64_t)r * g.K + d >= nti || d >= g.K) break;
            uint32_t t = ct;
            uint64_t j = cj;
            advance(g, t, j, r * g.K + d);
            if (INDEX && item_indexed(g, j))
                idx_pos[(uint64_t)t * g.nb0 + item_block(g, j)] = G0 + ioff[r];
            if (j == 0) {
                const uint64_t off = prefix + ioff[r];
                uint32_t *h = reinterpret_cast<uint32_t *>(out + g.es + 8ull * t);
                h[0] = (uint32_t)off;
                h[1] = (uint32_t)(off >> 32);
            }
        }
    }
    const bool all_buffered = tI0 + (uint64_t)nbuf * g.K >= tIN;
    const uint32_t buffered = all_buffered ? run : L.s_rbase[nbuf];
    copy_out_bytes(out, G0, buffered, L.ob, tid);
    // Rounds that did not fit: encode again and write them one at a time.
    for (uint32_t r = nbuf; r < g.R; ++r) {
        const uint64_t ri0 = tI0 + (uint64_t)r * g.K;
        if (ri0 >= tIN) break;
        const uint32_t nit = (uint32_t)min<uint64_t>(g.K, tIN - ri0);
        uint32_t rt0;
        uint64_t rj0;
        locate(g, ri0, rt0, rj0);
        const Oct s = octet_of(g, rt0, rj0, nit, d, o);
        OctIn<V> in;
        octet_load(x, s, in);
        __syncthreads();  // ob is free (the previous copy has read it)
        uint32_t ib;
        const uint32_t rt = general_round<V>(s, in, inv, d, o, g.K, L, 0, true, ib);
        // The round was sized once already; a different size here would be a
        // defect, so it is reported instead of written.
        const uint32_t rend = (ri0 + g.K < tIN) ? L.s_rbase[r + 1] : run;
        if (L.s_rbase[r] + rt != rend || rt > (uint32_t)kOutCap) {
            if (tid == 0) raise_error(ctl, kErrCapacity);
            return;
        }
        copy_out_bytes(out, G0 + L.s_rbase[r], rt, L.ob, tid);
    }
}

// Float fields: seven workgroups per CU (at most 72 VGPRs, no spills; the
// LDS of CompressLds is under 20 KiB); at 64 VGPRs the simple path spilled.
// Double fields keep more registers (their octets are twice as wide) and run
// fewer workgroups.
template <typename V, bool INDEX, int SPB>
__global__ __launch_bounds__(kWG, sizeof(V) == 4 ? 7 : 1) void k_compress(const V *__restrict__ x,
                                                  const V *__restrict__ d_eb, V eb_host,
                                                  Geom g, uint8_t *__restrict__ out, uint64_t cap,
                                                  uint64_t *__restrict__ status,
                                                  uint32_t *__restrict__ ctl,
                                                  uint64_t *__restrict__ d_size,
                                                  uint64_t *__restrict__ idx_pos,
                                                  uint32_t *__restrict__ idx_carry,
                                                  TileOrder ord)
{
    __shared__ CompressLds L;
    compress_tile<V, INDEX, SPB>(x, d_eb, eb_host, g, out, cap, status, ctl, d_size, idx_pos,
                                 idx_carry, tile_of_block(ord, blockIdx.x), L);
}

// ---------------------------------------------------------------------------
// Decompress
// ---------------------------------------------------------------------------

// Copies st[a0, b) into buf[0, b - a0) (a0 16-byte aligned; the last vector
// may copy up to 15 bytes past b, never past len), 16 bytes per thread per
// step.  A thread's first kStageVec vectors are all loaded before any is
// stored, from an always-valid address (an idle slot reads a vector it then
// drops), so their latencies overlap: a loop that stores each vector before
// loading the next waits for every load in turn.
constexpr int kStageVec = 6;  // 24 KiB per workgroup: a whole parse tile (16 KiB + overlap)
__device__ __forceinline__ void stage_bytes(const uint8_t *__restrict__ st, uint64_t len,
                                            uint64_t a0, uint64_t b, uint8_t *buf)
{
    typedef uint32_t u4v __attribute__((ext_vector_type(4)));
    static_assert(kStageVec == 6, "the barrier below names every staged vector");
    uint64_t q = a0 + 16ull * threadIdx.x;
    if (len >= 16) {
        const uint64_t safe = a0 + 16 <= len ? a0 : len - 16;
        u4v v[kStageVec];
#pragma unroll
        for (int k = 0; k < kStageVec; ++k) {
            const uint64_t qk = q + 16ull * kWG * k;
            const uint64_t at = (qk < b && qk + 16 <= len) ? qk : safe;
            __builtin_memcpy(&v[k], st + at, 16);
        }
        // An empty statement that takes every loaded vector: the compiler
        // cannot sink a load into the conditional store that uses it (which
        // would put a wait between one load and the next).
        asm volatile("" ::"v"(v[0]), "v"(v[1]), "v"(v[2]), "v"(v[3]), "v"(v[4]), "v"(v[5]));
#pragma unroll
        for (int k = 0; k < kStageVec; ++k) {
            const uint64_t qk = q + 16ull * kWG * k;
            if (qk < b) {
                if (qk + 16 <= len) {
                    *reinterpret_cast<u4v *>(buf + (qk - a0)) = v[k];
                } else {
                    for (int j = 0; j < 16; ++j)
                        if (qk + j < len) buf[qk - a0 + j] = st[qk + j];
                }
            }
        }
        q += 16ull * kWG * kStageVec;
    }
    for (; q < b; q += 16ull * kWG) {
        if (q + 16 <= len) {
            *reinterpret_cast<uint4 *>(buf + (q - a0)) = *reinterpret_cast<const uint4 *>(st + q);
        } else {
            for (int k = 0; k < 16; ++k)
                if (q + k < len) buf[q - a0 + k] = st[q + k];
        }
    }
}

__device__ __forceinline__ uint64_t read_u64_le4(const uint8_t *p)
{
    const uint32_t *w = reinterpret_cast<const uint32_t *>(p);
    return (uint64_t)w[0] | ((uint64_t)w[1] << 32);
}

// Reads the chunk offsets, validates them and lays out the parse tiles:
// ctab = [cstart[T] | aend[T] | ttab[T+1]] where chunk t's blocks occupy
// [cstart[t] + 4, aend[t]) (mode 2: its units [cstart[t], aend[t])) and its
// parse tiles are [ttab[t], ttab[t+1]).
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
        const uint64_t off = read_u64_le4(st + g.es + 8ull * t);
        const uint64_t cs = g.body + off;
        ttab[t] = tiles;
        cstart[t] = cs;
        aend[t] = cs;
        if (nt == 0) {
            if (off != 0) bad = true;
            continue;
        }
        const uint64_t ce = (t + 1 < g.T) ? g.body + read_u64_le4(st + g.es + 8ull * (t + 1)) : len;
        // The parsed area: blocks after the chunk's head (mode 1), or all of
        // the chunk's units (mode 2, at least one of 5 bytes).
        const uint64_t a0 = g.mode == 2 ? cs : cs + 4;
        if (off > len || a0 + (g.mode == 2 ? 5 : 0) > ce || a0 > ce || ce > len) {
            bad = true;
            break;
        }
        aend[t] = ce;
        const uint64_t nb = (t == g.T - 1) ? g.nblast : g.nb0;
        const uint64_t alen = ce - a0;
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

// LDS bitmap bits set with ds_or: a walk's next width-byte read does not wait
// for the word's old value.
struct BmAtomicOr {
    __device__ void operator()(uint32_t *w, uint32_t m) const { atomicOr(w, m); }
};

// Recovers the start of every SZp block of a stream that carries no index.
// Parse tile i of chunk t covers bytes [s, te) of the chunk's block area and
// is split into four sub-tiles of U bytes, one per wave.  The "entry" of a
// range is the first block starting in it, its "exit" the first block
// starting at or after its end.
//  1. Each wave finds one speculative chain through its sub-tile: its 64
//     lanes try 64 candidates c in [u, u + W) (W = longest block) at a time,
//     and the lowest candidate whose chain reaches the sub-tile end wins; its
//     block starts are marked.  Runs of zero bytes (b = 0 blocks) are
//     stepped over a word at a time.
//  2. For the last sub-tile only, every chain starting in [u, u + W) must
//     join the speculative chain, die on a width byte > 32 or leave at the
//     same exit (all threads check candidates): then the tile's exit is
//     known without its entry and is published before the predecessor tile
//     finishes.
//  3. One thread resolves the true chain: the tile's entry (the chunk's first
//     block, or the predecessor's published exit) is walked until it joins
//     the sub-tile's speculative chain, then the next sub-tile starts at that
//     chain's exit.
//  4. Block counts are scanned with a decoupled look-back (the whole
//     workgroup) over the chunk's tiles and every block's stream position is
//     written to blkpos[].
// A stream whose chains never synchronise is still parsed exactly, serially.
__global__ __launch_bounds__(kWG) void k_parse(const uint8_t *__restrict__ st, uint64_t len,
                                               Geom g, uint32_t U, uint32_t W,
                                               const uint64_t *__restrict__ ctab,
                                               uint64_t *__restrict__ exit_st,
                                               uint64_t *__restrict__ cnt_st,
                                               uint64_t *__restrict__ blkpos,
                                               uint32_t *__restrict__ ctl)
{
    extern __shared__ __align__(16) uint8_t dyn[];
    const uint32_t S = 4 * U, NW = S / 32;
    uint8_t *buf = dyn;
    uint32_t *spec = reinterpret_cast<uint32_t *>(dyn + ((S + 64 + 15) & ~15u));
    uint32_t *pre = spec + NW;
    uint16_t *rspec = reinterpret_cast<uint16_t *>(pre + NW);  // spec bits before word, per sub-tile
    uint16_t *rpre = rspec + NW;
    __shared__ uint64_t s_exit[4];
    __shared__ uint32_t s_cnt[4], s_npre[4], s_mrank[4];
    __shared__ int s_merged[4], s_ok[4], s_broken[4];
    __shared__ int s_nonsync;
    __shared__ uint64_t s_prefix, s_lastpos;
    __shared__ uint64_t s_lb[8];

    const int tid = threadIdx.x, lane = __lane_id(), wv = tid >> 6;
    const uint64_t gt = blockIdx.x;
    __builtin_amdgcn_s_setprio(2);  // until the tile's block count is published
    const uint64_t *cstart = ctab, *aend = ctab + g.T, *ttab = ctab + 2 * g.T;
    if (gt >= ttab[g.T]) return;
    uint32_t lo = 0, hi = g.T - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (ttab[mid] <= gt) lo = mid; else hi = mid - 1;
    }
    const uint32_t t = lo;
    const uint64_t first = ttab[t];
    const uint64_t i = gt - first;
    const uint32_t mode = g.mode;
    const uint64_t as = cstart[t] + (mode == 2 ? 0 : 4), ae = aend[t];
    const uint64_t s = as + i * S;
    const uint64_t te = min(s + S, ae);
    const uint32_t span = (uint32_t)(te - s);
    const uint64_t s0 = s & ~15ull;
    const uint32_t sh = (uint32_t)(s - s0);  // buf[p + sh] holds tile byte p
    const uint32_t B = g.B;

    // The tile's bytes and 16 more (a mode-2 unit starting in the last 4
    // bytes has its width byte beyond te; the walks' zero runs may look 3
    // bytes past a sub-tile).
    const uint64_t tl = min(te + 16, len);
    stage_bytes(st, len, s0, tl, buf);
    for (uint32_t k = tid; k < NW; k += kWG) {
        spec[k] = 0;
        pre[k] = 0;
    }
    if (tid == 0) {
        s_nonsync = 0;
        s_lastpos = ~0ull;
    }
    __syncthreads();
    const uint8_t *tb = buf + sh;  // tile byte p is tb[p]

    // 1. Speculative chain of each sub-tile (all lanes of wave w).
    const uint32_t u = wv * U, v = min(u + U, span);
    {
        bool ok = false;
        uint32_t cnt = 0, ex = v;
        const uint32_t cmax = min(u + W, v);
        for (uint32_t c0 = u; c0 < cmax; c0 += 64) {  // wave-uniform loop
            const uint32_t c = c0 + lane;
            uint32_t e = 0;
            const bool leaves = c < cmax && parse_chain_leaves(tb, c, v, B, e, mode);
            const uint64_t m = __ballot(leaves);
            if (m) {
                const int w = __ffsll((long long)m) - 1;  // the lowest candidate that leaves
                uint32_t n = 0;
                if (lane == w) n = parse_mark_chain(tb, spec, c, v, B, mode, BmAtomicOr());
                cnt = __builtin_amdgcn_readlane(n, w);
                ex = __builtin_amdgcn_readlane(e, w);
                ok = true;
                break;
            }
        }
        if (lane == 0) {
            s_ok[wv] = ok;
            s_cnt[wv] = ok ? cnt : 0;
            s_exit[wv] = s + (ok ? ex : v);
        }
    }
    __syncthreads();

    // 2. Does the last sub-tile's exit depend on its entry?
    const uint32_t lastw = span ? (span - 1) / U : 0;
    if (span) {  // all threads share the last sub-tile's candidates
        const uint32_t ul = lastw * U, vl = min(ul + U, span);
        if (!s_ok[lastw]) {
            if (tid == 0) s_nonsync = 1;
        } else {
            const uint32_t cmax = min(ul + W, vl);
            const uint64_t sx = s_exit[lastw];
            for (uint32_t c = ul + tid; c < cmax; c += kWG) {
                const uint32_t leave = parse_cand_walk(tb, spec, c, vl, B, mode);
                if (leave && s + leave != sx) s_nonsync = 1;
            }
        }
    }
    __syncthreads();
    if (tid == 0 && !s_nonsync && span) st_status(&exit_st[gt], kFlagIncl | s_exit[lastw]);

    // Spec ranks: spec bits before each word, within its sub-tile.
    {
        const uint32_t wpw = U / 32;  // words per sub-tile
        const uint32_t w0 = wv * wpw;
        uint32_t run = 0;
        for (uint32_t k0 = 0; k0 < wpw; k0 += 64) {  // every lane runs every step
            const uint32_t k = k0 + lane;
            const uint32_t c = k < wpw ? (uint32_t)__builtin_popcount(spec[w0 + k]) : 0u;
            const uint32_t incl = wave_incl_scan_u32(c, lane);
            if (k < wpw) rspec[w0 + k] = (uint16_t)(run + incl - c);
            run += __shfl(incl, 63);
        }
    }
    __syncthreads();

    // 3. True chain, sub-tile by sub-tile.
    if (tid == 0) {
        uint64_t e = as;
        if (i > 0) e = wait_status(&exit_st[gt - 1], ctl) & kValMask;
        bool broken = e < s;  // an exit before this tile: corrupt stream
        for (uint32_t w = 0; w < 4; ++w) {
            const uint32_t uw = w * U, vw = min(uw + U, span);
            s_npre[w] = 0;
            s_merged[w] = 0;
            s_mrank[w] = 0;
            s_broken[w] = broken;
            if (uw >= span || broken || e >= s + vw) continue;  // no block starts here
            uint32_t p = (uint32_t)(e - s), npre = 0;
            const int res = parse_true_walk(tb, spec, pre, p, vw, B, npre, mode, BmAtomicOr());
            s_npre[w] = npre;
            if (res == 1) {
                s_merged[w] = 1;
                s_mrank[w] = parse_rank(spec, rspec, p);
                e = s_ok[w] ? s_exit[w] : s + vw;
            } else if (res == 2) {
                broken = true;
                s_broken[w] = 1;
            } else {
                e = s + p;
            }
        }
        if (s_nonsync && span) st_status(&exit_st[gt], kFlagIncl | (broken ? te : e));
        uint32_t total = 0;
        for (uint32_t w = 0; w < 4; ++w) {
            const uint32_t c = s_npre[w] + (s_merged[w] ? s_cnt[w] - s_mrank[w] : 0u);
            s_cnt[w] = c;  // blocks of sub-tile w on the true chain
            total += c;
        }
        s_prefix = total;
    }
    __syncthreads();
    const uint32_t tile_cnt = (uint32_t)s_prefix;
    __syncthreads();
    {
        if (tid < 64) lookback_publish(cnt_st, gt, first, tile_cnt);
        __builtin_amdgcn_s_setprio(0);
        const uint64_t pre_cnt = lookback_sum_wg(cnt_st, gt, first, tile_cnt, ctl, true, s_lb);
        if (tid == 0) s_prefix = pre_cnt;
    }
    // Pre ranks: pre bits before each word, within its sub-tile.
    {
        const uint32_t wpw = U / 32;
        const uint32_t w0 = wv * wpw;
        uint32_t run = 0;
        for (uint32_t k0 = 0; k0 < wpw; k0 += 64) {
            const uint32_t k = k0 + lane;
            const uint32_t c = k < wpw ? (uint32_t)__builtin_popcount(pre[w0 + k]) : 0u;
            const uint32_t incl = wave_incl_scan_u32(c, lane);
            if (k < wpw) rpre[w0 + k] = (uint16_t)(run + incl - c);
            run += __shfl(incl, 63);
        }
    }
    __syncthreads();

    // 4. Block positions.
    const uint64_t prefix = s_prefix;
    const uint64_t nbt = (t == g.T - 1) ? g.nblast : g.nb0;
    const uint64_t gbase = (uint64_t)t * g.nb0;
    uint64_t wbase[4];
    {
        uint64_t acc = prefix;
        for (int w = 0; w < 4; ++w) {
            wbase[w] = acc;
            acc += s_cnt[w];
        }
    }
    for (uint32_t k = tid; k < NW; k += kWG) {
        const uint32_t w = (k * 32) / U;
        uint32_t m = pre[k];
        while (m) {
            const uint32_t bit = __builtin_ctz(m);
            m &= m - 1;
            const uint32_t p = k * 32 + bit;
            const uint64_t bi = wbase[w] + parse_rank(pre, rpre, p);
            if (bi < nbt) {
                blkpos[gbase + bi] = s + p;
                if (bi == nbt - 1) s_lastpos = s + p;
            }
        }
        if (s_merged[w]) {
            m = spec[k];
            while (m) {
                const uint32_t bit = __builtin_ctz(m);
                m &= m - 1;
                const uint32_t p = k * 32 + bit;
                const uint32_t rk = parse_rank(spec, rspec, p);
                if (rk < s_mrank[w]) continue;
                const uint64_t bi = wbase[w] + s_npre[w] + (rk - s_mrank[w]);
                if (bi < nbt) {
                    blkpos[gbase + bi] = s + p;
                    if (bi == nbt - 1) s_lastpos = s + p;
                }
            }
        }
    }
    __syncthreads();
    if (tid == 0) {
        bool bad = false;
        for (uint32_t w = 0; w < 4; ++w)
            if (s_broken[w] && wbase[w] + s_npre[w] < nbt) bad = true;
        if (gt + 1 == ttab[t + 1] && prefix + tile_cnt < nbt) bad = true;  // truncated chunk
        if (s_lastpos != ~0ull) {
            // The last block is shorter when (n_t - 1) % B != 0 (mode 2: its
            // unit when n_t % B != 0); its end must be the chunk end (exact
            // for all but the last chunk).
            const uint64_t pl = s_lastpos;
            uint64_t end;
            uint32_t b;
            if (mode == 2) {
                const uint64_t Ll = chunk_len(g, t) - (nbt - 1) * (uint64_t)g.B - 1;
                b = pl + 4 < len ? st[pl + 4] : 255u;
                end = pl + 4 + block_bytes(b, (uint32_t)Ll);
                if (!Ll && pl + 4 < len) b = 0;  // a 1-value unit's width byte is not read
            } else {
                const uint64_t Ll = chunk_len(g, t) - 1 - (nbt - 1) * (uint64_t)g.B;
                b = st[pl];
                end = pl + block_bytes(b, (uint32_t)Ll);
            }
            if (b > 32 || (t + 1 < g.T ? end != ae : end > len)) bad = true;
            if (t + 1 == g.T) ctl[4] = (uint32_t)end, ctl[5] = (uint32_t)(end >> 32);
        }
        if (bad) raise_error(ctl, kErrStream);
    }
}

// Writes one thread's eight output values.  NT: non-temporal stores, so
// that an output far larger than the memory-side cache does not leave it
// full of dirty lines for the next pass over the field.  Measured on the
// 512^3 field: the look-back decode 330 -> 286 us for float (the float64
// foreign-stream decode was slower with them: 1.28 -> 0.83 TB/s, so float64
// keeps ordinary stores); the pipelined decode, which waits for its stores
// before staging the next tile, 161 -> 218 us, so it keeps ordinary stores.
template <bool NT, typename V>
__device__ __forceinline__ void store_octet_out(V *p, const V y[8])
{
    if (NT) {
        // Octets start at any 4-byte boundary (blocks start at a chunk's
        // element lo + 1): a vector type of 4-byte alignment keeps the access
        // well defined.  (16-byte-aligned non-temporal vectors shifted over
        // DPP were measured slower and wrote more: 1.83x the output bytes
        // against 1.28x, 355 against 285 us on a 512^3 decode.)
        typedef uint32_t u4 __attribute__((ext_vector_type(4), aligned(4)));
        constexpr int nv = (int)sizeof(V) / 2;  // 16-byte pieces
        u4 w[nv];
        __builtin_memcpy(w, y, 8 * sizeof(V));
        u4 *q = reinterpret_cast<u4 *>(p);
#pragma unroll
        for (int k = 0; k < nv; ++k) __builtin_nontemporal_store(w[k], q + k);
    } else {
        __builtin_memcpy(p, y, 8 * sizeof(V));
    }
}

// Writes tot bytes (a multiple of 4) from LDS src to dst[0, tot) with
// non-temporal stores: 16-byte stores for the aligned middle, each store
// instruction of a wave covering 1 KiB of whole lines, and 4-byte stores for
// the ragged ends, which share lines with the neighbouring tiles.
__device__ __forceinline__ void copy_out_nt(uint8_t *__restrict__ dst, uint32_t tot, const uint8_t *src,
                                            int tid)
{
    const uintptr_t G0 = (uintptr_t)dst, G1 = G0 + tot;
    const uintptr_t a0 = (G0 + 15) & ~(uintptr_t)15, a1 = G1 & ~(uintptr_t)15;
    const uint32_t head = (uint32_t)(a0 - G0) / 4, tail = (uint32_t)(G1 - a1) / 4;
    if ((uint32_t)tid < head) {
        uint32_t w;
        __builtin_memcpy(&w, src + 4 * tid, 4);
        __builtin_nontemporal_store(w, reinterpret_cast<uint32_t *>(dst) + tid);
    } else if ((uint32_t)tid >= 4 && (uint32_t)tid < 4 + tail) {
        const uint32_t k = (uint32_t)(a1 - G0) / 4 + (uint32_t)tid - 4;
        uint32_t w;
        __builtin_memcpy(&w, src + 4 * k, 4);
        __builtin_nontemporal_store(w, reinterpret_cast<uint32_t *>(dst) + k);
    }
    typedef uint32_t u4 __attribute__((ext_vector_type(4)));
    const uint32_t nv = (uint32_t)((a1 - a0) >> 4);
    const uint8_t *s = src + (a0 - G0);
    u4 *d = reinterpret_cast<u4 *>(dst + (a0 - G0));
#pragma unroll 4
    for (uint32_t v = tid; v < nv; v += kWG) {
        u4 val;
        __builtin_memcpy(&val, s + 16 * v, 16);  // any 4-byte LDS offset
        __builtin_nontemporal_store(val, d + v);
    }
}

// Unpacks one slot (SZp/src/szpd_float.cc:81-110, SZp/src/szp_TypeManager.cc:436-944).
// The slot's item starts at buf[lp]; lend bounds the bytes present (reads may
// run up to 32 bytes past it, inside the buffer's slack).  dl[i] = inclusive
// prefix of the slot's deltas, or the absolute head value in dl[0] for a head
// slot (h = 1).  Returns the number of values (0 for idle slots); sets bad on
// an invalid width or a block that overruns the bytes present.  FULL: the
// caller guarantees a full octet (8 * o + 8 <= L), so no value is masked.
template <bool FULL = false>
__device__ __forceinline__ uint32_t decode_slot(uint32_t L, bool head, uint32_t o,
                                                const uint8_t *buf, uint32_t lp, uint32_t lend,
                                                uint32_t dl[8], uint32_t &h, bool &bad)
{
    h = 0;
    if (!FULL) {
#pragma unroll
        for (int i = 0; i < 8; ++i) dl[i] = 0;
        if (head) {
            if (o) return 0;
            h = 1;
            __builtin_memcpy(&dl[0], buf + lp, 4);
            return 1;
        }
        if (8 * o >= L) return 0;
    }
    const uint32_t m = FULL ? 8u : min(8u, L - 8 * o);
    // A bad width or a block past the bytes present is reported and decoded
    // as width 0; every path below writes all of dl (no early exit, so the
    // compiler need not clear dl at each branch).
    const uint32_t b0 = buf[lp];
    const bool okb = b0 <= 32 && lp + block_bytes(b0, L) <= lend;
    if (!okb) bad = true;
    const uint32_t b = okb ? b0 : 0u;
    const uint32_t S8 = (L + 7) >> 3;
    const uint32_t sg = buf[lp + 1 + o];
    const uint32_t nb = b >> 3, r = b & 7;
    const uint8_t *p1 = buf + lp + 1 + S8 + 8 * o * nb;
    uint32_t v[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) v[i] = 0;
    if (nb == 1) {
        const uint2 w = lds_ld64(p1);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            v[i] = (w.x >> (8 * i)) & 0xFFu;
            v[i + 4] = (w.y >> (8 * i)) & 0xFFu;
        }
    } else if (nb == 2) {
        const uint4 w = lds_ld128(p1);
        v[0] = w.x & 0xFFFFu; v[1] = w.x >> 16;
        v[2] = w.y & 0xFFFFu; v[3] = w.y >> 16;
        v[4] = w.z & 0xFFFFu; v[5] = w.z >> 16;
        v[6] = w.w & 0xFFFFu; v[7] = w.w >> 16;
    } else if (nb == 3) {
        const uint4 a = lds_ld128(p1);
        const uint2 c = lds_ld64(p1 + 16);
        v[0] = a.x & 0xFFFFFFu;
        v[1] = __builtin_amdgcn_alignbyte(a.y, a.x, 3) & 0xFFFFFFu;
        v[2] = __builtin_amdgcn_alignbyte(a.z, a.y, 2) & 0xFFFFFFu;
        v[3] = a.z >> 8;
        v[4] = a.w & 0xFFFFFFu;
        v[5] = __builtin_amdgcn_alignbyte(c.x, a.w, 3) & 0xFFFFFFu;
        v[6] = __builtin_amdgcn_alignbyte(c.y, c.x, 2) & 0xFFFFFFu;
        v[7] = c.y >> 8;
    } else if (nb == 4) {
        const uint4 a = lds_ld128(p1), c = lds_ld128(p1 + 16);
        v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
        v[4] = c.x; v[5] = c.y; v[6] = c.z; v[7] = c.w;
    }
    uint64_t A = 0;
    if (r) {
        const uint2 w = lds_ld64(buf + lp + 1 + S8 + nb * L + o * r);
        A = __builtin_bswap64(((uint64_t)w.y << 32) | w.x);  // stream bits MSB-first
    }
    const uint32_t mask = (1u << r) - 1;
    uint32_t run = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        // With r == 0 both A and mask are 0, so the shift amount (mod 64) is moot.
        const uint32_t low = (uint32_t)(A >> ((64u - r * (i + 1)) & 63u)) & mask;
        const uint32_t mag = (v[i] << r) | low;
        // s = ~0 when sign bit 7 - i is set; (mag ^ s) - s is then 0 - mag.
        const uint32_t s = (uint32_t)((int32_t)(sg << (24 + i)) >> 31);
        if (FULL || (uint32_t)i < m) run += (mag ^ s) - s;
        dl[i] = run;
    }
    return m;
}

// Segmented scan of slot aggregates (h, v) in slot order: a head restarts
// the running value.  Returns the slot's exclusive (eh, ev) and the round's
// total (rh, rv).
__device__ __forceinline__ void seg_scan_round(uint32_t h, uint32_t v, uint32_t *s_wh,
                                               uint32_t *s_wv, uint32_t &eh, uint32_t &ev,
                                               uint32_t &rh, uint32_t &rv)
{
    const int lane = __lane_id(), wv = threadIdx.x >> 6;
#pragma unroll
    for (int dd = 1; dd < 64; dd <<= 1) {
        const uint32_t ph = __shfl_up(h, dd), pv = __shfl_up(v, dd);
        if (lane >= dd) {
            v = h ? v : pv + v;
            h |= ph;
        }
    }
    if (lane == 63) {
        s_wh[wv] = h;
        s_wv[wv] = v;
    }
    __syncthreads();
    eh = __shfl_up(h, 1);
    ev = __shfl_up(v, 1);
    if (lane == 0) {
        eh = 0;
        ev = 0;
    }
    uint32_t wh = 0, wvv = 0;
    for (int k = 0; k < wv; ++k) {
        wvv = s_wh[k] ? s_wv[k] : wvv + s_wv[k];
        wh |= s_wh[k];
    }
    ev = eh ? ev : wvv + ev;
    eh |= wh;
    rh = 0;
    rv = 0;
    for (int k = 0; k < kWG / 64; ++k) {
        rv = s_wh[k] ? s_wv[k] : rv + s_wv[k];
        rh |= s_wh[k];
    }
    __syncthreads();
}

// Copies stream bytes [a, b) to buf[0, b - a) (16-byte loads, 16-byte aligned
// LDS destination offset base_sh = a & 15 so buf[p - (a & ~15)] holds p).
// Copies st[a & ~15, b) into buf, 16 bytes per thread per step (decode
// tiles; the staged form above measured slower here: 287 -> 304 us on the
// look-back decode, whose tiles need about three vectors per thread).
__device__ __forceinline__ void load_bytes(const uint8_t *__restrict__ st, uint64_t len,
                                           uint64_t a, uint64_t b, uint8_t *buf)
{
    const uint64_t a0 = a & ~15ull;
    for (uint64_t q = a0 + 16ull * threadIdx.x; q < b; q += 16ull * kWG) {
        if (q + 16 <= len) {
            *reinterpret_cast<uint4 *>(buf + (q - a0)) = *reinterpret_cast<const uint4 *>(st + q);
        } else {
            for (int k = 0; k < 16; ++k)
                if (q + k < len) buf[q - a0 + k] = st[q + k];
        }
    }
}

// Fast path of k_decode for a simple tile (full blocks of one chunk, SPB =
// B / 8 a power of two up to 32), the mirror of compress_simple_fast: thread
// tid decodes elements r * 2048 + 8 * tid + [0, 8) of every round r.  The
// block positions of the thread's items and the tile's byte range (its first
// block and the item after it) are loaded together; the tile's bytes go to
// LDS once; all rounds are unpacked into registers; one wave scan per round
// and one barrier give every octet its Lorenzo prefix; the tile's carry comes
// from the side index or from a look-back over the tiles' delta sums; values
// fl32((float)q * eb) are stored straight to HBM.  Tiles whose bytes exceed
// the LDS buffer are read one round at a time.  Returns false, before any
// side effect, for the field's last tile (its end is not in the index); the
// caller then takes the general path.
template <typename V, int SPB, bool INDEXED>
__device__ __forceinline__ bool decode_simple_fast(
    const uint8_t *__restrict__ st, uint64_t len, const Geom &g,
    const uint64_t *__restrict__ blkpos, const uint32_t *__restrict__ idx_carry,
    V *__restrict__ out, uint64_t *status, uint32_t *ctl, uint64_t tile, uint64_t tIN,
    uint32_t ct, uint64_t cj, uint32_t nti, V eb, uint8_t *buf, uint32_t *s_wsum,
    uint64_t *s_lb, int *s_bad)
{
    constexpr uint32_t K = 256 / SPB, B = 8 * SPB;
    const int tid = threadIdx.x, lane = __lane_id(), wv = tid >> 6;
    const uint32_t d = (uint32_t)tid / SPB, o = (uint32_t)tid % SPB;
    if (tIN >= g.nitems) return false;
    // Not the field's last tile, so the tile holds whole rounds (nti = K * R)
    // and the round tests below are uniform.
    // The side index's carry is loaded with the positions, not after the
    // unpack, so that its latency is not added to the tile's.
    const uint32_t icarry = INDEXED ? idx_carry[tile] : 0u;
    const uint64_t pb = (uint64_t)ct * g.nb0 + (cj - 1);  // blkpos index of the first block
    uint64_t pos[kMaxRounds];
#pragma unroll
    for (int r = 0; r < kMaxRounds; ++r)
        pos[r] = ((uint32_t)r * K < nti) ? blkpos[pb + r * K + d] : 0;
    const uint64_t P0 = blkpos[pb];
    uint32_t t1;
    uint64_t j1;
    locate(g, tIN, t1, j1);
    const uint64_t P1 = j1 ? blkpos[(uint64_t)t1 * g.nb0 + (j1 - 1)]
                           : g.body + read_u64_le4(st + g.es + 8ull * t1);
    if (tid == 0) *s_bad = 0;
    bool bad = !(P0 < P1 && P1 <= len);
    uint32_t dl[kMaxRounds][8];
#pragma unroll
    for (int r = 0; r < kMaxRounds; ++r)
#pragma unroll
        for (int i = 0; i < 8; ++i) dl[r][i] = 0;
    if (!bad && P1 - (P0 & ~15ull) <= (uint64_t)kDecTileBytes) {
        // The whole tile's bytes at once.
        const uint64_t tbase = P0 & ~15ull;
        load_bytes(st, len, P0, P1, buf);
        __syncthreads();
        const uint32_t lend = (uint32_t)(P1 - tbase);
#pragma unroll
        for (int r = 0; r < kMaxRounds; ++r) {
            if ((uint32_t)r * K < nti) {
                uint32_t h;
                if (pos[r] < P0 || pos[r] >= P1)
                    bad = true;
                else
                    decode_slot<true>(B, false, o, buf, (uint32_t)(pos[r] - tbase), lend, dl[r], h,
                                      bad);
            }
        }
    } else if (!bad) {
        // A tile too large for the buffer (barely compressible data): one
        // round's bytes at a time (a round holds at most K * block_bytes(32,
        // B) bytes, well below the buffer).
#pragma unroll
        for (int r = 0; r < kMaxRounds; ++r) {
            if ((uint32_t)r * K >= nti) break;
            const uint64_t Pr = blkpos[pb + r * K];
            const uint64_t Pn = (uint32_t)(r + 1) * K < nti ? blkpos[pb + (r + 1) * K] : P1;
            const uint64_t rbase = Pr & ~15ull;
            const bool ok = Pr < Pn && Pn <= P1 && Pn - rbase <= (uint64_t)kDecTileBytes;
            __syncthreads();  // the previous round is done with buf
            if (ok) load_bytes(st, len, Pr, Pn, buf);
            __syncthreads();
            if (!ok) {
                bad = true;
            } else if ((uint32_t)r * K < nti) {
                uint32_t h;
                if (pos[r] < Pr || pos[r] >= Pn)
                    bad = true;
                else
                    decode_slot<true>(B, false, o, buf, (uint32_t)(pos[r] - rbase), (uint32_t)(Pn - rbase),
                                dl[r], h, bad);
            }
        }
    }
    if (bad) *s_bad = 1;
    uint32_t ex[kMaxRounds];
#pragma unroll
    for (int r = 0; r < kMaxRounds; ++r) {
        const uint32_t v = dl[r][7];
        const uint32_t incl = wave_incl_scan_u32(v, lane);
        ex[r] = incl - v;
        const uint32_t wt = __builtin_amdgcn_readlane(incl, 63);
        if (lane == 0) s_wsum[r * 4 + wv] = wt;
    }
    __syncthreads();
    uint32_t agg = 0;
#pragma unroll
    for (int k = 0; k < kMaxRounds * 4; ++k) agg += s_wsum[k];
    if (*s_bad) {
        if (!INDEXED && tid == 0) publish_abort_d(status, g, tile);
        if (tid == 0) raise_error(ctl, kErrStream);
        return true;
    }
    uint32_t carry;
    if (INDEXED) {
        carry = icarry;
    } else {
        // The tile holds no chunk head: its aggregate is a plain delta sum.
        if (tid < 64) lookback_publish_d(status, g, tile, kFlagAgg | agg, ctl);
        __builtin_amdgcn_s_setprio(0);
        carry = lookback_decode(status, g, tile, agg, ctl, s_lb);
    }
    const uint64_t Et = item_elem(g, ct, cj);
    uint32_t base = carry;
    if (sizeof(V) == 4) {
        // Float: two rounds at a time through LDS (buf is free: every slot
        // was unpacked before the barrier above), then whole-line
        // non-temporal stores.  A round spans 2048 values (8 KiB).
        static_assert(2 * 2048 * 4 <= kDecTileBytes, "two rounds of values fit the tile buffer");
#pragma unroll
        for (int r = 0; r < kMaxRounds; ++r) {
            uint32_t before = base;
#pragma unroll
            for (int w = 0; w < 4; ++w) before += w < wv ? s_wsum[r * 4 + w] : 0u;
#pragma unroll
            for (int w = 0; w < 4; ++w) base += s_wsum[r * 4 + w];
            if (r % 2 == 0) __syncthreads();  // buf: unpacked, or the previous pair written out
            if ((uint32_t)r * K < nti) {
                const uint32_t q0 = before + ex[r];
                V y[8];
#pragma unroll
                for (int i = 0; i < 8; ++i) y[i] = (V)(int32_t)(q0 + dl[r][i]) * eb;
                __builtin_memcpy(buf + (r % 2) * 8192 + 32 * tid, y, 32);
            }
            if (r % 2 == 1 || r + 1 == kMaxRounds) {
                const int r0 = r - r % 2;
                const uint32_t nr = min<uint32_t>((uint32_t)(r - r0 + 1),
                                                  (uint32_t)max<int64_t>(0, (int64_t)nti / K - r0));
                __syncthreads();
                if (nr)
                    copy_out_nt(reinterpret_cast<uint8_t *>(out + Et + (uint64_t)r0 * 2048u), nr * 8192u,
                                buf, tid);
            }
        }
        return true;
    }
#pragma unroll
    for (int r = 0; r < kMaxRounds; ++r) {
        uint32_t before = base;
#pragma unroll
        for (int w = 0; w < 4; ++w) before += w < wv ? s_wsum[r * 4 + w] : 0u;
#pragma unroll
        for (int w = 0; w < 4; ++w) base += s_wsum[r * 4 + w];
        if ((uint32_t)r * K < nti) {
            const uint32_t q0 = before + ex[r];
            V y[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) y[i] = (V)(int32_t)(q0 + dl[r][i]) * eb;
            store_octet_out<false>(out + Et + (uint64_t)r * 2048u + 8u * (uint32_t)tid, y);
        }
    }
    return true;
}

// Fast path of k_decode for a fastMode-2 tile of whole units (unit_tile),
// the mirror of compress_units_fast: the tile's bytes go to LDS once; thread
// tid decodes block indices 8o .. 8o + 7 of its unit (o = tid % SPB; the
// unit's last lane has 7), a scan within the unit's SPB lanes gives the
// deltas before them, and the thread writes unit values 8o .. 8o + 7: value
// 8o is the head plus those deltas, value 8o + k the same plus its own first
// k deltas.  Units restart the Lorenzo chain, so no carry or look-back is
// needed.  Returns false, before any side effect, for the field's last tile
// (its end is not in the positions); the caller then takes the general path.
template <typename V, int SPB>
__device__ __forceinline__ bool decode_units_fast(
    const uint8_t *__restrict__ st, uint64_t len, const Geom &g,
    const uint64_t *__restrict__ blkpos, V *__restrict__ out, uint32_t *ctl, uint64_t tIN,
    uint32_t ct, uint64_t cj, uint32_t nti, V eb, uint8_t *buf, int *s_bad)
{
    constexpr uint32_t K = 256 / SPB, B = 8 * SPB;
    constexpr int kRounds = kMaxRounds / 2;
    const int tid = threadIdx.x, lane = __lane_id();
    const uint32_t d = (uint32_t)tid / SPB, o = (uint32_t)tid % SPB;
    const uint32_t nu = nti / 2;
    if (tIN >= g.nitems) return false;
    const uint64_t pb = (uint64_t)ct * g.nb0 + cj / 2;  // blkpos index of the tile's first unit
    uint64_t pos[kRounds];
#pragma unroll
    for (int r = 0; r < kRounds; ++r) pos[r] = (r * K + d < nu) ? blkpos[pb + r * K + d] : 0;
    const uint64_t P0 = blkpos[pb];
    uint32_t t1;
    uint64_t j1;
    locate(g, tIN, t1, j1);
    // The tile ends where the next unit (or the next chunk) starts.
    const uint64_t P1 = j1 ? blkpos[(uint64_t)t1 * g.nb0 + (j1 >> 1)]
                           : g.body + read_u64_le4(st + g.es + 8ull * t1);
    if (tid == 0) *s_bad = 0;
    const uint64_t tbase = P0 & ~15ull;
    bool bad = !(P0 < P1 && P1 <= len && P1 - tbase <= (uint64_t)kDecTileBytes);
    if (!bad) load_bytes(st, len, P0, P1, buf);
    __syncthreads();
    const uint32_t lend = bad ? 0u : (uint32_t)(P1 - tbase);
    uint32_t dl[kRounds][8];
    int32_t hq[kRounds];
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
        hq[r] = 0;
#pragma unroll
        for (int i = 0; i < 8; ++i) dl[r][i] = 0;
        if (r * K + d < nu && !bad) {
            if (pos[r] < P0 || pos[r] + 5 > P1) {
                bad = true;
            } else {
                const uint32_t lp = (uint32_t)(pos[r] - tbase);
                uint32_t hv, h;
                __builtin_memcpy(&hv, buf + lp, 4);
                hq[r] = (int32_t)hv;
                decode_slot(B - 1, false, o, buf, lp + 4, lend, dl[r], h, bad);
            }
        }
    }
    if (bad) *s_bad = 1;
    __syncthreads();
    if (*s_bad) {
        if (tid == 0) raise_error(ctl, kErrStream);
        return true;
    }
    const uint64_t Et = item_elem(g, ct, cj);
    const int gs = lane - (int)o;  // the unit's first lane
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
        const uint32_t v = dl[r][7];
        const uint32_t ex = wave_incl_scan_u32(v, lane) - v;
        const uint32_t before = ex - (uint32_t)__shfl((int)ex, gs);  // deltas of the unit's earlier lanes
        const uint32_t q0 = (uint32_t)__shfl(hq[r], gs) + before;
        if (r * K + d < nu) {
            V y[8];
            y[0] = (V)(int32_t)q0 * eb;
#pragma unroll
            for (int i = 1; i < 8; ++i) y[i] = (V)(int32_t)(q0 + dl[r][i - 1]) * eb;
            store_octet_out<sizeof(V) == 4>(out + Et + (uint64_t)r * 2048u + 8u * (uint32_t)tid, y);
        }
    }
    return true;
}

// Decodes tiles of g.R rounds of K items.  The item positions of the tile
// (chunk heads from the header, blocks from the parse) and, when they fit,
// all its stream bytes are loaded once.  Pass 1 unpacks every round to get
// its segmented delta sum; one look-back gives the tile's carry (the q of the
// element before the tile); pass 2 unpacks again and writes
// fl32((float)q * eb).
//
// With INDEXED, blkpos and idx_carry come from the compressor's side index:
// the carry is known up front, so the look-back and pass 1 are skipped.
// LDS of one decode workgroup.
struct DecodeLds {
    __align__(16) uint8_t buf[kDecTileBytes + 64];
    uint32_t s_pos[kWG * kMaxRounds];  // item positions - tile base
    uint32_t s_wh[kWG / 64], s_wv[kWG / 64];
    uint64_t s_base, s_end;
    uint32_t s_carry;
    int s_bad;
    uint32_t s_wsum[kMaxRounds * 4];
    uint64_t s_lb[8];
};

// One decode tile (all threads of the workgroup call it; every exit is
// uniform).
template <typename V, bool INDEXED, int SPB>
__device__ __forceinline__ void decode_tile(const uint8_t *__restrict__ st, uint64_t len, const Geom &g,
                            const uint64_t *__restrict__ blkpos,
                            const uint32_t *__restrict__ idx_carry, V *__restrict__ out,
                            uint64_t *__restrict__ status, uint32_t *__restrict__ ctl,
                            uint64_t tile, DecodeLds &L)
{
    uint8_t *buf = L.buf;
    uint32_t *s_pos = L.s_pos, *s_wh = L.s_wh, *s_wv = L.s_wv;
    uint64_t &s_base = L.s_base, &s_end = L.s_end;
    uint32_t &s_carry = L.s_carry;
    int &s_bad = L.s_bad;

    // Foreign streams: raised issue priority until the tile's delta sum is
    // published, as in the compress kernel (successors' look-backs wait for
    // it).
    if (!INDEXED) __builtin_amdgcn_s_setprio(2);
    const int tid = threadIdx.x, lane = __lane_id();
    const V eb = get_eb<V>(st);
    const uint64_t tI0 = tile * (uint64_t)g.K * g.R;
    const uint64_t tIN = min(tI0 + (uint64_t)g.K * g.R, g.nitems);
    const uint32_t nti = (uint32_t)(tIN - tI0);
    uint32_t t0;
    uint64_t j0;
    locate(g, tI0, t0, j0);
    if (SPB && simple_tile(g, t0, j0, nti) &&
        decode_simple_fast<V, (SPB ? SPB : 1), INDEXED>(st, len, g, blkpos, idx_carry, out, status,
                                                     ctl, tile, tIN, t0, j0, nti, eb, buf,
                                                     L.s_wsum, L.s_lb, &s_bad))
        return;
    if (SPB && unit_tile(g, t0, j0, nti) &&
        decode_units_fast<V, (SPB ? SPB : 1)>(st, len, g, blkpos, out, ctl, tIN, t0, j0, nti, eb,
                                              buf, &s_bad))
        return;

    // Item positions of the whole tile.
    uint64_t mypos[kMaxRounds];
#pragma unroll
    for (int k = 0; k < kMaxRounds; ++k) {
        const uint32_t it = tid + kWG * k;
        mypos[k] = 0;
        if (it < nti) {
            uint32_t t = t0;
            uint64_t j = j0;
            advance(g, t, j, it);
            // Chunk heads sit at the offsets of the stream header; a fastMode-2
            // block follows its unit's 4-byte head.
            if (g.mode == 2)
                mypos[k] = blkpos[(uint64_t)t * g.nb0 + (j >> 1)] + ((j & 1) ? 4u : 0u);
            else
                mypos[k] = j == 0 ? g.body + read_u64_le4(st + g.es + 8ull * t)
                                  : blkpos[(uint64_t)t * g.nb0 + (j - 1)];
            if (it == 0) s_base = mypos[k];
            if (it == nti - 1) {
                uint64_t e;
                if (item_head(g, j)) {
                    e = mypos[k] + 4;
                } else {
                    const uint32_t b = mypos[k] < len ? st[mypos[k]] : 255u;
                    e = b > 32 ? ~0ull : mypos[k] + block_bytes(b, item_len(g, t, j));
                }
                s_end = e;
            }
        }
    }
    __syncthreads();
    const uint64_t tbase = s_base & ~15ull, tend = s_end;
    const bool whole = tend <= len && s_base <= tend && tend - tbase <= (uint64_t)kDecTileBytes;
    bool bad = !(tend <= len && s_base <= tend);
#pragma unroll
    for (int k = 0; k < kMaxRounds; ++k) {
        const uint32_t it = tid + kWG * k;
        if (it < nti) {
            if (mypos[k] < tbase || mypos[k] - tbase >= (1ull << 32)) bad = true;
            s_pos[it] = (uint32_t)(mypos[k] - tbase);
        }
    }
    if (whole) load_bytes(st, len, s_base, tend, buf);
    if (tid == 0) {
        s_bad = 0;
        if (INDEXED) s_carry = idx_carry[tile];
    }
    __syncthreads();
    if (bad) s_bad = 1;
    __syncthreads();

    // Slot -> item within a round.
    const uint32_t d = ((uint32_t)tid * g.spb_magic) >> 16;
    const uint32_t o = (uint32_t)tid - d * g.spb;

    for (int pass = INDEXED ? 1 : 0; pass < 2 && !s_bad; ++pass) {
        uint32_t th = 0, tv = 0;
        uint32_t carry = 0;
        if (pass == 1) carry = s_carry;
        for (uint32_t r = 0; r < g.R; ++r) {
            const uint64_t ri0 = tI0 + (uint64_t)r * g.K;
            if (ri0 >= tIN) break;
            const uint64_t riN = min(ri0 + g.K, tIN);
            const uint32_t rfirst = r * g.K, rn = (uint32_t)(riN - ri0);
            // Bytes of this round: all of the tile, or this round alone.
            uint32_t rb = 0, rend = (uint32_t)(tend - tbase);
            if (!whole) {
                const uint64_t a = tbase + s_pos[rfirst];
                const uint64_t b = riN < tIN ? tbase + s_pos[rfirst + rn] : tend;
                if (b - (a & ~15ull) > (uint64_t)kDecTileBytes) {
                    if (tid == 0) s_bad = 1;
                    break;
                }
                __syncthreads();
                load_bytes(st, len, a, b, buf);
                __syncthreads();
                rb = (uint32_t)((a & ~15ull) - tbase);
                rend = (uint32_t)(b - tbase);
            }
            const bool active = d < g.K && d < rn;
            uint32_t t = t0, L = 0;
            uint64_t j = j0, es = 0;
            if (active) {
                advance(g, t, j, rfirst + d);
                L = item_len(g, t, j);
                es = item_elem(g, t, j);
            }
            uint32_t dl[8], h = 0, m = 0;
            bool b2 = false;
            if (active)
                m = decode_slot(L, item_head(g, j), o, buf, s_pos[rfirst + d] - rb, rend - rb, dl, h,
                                b2);
            if (b2) s_bad = 1;
            const uint32_t v = m ? (h ? dl[0] : dl[m - 1]) : 0u;
            uint32_t eh, ev, rh, rv;
            seg_scan_round(h, v, s_wh, s_wv, eh, ev, rh, rv);
            if (pass == 1 && m) {
                // carry (+) tile prefix of the round (+) slot prefix in the round
                uint32_t before = th ? tv : carry + tv;
                before = eh ? ev : before + ev;
                V *dst = out + es + (h ? 0u : 8u * o);
                if (h) {
                    dst[0] = (V)(int32_t)dl[0] * eb;
                } else if (m == 8) {
                    V y[8];
#pragma unroll
                    for (int i = 0; i < 8; ++i) y[i] = (V)(int32_t)(before + dl[i]) * eb;
                    __builtin_memcpy(dst, y, 8 * sizeof(V));
                } else {
#pragma unroll
                    for (int i = 0; i < 8; ++i)
                        if ((uint32_t)i < m) dst[i] = (V)(int32_t)(before + dl[i]) * eb;
                }
            }
            tv = rh ? rv : tv + rv;
            th |= rh;
        }
        if (pass == 1) break;
        __syncthreads();
        if (s_bad) break;
        // Look-back: a tile whose aggregate contains a head is absolute by itself.
        if (tid < 64) {
            const bool lb = tile > 0 && !item_head(g, j0);
            lookback_publish_d(status, g, tile, ((th || !lb) ? kFlagIncl : kFlagAgg) | tv, ctl);
            uint32_t c = 0;
            if (lb) {
                int64_t hi = (int64_t)tile - 1;
                while (true) {
                    const int64_t idx = hi - lane;
                    const uint64_t sx = idx >= 0 ? wait_status(&status[idx], ctl) : kFlagIncl;
                    const uint64_t stop = __ballot((sx >> 62) == 2);
                    if (stop) {
                        const int f = __ffsll((long long)stop) - 1;
                        c += wave_sum_u32(lane <= f ? (uint32_t)sx : 0u);
                        break;
                    }
                    c += wave_sum_u32((uint32_t)sx);
                    hi -= 64;
                }
                if (!th && lane == 0) lookback_finish_d(status, g, tile, c + tv);
            }
            if (lane == 0) s_carry = c;
        }
        __syncthreads();
    }
    __syncthreads();
    if (s_bad && tid == 0) {
        // Publish (harmlessly, if already published) so no successor waits.
        if (!INDEXED) publish_abort_d(status, g, tile);
        raise_error(ctl, kErrStream);
    }
}

template <typename V, bool INDEXED, int SPB>
__global__ __launch_bounds__(kWG) void k_decode(const uint8_t *__restrict__ st, uint64_t len,
                                                Geom g, const uint64_t *__restrict__ blkpos,
                                                const uint32_t *__restrict__ idx_carry,
                                                V *__restrict__ out,
                                                uint64_t *__restrict__ status,
                                                uint32_t *__restrict__ ctl, TileOrder ord)
{
    __shared__ DecodeLds L;
    const uint64_t tile = tile_of_block(ord, blockIdx.x);
    if (ld_ctl(&ctl[1]) & (kErrStream | kErrTimeout)) {
        // Successors may wait on this tile: publish before leaving.
        if (!INDEXED && threadIdx.x == 0) publish_abort_d(status, g, tile);
        return;
    }
    decode_tile<V, INDEXED, SPB>(st, len, g, blkpos, idx_carry, out, status, ctl, tile, L);
}

// ---------------------------------------------------------------------------
// Persistent, software-pipelined decode with the side index
// ---------------------------------------------------------------------------
//
// With the side index a tile needs nothing from any other tile, so one
// workgroup can decode a stream of tiles (taken by ticket, below) in a
// pipeline: while it unpacks the current tile from LDS, the stream bytes,
// block positions and carry of the next one are in flight into registers.
// The next tile's byte range is read at the top of the step and its bytes
// after the step's first barrier, so a tile's two dependent memory latencies
// (positions, then bytes) overlap the current tile's work instead of adding
// to it.  Simple tiles (full blocks of one chunk, B/8 a
// power of two up to 32) whose bytes fit kPipeBytes go through the pipeline;
// every other tile (chunk heads, the field's last tile, barely compressible
// tiles) is decoded in place by decode_tile.  No tile waits for another, so
// any grid size and any residency are correct.

constexpr int kPipeVec = 4;                                 // 16-byte stream vectors per thread
constexpr uint64_t kPipeBytes = 16ull * kWG * kPipeVec;     // a pipelined tile's byte span
static_assert(kPipeBytes + 32 <= (uint64_t)kDecTileBytes + 64, "the tile buffer holds a pipelined tile");

// One tile of the pipeline; uniform across the workgroup.
struct PipeTile {
    uint64_t tile;    // ~0 past the last tile
    uint64_t P0, P1;  // stream bytes of a simple tile: its first block .. the next item
    uint64_t pb;      // blkpos index of its first block
    uint64_t Et;      // its first element
    uint32_t nti;
    bool simple;      // simple, and not the field's last tile
};

// Locates tile `tile` and, for a simple tile, issues the loads of its byte
// range (used after the step's first barrier, by pipe_issue).
__device__ __forceinline__ PipeTile pipe_plan(const uint8_t *__restrict__ st, const Geom &g,
                                              const uint64_t *__restrict__ blkpos, uint64_t tile)
{
    PipeTile p;
    p.tile = tile < g.ntiles ? tile : ~0ull;
    p.P0 = p.P1 = p.pb = p.Et = 0;
    p.nti = 0;
    p.simple = false;
    if (p.tile == ~0ull) return p;
    const uint64_t tI0 = tile * (uint64_t)g.K * g.R;
    const uint64_t tIN = min(tI0 + (uint64_t)g.K * g.R, g.nitems);
    uint32_t ct;
    uint64_t cj;
    locate(g, tI0, ct, cj);
    p.nti = (uint32_t)(tIN - tI0);
    if (!simple_tile(g, ct, cj, p.nti) || tIN >= g.nitems) return p;
    p.simple = true;
    p.pb = (uint64_t)ct * g.nb0 + (cj - 1);
    p.Et = item_elem(g, ct, cj);
    p.P0 = blkpos[p.pb];
    uint32_t t1;
    uint64_t j1;
    locate(g, tIN, t1, j1);
    p.P1 = j1 ? blkpos[(uint64_t)t1 * g.nb0 + (j1 - 1)] : g.body + read_u64_le4(st + g.es + 8ull * t1);
    return p;
}

// Issues the loads of a planned simple tile: its stream bytes (thread tid
// holds the 16-byte vectors tid + kWG * v of the span from P0 rounded down
// to 16), the positions of its slots' blocks and its carry.  Returns false
// (no loads) when the tile goes to decode_tile instead.
template <int SPB>
__device__ __forceinline__ bool pipe_issue(const uint8_t *__restrict__ st, uint64_t len,
                                           const uint64_t *__restrict__ blkpos,
                                           const uint32_t *__restrict__ idx_carry,
                                           const PipeTile &p, uint4 (&rb)[kPipeVec],
                                           uint64_t (&pos)[kMaxRounds], uint32_t &carry)
{
    constexpr uint32_t K = 256 / SPB;
    if (p.tile == ~0ull || !p.simple) return false;
    const uint64_t a0 = p.P0 & ~15ull;
    if (!(p.P0 < p.P1 && p.P1 <= len && p.P1 - a0 <= kPipeBytes)) return false;
    const uint32_t tid = threadIdx.x, d = tid / SPB;
#pragma unroll
    for (int v = 0; v < kPipeVec; ++v) {
        const uint64_t q = a0 + 16ull * (tid + (uint32_t)kWG * v);
        rb[v] = make_uint4(0, 0, 0, 0);
        if (q < p.P1) {
            if (q + 16 <= len) {
                rb[v] = *reinterpret_cast<const uint4 *>(st + q);
            } else {
                uint32_t w[4] = {0, 0, 0, 0};
#pragma unroll
                for (int k = 0; k < 16; ++k)
                    if (q + k < len) w[k >> 2] |= (uint32_t)st[q + k] << (8 * (k & 3));
                rb[v] = make_uint4(w[0], w[1], w[2], w[3]);
            }
        }
    }
#pragma unroll
    for (int r = 0; r < kMaxRounds; ++r)
        pos[r] = blkpos[p.pb + r * K + d];  // whole tiles of kMaxRounds rounds
    carry = idx_carry[p.tile];
    return true;
}

// Unpacks a pipelined tile whose bytes and slot positions are in L and
// stores its values (the second half of decode_simple_fast).
template <typename V, int SPB>
__device__ __forceinline__ void pipe_decode(const PipeTile &p, uint32_t carry, V eb,
                                            V *__restrict__ out, uint32_t *__restrict__ ctl,
                                            DecodeLds &L)
{
    constexpr uint32_t K = 256 / SPB, B = 8 * SPB;
    const int tid = threadIdx.x, lane = __lane_id(), wv = tid >> 6;
    const uint32_t d = (uint32_t)tid / SPB, o = (uint32_t)tid % SPB;
    const uint32_t lend = (uint32_t)(p.P1 - (p.P0 & ~15ull));
    // The pipeline runs only on fields whose tiles are kMaxRounds whole
    // rounds (g.R == kMaxRounds), so every slot of every round is live and
    // the number of stores after the next tile's loads is fixed.
    bool bad = false;
    uint32_t dl[kMaxRounds][8];
#pragma unroll
    for (int r = 0; r < kMaxRounds; ++r) {
        const uint32_t lp0 = L.s_pos[r * K + d];
        if (lp0 == ~0u) bad = true;
        uint32_t h;
        decode_slot<true>(B, false, o, L.buf, lp0 == ~0u ? 0u : lp0, lend, dl[r], h, bad);
    }
    if (bad) raise_error(ctl, kErrStream);
    uint32_t ex[kMaxRounds];
#pragma unroll
    for (int r = 0; r < kMaxRounds; ++r) {
        const uint32_t v = dl[r][7];
        const uint32_t incl = wave_incl_scan_u32(v, lane);
        ex[r] = incl - v;
        const uint32_t wt = __builtin_amdgcn_readlane(incl, 63);
        if (lane == 0) L.s_wsum[r * 4 + wv] = wt;
    }
    __syncthreads();
    uint32_t base = carry;
#pragma unroll
    for (int r = 0; r < kMaxRounds; ++r) {
        uint32_t before = base;
#pragma unroll
        for (int w = 0; w < 4; ++w) before += w < wv ? L.s_wsum[r * 4 + w] : 0u;
#pragma unroll
        for (int w = 0; w < 4; ++w) base += L.s_wsum[r * 4 + w];
        const uint32_t q0 = before + ex[r];
        V y[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) y[i] = (V)(int32_t)(q0 + dl[r][i]) * eb;
        store_octet_out<false>(out + p.Et + (uint64_t)r * 2048u + 8u * (uint32_t)tid, y);
    }
}

// Tiles are handed out by class: workgroup b is in class c = b % C (C =
// min(G, kPipeClasses); the dispatcher places workgroups on the XCDs round
// robin, so a class is one XCD's workgroups) and its class's tiles are
// c, c + C, c + 2C, ...  Its first two tickets are fixed (its rank r = b / C
// in the class, then r + n for the class's n workgroups: tile b, then b + G
// when C divides G); every later one comes from the class's counter, so that
// a workgroup the CU's arbiter serves late takes fewer tiles instead of
// finishing last.  The ticket of the tile two steps ahead is asked for as a
// step's stream bytes are staged and read at the step's end.  A workgroup
// holds at most three unfinished tickets (the current tile, the next and the
// one asked for): when a class's counter runs dry its workgroups end that
// many of their own steps later, so fewer is a shorter tail.  A workgroup takes at most kPipeMaxSteps tickets; the
// grid is sized so that a class's workgroups can take all of its tiles that
// way (pipe_grid_size).
// Tiles the pipeline passes over are decoded by decode_tile: the general
// tiles (ord: chunk heads, short last blocks and the field's last tile, one
// per workgroup) first, so that the tickets even out their cost, and after
// the loop the simple tiles whose bytes exceed kPipeBytes (each workgroup
// lists its own in LDS).
// At most 128 VGPRs: four workgroups per CU, each with one tile in flight
// and one being unpacked.
template <typename V, int SPB>
__global__ __launch_bounds__(kWG, 4) void k_decode_pipe(const uint8_t *__restrict__ st, uint64_t len,
                                                     Geom g, const uint64_t *__restrict__ blkpos,
                                                     const uint32_t *__restrict__ idx_carry,
                                                     V *__restrict__ out, uint32_t *__restrict__ ctl,
                                                     TileOrder ord)
{
    __shared__ DecodeLds L;
    __shared__ uint32_t s_defer[kPipeMaxSteps];  // tickets of the deferred tiles
    __shared__ uint32_t s_tk, s_ndef;
    const uint32_t tid = threadIdx.x;
    const uint64_t G = gridDim.x;
    const V eb = get_eb<V>(st);
    const uint32_t C = (uint32_t)min<uint64_t>(G, kPipeClasses);
    const uint32_t c = blockIdx.x % C, r = blockIdx.x / C;
    const uint32_t nper = (uint32_t)((G - c + C - 1) / C);
    const uint64_t ctiles = c < g.ntiles ? (g.ntiles - c + C - 1) / C : 0;
    uint32_t *const ctr = ctl + kPipeCtr0 + kPipeCtrStride * c;
    auto tile_of = [&](uint32_t t) -> uint64_t { return t < ctiles ? c + (uint64_t)C * t : ~0ull; };
    if (blockIdx.x < ord.n) {
        decode_tile<V, true, SPB>(st, len, g, blkpos, idx_carry, out, nullptr, ctl, ord.t[blockIdx.x], L);
        __syncthreads();
    }
    // Tickets of the current tile and of the one whose loads go out at this
    // step; thread 0 holds `got`, the counter's answer to this step's
    // request when `asked`.
    uint32_t kcur = r, knxt = r + nper, got = 0;
    bool asked = false;
    if (tid == 0) s_ndef = 0;  // thread 0 alone lists deferred tiles
    PipeTile cur = pipe_plan(st, g, blkpos, tile_of(kcur));
    uint4 rb[kPipeVec];
    uint64_t pos[kMaxRounds];
    uint32_t carry = 0;
    bool have = pipe_issue<SPB>(st, len, blkpos, idx_carry, cur, rb, pos, carry);
    for (uint32_t step = 0; cur.tile != ~0ull; ++step) {
        constexpr uint32_t K = 256 / SPB;
        const uint32_t d = tid / SPB, o = tid % SPB;
        // The next tile's positions are read now, its bytes after the
        // barrier: the staging below covers the first round trip.
        const PipeTile nxt = pipe_plan(st, g, blkpos, tile_of(knxt));
        const uint32_t ccur = carry;
        if (have) {
            // This tile's bytes and slot positions go to LDS ...
            const uint64_t a0 = cur.P0 & ~15ull;
#pragma unroll
            for (int v = 0; v < kPipeVec; ++v) {
                const uint32_t e = 16u * (tid + (uint32_t)kWG * v);
                if (a0 + e < cur.P1) *reinterpret_cast<uint4 *>(L.buf + e) = rb[v];
            }
            if (o == 0) {
#pragma unroll
                for (int q = 0; q < kMaxRounds; ++q)
                    L.s_pos[q * K + d] = (pos[q] >= cur.P0 && pos[q] < cur.P1)
                                             ? (uint32_t)(pos[q] - a0)
                                             : ~0u;
            }
        }
        if (tid == 0) {
            if (!have && cur.simple) s_defer[s_ndef++] = kcur;
            // The ticket of the tile after next, read at the end of this
            // step: none once a ticket lies past the class's tiles (every
            // later one does too).  The address is offset by a register zero
            // the compiler cannot see through, which keeps its wave-uniform
            // atomic rewrite (one atomic for the wave, its answer read into a
            // scalar at once) away.
            asked = knxt < ctiles && step + 2 < kPipeMaxSteps;
            if (asked) {
                uint32_t z;
                asm volatile("v_mov_b32 %0, 0" : "=v"(z));
                got = atomicAdd(ctr + z, 1u);
            }
        }
        __syncthreads();
        // ... the next tile's loads go out, then this tile is unpacked.
        const bool have_n = pipe_issue<SPB>(st, len, blkpos, idx_carry, nxt, rb, pos, carry);
        if (have) pipe_decode<V, SPB>(cur, ccur, eb, out, ctl, L);
        if (tid == 0) s_tk = asked ? 2 * nper + got : ~0u;
        __syncthreads();  // L and s_tk are reused by the next step
        have = have_n;
        cur = nxt;
        kcur = knxt;
        knxt = __builtin_amdgcn_readfirstlane(s_tk);
    }
    __syncthreads();
    const uint32_t nd = s_ndef;
    for (uint32_t k = 0; k < nd; ++k) {
        decode_tile<V, true, SPB>(st, len, g, blkpos, idx_carry, out, nullptr, ctl, tile_of(s_defer[k]), L);
        __syncthreads();
    }
}

// fastMode 2 with B = 1: every unit is one value, which the compressor
// writes as [q i32][0] (SZp/src/szp_float.cc:817-862) while the reference's
// decompressor reads 4 bytes per unit and never the zero width byte
// (SZp/src/szpd_float.cc:472-483, the `current_block_size > 1` test), so
// value k of chunk t comes from the 4 bytes at the chunk's offset + 4k.  This
// kernel reproduces that decoding exactly, as a drop-in must.
template <typename V>
__global__ __launch_bounds__(kWG) void k_decode_ra_unit1(const uint8_t *__restrict__ st,
                                                         uint64_t len, Geom g,
                                                         V *__restrict__ out,
                                                         uint32_t *__restrict__ ctl)
{
    const V eb = get_eb<V>(st);
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < g.n; i += stride) {
        const uint32_t t = g.tbs ? (uint32_t)min<uint64_t>(i / g.tbs, g.T - 1) : g.T - 1;
        const uint64_t p = g.body + read_u64_le4(st + g.es + 8ull * t) + 4 * (i - chunk_lo(g, t));
        if (p + 4 > len) {
            raise_error(ctl, kErrStream);
            return;
        }
        const uint32_t q = (uint32_t)st[p] | ((uint32_t)st[p + 1] << 8) |
                           ((uint32_t)st[p + 2] << 16) | ((uint32_t)st[p + 3] << 24);
        out[i] = (V)(int32_t)q * eb;
    }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

uint64_t compress_bound(uint64_t n, int B, int T, int dtype, int mode)
{
    Geom g;
    if (!make_geom(n, B, T, &g, dtype == 1 ? 8u : 4u, (uint32_t)mode)) return 0;
    uint64_t total = g.body;
    const uint64_t nl = n - (uint64_t)(T - 1) * g.tbs;
    if (mode == 2) {
        // Per unit: 4 head bytes, a width byte and ceil((B - 1) / 8) sign
        // bytes at most; 4 bytes per other value.
        const uint64_t per_unit = 5 + ((uint64_t)B - 1 + 7) / 8;
        if (g.tbs) total += (uint64_t)(T - 1) * (g.nb0 * per_unit + 4 * (g.tbs - g.nb0));
        if (nl) total += g.nblast * per_unit + 4 * (nl - g.nblast);
        return total;
    }
    const uint64_t per_block = 1 + ((uint64_t)B + 7) / 8;
    if (g.tbs) total += (uint64_t)(T - 1) * (4 + g.nb0 * per_block + 4 * (g.tbs - 1));
    if (nl) total += 4 + g.nblast * per_block + 4 * (nl - 1);
    return total;
}

// Optional per-kernel device timing: when enabled, every launch below is
// bracketed by HIP events on its own stream and the elapsed times are summed
// per kernel when read back (szp_amd_profile_*).
namespace prof {

struct Rec {
    const char *name;
    hipEvent_t a, b;
};

std::mutex mu;
bool on = false;
std::string only;  // when not empty, only launches of this kernel are timed
std::vector<Rec> recs;
std::vector<hipEvent_t> pool;
std::vector<std::pair<std::string, std::pair<uint64_t, double>>> totals;

hipEvent_t take()
{
    if (!pool.empty()) {
        hipEvent_t e = pool.back();
        pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}

struct Scope {
    Rec r{nullptr, nullptr, nullptr};
    hipStream_t s;
    Scope(const char *name, hipStream_t st) : s(st)
    {
        std::lock_guard<std::mutex> lk(mu);
        if (!on || (!only.empty() && only != name)) return;
        r.name = name;
        r.a = take();
        r.b = take();
        (void)hipEventRecord(r.a, s);
    }
    ~Scope()
    {
        if (!r.name) return;
        (void)hipEventRecord(r.b, s);
        std::lock_guard<std::mutex> lk(mu);
        recs.push_back(r);
    }
};

void collect_locked()
{
    for (auto &r : recs) {
        float ms = 0.f;
        (void)hipEventSynchronize(r.b);
        (void)hipEventElapsedTime(&ms, r.a, r.b);
        bool found = false;
        for (auto &t : totals) {
            if (t.first == r.name) {
                t.second.first += 1;
                t.second.second += ms;
                found = true;
            }
        }
        if (!found) totals.push_back({std::string(r.name), {1, (double)ms}});
        pool.push_back(r.a);
        pool.push_back(r.b);
    }
    recs.clear();
}

}  // namespace prof

// ---------------------------------------------------------------------------
// Batched forms: the fields of a batch share one persistent launch.  The
// host lists every tile of every field in a queue (field by field; a field's
// tiles in ascending order for compress) and workgroups take `ticket`
// consecutive queue entries at a time from one atomic counter, running them
// in queue order.  A compress tile waits only for lower tiles of its own
// field, which sit earlier in the queue: taken by a running workgroup, or
// earlier in the same ticket.  So there is no residency assumption, and
// batches on concurrent streams cannot starve each other.
// ---------------------------------------------------------------------------

// One field of a batch, as the kernels see it (device memory).
struct BatchJob {
    Geom g;
    const float *x;      // compress input
    float *y;            // decompress output
    const uint8_t *st;   // decompress input
    uint8_t *out;        // compress output
    uint64_t cap;        // compress capacity / decompress stream length
    uint64_t *d_size;    // compressed size
    uint64_t *status;    // look-back words (compress)
    uint64_t *idx_pos;   // side index
    uint32_t *idx_carry;
    float eb;            // ABS: the bound; REL: written by k_batch_range_final
    float rel;
    uint32_t rel_mode;
    uint32_t seg0, nseg;  // range segments (REL)
};

struct QEntry {
    uint32_t job;
    uint32_t tile;
};

constexpr uint64_t kBatchSeg = 65536;  // elements per range segment

// Min/max of one segment of a REL field (index-free, as k_range_partial).
__global__ __launch_bounds__(kWG) void k_batch_range_partial(const BatchJob *__restrict__ jobs,
                                                             const uint32_t *__restrict__ smap,
                                                             RangePart<float> *__restrict__ parts)
{
    const uint32_t sg = blockIdx.x;
    const BatchJob &J = jobs[smap[sg]];
    const uint64_t n = J.g.n;
    const uint64_t s0 = (uint64_t)(sg - J.seg0) * kBatchSeg;
    const uint32_t len = (uint32_t)(min(s0 + kBatchSeg, n) - s0);
    const float *xs = J.x + s0;
    float mn = __builtin_nanf(""), mx = __builtin_nanf("");
    if ((reinterpret_cast<uintptr_t>(xs) & 15) == 0) {
        const uint32_t full = len / (4u * kWG) * (4u * kWG);
        for (uint32_t v = 4u * threadIdx.x; v < full; v += 4u * kWG) {
            const float4 q = *reinterpret_cast<const float4 *>(xs + v);
            mn = fminf(mn, fminf(fminf(q.x, q.y), fminf(q.z, q.w)));
            mx = fmaxf(mx, fmaxf(fmaxf(q.x, q.y), fmaxf(q.z, q.w)));
        }
        for (uint32_t i = full + threadIdx.x; i < len; i += kWG) {
            mn = fminf(mn, xs[i]);
            mx = fmaxf(mx, xs[i]);
        }
    } else {
        for (uint32_t i = threadIdx.x; i < len; i += kWG) {
            mn = fminf(mn, xs[i]);
            mx = fmaxf(mx, xs[i]);
        }
    }
    range_reduce_wg(mn, mx);
    if (threadIdx.x == 0) parts[sg] = RangePart<float>{mn, mx};
}

// eb of every REL field from its segments (one workgroup per field).
__global__ __launch_bounds__(kWG) void k_batch_range_final(BatchJob *__restrict__ jobs,
                                                           const RangePart<float> *__restrict__ parts)
{
    BatchJob &J = jobs[blockIdx.x];
    if (!J.rel_mode) return;
    float mn = __builtin_nanf(""), mx = __builtin_nanf("");
    for (uint32_t k = threadIdx.x; k < J.nseg; k += kWG) {
        const RangePart<float> p = parts[J.seg0 + k];
        mn = fminf(mn, p.mn);
        mx = fmaxf(mx, p.mx);
    }
    range_reduce_wg(mn, mx);
    if (threadIdx.x == 0) {
        float eb;
        if (isnan(J.x[0])) {
            eb = J.x[0];
        } else {
            const float range = (mx == mn) ? mx - mx : mx - mn;
            eb = range * J.rel;
        }
        J.eb = eb;
    }
}

// Takes the next ticket: the first queue entry of `ticket` consecutive ones.
__device__ __forceinline__ uint64_t take_ticket(uint64_t *queue, uint32_t ticket, uint64_t *s_t)
{
    __syncthreads();  // the previous ticket's tiles are done with LDS and s_t
    if (threadIdx.x == 0) *s_t = atomicAdd(reinterpret_cast<unsigned long long *>(queue), ticket);
    __syncthreads();
    // The same value in every lane: made uniform, so that the queue entry and
    // its job's geometry are loaded into scalar registers.
    const uint64_t t = *s_t;
    return ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(t >> 32)) << 32) |
           (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)t);
}

// Five workgroups per CU: in the persistent loop the compiler holds the tile
// body's lane-derived values (hoisted out of the loop) across tiles, and at
// 72 VGPRs (seven per CU) it spilled about 200 bytes per lane.
template <bool INDEX, int SPB>
__global__ __launch_bounds__(kWG, 5) void k_batch_compress(const BatchJob *__restrict__ jobs,
                                                        const QEntry *__restrict__ qmap,
                                                        uint64_t total, uint64_t *queue,
                                                        uint32_t ticket, uint32_t *ctl)
{
    __shared__ CompressLds L;
    __shared__ uint64_t s_t;
    for (uint64_t t0 = take_ticket(queue, ticket, &s_t); t0 < total;
         t0 = take_ticket(queue, ticket, &s_t)) {
        const uint64_t t1 = min(t0 + ticket, total);
        for (uint64_t q = t0; q < t1; ++q) {
            const QEntry e = qmap[q];
            const BatchJob &J = jobs[e.job];
            const Geom g = J.g;
            compress_tile<float, INDEX, SPB, false>(J.x, &J.eb, 0.f, g, J.out, J.cap, J.status, ctl,
                                                    J.d_size, J.idx_pos, J.idx_carry, e.tile, L);
            __syncthreads();
        }
    }
}

// One tile per workgroup (grid = the batch's tiles): a workgroup takes one
// ticket and exits, so no loop holds hoisted values across tiles and the
// kernel keeps the single-field kernel's seven workgroups per CU.  The
// ticket still orders the tiles: a workgroup takes it once resident, and a
// tile waits only for lower tickets, taken by workgroups already running.
template <bool INDEX, int SPB>
__global__ __launch_bounds__(kWG, 7) void k_batch_compress_once(const BatchJob *__restrict__ jobs,
                                                             const QEntry *__restrict__ qmap,
                                                             uint64_t total, uint64_t *queue,
                                                             uint32_t *ctl)
{
    __shared__ CompressLds L;
    __shared__ uint64_t s_t;
    if (threadIdx.x == 0) s_t = atomicAdd(reinterpret_cast<unsigned long long *>(queue), 1ull);
    __syncthreads();
    const uint64_t q = ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(s_t >> 32)) << 32) |
                       (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)s_t);
    if (q >= total) return;
    const QEntry e = qmap[q];
    const BatchJob &J = jobs[e.job];
    const Geom g = J.g;
    compress_tile<float, INDEX, SPB, false>(J.x, &J.eb, 0.f, g, J.out, J.cap, J.status, ctl,
                                            J.d_size, J.idx_pos, J.idx_carry, e.tile, L);
}

template <int SPB>
__global__ __launch_bounds__(kWG) void k_batch_decode(const BatchJob *__restrict__ jobs,
                                                      const QEntry *__restrict__ qmap,
                                                      uint64_t total, uint64_t *queue,
                                                      uint32_t ticket, uint32_t *ctl)
{
    __shared__ DecodeLds L;
    __shared__ uint64_t s_t;
    __shared__ int s_stop;
    for (uint64_t t0 = take_ticket(queue, ticket, &s_t); t0 < total;
         t0 = take_ticket(queue, ticket, &s_t)) {
        const uint64_t t1 = min(t0 + ticket, total);
        for (uint64_t q = t0; q < t1; ++q) {
            if (threadIdx.x == 0) s_stop = (ld_ctl(&ctl[1]) & (kErrStream | kErrTimeout)) != 0;
            __syncthreads();
            const bool stop = s_stop;
            __syncthreads();  // every thread has read s_stop before it is written again
            if (stop) continue;  // uniform: a malformed stream stops the batch's decode
            const QEntry e = qmap[q];
            const BatchJob &J = jobs[e.job];
            const Geom g = J.g;
            decode_tile<float, true, SPB>(J.st, J.cap, g, J.idx_pos, J.idx_carry, J.y, nullptr,
                                          ctl, e.tile, L);
            __syncthreads();
        }
    }
}

#define SZP_LAUNCH(name, kern, grid, block, lds, stream, ...)                    \
    do {                                                                       \
        prof::Scope _scope(name, stream);                                      \
        hipLaunchKernelGGL(kern, grid, block, lds, stream, __VA_ARGS__);       \
    } while (0)

void profile_enable(int enable, const char *only)
{
    std::lock_guard<std::mutex> lk(prof::mu);
    prof::collect_locked();
    prof::totals.clear();
    prof::on = enable != 0;
    prof::only = only ? only : "";
}

int profile_count()
{
    std::lock_guard<std::mutex> lk(prof::mu);
    prof::collect_locked();
    return (int)prof::totals.size();
}

int profile_entry(int i, char *name, int cap, uint64_t *launches, double *total_ms)
{
    std::lock_guard<std::mutex> lk(prof::mu);
    prof::collect_locked();
    if (i < 0 || i >= (int)prof::totals.size()) return -1;
    const auto &t = prof::totals[i];
    if (name && cap > 0) {
        snprintf(name, (size_t)cap, "%s", t.first.c_str());
    }
    if (launches) *launches = t.second.first;
    if (total_ms) *total_ms = t.second.second;
    return 0;
}

namespace {

constexpr uint64_t kAlign = 256;

uint64_t align_up(uint64_t v) { return (v + kAlign - 1) & ~(kAlign - 1); }

// Look-back words of a compress launch: a status per tile, then a status per
// group of tiles, kGroupStride words apart.
uint64_t compress_status_words(const Geom &g)
{
    return g.ntiles ? g.ntiles + kGroupStride * lb_groups(g.ntiles) : 1;
}

struct CompressWs {
    uint32_t *ctl;
    void *eb;     // the REL bound, float or double
    void *parts;  // RangePart<V>[kRangeParts]
    uint64_t *status;
};

uint64_t compress_ws_layout(const Geom &g, uint8_t *base, CompressWs *w)
{
    uint64_t off = 0;
    if (w) w->ctl = reinterpret_cast<uint32_t *>(base + off);
    off += align_up(64);
    if (w) w->eb = base + off;
    off += align_up(16);
    if (w) w->parts = base + off;
    off += align_up(sizeof(RangePart<double>) * kRangeParts);
    if (w) w->status = reinterpret_cast<uint64_t *>(base + off);
    off += align_up(8 * compress_status_words(g));
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

// Longest step of the parse: a block of B values at width 32 (mode 2: a
// unit, its 4-byte head and a block of B - 1 values).
uint32_t parse_max_step(int B, int mode)
{
    return mode == 2 ? 4 + block_bytes(32, (uint32_t)B - 1) : block_bytes(32, (uint32_t)B);
}

// Sub-tile bytes of the parse: at least the longest step, so that a
// sub-tile's entry always lies inside it.
uint32_t parse_subtile_bytes(int B, int mode)
{
    const uint32_t W = parse_max_step(B, mode);
    const uint32_t U = W > 4096 ? W : 4096;
    return (U + 255) & ~255u;
}

uint32_t parse_tile_bytes(int B, int mode) { return 4 * parse_subtile_bytes(B, mode); }

size_t parse_lds_bytes(int B, int mode)
{
    const uint32_t S = parse_tile_bytes(B, mode), NW = S / 32;
    return ((S + 64 + 15) & ~15u) + 2ull * NW * 4 + 2ull * NW * 2;
}

uint64_t decompress_ws_layout(const Geom &g, uint64_t len, uint8_t *base, DecompressWs *w)
{
    const uint32_t S = parse_tile_bytes((int)g.B, (int)g.mode);
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
    // Decode look-back: a status per tile, a tile's first word, then the
    // group statuses.
    if (w) w->dec_st = reinterpret_cast<uint64_t *>(base + off);
    off += 8 * (g.ntiles ? 2 * g.ntiles + kGroupStride * lb_groups(g.ntiles) : 1);
    if (w) w->nclr = (off - clr0) / 8;
    off = align_up(off);
    if (w) w->blkpos = reinterpret_cast<uint64_t *>(base + off);
    off += align_up(8 * (nblocks ? nblocks : 1));
    return off;
}

}  // namespace

uint64_t index_bytes(uint64_t n, int B, int T, int mode)
{
    Geom g;
    if (!make_geom(n, B, T, &g, 4, (uint32_t)mode)) return 0;
    const uint64_t nblocks = (uint64_t)(g.T - 1) * g.nb0 + g.nblast;
    return align_up(8 * (nblocks ? nblocks : 1)) + align_up(4 * (g.ntiles ? g.ntiles : 1));
}

static void index_layout(const Geom &g, uint8_t *base, uint64_t **pos, uint32_t **carry)
{
    const uint64_t nblocks = (uint64_t)(g.T - 1) * g.nb0 + g.nblast;
    *pos = reinterpret_cast<uint64_t *>(base);
    *carry = reinterpret_cast<uint32_t *>(base + align_up(8 * (nblocks ? nblocks : 1)));
}

uint64_t compress_workspace_bytes(uint64_t n, int B, int T, int mode)
{
    Geom g;
    if (!make_geom(n, B, T, &g, 4, (uint32_t)mode)) return 0;
    return compress_ws_layout(g, nullptr, nullptr);
}

uint64_t decompress_workspace_bytes(uint64_t stream_bytes, uint64_t n, int B, int T, int mode)
{
    Geom g;
    if (!make_geom(n, B, T, &g, 4, (uint32_t)mode)) return 0;
    return decompress_ws_layout(g, stream_bytes, nullptr, nullptr);
}

uint32_t *workspace_error_word(void *d_work)
{
    return reinterpret_cast<uint32_t *>(d_work) + 1;
}

// Status codes shared with the public header (szp_amd_status).
enum { kOk = 0, kArg = 1, kCap = 2, kHip = 3 };

// Look-back kernels (k_compress, and k_parse + k_decode without the side
// index) take their tiles in workgroup order, and a tile waits for the tiles
// before it: a launch makes progress because its lowest unfinished tile is
// always resident.  Two such launches on two streams could fill the device
// with each other's waiting workgroups (every wait is bounded, so that would
// end in SZP_AMD_ERR_TIMEOUT, not a hang).  The library therefore orders its
// look-back launches across streams, per device: a launch on stream s first
// waits for the previous look-back launch when that was on another stream.
// Launches on one stream are ordered anyway; a stream being captured into a
// graph is left alone (the graph replays on one stream).
class LookbackOrder {
  public:
    explicit LookbackOrder(hipStream_t s) : s_(s)
    {
        int dev = 0;
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kDevs ||
            hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone)
            return;
        d_ = &slots()[dev];
        d_->mu.lock();
        if (!d_->ev && hipEventCreateWithFlags(&d_->ev, hipEventDisableTiming) != hipSuccess)
            d_->ev = nullptr;
        if (d_->ev && d_->valid && d_->last != s) (void)hipStreamWaitEvent(s, d_->ev, 0);
    }
    // Creates the device's event ahead of its first ordered launch (so that
    // a later launch creates nothing).  Returns false on a HIP error.
    static bool prepare(int dev)
    {
        if (dev < 0 || dev >= kDevs) return false;
        Slot &d = slots()[dev];
        std::lock_guard<std::mutex> lk(d.mu);
        if (!d.ev && hipEventCreateWithFlags(&d.ev, hipEventDisableTiming) != hipSuccess) {
            d.ev = nullptr;
            return false;
        }
        return true;
    }
    ~LookbackOrder()
    {
        if (!d_) return;
        if (d_->ev && hipEventRecord(d_->ev, s_) == hipSuccess) {
            d_->last = s_;
            d_->valid = true;
        }
        d_->mu.unlock();
    }
    LookbackOrder(const LookbackOrder &) = delete;
    LookbackOrder &operator=(const LookbackOrder &) = delete;

  private:
    static constexpr int kDevs = 64;
    struct Slot {
        std::mutex mu;
        hipEvent_t ev = nullptr;
        hipStream_t last = nullptr;
        bool valid = false;
    };
    static Slot *slots()
    {
        static Slot s[kDevs];
        return s;
    }
    hipStream_t s_;
    Slot *d_ = nullptr;
};

// k_compress with the fast simple-tile path compiled for SPB = B / 8 when
// that is a power of two up to 32 (B = 8 ... 256), else without it.
template <typename V, bool INDEX>
static void launch_k_compress(const Geom &g, hipStream_t stream, const V *d_x, const V *d_eb,
                              V eb_host, uint8_t *d_out, uint64_t cap, uint64_t *status,
                              uint32_t *ctl, uint64_t *d_size, uint64_t *ipos, uint32_t *icarry)
{
    const dim3 grid((uint32_t)g.ntiles), block(kWG);
    const uint32_t spb = (g.B & 7) ? 0u : g.B / 8;
    const TileOrder ord = general_tiles(g);
    switch (spb) {
#define SZP_COMPRESS_CASE(S)                                                                     \
    case S:                                                                                      \
        SZP_LAUNCH("k_compress", (k_compress<V, INDEX, S>), grid, block, 0, stream, d_x, d_eb,   \
                   eb_host, g, d_out, cap, status, ctl, d_size, ipos, icarry, ord);              \
        break;
        SZP_COMPRESS_CASE(1)
        SZP_COMPRESS_CASE(2)
        SZP_COMPRESS_CASE(4)
        SZP_COMPRESS_CASE(8)
        SZP_COMPRESS_CASE(16)
        SZP_COMPRESS_CASE(32)
#undef SZP_COMPRESS_CASE
    default:
        SZP_LAUNCH("k_compress", (k_compress<V, INDEX, 0>), grid, block, 0, stream, d_x, d_eb, eb_host,
                   g, d_out, cap, status, ctl, d_size, ipos, icarry, ord);
    }
}

// k_decode with the fast simple-tile path compiled for SPB = B / 8 when that
// is a power of two up to 32.
template <typename V, bool INDEXED>
static void launch_k_decode(const char *name, const Geom &g, hipStream_t stream,
                            const uint8_t *d_stream, uint64_t len, const uint64_t *blkpos,
                            const uint32_t *carry, V *d_out, uint64_t *status, uint32_t *ctl)
{
    const dim3 grid((uint32_t)g.ntiles), block(kWG);
    const uint32_t spb = (g.B & 7) ? 0u : g.B / 8;
    // General tiles and the last tile (whose end the fast path does not
    // know) go first: they take the slower general path, and successors
    // look back over them (without the side index) or the kernel would end
    // on them.
    const TileOrder ord = general_tiles(g, true);
    switch (spb) {
#define SZP_DECODE_CASE(S)                                                                       \
    case S:                                                                                      \
        SZP_LAUNCH(name, (k_decode<V, INDEXED, S>), grid, block, 0, stream, d_stream, len, g,    \
                   blkpos, carry, d_out, status, ctl, ord);                                      \
        break;
        SZP_DECODE_CASE(1)
        SZP_DECODE_CASE(2)
        SZP_DECODE_CASE(4)
        SZP_DECODE_CASE(8)
        SZP_DECODE_CASE(16)
        SZP_DECODE_CASE(32)
#undef SZP_DECODE_CASE
    default:
        SZP_LAUNCH(name, (k_decode<V, INDEXED, 0>), grid, block, 0, stream, d_stream, len, g, blkpos,
                   carry, d_out, status, ctl, ord);
    }
}

template <typename V>
static int launch_compress_t(int mode, const V *d_x, uint64_t n, int err_mode, double abs_eb,
                             double rel, int B, int T, uint8_t *d_out, uint64_t cap,
                             uint64_t *d_size, void *d_work, uint64_t work_bytes,
                             hipStream_t stream, void *d_index, uint64_t idx_bytes)
{
    Geom g;
    if (!make_geom(n, B, T, &g, sizeof(V), (uint32_t)mode) || !d_out || !d_size || !d_work)
        return kArg;
    if (err_mode != 0 && err_mode != 1) return kArg;
    if (err_mode == 1 && n == 0) return kArg;
    if (n && !d_x) return kArg;
    if ((reinterpret_cast<uintptr_t>(d_out) & 15) || (reinterpret_cast<uintptr_t>(d_work) & 255))
        return kArg;
    if (cap < g.body) return kCap;
    if (d_index && ((reinterpret_cast<uintptr_t>(d_index) & 255) ||
                    idx_bytes < index_bytes(n, B, T, mode)))
        return d_index && idx_bytes < index_bytes(n, B, T, mode) ? kCap : kArg;
    CompressWs w;
    if (compress_ws_layout(g, reinterpret_cast<uint8_t *>(d_work), &w) > work_bytes) return kCap;

    const uint64_t nst = compress_status_words(g);
    const uint32_t cgrid = (uint32_t)std::min<uint64_t>((nst + kWG - 1) / kWG, 4096);
    LookbackOrder order(stream);
    // REL: the range pass clears these words.
    if (err_mode != 1)
        SZP_LAUNCH("k_clear", k_clear, dim3(cgrid), dim3(kWG), 0, stream, w.ctl, w.status, nst);
    const V *d_eb = nullptr;
    // ABS: cast to the value type (float: SZp/src/szp.cc:32; double: :62).
    const V eb_host = (V)abs_eb;
    const bool vec = (reinterpret_cast<uintptr_t>(d_x) & 15) == 0;
    if (err_mode == 1) {
        uint64_t seg = (n + kRangeParts - 1) / kRangeParts;
        seg = (seg + 4095) & ~4095ull;  // a multiple of 4 * kWG * kRangeUnroll
        const int np = (int)((n + seg - 1) / seg);
        RangePart<V> *parts = reinterpret_cast<RangePart<V> *>(w.parts);
        V *d_ebw = reinterpret_cast<V *>(w.eb);
        if (vec)
            SZP_LAUNCH("k_range_partial", (k_range_partial<V, true>), dim3(np), dim3(kWG), 0, stream,
                       d_x, n, seg, parts, w.ctl, w.status, nst);
        else
            SZP_LAUNCH("k_range_partial", (k_range_partial<V, false>), dim3(np), dim3(kWG), 0, stream,
                       d_x, n, seg, parts, w.ctl, w.status, nst);
        SZP_LAUNCH("k_range_final", k_range_final<V>, dim3(1), dim3(kWG), 0, stream, d_x, parts, np,
                   (V)rel, d_ebw);
        d_eb = d_ebw;
    }
    if (n == 0) {
        SZP_LAUNCH("k_empty_stream", k_empty_stream<V>, dim3(1), dim3(kWG), 0, stream, d_out, d_eb, eb_host,
                           g.T, d_size);
    } else if (d_index) {
        uint64_t *ipos;
        uint32_t *icarry;
        index_layout(g, reinterpret_cast<uint8_t *>(d_index), &ipos, &icarry);
        launch_k_compress<V, true>(g, stream, d_x, d_eb, eb_host, d_out,
                                   std::min<uint64_t>(cap, kMaxCompressStream), w.status, w.ctl,
                                   d_size, ipos, icarry);
    } else {
        launch_k_compress<V, false>(g, stream, d_x, d_eb, eb_host, d_out,
                                    std::min<uint64_t>(cap, kMaxCompressStream), w.status, w.ctl,
                                    d_size, nullptr, nullptr);
    }
    return hipGetLastError() == hipSuccess ? kOk : kHip;
}

int launch_compress(int dtype, const void *d_x, uint64_t n, int err_mode, double abs_eb,
                    double rel, int B, int T, uint8_t *d_out, uint64_t cap, uint64_t *d_size,
                    void *d_work, uint64_t work_bytes, hipStream_t stream, void *d_index,
                    uint64_t idx_bytes, int mode)
{
    if (dtype == 0)
        return launch_compress_t<float>(mode, reinterpret_cast<const float *>(d_x), n, err_mode,
                                        abs_eb, rel, B, T, d_out, cap, d_size, d_work, work_bytes,
                                        stream, d_index, idx_bytes);
    if (dtype == 1)
        return launch_compress_t<double>(mode, reinterpret_cast<const double *>(d_x), n, err_mode,
                                         abs_eb, rel, B, T, d_out, cap, d_size, d_work,
                                         work_bytes, stream, d_index, idx_bytes);
    return kArg;
}

// ---- compress with no allocation ----
//
// The workspace of launch_compress_noalloc is a file-scope device array, one
// per device (each device holds its own copy until the process ends).  In host
// code `scratch` names a host-side stand-in, not a device address: host code
// uses only the address hipGetSymbolAddress returns for the current device
// (noalloc_scratch), looked up once per device on its first call.
static __device__ __align__(256) uint8_t scratch[kNoallocScratchBytes];

// A device-side error of the call (ctl[1] != 0) makes its size read ~0.
__global__ void k_size_status(const uint32_t *ctl, uint64_t *d_size)
{
    if (threadIdx.x == 0 && ctl[1]) *d_size = ~0ull;
}

namespace {
std::mutex g_scratch_mu;
void *g_scratch[64];
}  // namespace

// The current device's scratch address (set up on its first call), or null.
static void *noalloc_scratch()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return nullptr;
        if (!LookbackOrder::prepare(dev)) return nullptr;
        g_scratch[dev] = p;
    }
    return g_scratch[dev];
}

int launch_compress_noalloc(const float *d_x, uint64_t n, int err_mode, double abs_eb,
                            double rel, int B, int T, uint8_t *d_out, uint64_t cap,
                            uint64_t *d_size, hipStream_t stream)
{
    const uint64_t need = compress_workspace_bytes(n, B, T, 1);
    if (!need) return kArg;
    if (need > kNoallocScratchBytes) return kCap;
    void *ws = noalloc_scratch();
    if (!ws) return kHip;
    const int rc = launch_compress(0, d_x, n, err_mode, abs_eb, rel, B, T, d_out, cap, d_size, ws,
                                   kNoallocScratchBytes, stream);
    if (rc) return rc;
    SZP_LAUNCH("k_size_status", k_size_status, dim3(1), dim3(64), 0, stream,
               reinterpret_cast<const uint32_t *>(ws), d_size);
    return hipGetLastError() == hipSuccess ? kOk : kHip;
}

template <typename V>
static int launch_decompress_t(int mode, const uint8_t *d_stream, uint64_t len, uint64_t n, int B,
                               int T, V *d_out, void *d_work, uint64_t work_bytes,
                               hipStream_t stream)
{
    Geom g;
    if (!make_geom(n, B, T, &g, sizeof(V), (uint32_t)mode) || !d_stream || !d_work) return kArg;
    if (n && !d_out) return kArg;
    if ((reinterpret_cast<uintptr_t>(d_stream) & 15) || (reinterpret_cast<uintptr_t>(d_work) & 255))
        return kArg;
    DecompressWs w;
    if (decompress_ws_layout(g, len, reinterpret_cast<uint8_t *>(d_work), &w) > work_bytes)
        return kCap;
    if (mode == 2 && B == 1) {
        // Units of one value decode as the reference does (k_decode_ra_unit1).
        if (len < g.body) return kArg;
        SZP_LAUNCH("k_clear", k_clear, dim3(1), dim3(kWG), 0, stream, w.ctl, nullptr, 0);
        if (n)
            SZP_LAUNCH("k_decode_ra_unit1", k_decode_ra_unit1<V>,
                       dim3((uint32_t)std::min<uint64_t>((n + kWG - 1) / kWG, 8192)), dim3(kWG), 0,
                       stream, d_stream, len, g, d_out, w.ctl);
        return hipGetLastError() == hipSuccess ? kOk : kHip;
    }
    const uint32_t S = parse_tile_bytes(B, mode);
    const uint32_t W = parse_max_step(B, mode);
    const uint32_t cgrid = (uint32_t)std::min<uint64_t>((w.nclr + kWG - 1) / kWG, 4096);
    LookbackOrder order(stream);
    SZP_LAUNCH("k_dsetup", k_dsetup, dim3(cgrid ? cgrid : 1), dim3(kWG), 0, stream, d_stream, len, g,
                       S, w.ctab, w.exit_st, w.nclr, w.ctl);
    if (n == 0) return hipGetLastError() == hipSuccess ? kOk : kHip;
    const uint64_t nt1 = len / S + 1 + g.T;
    SZP_LAUNCH("k_parse", k_parse, dim3((uint32_t)nt1), dim3(kWG), parse_lds_bytes(B, mode), stream,
               d_stream, len, g, parse_subtile_bytes(B, mode), W, w.ctab, w.exit_st, w.cnt_st, w.blkpos,
               w.ctl);
    launch_k_decode<V, false>("k_decode", g, stream, d_stream, len, w.blkpos, nullptr, d_out,
                              w.dec_st, w.ctl);
    return hipGetLastError() == hipSuccess ? kOk : kHip;
}

int launch_decompress(int dtype, const uint8_t *d_stream, uint64_t len, uint64_t n, int B, int T,
                      void *d_out, void *d_work, uint64_t work_bytes, hipStream_t stream, int mode)
{
    if (dtype == 0)
        return launch_decompress_t<float>(mode, d_stream, len, n, B, T,
                                          reinterpret_cast<float *>(d_out), d_work, work_bytes,
                                          stream);
    if (dtype == 1)
        return launch_decompress_t<double>(mode, d_stream, len, n, B, T,
                                           reinterpret_cast<double *>(d_out), d_work, work_bytes,
                                           stream);
    return kArg;
}

// Workgroups of k_decode_pipe<V, SPB> the device holds at once (occupancy
// times CUs), found once per process; any grid size is correct for it, this
// one keeps every CU busy without queued workgroups.  SZP_AMD_PIPE_GRID=g
// (g > 0) caps it, so that a small field still runs many pipeline steps per
// workgroup (a test and tuning knob).
template <typename V, int SPB>
static uint32_t pipe_grid()
{
    const char *cap = getenv("SZP_AMD_PIPE_GRID");
    const long capv = cap ? strtol(cap, nullptr, 10) : 0;
    static const uint32_t grid = [] {
        int dev = 0, cus = 0, per_cu = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_decode_pipe<V, SPB>, kWG, 0) !=
                hipSuccess ||
            cus <= 0 || per_cu <= 0)
            return 1024u;
        return (uint32_t)(cus * per_cu);
    }();
    return capv > 0 && (uint64_t)capv < grid ? (uint32_t)capv : grid;
}

template <typename V>
static int launch_decompress_indexed_t(int mode, const uint8_t *d_stream, uint64_t len,
                                       const void *d_index, uint64_t idx_bytes, uint64_t n, int B,
                                       int T, V *d_out, void *d_work, uint64_t work_bytes,
                                       hipStream_t stream)
{
    Geom g;
    if (!make_geom(n, B, T, &g, sizeof(V), (uint32_t)mode) || !d_stream || !d_work || !d_index)
        return kArg;
    if (n && !d_out) return kArg;
    if ((reinterpret_cast<uintptr_t>(d_stream) & 15) || (reinterpret_cast<uintptr_t>(d_work) & 255) ||
        (reinterpret_cast<uintptr_t>(d_index) & 255))
        return kArg;
    if (idx_bytes < index_bytes(n, B, T, mode) || work_bytes < 256) return kCap;
    if (len < g.body) return kArg;
    uint32_t *ctl = reinterpret_cast<uint32_t *>(d_work);
    // The whole 256-byte control block: error word and the pipelined
    // decode's ticket counters (words 8..63, cleared as 28 status words).
    SZP_LAUNCH("k_clear", k_clear, dim3(1), dim3(kWG), 0, stream, ctl,
               reinterpret_cast<uint64_t *>(ctl + 8), 28);
    if (n == 0) return hipGetLastError() == hipSuccess ? kOk : kHip;
    if (mode == 2 && B == 1) {
        // As in the parse path: the reference's decoding of units of one value.
        SZP_LAUNCH("k_decode_ra_unit1", k_decode_ra_unit1<V>,
                   dim3((uint32_t)std::min<uint64_t>((n + kWG - 1) / kWG, 8192)), dim3(kWG), 0,
                   stream, d_stream, len, g, d_out, ctl);
        return hipGetLastError() == hipSuccess ? kOk : kHip;
    }
    uint64_t *ipos;
    uint32_t *icarry;
    index_layout(g, reinterpret_cast<uint8_t *>(const_cast<void *>(d_index)), &ipos, &icarry);
    const uint32_t spb = (g.B & 7) ? 0u : g.B / 8;
    // The pipelined kernel needs every general tile in the list (at most
    // kMaxGenTiles of them; the last tile is always one) and tiles of
    // kMaxRounds whole rounds (chunks not too short).
    const TileOrder ord = general_tiles(g, true);
    if (mode == 1 && spb && spb <= 32 && !(spb & (spb - 1)) && ord.n && g.R == (uint32_t)kMaxRounds) {
        switch (spb) {
#define SZP_PIPE_CASE(S)                                                                         \
    case S: {                                                                                    \
        const uint64_t G = pipe_grid_size(g, ord, pipe_grid<V, S>());                              \
        if (G > 0xffffffffu || !pipe_grid_covers(g, G)) return kArg;                             \
        const dim3 grid((uint32_t)G);                                                            \
        SZP_LAUNCH("k_decode_indexed", (k_decode_pipe<V, S>), grid, dim3(kWG), 0, stream,        \
                   d_stream, len, g, ipos, icarry, d_out, ctl, ord);                             \
        break;                                                                                   \
    }
            SZP_PIPE_CASE(1)
            SZP_PIPE_CASE(2)
            SZP_PIPE_CASE(4)
            SZP_PIPE_CASE(8)
            SZP_PIPE_CASE(16)
            SZP_PIPE_CASE(32)
#undef SZP_PIPE_CASE
        }
        return hipGetLastError() == hipSuccess ? kOk : kHip;
    }
    launch_k_decode<V, true>("k_decode_indexed", g, stream, d_stream, len, ipos, icarry, d_out,
                             nullptr, ctl);
    return hipGetLastError() == hipSuccess ? kOk : kHip;
}

int launch_decompress_indexed(int dtype, const uint8_t *d_stream, uint64_t len,
                              const void *d_index, uint64_t idx_bytes, uint64_t n, int B, int T,
                              void *d_out, void *d_work, uint64_t work_bytes, hipStream_t stream,
                              int mode)
{
    if (dtype == 0)
        return launch_decompress_indexed_t<float>(mode, d_stream, len, d_index, idx_bytes, n, B,
                                                  T, reinterpret_cast<float *>(d_out), d_work,
                                                  work_bytes, stream);
    if (dtype == 1)
        return launch_decompress_indexed_t<double>(mode, d_stream, len, d_index, idx_bytes, n, B,
                                                   T, reinterpret_cast<double *>(d_out), d_work,
                                                   work_bytes, stream);
    return kArg;
}

// ---------------------------------------------------------------------------
// Batched forms (host side)
// ---------------------------------------------------------------------------

namespace {

struct BatchLayout {
    uint64_t jobs, qmap, smap, parts, status, nstatus, tiles, segs, total;
};

// Workspace of a batch: [ctl | queue counter + look-back words | jobs | queue
// map | segment map | range partials].
bool batch_layout(const BatchField *f, int F, int B, int T, int decode, BatchLayout *L)
{
    if (F < 1 || !f) return false;
    uint64_t tiles = 0, segs = 0, nstat = 1;
    for (int k = 0; k < F; ++k) {
        Geom g;
        if (f[k].n == 0 || !make_geom(f[k].n, B, T, &g)) return false;
        tiles += g.ntiles;
        if (!decode) {
            nstat += compress_status_words(g);
            if (f[k].err_mode == 1) segs += (f[k].n + kBatchSeg - 1) / kBatchSeg;
        }
    }
    if (tiles >= (1ull << 32)) return false;
    uint64_t off = align_up(64);
    L->status = off;
    L->nstatus = nstat;
    off += align_up(8 * nstat);
    L->jobs = off;
    off += align_up(sizeof(BatchJob) * (uint64_t)F);
    L->qmap = off;
    off += align_up(sizeof(QEntry) * tiles);
    L->smap = off;
    off += align_up(4 * (segs ? segs : 1));
    L->parts = off;
    off += align_up(sizeof(RangePart<float>) * (segs ? segs : 1));
    L->tiles = tiles;
    L->segs = segs;
    L->total = off;
    return true;
}

template <typename K>
uint32_t resident_grid(K kernel)
{
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kWG, 0) != hipSuccess)
        return 1024;
    const int g = cus * (per_cu > 0 ? per_cu : 1);
    return g > 0 ? (uint32_t)g : 1024u;
}

}  // namespace

uint64_t batch_workspace_bytes(const BatchField *f, int F, int B, int T, int decode)
{
    BatchLayout L;
    return batch_layout(f, F, B, T, decode, &L) ? L.total : 0;
}

int batch_prepare(const BatchField *f, int F, int B, int T, int decode, uint64_t *d_sizes,
                  void *d_work, uint64_t work_bytes, BatchPlan *plan)
{
    BatchLayout L;
    if (!plan || !d_work || (reinterpret_cast<uintptr_t>(d_work) & 255)) return kArg;
    if (!batch_layout(f, F, B, T, decode, &L)) return kArg;
    if (L.total > work_bytes) return kCap;
    if (!decode && !d_sizes) return kArg;
    const bool with_index = f[0].index != nullptr;
    uint8_t *base = reinterpret_cast<uint8_t *>(d_work);
    std::vector<BatchJob> jobs((size_t)F);
    std::vector<QEntry> qmap;
    std::vector<uint32_t> smap;
    qmap.reserve(L.tiles);
    uint64_t st_off = L.status + 8;  // after the queue counter
    uint64_t seg = 0;
    for (int k = 0; k < F; ++k) {
        const BatchField &fk = f[k];
        Geom g;
        make_geom(fk.n, B, T, &g);
        if (!fk.in || !fk.out) return kArg;
        if ((fk.index != nullptr) != with_index || (decode && !with_index)) return kArg;
        if (fk.index && (reinterpret_cast<uintptr_t>(fk.index) & 255)) return kArg;
        BatchJob &J = jobs[(size_t)k];
        memset(&J, 0, sizeof(J));
        J.g = g;
        if (fk.index) index_layout(g, reinterpret_cast<uint8_t *>(fk.index), &J.idx_pos, &J.idx_carry);
        if (decode) {
            if ((reinterpret_cast<uintptr_t>(fk.in) & 15) || fk.in_bytes < g.body) return kArg;
            J.st = reinterpret_cast<const uint8_t *>(fk.in);
            J.y = reinterpret_cast<float *>(fk.out);
            J.cap = fk.in_bytes;
        } else {
            if ((reinterpret_cast<uintptr_t>(fk.out) & 15) || fk.out_cap < g.body) return kArg;
            if (fk.err_mode != 0 && fk.err_mode != 1) return kArg;
            J.x = reinterpret_cast<const float *>(fk.in);
            J.out = reinterpret_cast<uint8_t *>(fk.out);
            J.cap = std::min<uint64_t>(fk.out_cap, kMaxCompressStream);
            J.d_size = d_sizes + k;
            J.status = reinterpret_cast<uint64_t *>(base + st_off);
            st_off += 8 * compress_status_words(g);
            J.rel_mode = fk.err_mode == 1;
            J.eb = (float)fk.abs_eb;  // SZp/src/szp.cc:32
            J.rel = (float)fk.rel;
            if (J.rel_mode) {
                J.seg0 = (uint32_t)seg;
                J.nseg = (uint32_t)((fk.n + kBatchSeg - 1) / kBatchSeg);
                for (uint32_t s2 = 0; s2 < J.nseg; ++s2) smap.push_back((uint32_t)k);
                seg += J.nseg;
            }
        }
        // Queue order of this field.  Compress: ascending, so that the tiles
        // of one ticket (taken together, run in order) never wait for a
        // later tile of the same ticket.  Decode (no waits with the side
        // index): the general tiles and the last tile first.
        TileOrder ord;
        ord.n = 0;
        if (decode) ord = general_tiles(g, true);
        for (uint64_t b = 0; b < g.ntiles; ++b)
            qmap.push_back(QEntry{(uint32_t)k, (uint32_t)tile_of_block(ord, b)});
    }
    if (hipMemcpy(base + L.jobs, jobs.data(), sizeof(BatchJob) * jobs.size(), hipMemcpyHostToDevice) !=
            hipSuccess ||
        hipMemcpy(base + L.qmap, qmap.data(), sizeof(QEntry) * qmap.size(), hipMemcpyHostToDevice) !=
            hipSuccess ||
        (!smap.empty() && hipMemcpy(base + L.smap, smap.data(), 4 * smap.size(),
                                    hipMemcpyHostToDevice) != hipSuccess))
        return kHip;
    plan->work = d_work;
    plan->jobs = L.jobs;
    plan->qmap = L.qmap;
    plan->smap = L.smap;
    plan->parts = L.parts;
    plan->status = L.status;
    plan->nstatus = L.nstatus;
    plan->tiles = L.tiles;
    plan->segs = L.segs;
    plan->F = (uint32_t)F;
    plan->B = (uint32_t)B;
    plan->T = (uint32_t)T;
    plan->spb = (B & 7) ? 0u : (uint32_t)B / 8;
    if (plan->spb & (plan->spb - 1) || plan->spb > 32) plan->spb = 0;
    plan->index = with_index;
    plan->decode = decode != 0;
    return kOk;
}

int batch_launch(const BatchPlan &P, uint32_t ticket, uint32_t grid, hipStream_t stream)
{
    if (!P.work || !P.F || ticket == 0) return kArg;
    if (grid == kBatchGridPerTile && (P.decode || P.tiles > 0x7fffffffu)) return kArg;
    // A compress tile waits until every lower tile of its field has started;
    // a workgroup starts the later tiles of a ticket only after the earlier
    // ones finish, so tickets of several compress tiles would chain the
    // workgroups one after another.  Compress takes one tile per ticket.
    if (!P.decode) ticket = 1;
    uint8_t *base = reinterpret_cast<uint8_t *>(P.work);
    uint32_t *ctl = reinterpret_cast<uint32_t *>(base);
    uint64_t *status = reinterpret_cast<uint64_t *>(base + P.status);
    BatchJob *jobs = reinterpret_cast<BatchJob *>(base + P.jobs);
    const QEntry *qmap = reinterpret_cast<const QEntry *>(base + P.qmap);
    const uint32_t cgrid = (uint32_t)std::min<uint64_t>((P.nstatus + kWG - 1) / kWG, 4096);
    SZP_LAUNCH("k_clear", k_clear, dim3(cgrid), dim3(kWG), 0, stream, ctl, status, P.nstatus);
    const dim3 block(kWG);
    if (P.decode) {
        switch (P.spb) {
#define SZP_BATCH_DECODE_CASE(S)                                                                 \
    case S: {                                                                                    \
        const uint32_t gr = grid ? grid : resident_grid(k_batch_decode<S>);                      \
        SZP_LAUNCH("k_batch_decode", k_batch_decode<S>, dim3(gr), block, 0, stream, jobs, qmap,  \
                   P.tiles, status, ticket, ctl);                                                \
        break;                                                                                   \
    }
            SZP_BATCH_DECODE_CASE(0)
            SZP_BATCH_DECODE_CASE(1)
            SZP_BATCH_DECODE_CASE(2)
            SZP_BATCH_DECODE_CASE(4)
            SZP_BATCH_DECODE_CASE(8)
            SZP_BATCH_DECODE_CASE(16)
            SZP_BATCH_DECODE_CASE(32)
#undef SZP_BATCH_DECODE_CASE
        }
        return hipGetLastError() == hipSuccess ? kOk : kHip;
    }
    if (P.segs) {
        SZP_LAUNCH("k_batch_range_partial", k_batch_range_partial, dim3((uint32_t)P.segs), block, 0,
                   stream, jobs, reinterpret_cast<const uint32_t *>(base + P.smap),
                   reinterpret_cast<RangePart<float> *>(base + P.parts));
        SZP_LAUNCH("k_batch_range_final", k_batch_range_final, dim3(P.F), block, 0, stream, jobs,
                   reinterpret_cast<const RangePart<float> *>(base + P.parts));
    }
#define SZP_BATCH_COMPRESS_CASE(I, S)                                                            \
    case S: {                                                                                    \
        if (grid == kBatchGridPerTile) {                                                         \
            SZP_LAUNCH("k_batch_compress", (k_batch_compress_once<I, S>), dim3(P.tiles), block,  \
                       0, stream, jobs, qmap, P.tiles, status, ctl);                             \
        } else {                                                                                 \
            const uint32_t gr = grid ? grid : resident_grid(k_batch_compress<I, S>);             \
            SZP_LAUNCH("k_batch_compress", (k_batch_compress<I, S>), dim3(gr), block, 0, stream, \
                       jobs, qmap, P.tiles, status, ticket, ctl);                                \
        }                                                                                        \
        break;                                                                                   \
    }
    if (P.index) {
        switch (P.spb) {
            SZP_BATCH_COMPRESS_CASE(true, 0)
            SZP_BATCH_COMPRESS_CASE(true, 1)
            SZP_BATCH_COMPRESS_CASE(true, 2)
            SZP_BATCH_COMPRESS_CASE(true, 4)
            SZP_BATCH_COMPRESS_CASE(true, 8)
            SZP_BATCH_COMPRESS_CASE(true, 16)
            SZP_BATCH_COMPRESS_CASE(true, 32)
        }
    } else {
        switch (P.spb) {
            SZP_BATCH_COMPRESS_CASE(false, 0)
            SZP_BATCH_COMPRESS_CASE(false, 1)
            SZP_BATCH_COMPRESS_CASE(false, 2)
            SZP_BATCH_COMPRESS_CASE(false, 4)
            SZP_BATCH_COMPRESS_CASE(false, 8)
            SZP_BATCH_COMPRESS_CASE(false, 16)
            SZP_BATCH_COMPRESS_CASE(false, 32)
        }
    }
#undef SZP_BATCH_COMPRESS_CASE
    return hipGetLastError() == hipSuccess ? kOk : kHip;
}

uint32_t batch_resident_grid(const BatchPlan &P)
{
    // Occupancy does not depend on the template arguments' SPB beyond
    // registers, which the SPB = 16 forms bound.
    if (P.decode) return resident_grid(k_batch_decode<16>);
    return P.index ? resident_grid(k_batch_compress<true, 16>) : resident_grid(k_batch_compress<false, 16>);
}

}  // namespace szp_amd
