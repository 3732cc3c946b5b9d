// Geometry of the item/tile decomposition shared by the HIP kernels and the
// host-compiled index checks (tests/cpu/geom_check.cc).  Everything here is
// plain integer arithmetic so that the same code can run on the CPU.
#pragma once

#include <stdint.h>
#include <string.h>

#ifdef __HIPCC__
#define SZP_HD __host__ __device__ __forceinline__
#else
#define SZP_HD inline
#endif

namespace szp_amd {

constexpr int kWG = 256;                 // threads per workgroup (4 wave64)
constexpr int kMaxBlock = 2048;          // largest SZp block size handled on the GPU
constexpr int kOutBufBytes = 8704 + 64;  // LDS bytes for one round's encoded blocks
constexpr int kMaxRounds = 4;            // rounds of K items per tile
constexpr int kOutCap = 17408;           // LDS bytes buffering a compress tile before placement
constexpr int kDecTileBytes = 20480;     // LDS bytes holding a decode tile's stream bytes

// Geometry of the item sequence of one field.  The stream body is the
// concatenation, in element order, of "items", each either a head (one raw
// quantized value, 4 bytes) or a block (a width byte and the packed deltas of
// its values against their predecessors).
//   fastMode 1 (SZp/src/szp_float.cc:201-275): each non-empty chunk is its
//     head followed by blocks of B values (the last one shorter).
//   fastMode 2, "random access" (SZp/src/szp_float.cc:752-870): each chunk is
//     cut into units of B values from its start; a unit is the head of its
//     first value followed by one block of the other L = len - 1 values
//     (L = 0 still writes its zero width byte).
// A tile is R rounds of K consecutive items; every item owns spb = ceil(B/8)
// thread slots ("octets").
struct Geom {
    uint64_t n;        // elements
    uint64_t tbs;      // floor(n / T), the size of chunks 0..T-2
    uint64_t c0;       // items in chunks 0..T-2 (0 when tbs == 0)
    uint64_t clast;    // items in chunk T-1
    uint64_t nitems;   // (T-1)*c0 + clast
    uint64_t nb0;      // SZp blocks (mode 2: units) in chunks 0..T-2
    uint64_t nblast;   // SZp blocks (mode 2: units) in chunk T-1
    uint64_t ntiles;   // ceil(nitems / (K * R))
    uint64_t body;     // es + 8T: offset of chunk 0 in the stream
    uint32_t es;       // bytes of the error bound at the stream's start (4 float, 8 double)
    uint32_t mode;     // 1: fastMode 1, 2: fastMode 2 (random access)
    uint32_t T;        // chunks (the reference's OpenMP team size)
    uint32_t B;        // SZp block size
    uint32_t spb;      // slots per item
    uint32_t K;        // items per round
    uint32_t spb_magic;  // slot / spb == (slot * spb_magic) >> 16 for slot < 256
    uint32_t R;          // rounds per tile (1 when chunks are smaller than a tile)
};

// Returns false when (n, B, T, mode) is outside what the kernels handle.  es
// is the size of the value type: the stream opens with the error bound in
// that type (SZp/src/szp_float.cc:165, SZp/src/szp_double.cc:165).
inline bool make_geom(uint64_t n, int B, int T, Geom *g, uint32_t es = 4, uint32_t mode = 1)
{
    if (B < 1 || B > kMaxBlock || T < 1 || (mode != 1 && mode != 2)) return false;
    const uint64_t b = (uint64_t)B;
    auto blocks_of = [&](uint64_t m) -> uint64_t {
        return mode == 2 ? (m + b - 1) / b : (m ? (m - 1 + b - 1) / b : 0);
    };
    auto items_of = [&](uint64_t m) -> uint64_t {
        return mode == 2 ? 2 * blocks_of(m) : (m ? 1 + blocks_of(m) : 0);
    };
    g->mode = mode;
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
    // Tiles of kMaxRounds rounds, unless chunks are so short that a tile would
    // hold more chunk heads than one workgroup records (at most kWG).
    g->R = (g->c0 == 0 || g->c0 * (kWG - 1) >= (uint64_t)g->K * kMaxRounds) ? kMaxRounds : 1;
    const uint64_t per_tile = (uint64_t)g->K * g->R;
    g->ntiles = (g->nitems + per_tile - 1) / per_tile;
    g->es = es;
    g->body = es + 8ull * (uint64_t)T;
    return true;
}

// Bytes of one block of L values at width b (SZp/src/szp_TypeManager.cc:20-31).
SZP_HD uint32_t block_bytes(uint32_t b, uint32_t L)
{
    if (!b) return 1;
    return 1 + ((L + 7) >> 3) + (b >> 3) * L + (((b & 7) * L + 7) >> 3);
}

// Item -> (chunk t, index j within the chunk; j == 0 is the chunk's first head).
SZP_HD void locate(const Geom &g, uint64_t item, uint32_t &t, uint64_t &j)
{
    const uint64_t lead = (uint64_t)(g.T - 1) * g.c0;
    if (item < lead) {
        t = (uint32_t)(item / g.c0);
        j = item - (uint64_t)t * g.c0;
    } else {
        t = g.T - 1;
        j = item - lead;
    }
}

// Moves the cursor (t, j) forward by d items.
SZP_HD void advance(const Geom &g, uint32_t &t, uint64_t &j, uint32_t d)
{
    j += d;
    while (t < g.T - 1 && j >= g.c0) {
        j -= g.c0;
        ++t;
    }
}

SZP_HD uint64_t chunk_lo(const Geom &g, uint32_t t) { return (uint64_t)t * g.tbs; }

SZP_HD uint64_t chunk_len(const Geom &g, uint32_t t)
{
    return t == g.T - 1 ? g.n - chunk_lo(g, t) : g.tbs;
}

// Item j of a chunk is a head (a raw quantized value).
SZP_HD bool item_head(const Geom &g, uint64_t j) { return g.mode == 2 ? !(j & 1) : j == 0; }

// Items whose stream position the block index holds (blocks in mode 1, units
// -- at their head -- in mode 2), and their number within the chunk.
SZP_HD bool item_indexed(const Geom &g, uint64_t j) { return g.mode == 2 ? !(j & 1) : j != 0; }
SZP_HD uint64_t item_block(const Geom &g, uint64_t j) { return g.mode == 2 ? j >> 1 : j - 1; }

// First element of item (t, j).
SZP_HD uint64_t item_elem(const Geom &g, uint32_t t, uint64_t j)
{
    if (g.mode == 2) return chunk_lo(g, t) + (j >> 1) * (uint64_t)g.B + (j & 1);
    return chunk_lo(g, t) + (j ? 1 + (j - 1) * (uint64_t)g.B : 0);
}

// Elements of item (t, j).
SZP_HD uint32_t item_len(const Geom &g, uint32_t t, uint64_t j)
{
    if (g.mode == 2) {
        if (!(j & 1)) return 1;
        const uint64_t rem = chunk_len(g, t) - (j >> 1) * (uint64_t)g.B;
        return (rem < g.B ? (uint32_t)rem : g.B) - 1;
    }
    if (!j) return 1;
    const uint64_t rem = chunk_len(g, t) - 1 - (j - 1) * (uint64_t)g.B;
    return rem < g.B ? (uint32_t)rem : g.B;
}

// First element of item `item` (n for item == nitems).
SZP_HD uint64_t item_start(const Geom &g, uint64_t item)
{
    if (item >= g.nitems) return g.n;
    uint32_t t;
    uint64_t j;
    locate(g, item, t, j);
    return item_elem(g, t, j);
}

// A tile whose nti items starting at cursor (ct, cj) are all full blocks of
// chunk ct, with B a multiple of 8: every octet then holds 8 values.
SZP_HD bool simple_tile(const Geom &g, uint32_t ct, uint64_t cj, uint32_t nti)
{
    if (g.mode != 1) return false;
    const uint64_t nfull = (chunk_len(g, ct) - 1) / g.B;
    return (g.B & 7) == 0 && cj >= 1 && nti >= 1 && cj + nti - 1 <= nfull;
}

// A fastMode-2 tile of whole units of B = 8 * spb values, all of chunk ct:
// it starts at a unit's head (cj even) and holds nti / 2 full units.  Its
// elements are contiguous from item_elem(g, ct, cj), unit d of fast round r
// (K units per round) at r * 2048 + d * B.
SZP_HD bool unit_tile(const Geom &g, uint32_t ct, uint64_t cj, uint32_t nti)
{
    if (g.mode != 2 || (g.B & 7) || (cj & 1) || (nti & 1) || !nti) return false;
    const uint64_t nfull = chunk_len(g, ct) / g.B;  // the chunk's full units
    return cj / 2 + nti / 2 <= nfull;
}

// Element offset, from a simple tile's first element, of the octet of slot
// (d, o) in round r.
SZP_HD uint32_t simple_octet_off(const Geom &g, uint32_t r, uint32_t d, uint32_t o)
{
    return (r * g.K + d) * g.B + 8 * o;
}

// The simple compress path's wave mapping (spb a power of two up to 32): lane
// l of wave w holds, in step s, octet l % spb of the tile's block
// simple_wave_block(...), i.e. the tile's elements simple_wave_off(...) +
// [0, 8): wave w covers the tile's w-th quarter (K blocks, 2048 elements) in
// four steps of 512.
SZP_HD uint32_t simple_wave_block(const Geom &g, uint32_t w, uint32_t s, uint32_t l)
{
    return w * g.K + s * (64 / g.spb) + l / g.spb;
}
SZP_HD uint32_t simple_wave_off(uint32_t w, uint32_t s, uint32_t l)
{
    return 2048 * w + 512 * s + 8 * l;
}

// One thread's octet of a round: up to 8 consecutive elements of one item.
// ---- dispatch order of the compress tiles ----
// Tiles that are not simple (a chunk head, a chunk's short last block) take
// the general path, which publishes its byte count for the look-back later
// than a simple tile; every later tile waits for it.  They are therefore
// dispatched first: workgroup b < n runs general tile t[b], the others run
// the remaining tiles in increasing order.  n = 0 keeps blockIdx order.
constexpr int kMaxGenTiles = 64;

struct TileOrder {
    uint32_t n;
    uint32_t t[kMaxGenTiles];  // ascending
};

// Tiles holding a chunk head or a chunk's short last block, and with
// with_last the field's last tile (ascending, each once).  n = 0 when there
// are more than kMaxGenTiles, or when every tile takes the general path
// anyway (B % 8 != 0, or fastMode 2).
inline TileOrder general_tiles(const Geom &g, bool with_last = false)
{
    TileOrder o;
    o.n = 0;
    if ((g.B & 7) || g.mode != 1) return o;
    const uint64_t per_tile = (uint64_t)g.K * g.R;
    auto add = [&](uint64_t tile) -> bool {
        if (o.n && o.t[o.n - 1] == tile) return true;
        if (o.n == (uint32_t)kMaxGenTiles) return false;
        o.t[o.n++] = (uint32_t)tile;
        return true;
    };
    for (uint32_t t = 0; t < g.T; ++t) {
        const uint64_t s = (uint64_t)t * g.c0;
        const uint64_t c = t + 1 < g.T ? g.c0 : g.clast;
        if (!c) continue;
        bool ok = add(s / per_tile);
        if (ok && (chunk_len(g, t) - 1) % g.B) ok = add((s + c - 1) / per_tile);
        if (!ok) {
            o.n = 0;
            return o;
        }
    }
    if (with_last && g.ntiles && !add(g.ntiles - 1)) o.n = 0;
    return o;
}

// ---- the pipelined decode with the side index ----
// A workgroup takes at most kPipeMaxSteps tiles (its LDS list of deferred
// tiles has that many entries).  Its tiles come from one of kPipeClasses
// ticket counters, ctl[kPipeCtr0 + kPipeCtrStride * c] (inside the 256-byte
// control block the caller's workspace starts with).
constexpr uint32_t kPipeMaxSteps = 2048;
constexpr uint32_t kPipeClasses = 8;
constexpr uint32_t kPipeCtr0 = 8, kPipeCtrStride = 7;
static_assert(kPipeCtr0 + kPipeCtrStride * (kPipeClasses - 1) < 64, "counters inside the control block");

// True when grid G lets the classes' workgroups, taking kPipeMaxSteps tiles
// each, take every tile: floor(G / C) * kPipeMaxSteps >= ceil(ntiles / C)
// for C = min(G, kPipeClasses) classes.
inline bool pipe_grid_covers(const Geom &g, uint64_t G)
{
    if (G == 0) return g.ntiles == 0;
    const uint64_t C = G < kPipeClasses ? G : kPipeClasses;
    return (G / C) * kPipeMaxSteps >= (g.ntiles + C - 1) / C;
}

// Grid of the pipelined decode: the resident workgroups, at least one per
// general tile (each decodes one), never more than the tiles unless the
// tickets need more workgroups (pipe_grid_covers).
inline uint64_t pipe_grid_size(const Geom &g, const TileOrder &ord, uint32_t resident)
{
    uint64_t G = g.ntiles < resident ? g.ntiles : resident;
    if (G < ord.n) G = ord.n;
    if (!pipe_grid_covers(g, G)) {
        const uint64_t per_class = (g.ntiles + kPipeClasses - 1) / kPipeClasses;
        const uint64_t need = kPipeClasses * ((per_class + kPipeMaxSteps - 1) / kPipeMaxSteps);
        if (G < need) G = need;
    }
    return G;
}

// The tile that workgroup b runs.
SZP_HD uint64_t tile_of_block(const TileOrder &o, uint64_t b)
{
    if (o.n == 0) return b;
    if (b < o.n) return o.t[b];
    // m-th tile outside the list: the first k with t[k] - k > m gives m + k.
    const uint64_t m = b - o.n;
    uint32_t lo = 0, hi = o.n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if ((uint64_t)o.t[mid] - mid > m) hi = mid; else lo = mid + 1;
    }
    return m + lo;
}

struct Oct {
    uint64_t e0;  // first element
    uint32_t m;   // elements in the octet (0: nothing to do)
    uint32_t L;   // elements of the item
    uint32_t t;   // chunk of the item
    bool live;    // the slot's item exists in this round
    bool head;    // the item is a head
    bool first;   // the item is its chunk's first head (its offset goes in the header)
};

SZP_HD Oct octet_of(const Geom &g, uint32_t ct, uint64_t cj, uint32_t nit, uint32_t d, uint32_t o)
{
    Oct s{0, 0, 0, 0, false, false, false};
    if (d >= g.K || d >= nit) return s;
    uint32_t t = ct;
    uint64_t j = cj;
    advance(g, t, j, d);
    s.live = true;
    s.t = t;
    s.head = item_head(g, j);
    s.first = j == 0;
    s.L = item_len(g, t, j);
    const uint64_t es = item_elem(g, t, j);
    if (s.head) {
        s.e0 = es;
        s.m = o == 0 ? 1u : 0u;
    } else {
        s.e0 = es + 8ull * o;
        s.m = 8 * o < s.L ? (s.L - 8 * o < 8u ? s.L - 8 * o : 8u) : 0u;
    }
    return s;
}

// ---- bitmap helpers (one bit per stream byte of a parse tile) ----
SZP_HD bool bm_test(const uint32_t *bm, uint32_t p)
{
    return (bm[p >> 5] >> (p & 31)) & 1u;
}
SZP_HD void bm_set(uint32_t *bm, uint32_t p) { bm[p >> 5] |= 1u << (p & 31); }
SZP_HD void bm_clear(uint32_t *bm, uint32_t p) { bm[p >> 5] &= ~(1u << (p & 31)); }
// Bits [p, p + k) as a mask of word w (k may be 0).
SZP_HD uint32_t bm_mask(uint32_t w, uint32_t p, uint32_t k)
{
    const uint32_t lo = w << 5, a = p > lo ? p - lo : 0u, e = p + k - lo;
    const uint32_t b = e < 32u ? e : 32u;
    if (a >= b) return 0u;
    const uint32_t hi = b == 32u ? 0xFFFFFFFFu : ((1u << b) - 1u);
    return hi & ~((1u << a) - 1u);
}
SZP_HD void bm_set_range(uint32_t *bm, uint32_t p, uint32_t k)
{
    if (!k) return;
    for (uint32_t w = p >> 5; w <= (p + k - 1) >> 5; ++w) bm[w] |= bm_mask(w, p, k);
}
// How the chain walks below set bitmap bits: OR a mask into a word.  The
// kernel passes an LDS atomic OR (its single-lane walks then do not wait for
// each word's old value); the host emulation a plain OR.
struct BmOr {
    SZP_HD void operator()(uint32_t *w, uint32_t m) const { *w |= m; }
};
template <typename OR>
SZP_HD void bm_set_by(uint32_t *bm, uint32_t p, OR op)
{
    op(&bm[p >> 5], 1u << (p & 31));
}
template <typename OR>
SZP_HD void bm_set_range_by(uint32_t *bm, uint32_t p, uint32_t k, OR op)
{
    if (!k) return;
    for (uint32_t w = p >> 5; w <= (p + k - 1) >> 5; ++w) op(&bm[w], bm_mask(w, p, k));
}
SZP_HD void bm_clear_range(uint32_t *bm, uint32_t p, uint32_t k)
{
    if (!k) return;
    for (uint32_t w = p >> 5; w <= (p + k - 1) >> 5; ++w) bm[w] &= ~bm_mask(w, p, k);
}
// First set bit in [p, p + k), or p + k.
SZP_HD uint32_t bm_first(const uint32_t *bm, uint32_t p, uint32_t k)
{
    if (!k) return p;
    for (uint32_t w = p >> 5; w <= (p + k - 1) >> 5; ++w) {
        const uint32_t m = bm[w] & bm_mask(w, p, k);
        if (m) return (w << 5) + __builtin_ctz(m);
    }
    return p + k;
}

// Zero bytes starting at buf[p], at most v - p (b == 0 blocks are single
// zero bytes, so a run of k zeros on a chain is k consecutive blocks).
// Reads up to 3 bytes past v, inside the buffer's slack.
SZP_HD uint32_t zero_run(const uint8_t *buf, uint32_t p, uint32_t v)
{
    uint32_t k = 0;
    while (p + k < v) {
        uint32_t w;
        __builtin_memcpy(&w, buf + p + k, 4);
        if (w) {
            k += (uint32_t)__builtin_ctz(w) >> 3;
            break;
        }
        k += 4;
    }
    return k < v - p ? k : v - p;
}

// Length of a block with width byte b and B values.
SZP_HD uint32_t blen(uint32_t b, uint32_t B) { return block_bytes(b, B); }

// A chain step of the parse.  Mode 1 steps over one block (width byte at p);
// mode 2 over one unit: 4 head bytes, then a block of B - 1 values whose width
// byte is at p + 4.  Returns the width byte (> 32: no item can start at p)
// and sets len to the step.
SZP_HD uint32_t parse_step(const uint8_t *tb, uint32_t p, uint32_t B, uint32_t mode,
                           uint32_t &len)
{
    if (mode == 2) {
        const uint32_t b = tb[p + 4];
        len = 4 + blen(b, B - 1);
        return b;
    }
    const uint32_t b = tb[p];
    len = blen(b, B);
    return b;
}

}  // namespace szp_amd

namespace szp_amd {

// ---- per-thread walks of the block-boundary parse (used by k_parse and by
// the host emulation in tests/cpu/parse_check.cc) ----
//
// tb[p] is byte p of the parse tile; [u, v) is one sub-tile; a chain's
// positions are tile offsets.  spec/pre are bitmaps over the tile.
//
// The speculative chain of sub-tile [u, v) is the chain of the first
// candidate c in [u, min(u + W, v)) that leaves the sub-tile (reaches v)
// without meeting a width byte > 32.  On the GPU the 64 lanes of a wave try
// 64 candidates at once with parse_chain_leaves and the lowest one that
// leaves marks its chain with parse_mark_chain; the host emulation tries the
// candidates in order with the same two functions.

// Walks the chain from c; true when it reaches v (exit = its first position
// >= v), false when it meets a width byte > 32.
SZP_HD bool parse_chain_leaves(const uint8_t *tb, uint32_t c, uint32_t v, uint32_t B,
                               uint32_t &exit, uint32_t mode = 1)
{
    uint32_t p = c;
    while (p < v) {
        uint32_t len;
        const uint32_t b = parse_step(tb, p, B, mode, len);
        if (b > 32) return false;
        p += (b || mode == 2) ? len : zero_run(tb, p, v);
    }
    exit = p;
    return true;
}

// Marks the block starts of the chain from c (which leaves the sub-tile) in
// spec; returns their count.
template <typename OR = BmOr>
SZP_HD uint32_t parse_mark_chain(const uint8_t *tb, uint32_t *spec, uint32_t c, uint32_t v,
                                 uint32_t B, uint32_t mode = 1, OR op = OR())
{
    uint32_t p = c, n = 0;
    while (p < v) {
        uint32_t len;
        const uint32_t b = parse_step(tb, p, B, mode, len);
        if (mode == 2) {
            bm_set_by(spec, p, op);
            ++n;
            p += len;
        } else if (b == 0) {
            const uint32_t k = zero_run(tb, p, v);
            bm_set_range_by(spec, p, k, op);
            n += k;
            p += k;
        } else {
            bm_set_by(spec, p, op);
            ++n;
            p += len;
        }
    }
    return n;
}

// Chain from candidate c: 0 if it joins spec or dies inside [u, v), else the
// position (>= v) where it leaves.
SZP_HD uint32_t parse_cand_walk(const uint8_t *tb, const uint32_t *spec, uint32_t c, uint32_t v,
                                uint32_t B, uint32_t mode = 1)
{
    uint32_t p = c;
    while (true) {
        if (bm_test(spec, p)) return 0;
        uint32_t len;
        const uint32_t b = parse_step(tb, p, B, mode, len);
        if (b > 32) return 0;
        if (b == 0 && mode == 1) {
            const uint32_t k = zero_run(tb, p, v);
            if (bm_first(spec, p, k) < p + k) return 0;
            p += k;
        } else {
            p += len;
        }
        if (p >= v) return p;
    }
}

// True chain from p in [u, v): marks its blocks before it joins spec in pre.
// Returns 1 merged (p = the join position), 2 broken (width byte > 32 at p),
// 0 left the sub-tile (p >= v).  npre counts the marked blocks.
template <typename OR = BmOr>
SZP_HD int parse_true_walk(const uint8_t *tb, const uint32_t *spec, uint32_t *pre, uint32_t &p,
                           uint32_t v, uint32_t B, uint32_t &npre, uint32_t mode = 1, OR op = OR())
{
    npre = 0;
    while (p < v) {
        if (bm_test(spec, p)) return 1;
        uint32_t len;
        const uint32_t b = parse_step(tb, p, B, mode, len);
        if (b > 32) return 2;
        if (b == 0 && mode == 1) {
            const uint32_t k = zero_run(tb, p, v);
            const uint32_t q = bm_first(spec, p, k);
            bm_set_range_by(pre, p, q - p, op);
            npre += q - p;
            if (q < p + k) {
                p = q;
                return 1;
            }
            p += k;
        } else {
            bm_set_by(pre, p, op);
            ++npre;
            p += len;
        }
    }
    return 0;
}

// Rank of p among the set bits of bm within its sub-tile (rk holds, per word,
// the set bits of the sub-tile's earlier words).
SZP_HD uint32_t parse_rank(const uint32_t *bm, const uint16_t *rk, uint32_t p)
{
    const uint32_t w = p >> 5;
    return rk[w] + (uint32_t)__builtin_popcount(bm[w] & ((1u << (p & 31)) - 1u));
}

}  // namespace szp_amd
