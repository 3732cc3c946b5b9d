// Geometry of the item/tile decomposition shared by the HIP kernels and the
// host-compiled index checks (tests/cpu/geom_check.cc).  Everything here is
// plain integer arithmetic so that the same code can run on the CPU.
#pragma once

#include <stdint.h>

#ifdef __HIPCC__
#define SZP_HD __host__ __device__ __forceinline__
#else
#define SZP_HD inline
#endif

namespace szp_amd {

constexpr int kWG = 256;                 // threads per workgroup (4 wave64)
constexpr int kMaxBlock = 2048;          // largest SZp block size handled on the GPU
constexpr int kOutBufBytes = 8704 + 64;  // LDS bytes for one round's encoded blocks
constexpr int kMaxRounds = 8;            // rounds of K items per tile
constexpr int kOutCap = 32768;           // LDS bytes buffering a compress tile before placement
constexpr int kDecTileBytes = 24576;     // LDS bytes holding a decode tile's stream bytes

// Geometry of the item sequence of one field.  The stream body is the
// concatenation, in element order, of "items": the raw head value of each
// non-empty chunk (4 bytes) followed by that chunk's SZp blocks.  A tile is R
// rounds of K consecutive items; every item owns spb = ceil(B/8) thread slots
// ("octets").
struct Geom {
    uint64_t n;        // elements
    uint64_t tbs;      // floor(n / T), the size of chunks 0..T-2
    uint64_t c0;       // items in chunks 0..T-2 (0 when tbs == 0)
    uint64_t clast;    // items in chunk T-1
    uint64_t nitems;   // (T-1)*c0 + clast
    uint64_t nb0;      // SZp blocks in chunks 0..T-2
    uint64_t nblast;   // SZp blocks in chunk T-1
    uint64_t ntiles;   // ceil(nitems / (K * R))
    uint64_t body;     // 4 + 8T: offset of chunk 0 in the stream
    uint32_t T;        // chunks (the reference's OpenMP team size)
    uint32_t B;        // SZp block size
    uint32_t spb;      // slots per item
    uint32_t K;        // items per round
    uint32_t spb_magic;  // slot / spb == (slot * spb_magic) >> 16 for slot < 256
    uint32_t R;          // rounds per tile (1 when chunks are smaller than a tile)
};

// Returns false when (n, B, T) is outside what the kernels handle.
inline bool make_geom(uint64_t n, int B, int T, Geom *g)
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
    // Tiles of kMaxRounds rounds, unless chunks are so short that a tile would
    // hold more chunk heads than one workgroup records (at most kWG).
    g->R = (g->c0 == 0 || g->c0 * (kWG - 1) >= (uint64_t)g->K * kMaxRounds) ? kMaxRounds : 1;
    const uint64_t per_tile = (uint64_t)g->K * g->R;
    g->ntiles = (g->nitems + per_tile - 1) / per_tile;
    g->body = 4 + 8ull * (uint64_t)T;
    return true;
}

// Bytes of one block of L values at width b (SZp/src/szp_TypeManager.cc:20-31).
SZP_HD uint32_t block_bytes(uint32_t b, uint32_t L)
{
    if (!b) return 1;
    return 1 + ((L + 7) >> 3) + (b >> 3) * L + (((b & 7) * L + 7) >> 3);
}

// Item -> (chunk t, index j within the chunk; j == 0 is the chunk's head).
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

// First element of item (t, j).
SZP_HD uint64_t item_elem(const Geom &g, uint32_t t, uint64_t j)
{
    return chunk_lo(g, t) + (j ? 1 + (j - 1) * (uint64_t)g.B : 0);
}

// Elements of item (t, j).
SZP_HD uint32_t item_len(const Geom &g, uint32_t t, uint64_t j)
{
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
    const uint64_t nfull = (chunk_len(g, ct) - 1) / g.B;
    return (g.B & 7) == 0 && cj >= 1 && nti >= 1 && cj + nti - 1 <= nfull;
}

// Element offset, from a simple tile's first element, of the octet of slot
// (d, o) in round r.
SZP_HD uint32_t simple_octet_off(const Geom &g, uint32_t r, uint32_t d, uint32_t o)
{
    return (r * g.K + d) * g.B + 8 * o;
}

// One thread's octet of a round: up to 8 consecutive elements of one item.
struct Oct {
    uint64_t e0;  // first element
    uint32_t m;   // elements in the octet (0: nothing to do)
    uint32_t L;   // elements of the item
    uint32_t t;   // chunk of the item
    bool live;    // the slot's item exists in this round
    bool head;    // the item is a chunk head
};

SZP_HD Oct octet_of(const Geom &g, uint32_t ct, uint64_t cj, uint32_t nit, uint32_t d, uint32_t o)
{
    Oct s{0, 0, 0, 0, false, false};
    if (d >= g.K || d >= nit) return s;
    uint32_t t = ct;
    uint64_t j = cj;
    advance(g, t, j, d);
    s.live = true;
    s.t = t;
    s.head = j == 0;
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

}  // namespace szp_amd
