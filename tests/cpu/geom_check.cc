// Host-side check of the kernels' index arithmetic: walks every tile, round
// and thread slot of the compress and decode kernels (the same szp_geom.h
// functions the kernels call) and asserts that every element a thread loads
// lies inside the field and inside the thread's own item, and that the
// octets of a round cover the round's elements exactly once.
// Usage: geom_check            (built-in geometry list; prints a summary)
#include <stdio.h>
#include <stdlib.h>
#include <stdint.h>
#include <vector>
#include <random>
#include <algorithm>

#include "szp_geom.h"

using namespace szp_amd;

static uint64_t g_checks = 0;
static int g_fail = 0;

#define CHECK(cond, ...)                                                          \
    do {                                                                          \
        ++g_checks;                                                               \
        if (!(cond)) {                                                            \
            if (g_fail < 20) {                                                    \
                fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__);              \
                fprintf(stderr, __VA_ARGS__);                                     \
                fprintf(stderr, "\n");                                            \
            }                                                                     \
            ++g_fail;                                                             \
        }                                                                         \
    } while (0)

// One compress tile: the element set each slot touches, round by round.
static void check_compress_tile(const Geom &g, uint64_t tile, std::vector<uint8_t> *cover)
{
    const uint64_t tI0 = tile * (uint64_t)g.K * g.R;
    const uint64_t tIN = tI0 + (uint64_t)g.K * g.R < g.nitems ? tI0 + (uint64_t)g.K * g.R : g.nitems;
    uint32_t ct;
    uint64_t cj;
    locate(g, tI0, ct, cj);
    const uint32_t nti = (uint32_t)(tIN - tI0);
    const bool simple = simple_tile(g, ct, cj, nti);
    if (unit_tile(g, ct, cj, nti) && g.spb <= 32 && (g.spb & (g.spb - 1)) == 0) {
        // The fastMode-2 unit fast path (instantiated for spb a power of two up
        // to 32): thread tid of fast round r loads the 8 values
        // Et + r * 2048 + 8 * tid + [0, 8) of unit r * K + tid / spb.
        CHECK(!simple, "a tile both simple and of units");
        const uint64_t Et = item_elem(g, ct, cj), lo = chunk_lo(g, ct);
        CHECK(Et == lo + (cj / 2) * g.B, "unit tile start");
        const uint32_t nu = nti / 2;
        CHECK((uint64_t)nu <= (uint64_t)g.K * (kMaxRounds / 2), "unit tile holds more than its fast rounds");
        for (uint32_t r = 0; r < (uint32_t)kMaxRounds / 2; ++r)
            for (uint32_t tid = 0; tid < (uint32_t)kWG; ++tid) {
                const uint32_t d = tid / g.spb;
                if (r * g.K + d >= nu) continue;
                const uint64_t e = Et + (uint64_t)r * 2048u + 8u * tid;
                const uint64_t u = (e - lo) / g.B;  // unit of the octet within the chunk
                CHECK(e + 8 <= g.n && e + 8 <= lo + chunk_len(g, ct), "unit octet past its chunk");
                CHECK((e + 7 - lo) / g.B == u && u == cj / 2 + r * g.K + d,
                      "unit octet outside its unit");
            }
    }
    if (simple && g.spb <= 32 && (g.spb & (g.spb - 1)) == 0) {
        // The simple path's wave mapping: every live lane's octet is its
        // block's octet l % spb, inside the field; idle lanes read the tile's
        // first octet.
        const uint64_t Et = item_elem(g, ct, cj);
        CHECK(Et >= 1 && Et + 8 <= g.n, "simple tile's first octet out of field");
        for (uint32_t w = 0; w < (uint32_t)kWG / 64; ++w)
            for (uint32_t s = 0; s < 4; ++s)
                for (uint32_t l = 0; l < 64; ++l) {
                    const uint32_t b = simple_wave_block(g, w, s, l);
                    if (b >= nti) continue;
                    const uint32_t off = simple_wave_off(w, s, l);
                    CHECK(off == simple_octet_off(g, b / g.K, b % g.K, l % g.spb),
                          "wave mapping (tile %llu w %u s %u l %u)", (unsigned long long)tile, w, s, l);
                    CHECK(Et + off + 8 <= g.n, "wave octet out of field");
                }
    }
    uint32_t rt = ct;
    uint64_t rj = cj;
    for (uint32_t r = 0; r < g.R; ++r) {
        const uint64_t ri0 = tI0 + (uint64_t)r * g.K;
        if (ri0 >= tIN) break;
        const uint32_t nit = (uint32_t)((tIN - ri0) < g.K ? tIN - ri0 : g.K);
        // The deferred path recomputes the cursor with locate(); it must agree.
        uint32_t lt;
        uint64_t lj;
        locate(g, ri0, lt, lj);
        CHECK(lt == rt && lj == rj, "cursor mismatch tile %llu round %u", (unsigned long long)tile, r);
        const uint64_t rE0 = item_elem(g, rt, rj), rE1 = item_start(g, ri0 + nit);
        CHECK(rE1 - rE0 <= 2048, "round spans %llu elements", (unsigned long long)(rE1 - rE0));
        for (uint32_t tid = 0; tid < (uint32_t)kWG; ++tid) {
            const uint32_t d = (tid * g.spb_magic) >> 16;
            const uint32_t o = tid - d * g.spb;
            CHECK(d == tid / g.spb, "slot magic");
            const Oct s = octet_of(g, rt, rj, nit, d, o);
            if (simple) {
                const uint64_t Et = item_elem(g, ct, cj);
                const bool live = d < g.K && r * g.K + d < nti;
                CHECK(live == s.live, "simple liveness differs from the general path");
                if (!live) continue;
                const uint32_t off = simple_octet_off(g, r, d, o);
                CHECK(s.m == 8 && !s.head && s.e0 == Et + off,
                      "simple octet (tile %llu r %u d %u o %u) off %u vs general e0 %llu m %u",
                      (unsigned long long)tile, r, d, o, off, (unsigned long long)s.e0, s.m);
                CHECK(Et + off >= 1 && Et + off + 8 <= g.n, "simple load out of field");
            }
            if (s.live && !s.head && s.L == 0) {
                // A fastMode-2 unit of one value: its block is empty.
                CHECK(g.mode == 2 && s.m == 0, "empty block outside fastMode 2");
            }
            if (!s.m) continue;
            CHECK(s.e0 + s.m <= g.n, "octet past the field end");
            if (!s.head) CHECK(s.e0 >= 1, "block octet at element 0");
            if (!s.head && g.mode == 2) {
                // The element before a block's octet lies in the same unit.
                uint32_t t2 = rt;
                uint64_t j2 = rj;
                advance(g, t2, j2, d);
                CHECK((j2 & 1) && s.e0 - 1 >= item_elem(g, t2, j2 - 1),
                      "block octet reads before its unit");
                CHECK(s.L + 1 <= g.B, "unit longer than B");
            }
            CHECK(s.e0 >= rE0 && s.e0 + s.m <= rE1, "octet outside its round");
            if (cover) {
                for (uint32_t i = 0; i < s.m; ++i) {
                    uint8_t &c = (*cover)[s.e0 + i];
                    CHECK(c == 0, "element %llu covered twice", (unsigned long long)(s.e0 + i));
                    c = 1;
                }
            }
        }
        advance(g, rt, rj, nit);
    }
}

// Dispatch order of the compress tiles: general tiles first, then the rest
// ascending; a bijection of [0, ntiles); listed tiles are exactly the tiles
// that are not simple (when the list is in use).
static void check_order(const Geom &g, bool with_last)
{
    const TileOrder o = general_tiles(g, with_last);
    for (uint32_t k = 0; k + 1 < o.n; ++k) CHECK(o.t[k] < o.t[k + 1], "general tiles not ascending");
    for (uint32_t k = 0; k < o.n; ++k) CHECK(o.t[k] < g.ntiles, "general tile past the end");
    const uint64_t per_tile = (uint64_t)g.K * g.R;
    auto is_simple = [&](uint64_t tile) {
        const uint64_t tI0 = tile * per_tile;
        const uint64_t tIN = tI0 + per_tile < g.nitems ? tI0 + per_tile : g.nitems;
        uint32_t ct;
        uint64_t cj;
        locate(g, tI0, ct, cj);
        return simple_tile(g, ct, cj, (uint32_t)(tIN - tI0));
    };
    auto listed = [&](uint64_t tile) {
        for (uint32_t k = 0; k < o.n; ++k)
            if (o.t[k] == tile) return true;
        return false;
    };
    const bool in_use = o.n > 0;
    if (g.ntiles <= (4ull << 20)) {
        std::vector<uint8_t> seen(g.ntiles, 0);
        uint64_t prev = 0;
        for (uint64_t b = 0; b < g.ntiles; ++b) {
            const uint64_t t = tile_of_block(o, b);
            CHECK(t < g.ntiles, "block %llu maps past the end", (unsigned long long)b);
            if (t >= g.ntiles) continue;
            CHECK(!seen[t], "tile %llu dispatched twice", (unsigned long long)t);
            seen[t] = 1;
            if (b > o.n) CHECK(t > prev, "non-general tiles out of order");
            if (b >= o.n) prev = t;
            const bool special = !is_simple(t) || (with_last && t == g.ntiles - 1);
            if (in_use) CHECK(listed(t) == special, "tile %llu listed %d simple %d",
                              (unsigned long long)t, (int)listed(t), (int)is_simple(t));
        }
    } else {
        std::mt19937_64 rng(g.ntiles);
        for (int k = 0; k < 20000; ++k) {
            const uint64_t b = k < 200 ? (uint64_t)k : rng() % g.ntiles;
            const uint64_t t = tile_of_block(o, b);
            CHECK(t < g.ntiles, "block maps past the end");
            const bool special = t < g.ntiles && (!is_simple(t) || (with_last && t == g.ntiles - 1));
            if (in_use && t < g.ntiles) CHECK(listed(t) == special, "listed vs simple");
        }
    }
    // Without a list, every tile takes blockIdx order.
    if (!in_use) CHECK(tile_of_block(o, 5) == 5, "identity order");
    // The pipelined decode's grid: workgroup w takes tiles of its class by
    // ticket (check_pipe_tickets) and the general tile o.t[w] for w < o.n.
    if (with_last && in_use && g.ntiles) {
        for (uint32_t resident : {1u, 3u, 1024u, 1u << 20}) {
            const uint64_t G = pipe_grid_size(g, o, resident);
            CHECK(G >= 1 && G <= g.ntiles, "pipe grid %llu for %llu tiles", (unsigned long long)G,
                  (unsigned long long)g.ntiles);
            CHECK(G >= o.n, "pipe grid below the general tiles");
            CHECK(pipe_grid_covers(g, G), "pipe grid leaves tiles untaken");
        }
    }
}

// The pipelined decode's tickets (k_decode_pipe), stepped workgroup by
// workgroup in a random order: every tile of every class must be decoded
// exactly once and no workgroup may take more than kPipeMaxSteps tiles.
static void check_pipe_tickets(uint64_t ntiles, uint64_t G, uint64_t seed)
{
    const uint32_t C = (uint32_t)std::min<uint64_t>(G, kPipeClasses);
    struct Wg { uint32_t kcur, knxt, step; bool done; };
    std::vector<uint32_t> ctr(C, 0), seen(ntiles, 0);
    std::vector<Wg> w(G);
    auto ctiles = [&](uint32_t c) -> uint64_t { return c < ntiles ? (ntiles - c + C - 1) / C : 0; };
    auto nper = [&](uint32_t c) -> uint32_t { return (uint32_t)((G - c + C - 1) / C); };
    for (uint64_t b = 0; b < G; ++b) {
        const uint32_t c = (uint32_t)(b % C), r = (uint32_t)(b / C), n = nper(c);
        w[b] = {r, r + n, 0, false};
    }
    std::mt19937_64 rng(seed);
    uint64_t live = G;
    while (live) {
        const uint64_t b = rng() % G;
        Wg &x = w[b];
        if (x.done) continue;
        const uint32_t c = (uint32_t)(b % C);
        if (x.kcur >= ctiles(c)) {
            x.done = true;
            --live;
            continue;
        }
        ++seen[c + (uint64_t)C * x.kcur];
        const bool asked = x.knxt < ctiles(c) && x.step + 2 < kPipeMaxSteps;
        const uint32_t k2 = asked ? 2 * nper(c) + ctr[c]++ : ~0u;
        x.kcur = x.knxt;
        x.knxt = k2;
        ++x.step;
        CHECK(x.step <= kPipeMaxSteps, "workgroup took more than kPipeMaxSteps tiles");
    }
    bool all = true;
    for (uint64_t t = 0; t < ntiles; ++t) {
        CHECK(seen[t] <= 1, "tile %llu decoded %u times", (unsigned long long)t, seen[t]);
        all = all && seen[t] == 1;
    }
    CHECK(all, "tickets left tiles undecoded (ntiles %llu, G %llu)",
                            (unsigned long long)ntiles, (unsigned long long)G);
}

static void check_geometry(uint64_t n, int B, int T, bool full, uint32_t mode = 1)
{
    Geom g;
    if (!make_geom(n, B, T, &g, 4, mode)) {
        CHECK(false, "make_geom rejected n=%llu B=%d T=%d", (unsigned long long)n, B, T);
        return;
    }
    CHECK((uint64_t)g.K * g.spb <= (uint64_t)kWG, "K * spb > kWG");
    CHECK(g.K * (g.B + 0ull) <= 2048 || g.B > 2048, "round holds more than 2048 elements");
    // Worst-case bytes of one round fit the decode round buffer.
    CHECK((uint64_t)g.K * block_bytes(32, g.B) <= (uint64_t)kOutBufBytes, "round bytes > buffer");
    check_order(g, false);
    check_order(g, true);
    std::vector<uint8_t> cover;
    if (full) cover.assign(n, 0);
    if (full) {
        for (uint64_t t = 0; t < g.ntiles; ++t) check_compress_tile(g, t, &cover);
        for (uint64_t e = 0; e < n; ++e) CHECK(cover[e] == 1, "element %llu not covered", (unsigned long long)e);
    } else {
        std::mt19937_64 rng(n ^ (uint64_t)B * 131 ^ (uint64_t)T);
        std::vector<uint64_t> tiles;
        for (uint64_t t = 0; t < g.ntiles && t < 300; ++t) tiles.push_back(t);
        for (uint64_t t = g.ntiles > 300 ? g.ntiles - 300 : 0; t < g.ntiles; ++t) tiles.push_back(t);
        for (int k = 0; k < 600 && g.ntiles; ++k) tiles.push_back(rng() % g.ntiles);
        for (uint64_t t : tiles) check_compress_tile(g, t, nullptr);
    }
}

int main()
{
    const int Bs[] = {1, 2, 3, 7, 8, 9, 16, 24, 31, 32, 64, 100, 127, 128, 255, 256, 512, 1000, 1024, 2048};
    const int Ts[] = {1, 2, 3, 7, 8, 16};
    const uint64_t ns[] = {1, 2, 7, 8, 9, 15, 16, 17, 40, 100, 1000, 4097, 20001, 65537, 70000, 300001};
    for (uint64_t n : ns)
        for (int B : Bs)
            for (int T : Ts) {
                check_geometry(n, B, T, true);
                check_geometry(n, B, T, true, 2);
            }
    // Full-size shapes, sampled tiles.
    check_geometry(134217728ull, 128, 16, false);
    check_geometry(134217728ull, 32, 16, false);
    check_geometry(67108864ull, 128, 8, false);
    check_geometry(4294967296ull, 64, 16, false);
    check_geometry(4294967296ull, 64, 1, false);
    check_geometry(4194307ull, 128, 8, false);
    check_geometry(134217728ull, 128, 16, false, 2);
    check_geometry(4294967296ull, 64, 16, false, 2);
    // Ticket coverage: small and ragged grids, a grid below the classes, the
    // 512^3 field's 16385 tiles on 1024 workgroups, 16384 tiles on 8, and grids that
    // cover exactly (every workgroup takes kPipeMaxSteps tiles).
    const uint64_t cases[][2] = {{1, 1},       {5, 3},    {9, 8},    {100, 13},   {1000, 7},
                                 {16385, 1024}, {16384, 8}, {16385, 17}, {32768, 16}, {32769, 24},
                                 {65536, 32}};
    for (const auto &cs : cases) {
        Geom g;
        g.ntiles = cs[0];
        CHECK(pipe_grid_covers(g, cs[1]), "case grid %llu does not cover %llu tiles",
              (unsigned long long)cs[1], (unsigned long long)cs[0]);
        for (uint64_t seed = 1; seed <= 3; ++seed) check_pipe_tickets(cs[0], cs[1], seed);
    }
    {
        Geom g;
        g.ntiles = 32769;
        CHECK(!pipe_grid_covers(g, 16), "a grid one tile short must be rejected");
    }
    printf("geom_check: %llu checks, %d failures\n", (unsigned long long)g_checks, g_fail);
    return g_fail ? 1 : 0;
}
