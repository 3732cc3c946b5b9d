// Host emulation of the block-boundary parse of a stream that carries no
// index (k_dsetup + k_parse in szp_amd/csrc/szp_kernels.hip).  The per-lane
// walks (parse_chain_leaves, parse_mark_chain, parse_cand_walk,
// parse_true_walk, parse_rank) are the szp_geom.h functions the kernel calls;
// this file replays the kernel's glue tile by tile in blockIdx order (a wave's
// 64 candidates at a time become one candidate at a time, lowest first, which
// picks the same chain), with every LDS array allocated at its exact size so
// that AddressSanitizer catches any access outside it.
//
// For each stream it compares the recovered block positions with a plain
// serial walk of the chunks, and the kernel's stream-error verdict with the
// serial walk's.  Streams are read from a case list written by the test:
//   parse_check <list>   lines: "<file> <n> <B> <T> <fastMode>"
// (the stream length is the file size, which may exceed the real stream).
// fastMode 2 streams are parsed unit by unit (a 4-byte head and a block).
#include <stdio.h>
#include <stdlib.h>
#include <stdint.h>
#include <string.h>
#include <algorithm>
#include <memory>
#include <vector>

#include "szp_geom.h"

using namespace szp_amd;

static uint64_t rd64(const uint8_t *p)
{
    uint64_t v;
    memcpy(&v, p, 8);
    return v;
}

// Longest chain step and sub-tile bytes, as the launch code sizes them.
static uint32_t max_step(uint32_t B, uint32_t mode)
{
    return mode == 2 ? 4 + block_bytes(32, B - 1) : block_bytes(32, B);
}

static uint32_t subtile_bytes(uint32_t W)
{
    const uint32_t U = W > 4096 ? W : 4096;
    return (U + 255) & ~255u;
}

struct Result {
    bool bad = false;
    std::vector<uint64_t> pos;     // block positions, chunk-major
    std::vector<uint8_t> written;  // times each position was written
};

// Serial walk: the positions of every chunk's blocks and whether the stream
// passes the same validity rules the kernel enforces (chunk offsets in range,
// width bytes <= 32 on the chain, last block ending at the chunk end, or
// inside the buffer for the last chunk).
static Result serial_parse(const uint8_t *st, uint64_t len, const Geom &g)
{
    Result r;
    const uint64_t nblocks = (uint64_t)(g.T - 1) * g.nb0 + g.nblast;
    r.pos.assign(nblocks, ~0ull);
    if (len < g.body) {
        r.bad = true;
        return r;
    }
    for (uint32_t t = 0; t < g.T; ++t) {
        const uint64_t nt = chunk_len(g, t);
        const uint64_t off = rd64(st + 4 + 8ull * t);
        const uint64_t cs = g.body + off;
        if (nt == 0) {
            if (off != 0) r.bad = true;
            continue;
        }
        const uint64_t ce = (t + 1 < g.T) ? g.body + rd64(st + 4 + 8ull * (t + 1)) : len;
        if (off > len || cs + 4 > ce || ce > len) {
            r.bad = true;
            return r;
        }
        const uint64_t nb = (t == g.T - 1) ? g.nblast : g.nb0;
        if (g.mode == 2) {
            // Units: [q i32][width byte][...]; a one-value unit's width byte
            // is not read.
            uint64_t p = cs;
            if (cs + 5 > ce) {
                r.bad = true;
                return r;
            }
            for (uint64_t j = 0; j < nb; ++j) {
                if (p + 4 >= len || p >= ce) {
                    r.bad = true;
                    return r;
                }
                const uint64_t L = ((j + 1 < nb) ? g.B : nt - (nb - 1) * (uint64_t)g.B) - 1;
                const uint32_t b = st[p + 4];
                if ((L || j + 1 < nb) && b > 32) {
                    r.bad = true;
                    return r;
                }
                r.pos[(uint64_t)t * g.nb0 + j] = p;
                p += 4 + block_bytes(L ? b : 0, (uint32_t)L);
            }
            if (nb && (t + 1 < g.T ? p != ce : p > len)) {
                r.bad = true;
                return r;
            }
            continue;
        }
        uint64_t p = cs + 4;
        for (uint64_t j = 0; j < nb; ++j) {
            if (p >= ce) {
                r.bad = true;
                return r;
            }
            const uint32_t b = st[p];
            if (b > 32) {
                r.bad = true;
                return r;
            }
            r.pos[(uint64_t)t * g.nb0 + j] = p;
            const uint64_t L = (j + 1 < nb) ? g.B : (nt - 1) - (nb - 1) * (uint64_t)g.B;
            p += block_bytes(b, (uint32_t)L);
        }
        if (nb && (t + 1 < g.T ? p != ce : p > len)) {
            r.bad = true;
            return r;
        }
    }
    return r;
}

static uint64_t g_tiles = 0, g_sync = 0;

// The kernel's glue, one workgroup (tile) at a time in blockIdx order.
static Result emulate(const uint8_t *st, uint64_t len, const Geom &g)
{
    Result r;
    const uint64_t nblocks = (uint64_t)(g.T - 1) * g.nb0 + g.nblast;
    r.pos.assign(nblocks, ~0ull);
    r.written.assign(nblocks, 0);
    const uint32_t B = g.B, mode = g.mode;
    const uint32_t W = max_step(B, mode);
    const uint32_t U = subtile_bytes(W), S = 4 * U, NW = S / 32;

    // k_dsetup
    std::vector<uint64_t> cstart(g.T), aend(g.T), ttab(g.T + 1);
    uint64_t tiles = 0;
    bool bad = len < g.body;
    for (uint32_t t = 0; t < g.T && !bad; ++t) {
        const uint64_t nt = chunk_len(g, t);
        const uint64_t off = rd64(st + 4 + 8ull * t);
        const uint64_t cs = g.body + off;
        ttab[t] = tiles;
        cstart[t] = cs;
        aend[t] = cs;
        if (nt == 0) {
            if (off != 0) bad = true;
            continue;
        }
        const uint64_t ce = (t + 1 < g.T) ? g.body + rd64(st + 4 + 8ull * (t + 1)) : len;
        const uint64_t a0 = mode == 2 ? cs : cs + 4;
        if (off > len || a0 + (mode == 2 ? 5 : 0) > ce || a0 > ce || ce > len) {
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
        r.bad = true;
        return r;
    }
    ttab[g.T] = tiles;
    std::vector<uint64_t> exit_st(tiles + 1, 0), cnt_incl(tiles + 1, 0);

    for (uint64_t gt = 0; gt < tiles; ++gt) {
        ++g_tiles;
        uint32_t t = 0;
        while (t + 1 < g.T && ttab[t + 1] <= gt) ++t;
        const uint64_t first = ttab[t];
        const uint64_t i = gt - first;
        const uint64_t as = cstart[t] + (mode == 2 ? 0 : 4), ae = aend[t];
        const uint64_t s = as + i * S;
        const uint64_t te = std::min<uint64_t>(s + S, ae);
        const uint32_t span = (uint32_t)(te - s);
        const uint64_t s0 = s & ~15ull;
        const uint32_t sh = (uint32_t)(s - s0);

        // LDS, exact sizes: the byte buffer, two bitmaps, two rank arrays.
        std::unique_ptr<uint8_t[]> buf(new uint8_t[S + 64]);
        std::unique_ptr<uint32_t[]> spec(new uint32_t[NW]), pre(new uint32_t[NW]);
        std::unique_ptr<uint16_t[]> rspec(new uint16_t[NW]), rpre(new uint16_t[NW]);
        memset(buf.get(), 0xA5, S + 64);  // stale LDS contents
        const uint64_t tl = std::min<uint64_t>(te + 16, len);
        for (uint64_t a = s0; a < tl; a += 16) {
            if (a + 16 <= len) {
                memcpy(buf.get() + (a - s0), st + a, 16);
            } else {
                for (int k = 0; k < 16; ++k)
                    if (a + k < len) buf[a - s0 + k] = st[a + k];
            }
        }
        for (uint32_t k = 0; k < NW; ++k) spec[k] = pre[k] = 0;
        const uint8_t *tb = buf.get() + sh;

        uint64_t s_exit[4];
        uint32_t s_cnt[4], s_npre[4], s_mrank[4];
        int s_merged[4], s_ok[4], s_broken[4];
        int s_nonsync = 0;
        uint64_t s_lastpos = ~0ull;

        // 1. speculative chains: the lowest candidate whose chain leaves
        for (uint32_t wv = 0; wv < 4; ++wv) {
            const uint32_t u = wv * U, v = std::min(u + U, span);
            uint32_t cnt = 0, ex = v;
            bool ok = false;
            const uint32_t cmax = std::min(u + W, v);
            for (uint32_t c = u; c < cmax && !ok; ++c) {
                uint32_t e = 0;
                if (parse_chain_leaves(tb, c, v, B, e, mode)) {
                    cnt = parse_mark_chain(tb, spec.get(), c, v, B, mode);
                    ex = e;
                    ok = true;
                }
            }
            s_ok[wv] = ok;
            s_cnt[wv] = ok ? cnt : 0;
            s_exit[wv] = s + (ok ? ex : v);
        }
        // 2. candidate check of the last sub-tile
        const uint32_t lastw = span ? (span - 1) / U : 0;
        if (span) {
            const uint32_t u = lastw * U, v = std::min(u + U, span);
            if (!s_ok[lastw]) {
                s_nonsync = 1;
            } else {
                const uint32_t cmax = std::min(u + W, v);
                for (uint32_t c = u; c < cmax; ++c) {
                    const uint32_t leave = parse_cand_walk(tb, spec.get(), c, v, B, mode);
                    if (leave && s + leave != s_exit[lastw]) s_nonsync = 1;
                }
            }
        }
        if (!s_nonsync && span) {
            exit_st[gt] = s_exit[lastw];
            ++g_sync;
        }
        // spec ranks
        for (uint32_t wv = 0; wv < 4; ++wv) {
            uint32_t run = 0;
            for (uint32_t k = 0; k < U / 32; ++k) {
                rspec[wv * (U / 32) + k] = (uint16_t)run;
                run += (uint32_t)__builtin_popcount(spec[wv * (U / 32) + k]);
            }
        }
        // 3. true chain
        uint32_t tile_cnt;
        {
            uint64_t e = as;
            if (i > 0) e = exit_st[gt - 1];
            bool broken = e < s;
            for (uint32_t w = 0; w < 4; ++w) {
                const uint32_t uw = w * U, vw = std::min(uw + U, span);
                s_npre[w] = 0;
                s_merged[w] = 0;
                s_mrank[w] = 0;
                s_broken[w] = broken;
                if (uw >= span || broken || e >= s + vw) continue;
                uint32_t p = (uint32_t)(e - s), npre = 0;
                const int res = parse_true_walk(tb, spec.get(), pre.get(), p, vw, B, npre, mode);
                s_npre[w] = npre;
                if (res == 1) {
                    s_merged[w] = 1;
                    s_mrank[w] = parse_rank(spec.get(), rspec.get(), p);
                    e = s_ok[w] ? s_exit[w] : s + vw;
                } else if (res == 2) {
                    broken = true;
                    s_broken[w] = 1;
                } else {
                    e = s + p;
                }
            }
            if (s_nonsync && span) exit_st[gt] = broken ? te : e;
            uint32_t total = 0;
            for (uint32_t w = 0; w < 4; ++w) {
                const uint32_t c = s_npre[w] + (s_merged[w] ? s_cnt[w] - s_mrank[w] : 0u);
                s_cnt[w] = c;
                total += c;
            }
            tile_cnt = total;
        }
        const uint64_t prefix = (gt == first) ? 0 : cnt_incl[gt - 1];
        cnt_incl[gt] = prefix + tile_cnt;
        // pre ranks
        for (uint32_t wv = 0; wv < 4; ++wv) {
            uint32_t run = 0;
            for (uint32_t k = 0; k < U / 32; ++k) {
                rpre[wv * (U / 32) + k] = (uint16_t)run;
                run += (uint32_t)__builtin_popcount(pre[wv * (U / 32) + k]);
            }
        }
        // 4. positions
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
        auto put = [&](uint64_t bi, uint64_t p) {
            if (bi < nbt) {
                r.pos[gbase + bi] = p;
                r.written[gbase + bi]++;
                if (bi == nbt - 1) s_lastpos = p;
            }
        };
        for (uint32_t k = 0; k < NW; ++k) {
            const uint32_t w = (k * 32) / U;
            uint32_t m = pre[k];
            while (m) {
                const uint32_t bit = __builtin_ctz(m);
                m &= m - 1;
                const uint32_t p = k * 32 + bit;
                put(wbase[w] + parse_rank(pre.get(), rpre.get(), p), s + p);
            }
            if (s_merged[w]) {
                m = spec[k];
                while (m) {
                    const uint32_t bit = __builtin_ctz(m);
                    m &= m - 1;
                    const uint32_t p = k * 32 + bit;
                    const uint32_t rk = parse_rank(spec.get(), rspec.get(), p);
                    if (rk < s_mrank[w]) continue;
                    put(wbase[w] + s_npre[w] + (rk - s_mrank[w]), s + p);
                }
            }
        }
        for (uint32_t w = 0; w < 4; ++w)
            if (s_broken[w] && wbase[w] + s_npre[w] < nbt) r.bad = true;
        if (gt + 1 == ttab[t + 1] && prefix + tile_cnt < nbt) r.bad = true;
        if (s_lastpos != ~0ull) {
            const uint64_t pl = s_lastpos;
            uint64_t end;
            uint32_t b;
            if (mode == 2) {
                const uint64_t Ll = chunk_len(g, t) - (nbt - 1) * (uint64_t)g.B - 1;
                b = pl + 4 < len ? st[pl + 4] : 255u;
                end = pl + 4 + block_bytes(b, (uint32_t)Ll);
                if (!Ll && pl + 4 < len) b = 0;
            } else {
                const uint64_t Ll = chunk_len(g, t) - 1 - (nbt - 1) * (uint64_t)g.B;
                b = st[pl];
                end = pl + block_bytes(b, (uint32_t)Ll);
            }
            if (b > 32 || (t + 1 < g.T ? end != ae : end > len)) r.bad = true;
        }
    }
    return r;
}

int main(int argc, char **argv)
{
    if (argc != 2) {
        fprintf(stderr, "usage: parse_check <case list>\n");
        return 2;
    }
    FILE *f = fopen(argv[1], "r");
    if (!f) return 2;
    char path[4096];
    unsigned long long n;
    int B, T, mode, ncase = 0, fail = 0, nbad = 0;
    while (fscanf(f, "%4095s %llu %d %d %d", path, &n, &B, &T, &mode) == 5) {
        FILE *sf = fopen(path, "rb");
        if (!sf) return 2;
        fseek(sf, 0, SEEK_END);
        const long len = ftell(sf);
        fseek(sf, 0, SEEK_SET);
        std::vector<uint8_t> st((size_t)len + 64, 0xEE);  // 64 bytes of slack: 16-byte loads
        if (fread(st.data(), 1, (size_t)len, sf) != (size_t)len) return 2;
        fclose(sf);
        Geom g;
        if (!make_geom(n, B, T, &g, 4, (uint32_t)mode)) {
            fprintf(stderr, "bad geometry %s\n", path);
            return 2;
        }
        ++ncase;
        const Result want = serial_parse(st.data(), (uint64_t)len, g);
        const Result got = emulate(st.data(), (uint64_t)len, g);
        nbad += want.bad;
        bool ok = want.bad == got.bad;
        if (ok && !want.bad) {
            for (size_t j = 0; j < want.pos.size(); ++j) {
                if (got.pos[j] != want.pos[j] || got.written[j] != 1) {
                    fprintf(stderr, "%s: block %zu at %llu, parsed %llu (written %d)\n", path, j,
                            (unsigned long long)want.pos[j], (unsigned long long)got.pos[j],
                            got.written[j]);
                    ok = false;
                    break;
                }
            }
        }
        if (!ok) {
            ++fail;
            fprintf(stderr, "FAIL %s n=%llu B=%d T=%d len=%ld serial_bad=%d parse_bad=%d\n", path,
                    n, B, T, len, (int)want.bad, (int)got.bad);
        }
    }
    fclose(f);
    printf("%d cases (%d invalid streams), %llu tiles (%llu synchronised), %d failures\n", ncase,
           nbad, (unsigned long long)g_tiles, (unsigned long long)g_sync, fail);
    return fail ? 1 : 0;
}
