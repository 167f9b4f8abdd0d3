/*
 * Device-side variation of the resident genome pool: one generation of the
 * flat postfix engine computed on the GPU from the parents in one pool arena
 * into the other (rc_pool_vary).
 *
 * Reference behaviour: Philox4x32-10 (genome.rs:214-300) and
 * Variation::next_generation (postfix.rs:500-893) with an empty archive
 * front: per island the best parent survives in slot 0 (the canonical
 * (error, cost, slot) order), a restarted island is refilled with random
 * programs, and every other slot is a child drawn from the counter-based
 * stream at (seed, island, generation, slot, attempt): a size-4 tournament,
 * a 50% crossover with a second tournament winner, one mutation (operator
 * change 25%, terminal resample 35%, subtree replacement 40%, falling back
 * through the other kinds) and a 50% second one, then the width, validity
 * and complexity filters and, with dedup, the generation-wide set of
 * program hashes (FNV-1a up to HALT).  64 failed attempts fall back to a
 * random program.  The reference runs this on the host.
 *
 * MI355X layout.  Every child is a pure function of its position except
 * for the dedup set, which is sequential in (island, slot) order.  So the
 * work is split into parallel passes and one ordered pass:
 *   vy_meta     one lane per parent: hash and node count (the keys);
 *   vy_seed     the set of parent hashes (open addressing, 64-bit CAS);
 *   vy_try      one lane per child: its first VY_TRY attempts (validity and
 *               hash), independent of every other child;
 *   vy_resolve  one lane walks the children in order against the set and
 *               picks each child's first admissible attempt, generating
 *               later attempts itself only when the first VY_TRY fail;
 *   vy_write    one lane per slot: elites, restarted islands and the chosen
 *               attempt of every child, regenerated from its position.
 * Without dedup vy_try / vy_resolve are skipped and vy_write takes each
 * child's first valid attempt.  Programs are at most VY_WMAX words wide.
 */
#include "rc_device.h"
#include "rc_internal.h"

#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

namespace {

constexpr int VY_WMAX = 128;  /* widest supported pool slot */
constexpr int VY_TRY = 8;     /* attempts precomputed per child */
constexpr int VY_ATTEMPTS = 64;
constexpr int VY_MAXOPS = 64;
constexpr uint32_t VY_HALT = 255;

/* ---- Philox4x32-10 --------------------------------------------------------- */

struct VyRng {
    uint32_t k0, k1;
    uint32_t c0, c1, c2, c3;
};

__device__ __forceinline__ VyRng vy_at(uint32_t seed, uint32_t island, uint32_t gen, uint32_t slot)
{
    return {seed, island, gen, slot, 0u, 0u};
}

/* Ten rounds of the Philox S-box on (c0..c3) with the key schedule; the
   first output word (the only one the engine reads), then the draw counter
   advances. */
__device__ __forceinline__ uint32_t vy_next(VyRng &r)
{
    uint32_t a = r.c0, b = r.c1, c = r.c2, d = r.c3, k0 = r.k0, k1 = r.k1;
#pragma unroll
    for (int round = 0; round < 10; ++round) {
        const uint32_t lo0 = 0xD2511F53u * a, hi0 = __umulhi(0xD2511F53u, a);
        const uint32_t lo1 = 0xCD9E8D57u * c, hi1 = __umulhi(0xCD9E8D57u, c);
        a = hi1 ^ b ^ k0;
        b = lo1;
        c = hi0 ^ d ^ k1;
        d = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    r.c3 += 1u;
    return a;
}

__device__ __forceinline__ double vy_unit(VyRng &r) { return (double)vy_next(r) / 4294967296.0; }
__device__ __forceinline__ uint32_t vy_below(VyRng &r, uint32_t n)
{
    return n == 0 ? 0u : vy_next(r) % n;
}

/* ---- programs -------------------------------------------------------------- */

__device__ __forceinline__ int vy_arity(uint32_t w)
{
    if ((w >= 6 && w <= 19) || w == 21 || w == 22 || w == 23 || w == 24 || w == 25 || w == 28 ||
        w == 29 || w == 33 || w == 42)
        return 1;
    if ((w >= 2 && w <= 5) || w == 20 || w == 26 || w == 27 || w == 30 || w == 31 ||
        (w >= 34 && w <= 41))
        return 2;
    return w == 32 ? 3 : -1;
}

/* A program under construction: words [0, len). */
struct VyProg {
    uint32_t w[VY_WMAX];
    int len;
};

/* Node spans of a program: start | end << 8 | arity << 16. */
struct VySpans {
    uint32_t s[VY_WMAX];
    int n;
};
__device__ __forceinline__ int sp_start(uint32_t v) { return (int)(v & 0xffu); }
__device__ __forceinline__ int sp_end(uint32_t v) { return (int)((v >> 8) & 0xffu); }
__device__ __forceinline__ int sp_arity(uint32_t v) { return (int)(v >> 16); }

/* The variation loop's scan: the nodes of a HALT-terminated program that
   reduces to one value (postfix.rs:895-939). */
__device__ bool vy_scan(const uint32_t *w, int len, VySpans &out)
{
    uint8_t st[VY_WMAX];
    int sp = 0, i = 0;
    bool halt = false;
    out.n = 0;
    while (i < len) {
        const uint32_t x = w[i];
        if (x == VY_HALT) {
            halt = true;
            break;
        }
        if (x == OP_CONST || x == OP_VAR) {
            if (i + 1 >= len) return false;
            out.s[out.n++] = (uint32_t)i | ((uint32_t)(i + 2) << 8);
            st[sp++] = (uint8_t)i;
            i += 2;
        } else {
            const int a = vy_arity(x);
            if (a <= 0 || sp < a) return false;
            const int start = st[sp - a];
            sp -= a;
            out.s[out.n++] = (uint32_t)start | ((uint32_t)(i + 1) << 8) | ((uint32_t)a << 16);
            st[sp++] = (uint8_t)start;
            i += 1;
        }
    }
    return halt && sp == 1 && out.n > 0;
}

/* The slot writer's parse (postfix.rs:268-312): node count, or -1. */
__device__ int vy_parse(const uint32_t *w, int len)
{
    int depth = 0, nodes = 0, i = 0;
    while (i < len) {
        const uint32_t x = w[i];
        if (x == VY_HALT) break;
        if (x == OP_CONST || x == OP_VAR) {
            if (i + 1 >= len || w[i + 1] == VY_HALT) return -1;
            ++depth;
            ++nodes;
            i += 2;
        } else {
            const int a = vy_arity(x);
            if (a <= 0 || depth < a) return -1;
            depth -= a - 1;
            ++nodes;
            i += 1;
        }
    }
    return (depth == 1 && i > 0) ? nodes : -1;
}

__device__ __forceinline__ uint64_t vy_hash(const uint32_t *w, int len)
{
    uint64_t h = 0xcbf29ce484222325ull;
    for (int i = 0; i < len; ++i) {
        h ^= (uint64_t)w[i];
        h *= 0x00000100000001b3ull;
        if (w[i] == VY_HALT) break;
    }
    return h;
}

__device__ __forceinline__ int vy_logical_len(const uint32_t *slot, int width)
{
    for (int i = 0; i < width; ++i)
        if (slot[i] == VY_HALT) return i + 1;
    return width;
}

/* Replace words [start, end) of p by rep[0, rlen); refused (p unchanged)
   when the result would be wider than width or shorter than 2 words. */
__device__ bool vy_splice(VyProg &p, int start, int end, const uint32_t *rep, int rlen, int width)
{
    const int nlen = p.len - (end - start) + rlen;
    if (nlen > width || nlen < 2) return false;
    const int tail = p.len - end;
    if (rlen > end - start) {
        for (int q = tail - 1; q >= 0; --q) p.w[start + rlen + q] = p.w[end + q];
    } else {
        for (int q = 0; q < tail; ++q) p.w[start + rlen + q] = p.w[end + q];
    }
    for (int q = 0; q < rlen; ++q) p.w[start + q] = rep[q];
    p.len = nlen;
    return true;
}

/* The engine's parameters and operator table as the kernels see them. */
struct VyCfg {
    int islands, pop, width, max_complexity, nvars, init_depth, dedup, nops;
    uint32_t seed, gen;
    uint32_t op_code[VY_MAXOPS];
    int by_arity[4][VY_MAXOPS];
    int n_arity[4];
};

__device__ void vy_terminal(const VyCfg &g, VyRng &r, uint32_t &a, uint32_t &b)
{
    if (g.nvars > 0 && vy_unit(r) < 0.65) {
        a = OP_VAR;
        b = vy_below(r, (uint32_t)g.nvars);
        return;
    }
    const double v = round((vy_unit(r) * 4.0 - 2.0) * 1000.0) / 1000.0;
    a = OP_CONST;
    b = __float_as_uint((float)v);
}

/* Append a random subtree of `budget` nodes and `depth` levels
   (postfix.rs:820-849), with an explicit frame stack in place of recursion
   so the draws come in the recursive order.  Words beyond `cap` are counted
   but not stored: such a subtree is wider than any slot and is refused by
   the caller, after the same draws the reference makes. */
__device__ void vy_subtree(const VyCfg &g, VyRng &r, int depth, int budget, uint32_t *out, int &len,
                           int cap)
{
    auto emit = [&](uint32_t w) {
        if (len < cap) out[len] = w;
        ++len;
    };
    struct Frame {
        int depth, budget, arity, k, remaining;
        uint32_t code;
    };
    Frame fs[24];
    int top = 0;
    fs[0] = {depth, budget, -1, 0, 0, 0u};
    while (top >= 0) {
        Frame &f = fs[top];
        if (f.arity < 0) {
            if (f.depth == 0 || f.budget <= 1 || vy_unit(r) < 0.15) {
                uint32_t a, b;
                vy_terminal(g, r, a, b);
                emit(a);
                emit(b);
                --top;
                continue;
            }
            int arity;
            if (vy_unit(r) < 0.35) arity = 1;
            else if (vy_unit(r) < 0.75) arity = 2;
            else arity = 3;
            if (g.n_arity[arity] == 0 || f.budget < arity + 1) {
                uint32_t a, b;
                vy_terminal(g, r, a, b);
                emit(a);
                emit(b);
                --top;
                continue;
            }
            f.code = g.op_code[g.by_arity[arity][vy_below(r, (uint32_t)g.n_arity[arity])]];
            f.arity = arity;
            f.k = 0;
            f.remaining = f.budget - 1 - arity;
        }
        if (f.k < f.arity) {
            const int share = max(f.remaining / (f.arity - f.k), 1);
            f.k += 1;
            fs[top + 1] = {f.depth - 1, share, -1, 0, 0, 0u};
            ++top;
            continue;
        }
        emit(f.code);
        --top;
    }
}

/* One edit of kind 0 (operator), 1 (terminal) or 2 (subtree) on the node
   spans computed before the edit (postfix.rs:739-817). */
__device__ bool vy_edit(const VyCfg &g, VyProg &p, const VySpans &sp, int kind, VyRng &r)
{
    int count = 0;
    for (int i = 0; i < sp.n; ++i) {
        const int a = sp_arity(sp.s[i]);
        count += (kind == 0 ? a > 0 : (kind == 1 ? a == 0 : true)) ? 1 : 0;
    }
    if (count == 0) return false;
    int target = (int)vy_below(r, (uint32_t)count), node = 0;
    for (int i = 0; i < sp.n; ++i) {
        const int a = sp_arity(sp.s[i]);
        if (!(kind == 0 ? a > 0 : (kind == 1 ? a == 0 : true))) continue;
        if (target == 0) {
            node = i;
            break;
        }
        --target;
    }
    const uint32_t s = sp.s[node];
    const int start = sp_start(s), end = sp_end(s), arity = sp_arity(s);
    if (kind == 0) {
        const int ar = min(arity, 3);
        if (g.n_arity[ar] < 2) return false;
        const uint32_t current = p.w[end - 1];
        int others = 0;
        for (int q = 0; q < g.n_arity[ar]; ++q) others += g.op_code[g.by_arity[ar][q]] != current;
        if (others == 0) return false;
        int pick = (int)vy_below(r, (uint32_t)others);
        for (int q = 0; q < g.n_arity[ar]; ++q) {
            const uint32_t code = g.op_code[g.by_arity[ar][q]];
            if (code == current) continue;
            if (pick == 0) {
                p.w[end - 1] = code;
                break;
            }
            --pick;
        }
        return true;
    }
    if (kind == 1) {
        uint32_t a, b;
        vy_terminal(g, r, a, b);
        p.w[start] = a;
        p.w[start + 1] = b;
        return true;
    }
    const int budget = (g.width - 1) - (p.len - end);
    if (budget < 2) return false;
    uint32_t tmp[VY_WMAX + 8];
    int tlen = 0;
    vy_subtree(g, r, g.init_depth, budget, tmp, tlen, VY_WMAX + 8);
    if (end - start + tlen > g.width) return false;
    return vy_splice(p, start, end, tmp, tlen, g.width);
}

__device__ void vy_mutate(const VyCfg &g, VyProg &p, VyRng &r)
{
    VySpans sp;
    if (!vy_scan(p.w, p.len, sp) || sp.n == 0) return;
    const double u = vy_unit(r);
    int order[3];
    if (u < 0.25) { order[0] = 0; order[1] = 2; order[2] = 1; }
    else if (u < 0.6) { order[0] = 1; order[1] = 0; order[2] = 2; }
    else { order[0] = 2; order[1] = 0; order[2] = 1; }
    for (int q = 0; q < 3; ++q)
        if (vy_edit(g, p, sp, order[q], r)) return;
}

/* The canonical key of parent s: non-NaN errors first, then the error's
   total order, then the node count, then the slot. */
__device__ __forceinline__ bool vy_less(const double *err, const uint32_t *cost, int a, int b)
{
    const double ea = err[a], eb = err[b];
    const int na = isnan(ea) ? 1 : 0, nb = isnan(eb) ? 1 : 0;
    if (na != nb) return na < nb;
    uint64_t ba = 0, bb = 0;
    if (!na) {
        const uint64_t x = (uint64_t)__double_as_longlong(ea), y = (uint64_t)__double_as_longlong(eb);
        ba = (x >> 63) ? ~x : (x | (1ull << 63));
        bb = (y >> 63) ? ~y : (y | (1ull << 63));
    }
    if (ba != bb) return ba < bb;
    if (cost[a] != cost[b]) return cost[a] < cost[b];
    return a < b;
}

__device__ int vy_tournament(const VyCfg &g, const double *err, const uint32_t *cost, int island,
                             VyRng &r)
{
    const int base = island * g.pop;
    int best = base + (int)vy_below(r, (uint32_t)g.pop);
    for (int t = 1; t < 4; ++t) {
        const int c = base + (int)vy_below(r, (uint32_t)g.pop);
        if (vy_less(err, cost, c, best)) best = c;
    }
    return best;
}

/* Attempt `attempt` at child (island, slot) into p; returns whether it
   passes the width / validity / complexity filters (postfix.rs:613-676). */
__device__ bool vy_attempt(const VyCfg &g, const uint32_t *parents, const double *err,
                           const uint32_t *cost, int island, int slot, int attempt, VyProg &p)
{
    VyRng r = vy_at(g.seed, (uint32_t)island, g.gen, (uint32_t)slot);
    r.c2 = (uint32_t)attempt;
    const uint32_t *src = parents + (size_t)vy_tournament(g, err, cost, island, r) * g.width;
    p.len = vy_logical_len(src, g.width);
    for (int i = 0; i < p.len; ++i) p.w[i] = src[i];
    if (vy_unit(r) < 0.5) {
        const uint32_t *oth = parents + (size_t)vy_tournament(g, err, cost, island, r) * g.width;
        const int olen = vy_logical_len(oth, g.width);
        VySpans a, b;
        if (vy_scan(p.w, p.len, a) && vy_scan(oth, olen, b) && a.n > 0 && b.n > 0) {
            const uint32_t sa = a.s[vy_below(r, (uint32_t)a.n)];
            const uint32_t sb = b.s[vy_below(r, (uint32_t)b.n)];
            vy_splice(p, sp_start(sa), sp_end(sa), oth + sp_start(sb), sp_end(sb) - sp_start(sb),
                      g.width);
        }
    }
    vy_mutate(g, p, r);
    if (vy_unit(r) < 0.5) vy_mutate(g, p, r);
    if (p.len > g.width) return false;
    VySpans sp;
    if (!vy_scan(p.w, p.len, sp)) return false;
    return sp.n <= g.max_complexity && vy_parse(p.w, p.len) >= 0;
}

/* A random program at (island, slot) after `salt` discarded draws
   (postfix.rs:585-604), never failing. */
__device__ void vy_random(const VyCfg &g, int island, int slot, int salt, VyProg &p)
{
    VyRng r = vy_at(g.seed, (uint32_t)island, g.gen, (uint32_t)slot);
    for (int q = 0; q < salt; ++q) (void)vy_next(r);
    const int budget = max(min(g.max_complexity, (g.width - 1) / 2), 1);
    for (int t = 0; t < 8; ++t) {
        p.len = 0;
        vy_subtree(g, r, g.init_depth, budget, p.w, p.len, VY_WMAX);
        if (p.len + 1 > g.width) continue; /* wider than a slot with its HALT */
        p.w[p.len++] = VY_HALT;
        const int nodes = vy_parse(p.w, p.len);
        if (nodes >= 0 && nodes <= g.max_complexity) return;
    }
    p.w[0] = OP_VAR;
    p.w[1] = 0;
    p.w[2] = VY_HALT;
    p.len = 3;
}

/* ---- kernels ---------------------------------------------------------------- */

__global__ void vy_meta(const uint32_t *__restrict__ parents, int n, int width,
                        uint64_t *__restrict__ phash, uint32_t *__restrict__ pcost)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const uint32_t *slot = parents + (size_t)s * width;
    const int len = vy_logical_len(slot, width);
    phash[s] = vy_hash(slot, len);
    const int nodes = vy_parse(slot, len);
    pcost[s] = nodes < 0 ? 0u : (uint32_t)min(nodes, 65535);
}

/* Open-addressing set of 64-bit hashes (0 marks an empty entry, so a hash
   of 0 is stored as 1). */
__device__ __forceinline__ uint64_t vy_key(uint64_t h) { return h == 0 ? 1ull : h; }

__global__ void vy_clear(uint64_t *__restrict__ table, int cap)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < cap) table[i] = 0ull;
}

__global__ void vy_seed(const uint64_t *__restrict__ phash, int n, uint64_t *table, int cap)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const uint64_t k = vy_key(phash[s]);
    unsigned long long *t = (unsigned long long *)table;
    for (int q = 0, i = (int)(k & (uint64_t)(cap - 1)); q < cap; ++q, i = (i + 1) & (cap - 1)) {
        const unsigned long long prev = atomicCAS(t + i, 0ull, (unsigned long long)k);
        if (prev == 0ull || prev == k) return;
    }
}

__global__ void vy_try(VyCfg g, const uint32_t *__restrict__ parents, const double *__restrict__ err,
                       const uint32_t *__restrict__ pcost, const unsigned char *__restrict__ restart,
                       uint64_t *__restrict__ thash, unsigned char *__restrict__ tok)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.islands * g.pop) return;
    const int island = c / g.pop, slot = c % g.pop;
    if (slot == 0 || (restart && restart[island])) return;
    VyProg p;
    for (int a = 0; a < VY_TRY; ++a) {
        const bool ok = vy_attempt(g, parents, err, pcost, island, slot, a, p);
        tok[(size_t)c * VY_TRY + a] = ok ? 1 : 0;
        thash[(size_t)c * VY_TRY + a] = ok ? vy_hash(p.w, p.len) : 0ull;
    }
}

/* The ordered pass: one lane, children in (island, slot) order.  Table
   accesses are volatile (coherent with the atomic inserts of vy_seed). */
__global__ void vy_resolve(VyCfg g, const uint32_t *__restrict__ parents,
                           const double *__restrict__ err, const uint32_t *__restrict__ pcost,
                           const unsigned char *__restrict__ restart,
                           const uint64_t *__restrict__ thash, const unsigned char *__restrict__ tok,
                           uint64_t *table, int cap, int *__restrict__ chosen)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    volatile uint64_t *t = table;
    auto find = [&](uint64_t k, int &at) {
        int i = (int)(k & (uint64_t)(cap - 1));
        for (int q = 0; q < cap; ++q, i = (i + 1) & (cap - 1)) {
            const uint64_t v = t[i];
            if (v == k) return true;
            if (v == 0ull) {
                at = i;
                return false;
            }
        }
        at = -1;
        return false;
    };
    VyProg p;
    for (int c = 0; c < g.islands * g.pop; ++c) {
        const int island = c / g.pop, slot = c % g.pop;
        if (slot == 0 || (restart && restart[island])) continue;
        int pick = -1;
        for (int a = 0; a < VY_ATTEMPTS && pick < 0; ++a) {
            bool ok;
            uint64_t h;
            if (a < VY_TRY) {
                ok = tok[(size_t)c * VY_TRY + a] != 0;
                h = thash[(size_t)c * VY_TRY + a];
            } else {
                ok = vy_attempt(g, parents, err, pcost, island, slot, a, p);
                h = ok ? vy_hash(p.w, p.len) : 0ull;
            }
            if (!ok) continue;
            int at = -1;
            if (find(vy_key(h), at)) continue;
            if (at >= 0) t[at] = vy_key(h);
            pick = a;
        }
        chosen[c] = pick;
    }
}

__global__ void vy_write(VyCfg g, const uint32_t *__restrict__ parents,
                         const double *__restrict__ err, const uint32_t *__restrict__ pcost,
                         const uint64_t *__restrict__ phash, const unsigned char *__restrict__ restart,
                         const int *__restrict__ chosen, uint32_t *__restrict__ kids,
                         unsigned short *__restrict__ klen, unsigned short *__restrict__ kcost,
                         uint64_t *__restrict__ khash, double *__restrict__ kerr)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.islands * g.pop) return;
    const int island = c / g.pop, slot = c % g.pop;
    uint32_t *dst = kids + (size_t)c * g.width;
    if (!(restart && restart[island]) && slot == 0) {
        const int base = island * g.pop;
        int best = base;
        for (int s = base + 1; s < base + g.pop; ++s)
            if (vy_less(err, pcost, s, best)) best = s;
        const uint32_t *src = parents + (size_t)best * g.width;
        for (int i = 0; i < g.width; ++i) dst[i] = src[i];
        klen[c] = (unsigned short)vy_logical_len(src, g.width);
        kcost[c] = (unsigned short)pcost[best];
        khash[c] = phash[best];
        kerr[c] = err[best];
        return;
    }
    VyProg p;
    if (restart && restart[island]) {
        vy_random(g, island, slot, 0, p);
    } else if (g.dedup) {
        const int a = chosen[c];
        if (a < 0 || !vy_attempt(g, parents, err, pcost, island, slot, a, p))
            vy_random(g, island, slot, VY_ATTEMPTS, p);
    } else {
        bool ok = false;
        for (int a = 0; a < VY_ATTEMPTS && !ok; ++a)
            ok = vy_attempt(g, parents, err, pcost, island, slot, a, p);
        if (!ok) vy_random(g, island, slot, VY_ATTEMPTS, p);
    }
    for (int i = 0; i < g.width; ++i) dst[i] = i < p.len ? p.w[i] : VY_HALT;
    klen[c] = (unsigned short)p.len;
    kcost[c] = (unsigned short)max(vy_parse(p.w, p.len), 0);
    khash[c] = vy_hash(p.w, p.len);
    kerr[c] = __builtin_nan("");
}

/* Host copy of the opcode arity table (gpu.rs:498-509). */
int vy_arity_host(uint32_t w)
{
    if ((w >= 6 && w <= 19) || w == 21 || w == 22 || w == 23 || w == 24 || w == 25 || w == 28 ||
        w == 29 || w == 33 || w == 42)
        return 1;
    if ((w >= 2 && w <= 5) || w == 20 || w == 26 || w == 27 || w == 30 || w == 31 ||
        (w >= 34 && w <= 41))
        return 2;
    return w == 32 ? 3 : -1;
}

size_t vy_take(size_t &cur, size_t bytes)
{
    const size_t o = (cur + 255) & ~(size_t)255;
    cur = o + bytes;
    return o;
}

} /* namespace */

/* The variation's device region, grown at the high-water mark. */
struct RcVaryState {
    unsigned char *dev = nullptr;
    size_t cap = 0;
};

void rc_vary_free(RcCtx *c)
{
    if (!c->vary) return;
    if (c->vary->dev) (void)hipFree(c->vary->dev);
    delete c->vary;
    c->vary = nullptr;
}

extern "C" int rc_pool_vary(void *ctx, int from, int islands, int pop, unsigned int seed,
                            unsigned int generation, int max_complexity, int nvars, int init_depth,
                            int dedup, const unsigned int *ops, int nops, const double *err,
                            const unsigned char *restart, unsigned short *out_len,
                            unsigned short *out_cost, unsigned long long *out_hash, double *out_err)
{
    RcCtx *c = (RcCtx *)ctx;
    if (!c || islands <= 0 || pop <= 0 || !ops || nops <= 0 || nops > VY_MAXOPS || !err ||
        max_complexity < 1 || nvars < 0 || nvars > 255 || init_depth < 0 || init_depth > 8)
        return -1;
    const int n = islands * pop;
    const int width = c->pool_width;
    if (n > c->pool_cap || width < 3 || width > VY_WMAX || !c->pool_raw[0] || !c->pool_raw[1])
        return -1;
    VyCfg g;
    memset(&g, 0, sizeof(g));
    g.islands = islands;
    g.pop = pop;
    g.width = width;
    g.max_complexity = max_complexity;
    g.nvars = nvars;
    g.init_depth = init_depth;
    g.dedup = dedup ? 1 : 0;
    g.seed = seed;
    g.gen = generation;
    for (int k = 0; k < nops; ++k) {
        const uint32_t code = ops[2 * k];
        const int ar = (int)ops[2 * k + 1];
        if (ar < 1 || ar > 3 || vy_arity_host(code) != ar) return -1;
        bool dup = false;
        for (int q = 0; q < g.nops; ++q) dup = dup || g.op_code[q] == code;
        if (dup) continue;
        g.by_arity[ar][g.n_arity[ar]++] = g.nops;
        g.op_code[g.nops++] = code;
    }
    int tcap = 1024;
    while (tcap < 4 * 2 * n) tcap <<= 1;
    size_t cur = 0;
    const size_t o_err = vy_take(cur, (size_t)n * 8);
    const size_t o_rs = vy_take(cur, (size_t)islands);
    const size_t o_phash = vy_take(cur, (size_t)n * 8);
    const size_t o_pcost = vy_take(cur, (size_t)n * 4);
    const size_t o_thash = vy_take(cur, (size_t)n * VY_TRY * 8);
    const size_t o_tok = vy_take(cur, (size_t)n * VY_TRY);
    const size_t o_table = vy_take(cur, (size_t)tcap * 8);
    const size_t o_chosen = vy_take(cur, (size_t)n * 4);
    const size_t o_klen = vy_take(cur, (size_t)n * 2);
    const size_t o_kcost = vy_take(cur, (size_t)n * 2);
    const size_t o_khash = vy_take(cur, (size_t)n * 8);
    const size_t o_kerr = vy_take(cur, (size_t)n * 8);
    const size_t end = vy_take(cur, 0);
    if (!c->vary) c->vary = new RcVaryState();
    RcVaryState *v = c->vary;
    if (end > v->cap) {
        if (v->dev) rc_dev_free(c, v->dev);
        v->dev = nullptr;
        v->cap = 0;
        if (rc_dev_alloc(c, (void **)&v->dev, end + end / 4) != 0) return -1;
        v->cap = end + end / 4;
    }
    unsigned char *D = v->dev;
    hipStream_t s = c->s_eval;
    const uint32_t *parents = c->pool_raw[from & 1];
    uint32_t *kids = c->pool_raw[(from & 1) ^ 1];
    double *d_err = (double *)(D + o_err);
    unsigned char *d_rs = restart ? D + o_rs : nullptr;
    uint64_t *d_phash = (uint64_t *)(D + o_phash);
    uint32_t *d_pcost = (uint32_t *)(D + o_pcost);
    uint64_t *d_thash = (uint64_t *)(D + o_thash);
    unsigned char *d_tok = D + o_tok;
    uint64_t *d_table = (uint64_t *)(D + o_table);
    int *d_chosen = (int *)(D + o_chosen);
    unsigned short *d_klen = (unsigned short *)(D + o_klen);
    unsigned short *d_kcost = (unsigned short *)(D + o_kcost);
    uint64_t *d_khash = (uint64_t *)(D + o_khash);
    double *d_kerr = (double *)(D + o_kerr);
    if (hipMemcpyAsync(d_err, err, (size_t)n * 8, hipMemcpyHostToDevice, s) != hipSuccess ||
        (restart &&
         hipMemcpyAsync(d_rs, restart, (size_t)islands, hipMemcpyHostToDevice, s) != hipSuccess))
        return -1;
    const unsigned blocks = (unsigned)((n + 63) / 64);
    vy_meta<<<blocks, 64, 0, s>>>(parents, n, width, d_phash, d_pcost);
    if (g.dedup) {
        vy_clear<<<(unsigned)((tcap + 255) / 256), 256, 0, s>>>(d_table, tcap);
        vy_seed<<<blocks, 64, 0, s>>>(d_phash, n, d_table, tcap);
        vy_try<<<blocks, 64, 0, s>>>(g, parents, d_err, d_pcost, d_rs, d_thash, d_tok);
        vy_resolve<<<1, 64, 0, s>>>(g, parents, d_err, d_pcost, d_rs, d_thash, d_tok, d_table, tcap,
                                    d_chosen);
    }
    vy_write<<<blocks, 64, 0, s>>>(g, parents, d_err, d_pcost, d_phash, d_rs, d_chosen, kids,
                                   d_klen, d_kcost, d_khash, d_kerr);
    if (hipGetLastError() != hipSuccess) return -1;
    /* The evaluator translates the pool from its host mirror. */
    std::vector<uint32_t> &mirror = c->pool_host[(from & 1) ^ 1];
    if (hipMemcpyAsync(mirror.data(), kids, (size_t)n * width * 4, hipMemcpyDeviceToHost, s) !=
            hipSuccess ||
        (out_len && hipMemcpyAsync(out_len, d_klen, (size_t)n * 2, hipMemcpyDeviceToHost, s) !=
                        hipSuccess) ||
        (out_cost && hipMemcpyAsync(out_cost, d_kcost, (size_t)n * 2, hipMemcpyDeviceToHost, s) !=
                         hipSuccess) ||
        (out_hash && hipMemcpyAsync(out_hash, d_khash, (size_t)n * 8, hipMemcpyDeviceToHost, s) !=
                         hipSuccess) ||
        (out_err && hipMemcpyAsync(out_err, d_kerr, (size_t)n * 8, hipMemcpyDeviceToHost, s) !=
                        hipSuccess) ||
        hipStreamSynchronize(s) != hipSuccess) {
        (void)hipGetLastError();
        return -1;
    }
    return 0;
}

extern "C" int rc_pool_download(void *ctx, int which, unsigned int *out, int nprog)
{
    RcCtx *c = (RcCtx *)ctx;
    if (!c || !out || nprog <= 0 || nprog > c->pool_cap || !c->pool_raw[which & 1]) return -1;
    const size_t words = (size_t)nprog * (size_t)c->pool_width;
    if (hipStreamSynchronize(c->s_eval) != hipSuccess ||
        hipMemcpy(out, c->pool_raw[which & 1], words * 4, hipMemcpyDeviceToHost) != hipSuccess) {
        (void)hipGetLastError();
        return -1;
    }
    return 0;
}
