/*
 * Oracle (test infrastructure only): the CPU restatement of the genome
 * pool's counter-based generator and the flat postfix variation engine.
 * Only tests/ load it, as the checker of the device variation.
 *
 *   or_philox        Philox4x32-10, genome.rs:214-300 (Salmon et al. 2011):
 *                    pinned by the Random123 known-answer vectors.
 *   or_vary          Variation::next_generation (postfix.rs:500-893) with an
 *                    empty archive front (front_p = 0): elitism, restarted
 *                    islands, tournament (k = 4), crossover, one or two
 *                    mutations, the width / validity / complexity filters,
 *                    the generation-wide dedup set and the random fallback.
 *                    Pinned by the Philox vectors and by the reference's own
 *                    invariants (postfix.rs:1019-1089: bounded and valid
 *                    children, position determinism, islands independent of
 *                    how many are evolved); the child streams themselves
 *                    have no recorded vectors, so beyond those the
 *                    restatement's parity is unpinned.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <array>
#include <unordered_set>
#include <vector>

namespace {

const uint32_t HALT = 255, CONSTW = 0, VARW = 1;

struct Philox {
    uint32_t key[2];
    uint32_t ctr[4];
    Philox(uint32_t seed, uint32_t island, uint32_t gen, uint32_t slot)
    {
        key[0] = seed;
        key[1] = island;
        ctr[0] = gen;
        ctr[1] = slot;
        ctr[2] = 0;
        ctr[3] = 0;
    }
    static void block(const uint32_t k_in[2], const uint32_t c_in[4], uint32_t out[4])
    {
        uint32_t c0 = c_in[0], c1 = c_in[1], c2 = c_in[2], c3 = c_in[3];
        uint32_t k0 = k_in[0], k1 = k_in[1];
        for (int round = 0; round < 10; ++round) {
            const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
            const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
            const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
            const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
            c0 = n0;
            c1 = (uint32_t)p1;
            c2 = n2;
            c3 = (uint32_t)p0;
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        out[0] = c0;
        out[1] = c1;
        out[2] = c2;
        out[3] = c3;
    }
    uint32_t first()
    {
        uint32_t o[4];
        block(key, ctr, o);
        ctr[3] += 1;
        return o[0];
    }
    double unit() { return (double)first() / 4294967296.0; }
    size_t below(size_t n) { return n == 0 ? 0 : (size_t)first() % n; }
};

int arity_of(uint32_t w)
{
    switch (w) {
        case 6: case 7: case 8: case 9: case 10: case 11: case 12: case 13: case 14: case 15:
        case 16: case 17: case 18: case 19: case 21: case 22: case 23: case 24: case 25:
        case 28: case 29: case 33: case 42:
            return 1;
        case 2: case 3: case 4: case 5: case 20: case 26: case 27: case 30: case 31: case 34:
        case 35: case 36: case 37: case 38: case 39: case 40: case 41:
            return 2;
        case 32:
            return 3;
        default:
            return -1;
    }
}

struct Node {
    size_t start, end;
    int arity;
};

/* The lenient scan of the variation loop: nodes of a HALT-terminated
   program that reduces to one value. */
bool scan(const std::vector<uint32_t> &p, std::vector<Node> &out)
{
    out.clear();
    std::vector<std::pair<size_t, size_t>> stack;
    bool halt = false;
    size_t i = 0;
    while (i < p.size()) {
        const uint32_t w = p[i];
        if (w == HALT) {
            halt = true;
            break;
        }
        if (w == CONSTW || w == VARW) {
            if (i + 1 >= p.size()) return false;
            out.push_back({i, i + 2, 0});
            stack.push_back({i, out.size() - 1});
            i += 2;
        } else {
            const int a = arity_of(w);
            if (a <= 0 || stack.size() < (size_t)a) return false;
            const size_t start = stack[stack.size() - (size_t)a].first;
            stack.resize(stack.size() - (size_t)a);
            out.push_back({start, i + 1, a});
            stack.push_back({start, out.size() - 1});
            i += 1;
        }
    }
    return halt && stack.size() == 1 && !out.empty();
}

/* The strict parse that guards a slot write: a leaf's operand is not HALT
   and the stack ends with one value. */
bool parse(const std::vector<uint32_t> &p, size_t &nodes)
{
    size_t depth = 0, count = 0, i = 0;
    while (i < p.size()) {
        const uint32_t w = p[i];
        if (w == HALT) break;
        if (w == CONSTW || w == VARW) {
            if (i + 1 >= p.size() || p[i + 1] == HALT) return false;
            ++depth;
            ++count;
            i += 2;
        } else {
            const int a = arity_of(w);
            if (a <= 0 || depth < (size_t)a) return false;
            depth -= (size_t)a - 1;
            ++count;
            i += 1;
        }
    }
    if (depth != 1 || i == 0) return false;
    nodes = count;
    return true;
}

uint64_t fnv(const std::vector<uint32_t> &p)
{
    uint64_t h = 0xcbf29ce484222325ull;
    for (uint32_t w : p) {
        h ^= (uint64_t)w;
        h *= 0x00000100000001b3ull;
        if (w == HALT) break;
    }
    return h;
}

size_t logical_len(const uint32_t *slot, size_t width)
{
    for (size_t i = 0; i < width; ++i)
        if (slot[i] == HALT) return i + 1;
    return width;
}

bool splice(std::vector<uint32_t> &p, const Node &s, const std::vector<uint32_t> &rep, size_t width)
{
    const size_t old_len = p.size();
    const size_t new_len = old_len - (s.end - s.start) + rep.size();
    if (new_len > width || new_len < 2) return false;
    std::vector<uint32_t> tail(p.begin() + (long)s.end, p.end());
    p.resize(s.start);
    p.insert(p.end(), rep.begin(), rep.end());
    p.insert(p.end(), tail.begin(), tail.end());
    return true;
}

struct Engine {
    size_t width, nvars, max_complexity, init_depth;
    uint32_t seed;
    bool dedup;
    std::vector<std::pair<uint32_t, int>> ops; /* (code, arity) in table order */
    std::vector<size_t> by_arity[4];

    std::array<uint32_t, 2> terminal(Philox &rng) const
    {
        if (nvars > 0 && rng.unit() < 0.65) return {VARW, (uint32_t)rng.below(nvars)};
        const double v = round((rng.unit() * 4.0 - 2.0) * 1000.0) / 1000.0;
        const float f = (float)v;
        uint32_t bits;
        memcpy(&bits, &f, 4);
        return {CONSTW, bits};
    }

    void subtree(Philox &rng, size_t depth, size_t budget, std::vector<uint32_t> &out) const
    {
        if (depth == 0 || budget <= 1 || rng.unit() < 0.15) {
            const auto t = terminal(rng);
            out.push_back(t[0]);
            out.push_back(t[1]);
            return;
        }
        size_t arity;
        if (rng.unit() < 0.35) arity = 1;
        else if (rng.unit() < 0.75) arity = 2;
        else arity = 3;
        const std::vector<size_t> &pool = by_arity[arity];
        if (pool.empty() || budget < arity + 1) {
            const auto t = terminal(rng);
            out.push_back(t[0]);
            out.push_back(t[1]);
            return;
        }
        const uint32_t code = ops[pool[rng.below(pool.size())]].first;
        const size_t remaining = budget - 1 - arity;
        for (size_t k = 0; k < arity; ++k) {
            const size_t share = std::max<size_t>(remaining / (arity - k), 1);
            subtree(rng, depth - 1, share, out);
        }
        out.push_back(code);
    }

    std::vector<uint32_t> random_program(Philox &rng) const
    {
        const size_t budget = std::max<size_t>(std::min(max_complexity, (width - 1) / 2), 1);
        std::vector<uint32_t> out;
        subtree(rng, init_depth, budget, out);
        out.push_back(HALT);
        return out;
    }

    bool edit(std::vector<uint32_t> &p, const std::vector<Node> &nodes, int kind, Philox &rng) const
    {
        auto eligible = [&](const Node &n) {
            return kind == 0 ? n.arity > 0 : (kind == 1 ? n.arity == 0 : true);
        };
        size_t count = 0;
        for (const Node &n : nodes) count += eligible(n) ? 1 : 0;
        if (count == 0) return false;
        size_t target = rng.below(count), idx = 0;
        for (size_t i = 0; i < nodes.size(); ++i) {
            if (!eligible(nodes[i])) continue;
            if (target == 0) {
                idx = i;
                break;
            }
            --target;
        }
        const Node n = nodes[idx];
        if (kind == 0) {
            const std::vector<size_t> &pool = by_arity[std::min(n.arity, 3)];
            if (pool.size() < 2) return false;
            const uint32_t current = p[n.end - 1];
            std::vector<uint32_t> others;
            for (size_t k : pool)
                if (ops[k].first != current) others.push_back(ops[k].first);
            if (others.empty()) return false;
            p[n.end - 1] = others[rng.below(others.size())];
            return true;
        }
        if (kind == 1) {
            const auto t = terminal(rng);
            p[n.start] = t[0];
            p[n.start + 1] = t[1];
            return true;
        }
        const size_t budget = (width - 1) - (p.size() - n.end);
        if (budget < 2) return false;
        std::vector<uint32_t> tmp;
        subtree(rng, init_depth, budget, tmp);
        if (n.end - n.start + tmp.size() > width) return false;
        return splice(p, n, tmp, width);
    }

    void mutate(std::vector<uint32_t> &p, Philox &rng) const
    {
        std::vector<Node> nodes;
        if (!scan(p, nodes) || nodes.empty()) return;
        const double r = rng.unit();
        int order[3];
        if (r < 0.25) { order[0] = 0; order[1] = 2; order[2] = 1; }
        else if (r < 0.6) { order[0] = 1; order[1] = 0; order[2] = 2; }
        else { order[0] = 2; order[1] = 0; order[2] = 1; }
        for (int kind : order)
            if (edit(p, nodes, kind, rng)) return;
    }

    void crossover(std::vector<uint32_t> &p, const std::vector<uint32_t> &other, Philox &rng) const
    {
        std::vector<Node> a, b;
        if (!scan(p, a) || !scan(other, b) || a.empty() || b.empty()) return;
        const Node sa = a[rng.below(a.size())];
        const Node sb = b[rng.below(b.size())];
        const std::vector<uint32_t> rep(other.begin() + (long)sb.start, other.begin() + (long)sb.end);
        splice(p, sa, rep, width);
    }
};

/* The canonical (error, cost, slot) key of a parent. */
struct Key {
    int nan;
    uint64_t bits;
    uint32_t cost;
    size_t index;
    bool operator<(const Key &o) const
    {
        if (nan != o.nan) return nan < o.nan;
        if (bits != o.bits) return bits < o.bits;
        if (cost != o.cost) return cost < o.cost;
        return index < o.index;
    }
};

}  // namespace

extern "C" {

void or_philox(const uint32_t *key, const uint32_t *ctr, uint32_t *out)
{
    Philox::block(key, ctr, out);
}

/* One generation of the flat variation engine.  parents: islands * pop slots
   of width words; err: their fitness; restart: per island (nullable); ops:
   nops (code, arity) pairs in table order.  Writes the children's words,
   length, node count, hash and error.  Returns the number of child attempts
   made, or -1 on bad arguments. */
long long or_vary(const uint32_t *parents, int islands, int pop, int width, const double *err,
                  const unsigned char *restart, uint32_t seed, uint32_t gen, int max_complexity,
                  int nvars, int init_depth, int dedup, const uint32_t *ops, int nops,
                  uint32_t *out_words, uint16_t *out_len, uint16_t *out_cost, uint64_t *out_hash,
                  double *out_err)
{
    if (!parents || islands <= 0 || pop <= 0 || width < 2 || nops <= 0) return -1;
    Engine e;
    e.width = (size_t)width;
    e.nvars = (size_t)nvars;
    e.max_complexity = (size_t)max_complexity;
    e.init_depth = (size_t)init_depth;
    e.seed = seed;
    e.dedup = dedup != 0;
    for (int k = 0; k < nops; ++k) {
        const uint32_t code = ops[2 * k];
        const int ar = (int)ops[2 * k + 1];
        if (ar < 1 || ar > 3) return -1;
        bool dup = false;
        for (auto &o : e.ops) dup = dup || o.first == code;
        if (dup) continue;
        e.by_arity[ar].push_back(e.ops.size());
        e.ops.push_back({code, ar});
    }
    const size_t n = (size_t)islands * (size_t)pop, W = (size_t)width;
    /* Parent metadata from their words. */
    std::vector<uint64_t> phash(n);
    std::vector<uint32_t> pcost(n);
    for (size_t s = 0; s < n; ++s) {
        const uint32_t *slot = parents + s * W;
        std::vector<uint32_t> prog(slot, slot + logical_len(slot, W));
        phash[s] = fnv(prog);
        size_t nodes = 0;
        pcost[s] = parse(prog, nodes) ? (uint32_t)std::min<size_t>(nodes, 65535) : 0;
    }
    auto key = [&](size_t s) {
        Key k;
        const double v = err[s];
        k.nan = isnan(v) ? 1 : 0;
        uint64_t b;
        memcpy(&b, &v, 8);
        k.bits = isnan(v) ? 0 : ((b >> 63) ? ~b : (b | (1ull << 63)));
        k.cost = pcost[s];
        k.index = s;
        return k;
    };
    auto tournament = [&](size_t island, Philox &rng) {
        const size_t base = island * (size_t)pop;
        size_t best = base + rng.below((size_t)pop);
        for (int t = 1; t < 4; ++t) {
            const size_t c = base + rng.below((size_t)pop);
            if (key(c) < key(best)) best = c;
        }
        return best;
    };
    for (size_t s = 0; s < n; ++s) {
        for (size_t w = 0; w < W; ++w) out_words[s * W + w] = HALT;
        out_len[s] = 0;
        out_cost[s] = 0;
        out_hash[s] = 0;
        out_err[s] = NAN;
    }
    auto write = [&](size_t s, const std::vector<uint32_t> &prog) {
        size_t nodes = 0;
        if (prog.size() < 2 || prog.size() > W || !parse(prog, nodes)) return false;
        for (size_t w = 0; w < W; ++w) out_words[s * W + w] = w < prog.size() ? prog[w] : HALT;
        out_len[s] = (uint16_t)prog.size();
        out_cost[s] = (uint16_t)std::min<size_t>(nodes, 65535);
        out_hash[s] = fnv(prog);
        return true;
    };
    auto fill_random = [&](size_t island, size_t slot, uint32_t salt) {
        Philox rng(seed, (uint32_t)island, gen, (uint32_t)slot);
        for (uint32_t q = 0; q < salt; ++q) rng.first();
        std::vector<uint32_t> best;
        for (int t = 0; t < 8; ++t) {
            std::vector<uint32_t> prog = e.random_program(rng);
            size_t nodes = 0;
            if (parse(prog, nodes) && nodes <= e.max_complexity && prog.size() <= W) {
                best = prog;
                break;
            }
        }
        if (best.empty()) best = {VARW, 0, HALT};
        const size_t s = island * (size_t)pop + slot;
        write(s, best);
        out_err[s] = NAN;
    };
    std::unordered_set<uint64_t> seen(phash.begin(), phash.end());
    long long attempts = 0;
    for (size_t i = 0; i < (size_t)islands; ++i) {
        if (restart && restart[i]) {
            for (size_t j = 0; j < (size_t)pop; ++j) fill_random(i, j, 0);
            continue;
        }
        size_t elite = i * (size_t)pop;
        for (size_t s = elite + 1; s < (i + 1) * (size_t)pop; ++s)
            if (key(s) < key(elite)) elite = s;
        const size_t d0 = i * (size_t)pop;
        for (size_t w = 0; w < W; ++w) out_words[d0 * W + w] = parents[elite * W + w];
        out_len[d0] = (uint16_t)logical_len(parents + elite * W, W);
        out_cost[d0] = (uint16_t)pcost[elite];
        out_hash[d0] = phash[elite];
        out_err[d0] = err[elite];
        for (size_t j = 1; j < (size_t)pop; ++j) {
            bool done = false;
            for (uint32_t attempt = 0; attempt < 64 && !done; ++attempt) {
                ++attempts;
                Philox rng(seed, (uint32_t)i, gen, (uint32_t)j);
                rng.ctr[2] = attempt;
                const uint32_t *ps = parents + tournament(i, rng) * W;
                std::vector<uint32_t> prog(ps, ps + logical_len(ps, W));
                if (rng.unit() < 0.5) {
                    const uint32_t *os = parents + tournament(i, rng) * W;
                    const std::vector<uint32_t> other(os, os + logical_len(os, W));
                    e.crossover(prog, other, rng);
                }
                e.mutate(prog, rng);
                if (rng.unit() < 0.5) e.mutate(prog, rng);
                std::vector<Node> nodes;
                if (prog.size() > W || !scan(prog, nodes)) continue;
                if (nodes.size() > e.max_complexity) continue;
                const uint64_t h = fnv(prog);
                if (e.dedup && !seen.insert(h).second) continue;
                done = write(i * (size_t)pop + j, prog);
            }
            if (!done) fill_random(i, j, 64);
        }
    }
    return attempts;
}

}  // extern "C"
