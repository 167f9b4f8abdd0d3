/*
 * Batched expression-bytecode evaluation on gfx950.
 *
 * Reference behaviour: rc_exec_row (eval.cu:404-573), eval_kernel /
 * eval_pool_kernel (759-828), score_kernel modes 0-4, 7-9 (846-1088),
 * predict_kernel (1093-1104) and the ABI at 3268-3450, 3514-3600, 3803-3880.
 *
 * MI355X design (not the reference's block-per-program layout):
 *  - the unit of work is one wavefront on one (program, row tile); a launch
 *    has nprog x tiles single-wave workgroups, so 4096 programs x 1M rows is
 *    ~1M waves and every CU stays busy regardless of the program mix;
 *  - the program stream is wave-uniform: instruction words come through the
 *    scalar cache (s_load), the operator switch is a uniform branch and the
 *    operand stack's depth is an SGPR, so the stack lives in VGPRs addressed
 *    with s_set_gpr_idx (no scratch, no LDS).  The top of stack is cached in
 *    a fixed register so a binary operator costs one indexed read;
 *  - each lane evaluates R rows per pass (rows base + r*64 + lane), which
 *    amortises the scalar dispatch over R rows and keeps every column read a
 *    coalesced 256-byte wave access;
 *  - workgroups are mapped XCD-aware: workgroup g runs on XCD g % 8, and XCD x
 *    is given the x-th eighth of the row tiles for every program, so a tile of
 *    rows is pulled into one XCD's L2 once and reused by all programs;
 *  - stack errors (underflow, overflow past 64, missing HALT) are static
 *    properties of a program, so the host checks them once and replaces an
 *    erroneous program by a NaN program; the kernel carries no per-op guard;
 *  - the host rebuilds each program's tree and re-emits it: input-free
 *    subtrees are evaluated once per batch by the same interpreter
 *    (rc_hoist_kernel) and enter the row streams as constants, and +, -, *, /
 *    with a leaf operand become one fused word, in the original operand
 *    order.  A C2-like batch (3-60 random nodes) dispatches about a third
 *    fewer words per row pass and skips a quarter of its transcendental
 *    calls, with every loss bit-identical to the plain postfix stream;
 *  - partial sums are per (program, tile) in f64 with a fixed wave butterfly,
 *    finished by a per-program pass in tile order: results depend only on the
 *    program and n, never on batch composition.
 */
#include "rc_device.h"
#include "rc_internal.h"
#include "rcsr_host.h"

#include <string.h>

#include <algorithm>
#include <cmath>

#include <mutex>
#include <vector>

namespace {

/* Evaluator internal stream: OP_CONST w, f32 bits | OP_VAR | (col << 16) |
   operator words | OP_HALT.  Unknown operators become RC_OP_NAN (sets the top
   of stack to NaN, the reference's unary default). */
constexpr uint32_t RC_OP_NAN = 254;
/* Fused words RC_OP_FUSED + (bop * 4 + left * 2 + var): tos = tos bop leaf
   (left = 0) or leaf bop tos (left = 1) for bop 0..3 = +, -, *, /; the leaf is
   input column (w >> 16) (var = 1) or the f32 word that follows (var = 0). */
constexpr uint32_t RC_OP_FUSED = 64;

/* Internal score pass: the per-program mean prediction that the correlation
   scores centre on (their first pass, eval.cu:930-960). */
constexpr int RC_SCORE_CORR_MEAN = 100;

constexpr int MODE_EVAL = 0;    /* sum of rho with 1e30 hard failures (rc_eval) */
constexpr int MODE_SCORE = 1;   /* f64 score objectives + bad flag */
constexpr int MODE_PREDICT = 2; /* raw predictions */

bool is_binary_op(uint32_t op)
{
    return (op >= OP_ADD && op <= OP_DIV) || op == OP_POW || op == OP_MOD || op == OP_ATAN2 ||
           op == OP_MIN || op == OP_MAX || (op >= OP_EQUAL && op <= OP_XOR);
}

bool is_unary_op(uint32_t op)
{
    return (op >= OP_SIN && op <= OP_GAUSS && op != OP_POW) || op == OP_STEP || op == OP_SINH ||
           op == OP_COSH || op == OP_FLOOR || op == OP_CEIL || op == OP_ROUND || op == OP_ERF ||
           op == OP_ERFC || op == OP_SIGN || op == OP_NOT;
}

/* One parsed node: a leaf (OP_CONST with its f32 bits, OP_VAR with its
   column) or an operator with its operands' node indices. */
struct ENode {
    uint32_t op;
    uint32_t arg;
    int kid[3];
};

/* Parse one program with the reference's static stack checks (rc_exec_row
   eval.cu:404-573 fails a row on underflow, overflow past 64 entries or a
   missing HALT; all three are properties of the program).  Nodes come out in
   postfix order, so a subtree is the contiguous range [first, root].  Returns
   false for a program the reference device would score as NaN on every row;
   *root is the bottom entry of the stack at HALT, the value the reference
   returns.  The nodes are appended to `nodes` (operand indices are indices
   into it); on failure the appended nodes are left for the caller to drop. */
bool parse_one(const uint32_t *w, size_t len, int nvars, std::vector<ENode> &nodes, int *root)
{
    int stk[RC_MAX_STACK];
    int depth = 0;
    size_t i = 0;
    for (;;) {
        if (i >= len) return false; /* no HALT in range */
        const uint32_t op = w[i++];
        ENode n = {op, 0u, {-1, -1, -1}};
        if (op == OP_CONST || op == OP_VAR) {
            if (i >= len) return false;
            const uint32_t arg = w[i++];
            if (depth >= RC_MAX_STACK) return false;
            if (op == OP_CONST || !((int)arg < nvars && arg < 65536u)) {
                /* A column outside the resident set reads NaN, not a stray
                   address. */
                n.op = OP_CONST;
                n.arg = (op == OP_CONST) ? arg : 0x7fc00000u;
            } else {
                n.arg = arg;
            }
        } else if (op == OP_HALT) {
            if (depth <= 0) return false;
            *root = stk[0];
            return true;
        } else if (op == OP_IF) {
            if (depth < 3) return false;
            depth -= 3;
            n.kid[0] = stk[depth];
            n.kid[1] = stk[depth + 1];
            n.kid[2] = stk[depth + 2];
        } else if (is_binary_op(op)) {
            if (depth < 2) return false;
            depth -= 2;
            n.kid[0] = stk[depth];
            n.kid[1] = stk[depth + 1];
        } else {
            if (depth < 1) return false;
            depth -= 1;
            n.op = is_unary_op(op) ? op : RC_OP_NAN;
            n.kid[0] = stk[depth];
        }
        stk[depth++] = (int)nodes.size();
        nodes.push_back(n);
    }
}

/* Emitter of one batch's streams.
 *
 * The main stream of a program re-emits the tree rooted at the value the
 * reference returns (operand stack entries above it are never read):
 *  - a subtree without inputs (constants and operators only) has the same
 *    value on every row.  It is emitted once, as a stand-alone program in the
 *    batch's hoist section, and its place in the main stream becomes a
 *    constant whose word the device fills in by running that program with the
 *    very interpreter that runs the rows (rc_hoist_kernel), so the value is
 *    bit-for-bit the one each row would have computed;
 *  - a unary operator applied to an input column (or to such a derived
 *    column) recurs across the programs of a batch, e.g. sin(x1) hundreds of
 *    times in 4096 random programs over 5 inputs.  The pairs (operator,
 *    column) used at least twice become derived columns, computed once over
 *    the rows by the interpreter's own unary code (rc_derive_kernel) and kept
 *    for the evaluator's lifetime (the f32 rows never change); the node then
 *    reads that column like an input;
 *  - +, -, *, / with a leaf operand (input or derived column, constant,
 *    hoisted constant) become one fused word that applies the operator to the
 *    top of stack and the leaf, in the original operand order (so results,
 *    NaN payloads included, are the unfused ones); when only the left operand
 *    is a leaf the right subtree is evaluated first, which only reorders
 *    independent, side-effect-free work.
 */
struct Emitter {
    const std::vector<ENode> *nodes = nullptr;
    std::vector<int> first, height, colid;
    std::vector<char> cst;
    std::vector<uint32_t> *code = nullptr;
    std::vector<uint32_t> hcode;               /* hoist section, offsets relative to it */
    std::vector<uint32_t> patch;               /* pairs (hoist offset, main word index) */
    int depth = 0, peak = 0;

    void analyse(const std::vector<ENode> &ns)
    {
        nodes = &ns;
        const size_t m = ns.size();
        first.assign(m, 0);
        height.assign(m, 1);
        cst.assign(m, 0);
        colid.assign(m, -1);
        for (size_t j = 0; j < m; ++j) {
            const ENode &n = ns[j];
            if (n.kid[0] < 0) {
                first[j] = (int)j;
                cst[j] = n.op == OP_CONST;
                if (n.op == OP_VAR) colid[j] = (int)n.arg;
                continue;
            }
            first[j] = first[n.kid[0]];
            bool c = true;
            int h = 0;
            for (int k = 0; k < 3 && n.kid[k] >= 0; ++k) {
                c = c && cst[n.kid[k]];
                h = std::max(h, height[n.kid[k]]);
            }
            cst[j] = c;
            height[j] = h + 1;
        }
    }

    bool leafish(int j) const { return (*nodes)[j].kid[0] < 0 || cst[j] || colid[j] >= 0; }

    void bump(int d)
    {
        depth += d;
        peak = std::max(peak, depth);
    }

    /* Plain postfix of subtree j into `out`. */
    void plain(int j, std::vector<uint32_t> &out) const
    {
        for (int q = first[j]; q <= j; ++q) {
            const ENode &n = (*nodes)[q];
            if (n.op == OP_CONST) {
                out.push_back(OP_CONST);
                out.push_back(n.arg);
            } else if (n.op == OP_VAR) {
                out.push_back(OP_VAR | (n.arg << 16));
            } else {
                out.push_back(n.op);
            }
        }
    }

    /* The constant word of constant-valued node j (a literal, or a hoisted
       subtree whose word the device fills in). */
    void const_word(int j)
    {
        const ENode &n = (*nodes)[j];
        if (n.kid[0] < 0) {
            code->push_back(n.arg);
            return;
        }
        patch.push_back((uint32_t)hcode.size());
        patch.push_back((uint32_t)code->size());
        code->push_back(0x7fc00000u);
        plain(j, hcode);
        hcode.push_back(OP_HALT);
    }

    void push_leaf(int j)
    {
        bump(1);
        if (colid[j] >= 0) {
            code->push_back(OP_VAR | ((uint32_t)colid[j] << 16));
        } else {
            code->push_back(OP_CONST);
            const_word(j);
        }
    }

    static int fused_index(uint32_t op)
    {
        return op == OP_ADD ? 0 : op == OP_SUB ? 1 : op == OP_MUL ? 2 : op == OP_DIV ? 3 : -1;
    }

    void fused(int bop, bool left, int leaf)
    {
        const bool var = colid[leaf] >= 0;
        const uint32_t h = RC_OP_FUSED + (uint32_t)(bop * 4 + (left ? 2 : 0) + (var ? 1 : 0));
        if (var) {
            code->push_back(h | ((uint32_t)colid[leaf] << 16));
        } else {
            code->push_back(h);
            const_word(leaf);
        }
    }

    void emit(int j)
    {
        const ENode &n = (*nodes)[j];
        if (leafish(j)) {
            push_leaf(j);
            return;
        }
        const int bop = fused_index(n.op);
        if (bop >= 0) {
            if (leafish(n.kid[1])) {
                emit(n.kid[0]);
                fused(bop, false, n.kid[1]);
                return;
            }
            if (leafish(n.kid[0])) {
                emit(n.kid[1]);
                fused(bop, true, n.kid[0]);
                return;
            }
        }
        int nk = 0;
        for (int k = 0; k < 3 && n.kid[k] >= 0; ++k, ++nk) emit(n.kid[k]);
        code->push_back(n.op);
        bump(1 - nk);
    }
};

/* Trees taller than this are emitted in their plain postfix form (the
   emitter recurses once per level). */
constexpr int RC_EMIT_MAX_HEIGHT = 2048;

/* Derived columns: passes of (assign known pairs, admit new ones). */
constexpr int RC_DERIVE_PASSES = 4;

uint64_t derive_key(uint32_t op, int col) { return ((uint64_t)op << 32) | (uint32_t)col; }

/* Open-addressing map from a nonzero derive key to an int (unary operator
   codes are nonzero, so key 0 marks an empty slot). */
struct KeyMap {
    std::vector<uint64_t> key;
    std::vector<int> val;
    size_t mask = 0;

    void reset(size_t n)
    {
        size_t cap = 16;
        while (cap < 2 * n) cap <<= 1;
        key.assign(cap, 0);
        val.assign(cap, 0);
        mask = cap - 1;
    }
    size_t slot(uint64_t k) const
    {
        size_t i = (size_t)((k * 0x9e3779b97f4a7c15ull) >> 17) & mask;
        while (key[i] != 0 && key[i] != k) i = (i + 1) & mask;
        return i;
    }
    const int *find(uint64_t k) const
    {
        const size_t i = slot(k);
        return key[i] == k ? &val[i] : nullptr;
    }
    int &at(uint64_t k)
    {
        const size_t i = slot(k);
        key[i] = k;
        return val[i];
    }
};

/* One contiguous range of a batch's programs: its parsed trees, their
   analysis and its share of the streams.  Node and stream indices are local
   to the chunk; the chunks are joined in program order. */
struct Chunk {
    int p0 = 0, p1 = 0;
    std::vector<ENode> all;
    std::vector<int> pstart, proot, depth;
    std::vector<int> un;       /* unary nodes of the live trees, in tree order */
    std::vector<uint32_t> end; /* end of each program's main stream in code */
    Emitter em;
    std::vector<uint32_t> code;
    KeyMap count;
    std::vector<uint64_t> seen;
    int peak = 1;

    void parse(const uint32_t *prog, const uint32_t *off, int nvars)
    {
        const int np = p1 - p0;
        all.reserve(off[p1] > off[p0] ? (size_t)(off[p1] - off[p0]) : 0);
        pstart.assign((size_t)np + 1, 0);
        proot.assign((size_t)np, -1);
        for (int q = 0; q < np; ++q) {
            pstart[q] = (int)all.size();
            const size_t a = off[p0 + q], b = off[p0 + q + 1];
            int root = -1;
            if (b > a && parse_one(prog + a, b - a, nvars, all, &root))
                proot[q] = root;
            else
                all.resize((size_t)pstart[q]);
        }
        pstart[np] = (int)all.size();
        em.code = &code;
        em.analyse(all);
        for (int q = 0; q < np; ++q) {
            if (proot[q] < 0) continue;
            for (int j = pstart[q]; j <= proot[q]; ++j) {
                const ENode &n = all[j];
                if (n.kid[0] >= 0 && n.kid[1] < 0 && n.op != RC_OP_NAN) un.push_back(j);
            }
        }
    }

    /* One pass of derived-column assignment: a unary-of-column node whose
       pair is known reads that column (so its parent, later in the order,
       is seen in the same pass); unknown pairs are counted when `tally`. */
    void derive_pass(const KeyMap &have, bool tally)
    {
        if (tally) count.reset(un.size());
        seen.clear();
        for (const int j : un) {
            if (em.colid[j] >= 0) continue;
            const int src = em.colid[all[j].kid[0]];
            if (src < 0) continue;
            const uint64_t key = derive_key(all[j].op, src);
            if (const int *col = have.find(key)) {
                em.colid[j] = *col;
            } else if (tally) {
                int &c = count.at(key);
                if (c++ == 0) seen.push_back(key);
            }
        }
    }

    void emit_all()
    {
        const int np = p1 - p0;
        depth.assign((size_t)np, 1);
        end.assign((size_t)np, 0);
        for (int q = 0; q < np; ++q) {
            const int root = proot[q];
            int d;
            if (root >= 0) {
                em.depth = 0;
                em.peak = 0;
                if (em.height[root] <= RC_EMIT_MAX_HEIGHT) {
                    em.emit(root);
                } else {
                    /* Plain postfix: the stack peak is the parse's. */
                    em.plain(root, code);
                    int dd = 0;
                    for (int j = em.first[root]; j <= root; ++j) {
                        int nk = 0;
                        while (nk < 3 && all[j].kid[nk] >= 0) ++nk;
                        dd += 1 - nk;
                        em.peak = std::max(em.peak, dd);
                    }
                }
                code.push_back(OP_HALT);
                d = std::max(em.peak, 1);
            } else {
                code.push_back(OP_CONST);
                code.push_back(0x7fc00000u);
                code.push_back(OP_HALT);
                d = 1;
            }
            depth[q] = d;
            peak = std::max(peak, d);
            end[q] = (uint32_t)code.size();
        }
    }
};

/* Admit the unknown pairs used at least twice across the chunks, most used
   first, then by operator and column, up to max_derived columns in all.
   Returns how many were admitted. */
int admit_derived(std::vector<Chunk> &ch, KeyMap &have, int nvars, std::vector<uint32_t> &derived,
                  int max_derived)
{
    size_t nseen = 0;
    for (const Chunk &c : ch) nseen += c.seen.size();
    KeyMap total;
    total.reset(nseen);
    std::vector<uint64_t> keys;
    for (const Chunk &c : ch)
        for (const uint64_t key : c.seen) {
            int &t = total.at(key);
            if (t == 0) keys.push_back(key);
            t += *c.count.find(key);
        }
    std::vector<std::pair<int, uint64_t>> cand;
    for (const uint64_t key : keys) {
        const int t = *total.find(key);
        if (t >= 2) cand.push_back({-t, key});
    }
    std::sort(cand.begin(), cand.end());
    int admitted = 0;
    for (const auto &c : cand) {
        if ((int)(derived.size() / 2) >= max_derived) break;
        have.at(c.second) = nvars + (int)(derived.size() / 2);
        derived.push_back((uint32_t)(c.second >> 32));
        derived.push_back((uint32_t)(c.second & 0xffffffffu));
        ++admitted;
    }
    return admitted;
}

} /* namespace */

/* Translate a batch.  code = [main streams | hoist section]; meta = [main
   code offset per program | program list grouped by stack class | pairs
   (hoist program offset, constant word index) for rc_hoist_kernel];
   cls_start receives the class boundaries in the program list.  derived
   (when given) holds the derived columns already computed, as (operator,
   source column) pairs for columns nvars, nvars + 1, ...; pairs the batch
   admits are appended and must be computed before the streams run.

   Derived columns: the unary-of-column nodes of the live trees are marked
   with their columns in passes; between passes the pairs used at least
   twice are admitted.  Programs are translated in chunks on host threads;
   the output depends only on the batch, never on the chunking. */
int rc_translate_programs(const uint32_t *prog, const uint32_t *off, int nprog, int nvars,
                          std::vector<uint32_t> &code, std::vector<uint32_t> &meta,
                          int cls_start[5], std::vector<uint32_t> *derived, int max_derived)
{
    code.clear();
    meta.assign((size_t)nprog * 2, 0);
    const int nch = std::max(1, std::min(8, nprog / 256));
    std::vector<Chunk> ch((size_t)nch);
    for (int i = 0; i < nch; ++i) {
        ch[i].p0 = (int)((long long)nprog * i / nch);
        ch[i].p1 = (int)((long long)nprog * (i + 1) / nch);
    }
    auto each = [&](const std::function<void(Chunk &)> &fn) {
        rc_parallel_for(nch, 1, [&](int lo, int hi) {
            for (int i = lo; i < hi; ++i) fn(ch[i]);
        });
    };
    each([&](Chunk &c) { c.parse(prog, off, nvars); });
    if (derived) {
        const int cap = std::min(max_derived, 65535 - nvars);
        size_t nun = 0;
        for (const Chunk &c : ch) nun += c.un.size();
        if (nun > 0) {
            KeyMap have;
            have.reset(derived->size() / 2 + std::min(nun, (size_t)std::max(0, cap)));
            for (size_t k = 0; k + 1 < derived->size(); k += 2)
                have.at(derive_key((*derived)[k], (int)(*derived)[k + 1])) = nvars + (int)(k / 2);
            for (int pass = 0; pass < RC_DERIVE_PASSES; ++pass) {
                const bool last = pass == RC_DERIVE_PASSES - 1;
                each([&](Chunk &c) { c.derive_pass(have, !last); });
                if (last || admit_derived(ch, have, nvars, *derived, cap) == 0) break;
            }
        }
    }
    each([&](Chunk &c) { c.emit_all(); });
    /* Join: main streams in program order, then the hoist section. */
    size_t main_words = 0, hoist_words = 0, npatch = 0;
    for (const Chunk &c : ch) {
        main_words += c.code.size();
        hoist_words += c.em.hcode.size();
        npatch += c.em.patch.size();
    }
    code.reserve(main_words + hoist_words);
    int peak = 1;
    for (const Chunk &c : ch) {
        const uint32_t base = (uint32_t)code.size();
        for (int q = 0; q < c.p1 - c.p0; ++q)
            meta[(size_t)(c.p0 + q)] = base + (q ? c.end[q - 1] : 0u);
        code.insert(code.end(), c.code.begin(), c.code.end());
        peak = std::max(peak, c.peak);
    }
    int k = 0;
    for (int cls = 0; cls < 4; ++cls) {
        cls_start[cls] = k;
        for (const Chunk &c : ch)
            for (int q = 0; q < c.p1 - c.p0; ++q) {
                const int d = c.depth[q];
                const int mine = d <= 8 ? 0 : d <= 16 ? 1 : d <= 32 ? 2 : 3;
                if (mine == cls) meta[(size_t)nprog + k++] = (uint32_t)(c.p0 + q);
            }
    }
    cls_start[4] = k;
    const uint32_t hbase = (uint32_t)code.size();
    meta.reserve(meta.size() + npatch);
    uint32_t mbase = 0, hofs = 0;
    for (const Chunk &c : ch) {
        code.insert(code.end(), c.em.hcode.begin(), c.em.hcode.end());
        for (size_t q = 0; q < c.em.patch.size(); q += 2) {
            meta.push_back(hbase + hofs + c.em.patch[q]);
            meta.push_back(mbase + c.em.patch[q + 1]);
        }
        mbase += (uint32_t)c.code.size();
        hofs += (uint32_t)c.em.hcode.size();
    }
    return peak;
}

/* ------------------------------------------------------------------------ */
/* Device interpreter.                                                       */

__device__ __noinline__ float rc_unary_call(unsigned op, float a) { return rc_unary_f32(op, a); }
__device__ __noinline__ float rc_binary_call(unsigned op, float a, float b)
{
    return rc_binary_f32(op, a, b);
}

/* A fused word's operator on the top of stack and a leaf value (an SGPR
   constant or a per-row load), in the program's operand order. */
template <int R, bool LEFT, typename Leaf>
__device__ __forceinline__ void rc_fused(uint32_t bop, float *tos, Leaf leaf)
{
    if (bop == 0) {
#pragma unroll
        for (int r = 0; r < R; ++r) tos[r] = LEFT ? leaf(r) + tos[r] : tos[r] + leaf(r);
    } else if (bop == 1) {
#pragma unroll
        for (int r = 0; r < R; ++r) tos[r] = LEFT ? leaf(r) - tos[r] : tos[r] - leaf(r);
    } else if (bop == 2) {
#pragma unroll
        for (int r = 0; r < R; ++r) tos[r] = LEFT ? leaf(r) * tos[r] : tos[r] * leaf(r);
    } else {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const float a = LEFT ? leaf(r) : tos[r];
            const float b = LEFT ? tos[r] : leaf(r);
            tos[r] = (fabsf(b) < 1e-12f) ? RC_F32_NAN : a / b;
        }
    }
}

/* Evaluate the program at `pc` for R rows of this lane: rows row0 +
   loff[r] / 4 (clamped in range by the caller). */
/* Score inputs beyond the rows: score weights (may be null), label signs,
   per-program mean predictions, the target mean, second per-item sums. */
struct RcScoreArgs {
    const double *w;
    const float *tsign;
    const double *pbar;
    double ybar;
    double *part2;
};

/* Column c of the evaluator's rows: an input (c < nv) or a derived column. */
__device__ __forceinline__ const float *rc_col(const float *cols, const float *dcols, int n, int nv,
                                               uint32_t c)
{
    return (c < (uint32_t)nv) ? cols + (size_t)c * (size_t)n : dcols + (size_t)(c - nv) * (size_t)n;
}

/* Row (row0 + loff / 4) of a column: a wave-uniform 64-bit base plus a
   32-bit per-lane byte offset within the row tile. */
__device__ __forceinline__ float rc_ld(const float *col, int row0, uint32_t loff)
{
    const char *b = (const char *)(col + row0);
    return *(const float *)(b + loff);
}

template <int S, int R>
__device__ __forceinline__ void rc_run(const uint32_t *__restrict__ pc,
                                       const float *__restrict__ cols,
                                       const float *__restrict__ dcols, int n, int nv,
                                       int row0, const uint32_t *loff, float *res)
{
    /* st[r][d]: the compiler keeps this in VGPRs (indexed by the uniform
       depth) as long as R * S <= 32; the per-class (S, R) choice keeps it so. */
    float tos[R];
    float st[R][S];
    int sp = 0;      /* entries below the top of stack */
    int depth = 0;
#pragma unroll
    for (int r = 0; r < R; ++r) tos[r] = 0.0f;
    for (;;) {
        const uint32_t w = *pc++;
        const uint32_t op = w & 0xffu;
        if (op == OP_VAR || op == OP_CONST) {
            if (depth > 0) {
#pragma unroll
                for (int r = 0; r < R; ++r) st[r][sp] = tos[r];
                ++sp;
            }
            ++depth;
            if (op == OP_VAR) {
                const float *col = rc_col(cols, dcols, n, nv, w >> 16);
#pragma unroll
                for (int r = 0; r < R; ++r) tos[r] = rc_ld(col, row0, loff[r]);
            } else {
                const float c = __uint_as_float(*pc++);
#pragma unroll
                for (int r = 0; r < R; ++r) tos[r] = c;
            }
        } else if (op >= RC_OP_FUSED && op < RC_OP_FUSED + 16u) {
            const uint32_t f = op - RC_OP_FUSED;
            const uint32_t bop = f >> 2;
            if (f & 1u) {
                const float *col = rc_col(cols, dcols, n, nv, w >> 16);
                float lv[R];
#pragma unroll
                for (int r = 0; r < R; ++r) lv[r] = rc_ld(col, row0, loff[r]);
                auto leaf = [&](int r) { return lv[r]; };
                if (f & 2u) rc_fused<R, true>(bop, tos, leaf);
                else rc_fused<R, false>(bop, tos, leaf);
            } else {
                const float c = __uint_as_float(*pc++);
                auto leaf = [=](int) { return c; };
                if (f & 2u) rc_fused<R, true>(bop, tos, leaf);
                else rc_fused<R, false>(bop, tos, leaf);
            }
        } else if (op == OP_MUL) {
            --sp; --depth;
#pragma unroll
            for (int r = 0; r < R; ++r) tos[r] = st[r][sp] * tos[r];
        } else if (op == OP_ADD) {
            --sp; --depth;
#pragma unroll
            for (int r = 0; r < R; ++r) tos[r] = st[r][sp] + tos[r];
        } else if (op == OP_SUB) {
            --sp; --depth;
#pragma unroll
            for (int r = 0; r < R; ++r) tos[r] = st[r][sp] - tos[r];
        } else if (op == OP_DIV) {
            --sp; --depth;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const float b = tos[r];
                tos[r] = (fabsf(b) < 1e-12f) ? RC_F32_NAN : st[r][sp] / b;
            }
        } else if (op == OP_HALT) {
            break;
        } else if (op == OP_NEG) {
#pragma unroll
            for (int r = 0; r < R; ++r) tos[r] = -tos[r];
        } else if (op == OP_SQUARE) {
#pragma unroll
            for (int r = 0; r < R; ++r) tos[r] = tos[r] * tos[r];
        } else if (op == OP_IF) {
            sp -= 2; depth -= 2;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const float cnd = st[r][sp];
                const float thn = st[r][sp + 1];
                tos[r] = (cnd > 0.0f) ? thn : tos[r];
            }
        } else if (op == OP_EXP) {
#pragma unroll
            for (int r = 0; r < R; ++r) tos[r] = expf(tos[r]);
        } else if (op == OP_LOG) {
#pragma unroll
            for (int r = 0; r < R; ++r) tos[r] = (tos[r] > 0.0f) ? logf(tos[r]) : RC_F32_NAN;
        } else if (op == OP_SIN) {
#pragma unroll
            for (int r = 0; r < R; ++r) tos[r] = sinf(tos[r]);
        } else if (op == OP_COS) {
#pragma unroll
            for (int r = 0; r < R; ++r) tos[r] = cosf(tos[r]);
        } else if (op == OP_POW || op == OP_MOD || op == OP_ATAN2 || op == OP_MIN ||
                   op == OP_MAX || (op >= OP_EQUAL && op <= OP_XOR)) {
            /* Rare binary operators: one out-of-line call per row keeps the
               interpreter's code small; the indices stay static. */
            --sp; --depth;
#pragma unroll
            for (int r = 0; r < R; ++r) tos[r] = rc_binary_call(op, st[r][sp], tos[r]);
        } else {
#pragma unroll
            for (int r = 0; r < R; ++r) tos[r] = rc_unary_call(op, tos[r]);
        }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) res[r] = (sp > 0) ? st[r][0] : tos[r];
}

/* Map workgroup g to (program, tile) so that XCD g % 8 owns a contiguous
   eighth of the tiles. Returns false for a padding workgroup. */
__device__ __forceinline__ bool rc_item_of(int g, int nprog, int tiles, int *p, int *t)
{
    if (tiles < 8) {
        if (g >= nprog * tiles) return false;
        *p = g / tiles;
        *t = g % tiles;
        return true;
    }
    const int x = g & 7;
    const int j = g >> 3;
    const int t0 = (int)(((long long)tiles * x) >> 3);
    const int t1 = (int)(((long long)tiles * (x + 1)) >> 3);
    const int tt = j / nprog;
    if (tt >= t1 - t0) return false;
    *t = t0 + tt;
    *p = j - tt * nprog;
    return true;
}

template <int S, int R, int MODE>
__global__ __launch_bounds__(64) void rc_eval_kernel(
    const float *__restrict__ cols, const float *__restrict__ dcols, int nv,
    const float *__restrict__ y, int n,
    const uint32_t *__restrict__ code, const uint32_t *__restrict__ meta,
    const uint32_t *__restrict__ plist, int ncls, int tiles, int tile_rows, int kind,
    double *__restrict__ part, int *__restrict__ part_bad, float *__restrict__ pred_out,
    RcScoreArgs sa)
{
    int j, t;
    if (!rc_item_of(blockIdx.x, ncls, tiles, &j, &t)) return;
    const int p = (int)plist[j];
    const int lane = threadIdx.x;
    const uint32_t *pc = code + meta[p];
    const int row0 = t * tile_rows;
    const int row1 = min(n, row0 + tile_rows);
    double acc = 0.0, acc2 = 0.0;
    int bad = 0;
    const bool corr = MODE == MODE_SCORE &&
                      (kind == RC_SCORE_CORRELATION || kind == RC_SCORE_CORRELATION_SIGNED);
    const bool classif = MODE == MODE_SCORE && kind >= RC_SCORE_LOG_LOSS &&
                         kind <= RC_SCORE_HINGE_SQUARED;
    const double pb = corr ? sa.pbar[p] : 0.0;
    for (int base = row0; base < row1; base += 64 * R) {
        uint32_t loff[R];
        bool ok[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int row = base + r * 64 + lane;
            ok[r] = row < row1;
            loff[r] = ok[r] ? (uint32_t)(row - row0) * 4u : 0u;
        }
        float yv[R];
        if (MODE != MODE_PREDICT) {
#pragma unroll
            for (int r = 0; r < R; ++r) yv[r] = rc_ld(y, row0, loff[r]);
        }
        float res[R];
        rc_run<S, R>(pc, cols, dcols, n, nv, row0, loff, res);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (!ok[r]) continue;
            const float pred = res[r];
            if (MODE == MODE_PREDICT) {
                pred_out[(size_t)p * (size_t)n + (size_t)(row0 + (loff[r] >> 2))] = pred;
                continue;
            }
            const bool hard = !isfinite(pred) || fabsf(pred) > 1e30f;
            if (MODE == MODE_EVAL) {
                if (hard) {
                    acc += 1e30;
                } else {
                    /* rc_row_loss (eval.cu:648-664) on the f32 residual. */
                    const float rr = pred - yv[r];
                    float l;
                    if (kind == RC_LOSS_LOG_SQUASHED) {
                        const float q = rr * rr;
                        l = isfinite(q) ? log1pf(q) : 2.0f * logf(fabsf(rr));
                    } else if (kind == RC_LOSS_ABSOLUTE) {
                        l = fabsf(rr);
                    } else if (kind == RC_LOSS_SQUARED) {
                        l = rr * rr;
                    } else {
                        l = RC_F32_NAN;
                    }
                    acc += (double)l;
                }
            } else {
                if (classif) {
                    /* eval.cu:884-901: invalid weight -> bad; zero weight ->
                       row ignored; then label and prediction must be finite. */
                    const int row = row0 + (int)(loff[r] >> 2);
                    const double wi = sa.w ? sa.w[row] : 1.0;
                    if (!isfinite(wi) || wi < 0.0) {
                        bad = 1;
                        continue;
                    }
                    if (wi == 0.0) continue;
                    const float tt = sa.tsign[row];
                    if (!isfinite(tt) || hard) {
                        bad = 1;
                        continue;
                    }
                    acc += wi * (double)rc_class_row_loss_f32(pred, tt, kind);
                    acc2 += wi;
                    continue;
                }
                if (hard) {
                    bad = 1;
                    continue;
                }
                if (kind == RC_SCORE_CORR_MEAN) {
                    acc += (double)pred;
                    continue;
                }
                if (corr) {
                    const double dp = (double)pred - pb;
                    const double dy = (double)yv[r] - sa.ybar;
                    acc += dp * dp;
                    acc2 += dp * dy;
                    continue;
                }
                const double rr = (double)pred - (double)yv[r];
                if (kind == RC_SCORE_ABSOLUTE) {
                    acc += fabs(rr);
                } else if (kind == RC_SCORE_LOG_SQUASHED) {
                    const double q = rr * rr;
                    acc += isfinite(q) ? log1p(q) : 2.0 * log(fabs(rr));
                } else if (kind == RC_SCORE_SIGNED) {
                    acc += rr;
                } else if (kind == RC_SCORE_MAXIMUM) {
                    acc = fmax(acc, fabs(rr));
                } else {
                    acc += rr * rr;
                }
            }
        }
    }
    if (MODE == MODE_PREDICT) return;
    double tot;
    if (MODE == MODE_SCORE && kind == RC_SCORE_MAXIMUM) {
        tot = acc;
        for (int off = 32; off > 0; off >>= 1) tot = fmax(tot, __shfl_xor(tot, off, 64));
    } else {
        tot = rc_wave_sum_f64(acc);
    }
    const int anybad = __any(bad) ? 1 : 0;
    const bool two = corr || classif;
    const double tot2 = two ? rc_wave_sum_f64(acc2) : 0.0;
    if (lane == 0) {
        const size_t item = (size_t)p * (size_t)tiles + (size_t)t;
        part[item] = tot;
        part_bad[item] = anybad;
        if (two) sa.part2[item] = tot2;
    }
}

/* Values of the batch's input-free subtrees: wave h runs hoist program
   hoist[2h] with the row interpreter (all lanes alike) and stores the result
   into constant word hoist[2h + 1] of the main streams. */
__global__ __launch_bounds__(64) void rc_hoist_kernel(uint32_t *__restrict__ code,
                                                      const uint32_t *__restrict__ hoist, int nhoist)
{
    const int h = blockIdx.x;
    if (h >= nhoist) return;
    const uint32_t loff[1] = {0u};
    float res[1];
    rc_run<RC_MAX_STACK, 1>(code + hoist[2 * h], nullptr, nullptr, 0, 0, 0, loff, res);
    if (threadIdx.x == 0) code[hoist[2 * h + 1]] = __float_as_uint(res[0]);
}

/* Derived column dst = op(column src) over the rows, with the operator code
   the interpreter's unary handlers use (rc_unary_f32 is the same arithmetic
   as its inline exp / log / sin / cos / neg / square). */
__global__ __launch_bounds__(256) void rc_derive_kernel(const float *__restrict__ cols,
                                                        float *__restrict__ dcols, int n, int nv,
                                                        uint32_t op, int src, int dst)
{
    const int row = blockIdx.x * 256 + threadIdx.x;
    if (row >= n) return;
    const float *s = rc_col(cols, dcols, n, nv, (uint32_t)src);
    dcols[(size_t)dst * (size_t)n + (size_t)row] = rc_unary_f32(op, s[row]);
}

/* Per-program finish in tile order (deterministic). */
__global__ void rc_eval_finish(const double *__restrict__ part, const int *__restrict__ part_bad,
                               const double *__restrict__ part2, int nprog, int tiles, int n,
                               int mode, int kind, double syy, float *__restrict__ out32,
                               double *__restrict__ out64, int *__restrict__ status,
                               double *__restrict__ pbar)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nprog) return;
    const bool two = mode == MODE_SCORE &&
                     (kind == RC_SCORE_CORRELATION || kind == RC_SCORE_CORRELATION_SIGNED ||
                      (kind >= RC_SCORE_LOG_LOSS && kind <= RC_SCORE_HINGE_SQUARED));
    double s = 0.0, s2 = 0.0;
    int bad = 0;
    for (int t = 0; t < tiles; ++t) {
        const size_t item = (size_t)p * (size_t)tiles + (size_t)t;
        if (mode == MODE_SCORE && kind == RC_SCORE_MAXIMUM) s = fmax(s, part[item]);
        else s += part[item];
        if (two) s2 += part2[item];
        bad |= part_bad[item];
    }
    if (mode == MODE_EVAL) {
        out32[p] = (float)s;
        return;
    }
    if (bad) {
        out64[p] = __builtin_inf();
        status[p] = RC_SCORE_NONFINITE;
        return;
    }
    status[p] = 0;
    if (kind == RC_SCORE_CORR_MEAN) {
        pbar[p] = s / (double)n;
    } else if (kind >= RC_SCORE_LOG_LOSS && kind <= RC_SCORE_HINGE_SQUARED) {
        /* Weighted mean; no positive weight is status 4 (eval.cu:917-925). */
        if (!(s2 > 0.0)) {
            out64[p] = __builtin_inf();
            status[p] = RC_SCORE_ALL_ZERO_WEIGHT;
        } else {
            out64[p] = s / s2;
        }
    } else if (kind == RC_SCORE_CORRELATION || kind == RC_SCORE_CORRELATION_SIGNED) {
        /* 1 - rho (signed) or 1 - |rho|; zero variance on either side is
           rho = 0 (eval.cu:1000-1008). */
        const double rho = (s <= 0.0 || syy <= 0.0) ? 0.0 : s2 / (sqrt(s) * sqrt(syy));
        out64[p] = kind == RC_SCORE_CORRELATION_SIGNED ? 1.0 - rho : 1.0 - fabs(rho);
    } else if (kind == RC_SCORE_AIC || kind == RC_SCORE_AICC || kind == RC_SCORE_BIC) {
        out64[p] = s;
    } else if (kind == RC_SCORE_SIGNED) {
        out64[p] = fabs(s) / (double)n;
    } else if (kind == RC_SCORE_MAXIMUM) {
        out64[p] = s;
    } else {
        out64[p] = s / (double)n;
    }
}

/* FNV-1a over a fixed-width slot up to and including HALT (eval.cu:2665-2681). */
__global__ void rc_hash_kernel(const uint32_t *__restrict__ prog, int width, int nprog,
                               unsigned long long *__restrict__ out)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nprog) return;
    const uint32_t *slot = prog + (size_t)p * (size_t)width;
    unsigned long long h = 0xcbf29ce484222325ULL;
    for (int q = 0; q < width; ++q) {
        const uint32_t w = slot[q];
        h ^= (unsigned long long)w;
        h *= 0x00000100000001b3ULL;
        if (w == OP_HALT) break;
    }
    out[p] = h;
}

/* ------------------------------------------------------------------------ */
/* Host side.                                                                */

namespace {

/* Row tiles are a multiple of 256 rows so every class (R = 4, 2, 1) walks
   whole passes; tiles depend only on n, so a program's sums never depend on
   the batch it came in. */
int tile_rows_for(int n, int tuned)
{
    const int whole = ((n + 255) / 256) * 256 > 0 ? ((n + 255) / 256) * 256 : 256;
    const int t = tuned > 0 ? tuned : 4096;
    return whole < t ? whole : t;
}

template <int S, int R>
void launch_class(int mode, dim3 grid, hipStream_t s, const RcCtx *c, const uint32_t *code,
                  const uint32_t *meta, const uint32_t *plist, int ncls, int tiles, int tile_rows,
                  int kind, float *pred, double *part, int *part_bad)
{
    const RcRows *r = &c->sets[0];
    const RcScoreArgs sa = {r->w64, c->d_tsign, c->d_pbar, r->ybar, c->d_part2};
    if (mode == MODE_EVAL)
        rc_eval_kernel<S, R, MODE_EVAL><<<grid, 64, 0, s>>>(
            r->cols, c->d_derived, r->v, r->y, r->n, code, meta, plist, ncls, tiles, tile_rows, kind,
            part, part_bad, pred, sa);
    else if (mode == MODE_SCORE)
        rc_eval_kernel<S, R, MODE_SCORE><<<grid, 64, 0, s>>>(
            r->cols, c->d_derived, r->v, r->y, r->n, code, meta, plist, ncls, tiles, tile_rows, kind,
            part, part_bad, pred, sa);
    else
        rc_eval_kernel<S, R, MODE_PREDICT><<<grid, 64, 0, s>>>(
            r->cols, c->d_derived, r->v, r->y, r->n, code, meta, plist, ncls, tiles, tile_rows, kind,
            part, part_bad, pred, sa);
}

/* The class launches of one row pass over a translated batch whose class
   starts are cls[0..4], with explicit partial-sum buffers (nprog * tiles
   items each). */
int launch_classes(const RcCtx *c, int mode, int kind, int nprog, const uint32_t *code,
                   const uint32_t *meta, const int *cls, int tiles, int tile_rows, hipStream_t s,
                   float *pred, double *part, int *part_bad)
{
    for (int q = 0; q < 4; ++q) {
        const int ncls = cls[q + 1] - cls[q];
        if (ncls <= 0) continue;
        const uint32_t *plist = meta + nprog + cls[q];
        size_t grid;
        if (tiles < 8) grid = (size_t)ncls * (size_t)tiles;
        else grid = ((size_t)(tiles + 7) / 8) * (size_t)ncls * 8;
        if (grid > 0x7fffffffu) return -1;
        dim3 g((unsigned)grid);
        if (q == 0)
            launch_class<8, 4>(mode, g, s, c, code, meta, plist, ncls, tiles, tile_rows, kind, pred,
                               part, part_bad);
        else if (q == 1)
            launch_class<16, 2>(mode, g, s, c, code, meta, plist, ncls, tiles, tile_rows, kind, pred,
                                part, part_bad);
        else if (q == 2)
            launch_class<32, 1>(mode, g, s, c, code, meta, plist, ncls, tiles, tile_rows, kind, pred,
                                part, part_bad);
        else
            launch_class<64, 1>(mode, g, s, c, code, meta, plist, ncls, tiles, tile_rows, kind, pred,
                                part, part_bad);
        if (hipGetLastError() != hipSuccess) return -1;
    }
    return 0;
}

int launch_rows(RcCtx *c, int mode, int kind, int nprog, const uint32_t *code,
                const uint32_t *meta, hipStream_t s, float *pred)
{
    RcRows *r = &c->sets[0];
    const int n = r->n;
    if (n <= 0) return -1;
    const int tile_rows = tile_rows_for(n, c->tune_eval_tile);
    const int tiles = (n + tile_rows - 1) / tile_rows;
    const size_t items = (size_t)nprog * (size_t)tiles;
    if (rc_grow(c, (void **)&c->d_part, &c->part_cap, items, sizeof(double)) != 0) return -1;
    if (rc_grow(c, (void **)&c->d_part_bad, &c->part_bad_cap, items, sizeof(int)) != 0) return -1;
    if (mode == MODE_SCORE &&
        (rc_grow(c, (void **)&c->d_part2, &c->part2_cap, items, sizeof(double)) != 0 ||
         rc_grow(c, (void **)&c->d_pbar, &c->pbar_cap, (size_t)nprog, sizeof(double)) != 0))
        return -1;
    if (launch_classes(c, mode, kind, nprog, code, meta, c->cls_start, tiles, tile_rows, s, pred,
                       c->d_part, c->d_part_bad) != 0)
        return -1;
    if (mode == MODE_PREDICT) return 0;
    rc_eval_finish<<<(nprog + 127) / 128, 128, 0, s>>>(c->d_part, c->d_part_bad, c->d_part2, nprog,
                                                       tiles, n, mode, kind, r->syy, c->d_out,
                                                       c->d_score, c->d_status, c->d_pbar);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

/* Upload a translated batch into the next arena slot. */
int upload_translated(RcCtx *c, const std::vector<uint32_t> &code, const std::vector<uint32_t> &meta,
                      int nprog)
{
    RcStats *st = &c->stats;
    double t = rc_now_ms();
    const int slot = c->arena_slot ^ 1;
    /* Capacity is per slot; a growth reallocates both slots. */
    if (code.size() > c->arena_code_words || meta.size() > c->arena_meta_words ||
        !c->arena_code[slot]) {
        (void)hipStreamSynchronize(c->s_eval);
        size_t wc = code.size() + code.size() / 4 + 256;
        size_t wm = meta.size() + meta.size() / 4 + 64;
        for (int i = 0; i < 2; ++i) {
            if (c->arena_code[i]) rc_dev_free(c, c->arena_code[i]);
            if (c->arena_meta[i]) rc_dev_free(c, c->arena_meta[i]);
            c->arena_code[i] = nullptr;
            c->arena_meta[i] = nullptr;
            if (rc_dev_alloc(c, (void **)&c->arena_code[i], wc * 4) != 0) return -1;
            if (rc_dev_alloc(c, (void **)&c->arena_meta[i], wm * 4) != 0) return -1;
        }
        c->arena_code_words = wc;
        c->arena_meta_words = wm;
    }
    const size_t nb = (code.size() + meta.size()) * 4;
    if (c->async_copy && rc_grow_pinned(c, (void **)&c->h_stage, &c->h_stage_bytes, nb) != 0)
        c->async_copy = 0;
    st->alloc_ms += rc_now_ms() - t;
    t = rc_now_ms();
    if (c->async_copy) {
        /* The staging buffer may still feed the previous copy. */
        (void)hipStreamSynchronize(c->s_eval);
        memcpy(c->h_stage, code.data(), code.size() * 4);
        memcpy(c->h_stage + code.size(), meta.data(), meta.size() * 4);
        if (hipMemcpyAsync(c->arena_code[slot], c->h_stage, code.size() * 4, hipMemcpyHostToDevice,
                           c->s_eval) != hipSuccess)
            return -1;
        if (hipMemcpyAsync(c->arena_meta[slot], c->h_stage + code.size(), meta.size() * 4,
                           hipMemcpyHostToDevice, c->s_eval) != hipSuccess)
            return -1;
    } else {
        if (hipMemcpy(c->arena_code[slot], code.data(), code.size() * 4, hipMemcpyHostToDevice) !=
                hipSuccess ||
            hipMemcpy(c->arena_meta[slot], meta.data(), meta.size() * 4, hipMemcpyHostToDevice) !=
                hipSuccess)
            return -1;
    }
    st->h2d_ms += rc_now_ms() - t;
    st->h2d_bytes += (unsigned long long)nb;
    c->arena_slot = slot;
    c->nprog = nprog;
    st->batch = nprog;
    return 0;
}

/* Derived columns: at most this many, and at most RC_DERIVED_BYTES of them. */
constexpr int RC_DERIVED_MAX = 256;
constexpr size_t RC_DERIVED_BYTES = (size_t)2 << 30;

int derived_limit(const RcCtx *c)
{
    const size_t n = (size_t)c->sets[0].n;
    if (n == 0) return 0;
    return (int)std::min<size_t>(RC_DERIVED_MAX, RC_DERIVED_BYTES / (4 * n));
}

/* Compute the derived columns the last translation admitted (pairs past
   c->derived_cols), growing the column block when needed. */
int derive_columns(RcCtx *c, const std::vector<uint32_t> &derived)
{
    RcRows *r = &c->sets[0];
    const int want = (int)(derived.size() / 2);
    const size_t colb = sizeof(float) * (size_t)r->n;
    if (want > c->derived_cap) {
        const int cap = std::min(derived_limit(c), ((want + 31) / 32) * 32);
        if (cap < want) return -1;
        float *nb = nullptr;
        if (rc_dev_alloc(c, (void **)&nb, colb * (size_t)cap) != 0) return -1;
        if (c->derived_cols > 0 &&
            hipMemcpyAsync(nb, c->d_derived, colb * (size_t)c->derived_cols, hipMemcpyDeviceToDevice,
                           c->s_eval) != hipSuccess) {
            rc_dev_free(c, nb);
            return -1;
        }
        /* A batch still in flight may read the old block. */
        if (hipStreamSynchronize(c->s_eval) != hipSuccess) return -1;
        if (c->d_derived) rc_dev_free(c, c->d_derived);
        c->d_derived = nb;
        c->derived_cap = cap;
    }
    const unsigned blocks = (unsigned)((r->n + 255) / 256);
    for (int k = c->derived_cols; k < want; ++k) {
        rc_derive_kernel<<<blocks, 256, 0, c->s_eval>>>(r->cols, c->d_derived, r->n, r->v,
                                                        derived[2 * k], (int)derived[2 * k + 1], k);
        if (hipGetLastError() != hipSuccess) return -1;
    }
    c->derived = derived;
    c->derived_cols = want;
    return 0;
}

/* Translate a batch, upload its streams, compute the derived columns it
   admits and fill in its hoisted constants.  The device work after the
   upload is timed with the batch's first evaluation (event tp). */
int translate_and_upload(RcCtx *c, const uint32_t *prog, const uint32_t *off, int nprog)
{
    std::vector<uint32_t> code, meta;
    std::vector<uint32_t> derived = c->derived;
    const int nv = c->sets[0].v;
    c->prep_timed = 0;
    c->max_stack = rc_translate_programs(prog, off, nprog, nv, code, meta, c->cls_start, &derived,
                                         derived_limit(c));
    if (upload_translated(c, code, meta, nprog) != 0) return -1;
    (void)hipEventRecord(c->tp, c->s_eval);
    if (derived.size() > c->derived.size() && derive_columns(c, derived) != 0) {
        /* No room for new columns: translate again with the computed ones. */
        (void)hipGetLastError();
        derived = c->derived;
        c->max_stack = rc_translate_programs(prog, off, nprog, nv, code, meta, c->cls_start,
                                             &derived, c->derived_cols);
        if (upload_translated(c, code, meta, nprog) != 0) return -1;
    }
    /* Fill in the hoisted constants before any launch reads the streams. */
    const size_t nhoist = (meta.size() - 2 * (size_t)nprog) / 2;
    if (nhoist > 0x7fffffffu) return -1;
    if (nhoist > 0) {
        const int slot = c->arena_slot;
        rc_hoist_kernel<<<(unsigned)nhoist, 64, 0, c->s_eval>>>(
            c->arena_code[slot], c->arena_meta[slot] + 2 * (size_t)nprog, (int)nhoist);
        if (hipGetLastError() != hipSuccess) return -1;
    }
    c->prep_timed = 1;
    return 0;
}

int ensure_outputs(RcCtx *c, int nprog)
{
    if (rc_grow(c, (void **)&c->d_out, &c->out_cap, (size_t)nprog, sizeof(float)) != 0) return -1;
    size_t cap = c->score_cap;
    if (rc_grow(c, (void **)&c->d_score, &cap, (size_t)nprog, sizeof(double)) != 0) return -1;
    size_t cap2 = c->score_cap;
    if (rc_grow(c, (void **)&c->d_status, &cap2, (size_t)nprog, sizeof(int)) != 0) return -1;
    c->score_cap = cap < cap2 ? cap : cap2;
    return 0;
}

float score_current(RcCtx *c, int nprog, int kind, const int *pcount, double *out, int *status)
{
    RcStats *st = &c->stats;
    const bool classif = kind >= RC_SCORE_LOG_LOSS && kind <= RC_SCORE_HINGE_SQUARED;
    /* A classification score needs the label signs (eval.cu:3664-3670). */
    if (!(kind >= 0 && kind <= 9) && !(classif && c->has_tsign)) {
        st->failed = 1;
        return -1.0f;
    }
    if (ensure_outputs(c, nprog) != 0) return -1.0f;
    double t = rc_now_ms();
    (void)hipEventRecord(c->t0, c->s_eval);
    const bool corr = kind == RC_SCORE_CORRELATION || kind == RC_SCORE_CORRELATION_SIGNED;
    if ((corr && launch_rows(c, MODE_SCORE, RC_SCORE_CORR_MEAN, nprog, c->arena_code[c->arena_slot],
                             c->arena_meta[c->arena_slot], c->s_eval, nullptr) != 0) ||
        launch_rows(c, MODE_SCORE, kind, nprog, c->arena_code[c->arena_slot],
                    c->arena_meta[c->arena_slot], c->s_eval, nullptr) != 0) {
        st->failed = 1;
        return -1.0f;
    }
    (void)hipEventRecord(c->t1, c->s_eval);
    st->launch_ms += rc_now_ms() - t;
    t = rc_now_ms();
    if (hipEventSynchronize(c->t1) != hipSuccess || hipGetLastError() != hipSuccess) {
        st->failed = 1;
        return -1.0f;
    }
    st->sync_ms += rc_now_ms() - t;
    t = rc_now_ms();
    if (hipMemcpy(out, c->d_score, sizeof(double) * (size_t)nprog, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(status, c->d_status, sizeof(int) * (size_t)nprog, hipMemcpyDeviceToHost) != hipSuccess) {
        st->failed = 1;
        return -1.0f;
    }
    st->d2h_ms += rc_now_ms() - t;
    st->d2h_bytes += (unsigned long long)nprog * 12ull;
    const int n = c->sets[0].n;
    if (kind == RC_SCORE_AIC || kind == RC_SCORE_AICC || kind == RC_SCORE_BIC) {
        /* The AIC family's parameter penalty, applied on the host so k may
           differ per candidate (eval.cu:3632-3645, metrics.rs). */
        for (int i = 0; i < nprog; ++i) {
            if (status[i] != 0) continue;
            const double dn = (double)n;
            const double dk = pcount ? (double)pcount[i] : 0.0;
            const double rss = out[i];
            const double base = (rss <= 0.0) ? -__builtin_inf() : dn * log(rss / dn);
            if (kind == RC_SCORE_AIC) {
                out[i] = base + 2.0 * dk;
            } else if (kind == RC_SCORE_BIC) {
                out[i] = base + dk * log(dn);
            } else if (!std::isfinite(base)) {
                out[i] = base;
            } else {
                const double denom = dn - dk - 1.0;
                out[i] = denom <= 0.0 ? __builtin_inf()
                                      : base + 2.0 * dk + 2.0 * dk * (dk + 1.0) / denom;
            }
        }
    }
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, c->prep_timed ? c->tp : c->t0, c->t1);
    c->prep_timed = 0;
    st->kernel_ms = ms;
    st->batch = nprog;
    st->wall_ms = rc_now_ms() - c->batch_start_ms;
    c->warm = 1;
    c->batches++;
    return ms;
}

/* ---- rc_eval_static: the evaluation entry point that allocates nothing ----

   Its device working memory (the translated streams, the per-(program,
   tile) partial sums and the outputs) lives in one file-scope device array.
   Each GPU holds its own copy at its own address, looked up once per device
   with hipGetSymbolAddress while that device is current (rc_create's set-up,
   or the first call on the device); host code uses only those addresses,
   never the array's name.  Every call on a GPU shares that GPU's array: the
   last call's contents stay, and calls that overlap on one GPU race.  No
   part of the array needs clearing: every partial sum and output is written
   before it is read. */
constexpr size_t RC_STATIC_BYTES = (size_t)32 << 20;
constexpr int RC_STATIC_MAX_DEVICES = 64;
static __device__ __align__(256) unsigned char scratch[RC_STATIC_BYTES];
void *g_static_ptr[RC_STATIC_MAX_DEVICES];
std::mutex g_static_mu;

size_t static_take(size_t &cur, size_t bytes)
{
    const size_t o = (cur + 255) & ~(size_t)255;
    cur = o + bytes;
    return o;
}

} /* namespace */

/* Look up this device's copy of the static array (the current device). */
int rc_static_scratch_init()
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= RC_STATIC_MAX_DEVICES) {
        (void)hipGetLastError();
        return -1;
    }
    std::lock_guard<std::mutex> lock(g_static_mu);
    if (g_static_ptr[dev]) return 0;
    void *ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) {
        (void)hipGetLastError();
        return -1;
    }
    g_static_ptr[dev] = ptr;
    return 0;
}

extern "C" {

float rc_eval_static(void *ctx, const unsigned int *prog, const unsigned int *off, int nprog,
                     float *out, int loss)
{
    RcCtx *c = (RcCtx *)ctx;
    if (!c || !prog || !off || !out || nprog <= 0 || c->pending) return -1.0f;
    RcRows *r = &c->sets[0];
    const int n = r->n;
    if (n <= 0) return -1.0f;
    if (rc_static_scratch_init() != 0) return -1.0f;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return -1.0f;
    unsigned char *base = (unsigned char *)g_static_ptr[dev];
    /* The batch's streams, with the derived columns the context already
       holds and no new ones (a new column would need device memory). */
    std::vector<uint32_t> code, meta;
    std::vector<uint32_t> derived = c->derived;
    int cls[5];
    (void)rc_translate_programs(prog, off, nprog, r->v, code, meta, cls, &derived, c->derived_cols);
    const int tile_rows = tile_rows_for(n, c->tune_eval_tile);
    const int tiles = (n + tile_rows - 1) / tile_rows;
    const size_t items = (size_t)nprog * (size_t)tiles;
    size_t cur = 0;
    const size_t o_code = static_take(cur, code.size() * 4);
    const size_t o_meta = static_take(cur, meta.size() * 4);
    const size_t o_part = static_take(cur, items * sizeof(double));
    const size_t o_bad = static_take(cur, items * sizeof(int));
    const size_t o_out = static_take(cur, (size_t)nprog * sizeof(float));
    if (cur > RC_STATIC_BYTES) return -2.0f; /* the batch does not fit the static array */
    uint32_t *d_code = (uint32_t *)(base + o_code);
    uint32_t *d_meta = (uint32_t *)(base + o_meta);
    double *d_part = (double *)(base + o_part);
    int *d_bad = (int *)(base + o_bad);
    float *d_out = (float *)(base + o_out);
    hipStream_t s = c->s_eval;
    (void)hipEventRecord(c->t0, s);
    if (hipMemcpyAsync(d_code, code.data(), code.size() * 4, hipMemcpyHostToDevice, s) !=
            hipSuccess ||
        hipMemcpyAsync(d_meta, meta.data(), meta.size() * 4, hipMemcpyHostToDevice, s) != hipSuccess)
        return -1.0f;
    const size_t nhoist = (meta.size() - 2 * (size_t)nprog) / 2;
    if (nhoist > 0x7fffffffu) return -1.0f;
    if (nhoist > 0) {
        rc_hoist_kernel<<<(unsigned)nhoist, 64, 0, s>>>(d_code, d_meta + 2 * (size_t)nprog,
                                                        (int)nhoist);
        if (hipGetLastError() != hipSuccess) return -1.0f;
    }
    if (launch_classes(c, MODE_EVAL, loss, nprog, d_code, d_meta, cls, tiles, tile_rows, s, nullptr,
                       d_part, d_bad) != 0)
        return -1.0f;
    rc_eval_finish<<<(nprog + 127) / 128, 128, 0, s>>>(d_part, d_bad, nullptr, nprog, tiles, n,
                                                       MODE_EVAL, loss, r->syy, d_out, nullptr,
                                                       nullptr, nullptr);
    (void)hipEventRecord(c->t1, s);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(out, d_out, (size_t)nprog * sizeof(float), hipMemcpyDeviceToHost, s) !=
            hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
        (void)hipGetLastError();
        return -1.0f;
    }
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, c->t0, c->t1);
    return ms;
}


int rch_eval_translate(const unsigned int *prog, const unsigned int *off, int nprog, int nvars,
                       unsigned int *code, int cap_code, unsigned int *meta, int cap_meta,
                       unsigned int *derived, int cap_derived, int *nderived, int *info)
{
    if (!prog || !off || nprog <= 0 || !info || (derived && (!nderived || *nderived < 0 ||
                                                             *nderived > cap_derived)))
        return -1;
    std::vector<uint32_t> cv, mv, dv;
    if (derived) dv.assign(derived, derived + 2 * (size_t)*nderived);
    int cls[5];
    const int peak = rc_translate_programs(prog, off, nprog, nvars, cv, mv, cls,
                                           derived ? &dv : nullptr, derived ? cap_derived : 0);
    info[0] = (int)cv.size();
    info[1] = (int)mv.size();
    for (int k = 0; k < 4; ++k) info[2 + k] = cls[k];
    info[6] = peak;
    if (!code || !meta || cv.size() > (size_t)cap_code || mv.size() > (size_t)cap_meta) return -2;
    memcpy(code, cv.data(), cv.size() * 4);
    memcpy(meta, mv.data(), mv.size() * 4);
    if (derived) {
        memcpy(derived, dv.data(), dv.size() * 4);
        *nderived = (int)(dv.size() / 2);
    }
    return 0;
}

int rc_set_programs(void *ctx, const unsigned int *prog, const unsigned int *off, int nprog)
{
    RcCtx *c = (RcCtx *)ctx;
    if (!c || !prog || !off || nprog <= 0) return -1;
    if (c->pending) {
        /* The reference refuses arena growth while a batch is in flight; the
           staging here is synchronised instead, and the overlap is counted. */
        c->overlapped_uploads++;
    }
    if (translate_and_upload(c, prog, off, nprog) != 0) {
        c->stats.failed = 1;
        return -1;
    }
    c->stats.wall_ms = rc_now_ms() - c->batch_start_ms;
    return 0;
}

int rc_eval_submit(void *ctx, int loss)
{
    RcCtx *c = (RcCtx *)ctx;
    if (!c || c->nprog <= 0 || c->pending) return -1;
    if (ensure_outputs(c, c->nprog) != 0) return -1;
    double t = rc_now_ms();
    (void)hipEventRecord(c->tq0, c->s_eval);
    if (launch_rows(c, MODE_EVAL, loss, c->nprog, c->arena_code[c->arena_slot],
                    c->arena_meta[c->arena_slot], c->s_eval, nullptr) != 0) {
        c->stats.failed = 1;
        return -1;
    }
    (void)hipEventRecord(c->tq1, c->s_eval);
    c->stats.launch_ms += rc_now_ms() - t;
    /* The batch takes its upload's prep event with it: a later
       rc_set_programs records into the other handle. */
    c->pending_prep = c->prep_timed;
    if (c->prep_timed) std::swap(c->tp, c->tq_prep);
    c->prep_timed = 0;
    c->pending = 1;
    c->pending_nprog = c->nprog;
    return 0;
}

float rc_eval_collect(void *ctx, float *out)
{
    RcCtx *c = (RcCtx *)ctx;
    if (!c || !out || !c->pending) return -1.0f;
    RcStats *st = &c->stats;
    double t = rc_now_ms();
    const hipError_t e = hipEventSynchronize(c->tq1);
    st->sync_ms += rc_now_ms() - t;
    c->pending = 0;
    if (e != hipSuccess || hipGetLastError() != hipSuccess) {
        st->failed = 1;
        return -1.0f;
    }
    t = rc_now_ms();
    if (hipMemcpy(out, c->d_out, sizeof(float) * (size_t)c->pending_nprog, hipMemcpyDeviceToHost) !=
        hipSuccess) {
        st->failed = 1;
        return -1.0f;
    }
    st->d2h_ms += rc_now_ms() - t;
    st->d2h_bytes += (unsigned long long)c->pending_nprog * 4ull;
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, c->pending_prep ? c->tq_prep : c->tq0, c->tq1) != hipSuccess) {
        st->failed = 1;
        return -1.0f;
    }
    st->kernel_ms = ms;
    st->batch = c->pending_nprog;
    st->wall_ms = rc_now_ms() - c->batch_start_ms;
    c->warm = 1;
    c->batches++;
    return ms;
}

float rc_eval(void *ctx, float *out, int loss)
{
    if (rc_eval_submit(ctx, loss) != 0) return -1.0f;
    return rc_eval_collect(ctx, out);
}

int rc_predict(void *ctx, const unsigned int *prog, const unsigned int *off, int nprog, float *out)
{
    RcCtx *c = (RcCtx *)ctx;
    if (!c || nprog <= 0 || !out) return -1;
    if (rc_set_programs(ctx, prog, off, nprog) != 0) return -1;
    c->prep_timed = 0;
    const size_t total = (size_t)nprog * (size_t)c->sets[0].n;
    float *d_pred = nullptr;
    if (hipMalloc((void **)&d_pred, sizeof(float) * (total ? total : 1)) != hipSuccess) return -1;
    int rc = launch_rows(c, MODE_PREDICT, 0, nprog, c->arena_code[c->arena_slot],
                         c->arena_meta[c->arena_slot], c->s_eval, d_pred);
    if (rc == 0 && hipStreamSynchronize(c->s_eval) != hipSuccess) rc = -1;
    if (rc == 0 && total &&
        hipMemcpy(out, d_pred, sizeof(float) * total, hipMemcpyDeviceToHost) != hipSuccess)
        rc = -1;
    (void)hipFree(d_pred);
    return rc;
}

float rc_score_batch(void *ctx, const unsigned int *prog, const unsigned int *off, int nprog, int kind,
                     const int *pcount, double *out, int *status)
{
    RcCtx *c = (RcCtx *)ctx;
    if (!c || nprog <= 0 || !out || !status) return -1.0f;
    if (rc_set_programs(ctx, prog, off, nprog) != 0) return -1.0f;
    return score_current(c, nprog, kind, pcount, out, status);
}

int rc_pool_reserve(void *ctx, int nprog, int width)
{
    RcCtx *c = (RcCtx *)ctx;
    if (!c || nprog <= 0 || width <= 0) return -1;
    if (c->pool_cap >= nprog && c->pool_width == width) return 0;
    for (int i = 0; i < 2; ++i) {
        if (c->pool_raw[i]) rc_dev_free(c, c->pool_raw[i]);
        c->pool_raw[i] = nullptr;
        if (rc_dev_alloc(c, (void **)&c->pool_raw[i], sizeof(uint32_t) * (size_t)nprog * width) != 0)
            return -1;
        c->pool_host[i].assign((size_t)nprog * width, OP_HALT);
    }
    if (c->d_hash) rc_dev_free(c, c->d_hash);
    c->d_hash = nullptr;
    if (rc_dev_alloc(c, (void **)&c->d_hash, sizeof(unsigned long long) * (size_t)nprog) != 0) return -1;
    c->pool_cap = nprog;
    c->pool_width = width;
    return 0;
}

int rc_pool_upload(void *ctx, int which, const unsigned int *prog, int nprog)
{
    RcCtx *c = (RcCtx *)ctx;
    if (!c || !prog || nprog <= 0 || nprog > c->pool_cap || !c->pool_raw[which & 1]) return -1;
    const size_t words = (size_t)nprog * (size_t)c->pool_width;
    double t = rc_now_ms();
    memcpy(c->pool_host[which & 1].data(), prog, words * 4);
    if (hipMemcpy(c->pool_raw[which & 1], prog, words * 4, hipMemcpyHostToDevice) != hipSuccess) return -1;
    c->stats.h2d_ms += rc_now_ms() - t;
    c->stats.h2d_bytes += words * 4;
    c->stats.batch = nprog;
    return 0;
}

static int pool_translate(RcCtx *c, int which, int nprog, int width)
{
    if (!c || nprog <= 0 || nprog > c->pool_cap || width != c->pool_width) return -1;
    std::vector<uint32_t> off((size_t)nprog + 1);
    for (int p = 0; p <= nprog; ++p) off[p] = (uint32_t)((size_t)p * width);
    return translate_and_upload(c, c->pool_host[which & 1].data(), off.data(), nprog);
}

float rc_pool_eval(void *ctx, int which, int nprog, int width, float *out, int loss)
{
    RcCtx *c = (RcCtx *)ctx;
    if (!c || !out || pool_translate(c, which, nprog, width) != 0) return -1.0f;
    return rc_eval(ctx, out, loss);
}

float rc_pool_score(void *ctx, int which, int nprog, int width, int kind, const int *pcount,
                    double *out, int *status)
{
    RcCtx *c = (RcCtx *)ctx;
    if (!c || !out || !status || pool_translate(c, which, nprog, width) != 0) return -1.0f;
    return score_current(c, nprog, kind, pcount, out, status);
}

int rc_pool_hash(void *ctx, int which, int nprog, int width, unsigned long long *out)
{
    RcCtx *c = (RcCtx *)ctx;
    if (!c || !out || !c->pool_raw[which & 1] || !c->d_hash || nprog <= 0 || nprog > c->pool_cap ||
        width != c->pool_width)
        return -1;
    rc_hash_kernel<<<(nprog + 127) / 128, 128, 0, c->s_eval>>>(c->pool_raw[which & 1], width, nprog,
                                                               c->d_hash);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(c->s_eval) != hipSuccess) return -1;
    if (hipMemcpy(out, c->d_hash, sizeof(unsigned long long) * (size_t)nprog, hipMemcpyDeviceToHost) !=
        hipSuccess)
        return -1;
    return 0;
}

} /* extern "C" */
