/*
 * Host-side program builders of the device arm (no GPU calls).
 *
 * C++ restatement of the reference host layer that turns an expression tree
 * into what the C ABI consumes:
 *   - rch_compile_program: gpu.rs:198-218 compile_program = lower (gpu.rs:320-345,
 *     asinh/acosh/atanh rewritten over native opcodes), the peak-stack gate
 *     (peak_stack gpu.rs:952-970, HOST_MAX_STACK = 56 at gpu.rs:134), emit
 *     (gpu.rs:349-438) and the HALT terminator;
 *   - rch_fit_plan: fit.rs:985-1056 gpu_fit_plan_reason = flatten
 *     (fit.rs:1101-1196, post-order with children first), the eligibility
 *     gates (no constants / p > 20 / m > 96 / unsupported operator) and the
 *     4-word records [op, kid0, kid1, slot] with c0 in slot order.
 * Trees arrive as pre-order node lists (op code, value); see rcsr_hip.h for
 * the operator codes (device opcodes plus 43/44/45 = asinh/acosh/atanh).
 */
#include <stdint.h>
#include <string.h>

#include <cmath>
#include <vector>

namespace {

enum {
    T_CONST = 0, T_VAR = 1, T_ADD = 2, T_SUB = 3, T_MUL = 4, T_DIV = 5, T_SIN = 6, T_COS = 7,
    T_TAN = 8, T_EXP = 9, T_LOG = 10, T_SQRT = 11, T_ABS = 12, T_TANH = 13, T_SIG = 14,
    T_SQUARE = 15, T_CUBE = 16, T_NEG = 17, T_GAUSS = 18, T_STEP = 19, T_POW = 20, T_SINH = 21,
    T_COSH = 22, T_FLOOR = 23, T_CEIL = 24, T_ROUND = 25, T_MOD = 26, T_ATAN2 = 27, T_ERF = 28,
    T_ERFC = 29, T_MIN = 30, T_MAX = 31, T_IF = 32, T_SIGN = 33, T_EQUAL = 34, T_LESS = 35,
    T_LESS_OR_EQUAL = 36, T_GREATER = 37, T_GREATER_OR_EQUAL = 38, T_AND = 39, T_OR = 40,
    T_XOR = 41, T_NOT = 42, T_ASINH = 43, T_ACOSH = 44, T_ATANH = 45, T_HALT = 255
};

const int HOST_MAX_STACK = 56;
const int GPU_FIT_MAX_CONSTANTS = 20;
const int GPU_FIT_MAX_NODES = 96;

struct Node {
    int op;
    double val;
    std::vector<Node> kids;
};

int arity(int op)
{
    switch (op) {
        case T_CONST: case T_VAR: return 0;
        case T_ADD: case T_SUB: case T_MUL: case T_DIV: case T_POW: case T_MOD: case T_ATAN2:
        case T_MIN: case T_MAX: case T_EQUAL: case T_LESS: case T_LESS_OR_EQUAL: case T_GREATER:
        case T_GREATER_OR_EQUAL: case T_AND: case T_OR: case T_XOR:
            return 2;
        case T_IF: return 3;
        default: return (op >= 0 && op <= T_ATANH) ? 1 : -1;
    }
}

bool parse(const int *ops, const double *vals, int nn, int &pos, Node &out)
{
    if (pos >= nn) return false;
    out.op = ops[pos];
    out.val = vals[pos];
    ++pos;
    int a = arity(out.op);
    if (a < 0) return false;
    out.kids.resize(a);
    for (int i = 0; i < a; ++i)
        if (!parse(ops, vals, nn, pos, out.kids[i])) return false;
    return true;
}

bool build(const int *ops, const double *vals, int nn, Node &root)
{
    int pos = 0;
    return parse(ops, vals, nn, pos, root) && pos == nn;
}

Node mk(int op, std::vector<Node> kids) { return Node{op, 0.0, std::move(kids)}; }
Node cst(double v) { return Node{T_CONST, v, {}}; }

/* gpu.rs:320-345: exact rewrites of the operators the kernel lacks. */
Node lower(const Node &n)
{
    if (n.op == T_CONST || n.op == T_VAR) return n;
    std::vector<Node> k;
    for (const Node &c : n.kids) k.push_back(lower(c));
    switch (n.op) {
        case T_ASINH: {
            const Node &x = k[0];
            Node sgn = mk(T_SUB, {mk(T_STEP, {x}), mk(T_STEP, {mk(T_NEG, {x})})});
            Node mag = mk(T_LOG, {mk(T_ADD, {mk(T_ABS, {x}),
                                             mk(T_SQRT, {mk(T_ADD, {mk(T_MUL, {x, x}), cst(1.0)})})})});
            return mk(T_MUL, {sgn, mag});
        }
        case T_ACOSH: {
            const Node &x = k[0];
            return mk(T_LOG, {mk(T_ADD, {x, mk(T_SQRT, {mk(T_MUL, {mk(T_SUB, {x, cst(1.0)}),
                                                                   mk(T_ADD, {x, cst(1.0)})})})})});
        }
        case T_ATANH: {
            const Node &x = k[0];
            return mk(T_MUL, {cst(0.5), mk(T_LOG, {mk(T_DIV, {mk(T_ADD, {cst(1.0), x}),
                                                              mk(T_SUB, {cst(1.0), x})})})});
        }
        default: return Node{n.op, n.val, k};
    }
}

/* gpu.rs:952-970: leaves push, unary is net zero, an n-ary operator pops n-1. */
void peak_walk(const Node &n, int &cur, int &peak)
{
    for (const Node &k : n.kids) peak_walk(k, cur, peak);
    if (n.kids.empty()) {
        ++cur;
        if (cur > peak) peak = cur;
    } else if (n.kids.size() > 1) {
        cur -= (int)n.kids.size() - 1;
    }
}

uint32_t f32_bits(double v)
{
    float f = (float)v;
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
}

/* gpu.rs:349-438: postfix emission; false for an operator with no opcode. */
bool emit(const Node &n, std::vector<uint32_t> &out)
{
    switch (n.op) {
        case T_CONST:
            out.push_back(T_CONST);
            out.push_back(f32_bits(n.val));
            return true;
        case T_VAR:
            out.push_back(T_VAR);
            out.push_back((uint32_t)n.val);
            return true;
        case T_ASINH: case T_ACOSH: case T_ATANH: return false;
        default: break;
    }
    if (n.op < 0 || n.op > T_NOT) return false;
    for (const Node &k : n.kids)
        if (!emit(k, out)) return false;
    out.push_back((uint32_t)n.op);
    return true;
}

/* fit.rs:890-920: operators with an analytic value/adjoint pair. */
bool fit_opcode_ok(int op)
{
    switch (op) {
        case T_CONST: case T_VAR: case T_ADD: case T_SUB: case T_MUL: case T_DIV: case T_SIN:
        case T_COS: case T_TAN: case T_EXP: case T_LOG: case T_SQRT: case T_ABS: case T_TANH:
        case T_SIG: case T_SQUARE: case T_CUBE: case T_NEG: case T_GAUSS: case T_POW:
        case T_SINH: case T_COSH: case T_ERF: case T_ERFC:
            return true;
        default: return false;
    }
}

struct FlatRec {
    int op;
    int k0, k1;
    uint32_t slot;
    double val;
};

int flatten(const Node &n, std::vector<FlatRec> &f, int &nconst, bool &ok)
{
    int kid[3] = {0, 0, 0};
    for (size_t i = 0; i < n.kids.size(); ++i) kid[i] = flatten(n.kids[i], f, nconst, ok);
    if (!fit_opcode_ok(n.op)) ok = false;
    FlatRec r{n.op, kid[0], kid[1], 0xffffffffu, n.val};
    if (n.op == T_CONST) r.slot = (uint32_t)nconst++;
    else if (n.op == T_VAR) r.slot = (uint32_t)n.val;
    f.push_back(r);
    return (int)f.size() - 1;
}

} /* namespace */

extern "C" {

/* Returns the number of words written (HALT included), or
   -1 unsupported operator, -2 peak stack above 56 (*peak holds it),
   -3 malformed tree or out buffer too small. */
int rch_compile_program(const int *ops, const double *vals, int nn, unsigned int *out, int cap,
                        int *peak)
{
    Node t;
    if (!ops || !vals || nn <= 0 || !build(ops, vals, nn, t)) return -3;
    Node low = lower(t);
    int cur = 0, pk = 0;
    peak_walk(low, cur, pk);
    if (peak) *peak = pk;
    if (pk > HOST_MAX_STACK) return -2;
    std::vector<uint32_t> words;
    if (!emit(low, words)) return -1;
    words.push_back(T_HALT);
    if ((int)words.size() > cap || !out) return -3;
    memcpy(out, words.data(), words.size() * 4);
    return (int)words.size();
}

/* Returns nconst (1..20) and writes m*4 words to out_words, the start
   constants to out_c0 and m to *nodes; or a GpuFitReject code:
   -1 no constants, -2 too many constants, -3 too many nodes,
   -5 unsupported operator, -6 malformed tree / buffer too small. */
int rch_fit_plan(const int *ops, const double *vals, int nn, unsigned int *out_words, int cap_words,
                 double *out_c0, int cap_c, int *nodes)
{
    Node t;
    if (!ops || !vals || nn <= 0 || !build(ops, vals, nn, t)) return -6;
    std::vector<FlatRec> f;
    int nconst = 0;
    bool ok = true;
    flatten(t, f, nconst, ok);
    if (nconst == 0) return -1;
    if (nconst > GPU_FIT_MAX_CONSTANTS) return -2;
    if ((int)f.size() > GPU_FIT_MAX_NODES) return -3;
    if (!ok) return -5;
    if ((int)f.size() * 4 > cap_words || nconst > cap_c || !out_words || !out_c0) return -6;
    for (size_t i = 0; i < f.size(); ++i) {
        out_words[4 * i + 0] = (uint32_t)f[i].op;
        out_words[4 * i + 1] = (uint32_t)f[i].k0;
        out_words[4 * i + 2] = (uint32_t)f[i].k1;
        out_words[4 * i + 3] = f[i].slot;
        if (f[i].op == T_CONST) out_c0[f[i].slot] = f[i].val;
    }
    if (nodes) *nodes = (int)f.size();
    return nconst;
}

} /* extern "C" */
