/*
 * Host compiler of the fit VM (encoding in rc_fitvm.h).
 *
 * Input: one candidate's post-order fit records [op, kid0, kid1, slot]
 * (fit.rs:1101-1196; the reference interprets them node by node in
 * ft_forward / ft_backward, eval.cu:1173-1317).  Output: a forward stream
 * and a reverse (adjoint) stream of word pairs in which every register slot
 * is assigned here, so the device interpreter only decodes and dispatches.
 *
 * The arithmetic per node is the reference's: the same f32 operation on the
 * same operand values; only operand order of ADD / MUL (exact in IEEE) and
 * the storage of intermediate values change.  The tape of local partials
 * keeps ft_forward's partial formulas; a partial that is a constant's value
 * is re-read from the constant instead of being stored.
 *
 * Programs that do not fit the VM's files (15 U slots, 15 B slots, 7 saved
 * adjoints, inputs 0..7) are reported as such and run on the general
 * interpreter instead.
 */
#include "rc_fitvm.h"

#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <vector>

namespace {

enum {
    O_CONST = 0, O_VAR = 1, O_ADD = 2, O_SUB = 3, O_MUL = 4, O_DIV = 5, O_SIN = 6, O_COS = 7,
    O_TAN = 8, O_EXP = 9, O_LOG = 10, O_SQRT = 11, O_ABS = 12, O_TANH = 13, O_SIG = 14,
    O_SQUARE = 15, O_CUBE = 16, O_NEG = 17, O_GAUSS = 18, O_POW = 20, O_SINH = 21, O_COSH = 22,
    O_ERF = 28, O_ERFC = 29
};

bool op_ok(uint32_t op)
{
    switch (op) {
        case O_CONST: case O_VAR: case O_ADD: case O_SUB: case O_MUL: case O_DIV: case O_SIN:
        case O_COS: case O_TAN: case O_EXP: case O_LOG: case O_SQRT: case O_ABS: case O_TANH:
        case O_SIG: case O_SQUARE: case O_CUBE: case O_NEG: case O_GAUSS: case O_POW:
        case O_SINH: case O_COSH: case O_ERF: case O_ERFC:
            return true;
        default: return false;
    }
}

bool is_bin(uint32_t op)
{
    return op == O_ADD || op == O_SUB || op == O_MUL || op == O_DIV || op == O_POW;
}

struct Nd {
    uint32_t op;
    int l, r;
    uint32_t slot;
    bool on;
};

enum Kind { K_OFF = 0, K_NODE = 1, K_CONST = 2 };

/* One forward instruction and what its reverse needs. */
struct Unit {
    uint32_t fx, fy;      /* forward word pair */
    uint32_t rkind;       /* reverse handler bit, 0 = none */
    uint32_t ropt;        /* R_GEN option bits (child kinds, unary) */
    uint32_t rop;         /* operator (R_GEN) */
    int ut, bt;           /* tape slots read by the reverse */
    int cslot;            /* constant whose value the reverse multiplies by */
    int xv;               /* input read by the reverse */
    int gl, gr;           /* gradient columns (constant slots) or -1 */
    int nodes_next;       /* children that continue the sweep: 0, 1 or 2 */
    bool ffused;          /* its forward word was folded into the previous one */
};

struct Compiler {
    std::vector<Nd> nd;
    int nvars = 0;
    bool vo = false;       /* value only: no tape */
    int usink = VM_U_SINK; /* the U sink slot */
    int depth = 0;
    std::vector<int> spill;
    bool ufree[VM_U_SLOTS];
    int ucount = 0, bnext = 0, umax = 0;
    bool overflow = false;
    std::vector<Unit> units;

    int ualloc()
    {
        for (int s = 0; s < usink; ++s)
            if (ufree[s]) {
                ufree[s] = false;
                umax = std::max(umax, s + 1);
                return s;
            }
        overflow = true;
        return usink;
    }
    /* A tape slot: none in value-only code. */
    int talloc() { return vo ? usink : ualloc(); }
    int balloc()
    {
        if (vo) return VM_B_SINK;
        if (bnext >= VM_B_SINK) {
            overflow = true;
            return VM_B_SINK;
        }
        return bnext++;
    }
    bool leaf(int i) const { return nd[i].op == O_CONST || nd[i].op == O_VAR; }
    Kind kind(int i) const
    {
        if (nd[i].op == O_CONST) return K_CONST;
        return nd[i].on ? K_NODE : K_OFF;
    }
    static uint32_t un_bits(uint32_t op)
    {
        return op == O_SIN ? VMF_SIN : op == O_COS ? VMF_COS : op == O_EXP ? VMF_EXP : 0u;
    }
    Unit blank() const
    {
        Unit u;
        memset(&u, 0, sizeof(u));
        u.ut = u.bt = -1;
        u.cslot = u.xv = u.gl = u.gr = -1;
        return u;
    }
    uint32_t yf(int cofs_slot, int us, int ur, int bs, int xv, uint32_t op) const
    {
        uint32_t y = 0;
        if (cofs_slot >= 0) y |= (uint32_t)(cofs_slot * 4) & 0xffu;
        y |= ((uint32_t)us & 0xfu) << 8;
        y |= ((uint32_t)ur & 0xfu) << 12;
        y |= ((uint32_t)bs & 0xfu) << 16;
        y |= ((uint32_t)(xv < 0 ? 0 : xv) & 0x7u) << 20;
        y |= (op & 0x1fu) << 23;
        return y;
    }

    void leaf_push(int i)
    {
        const int us = depth > 0 ? ualloc() : usink;
        if (depth > 0) spill.push_back(us);
        Unit u = blank();
        if (nd[i].op == O_CONST) {
            u.fx = VMF_PUSHC;
            u.fy = yf((int)nd[i].slot, us, 0, VM_B_SINK, -1, 0);
        } else {
            u.fx = VMF_PUSHV;
            u.fy = yf(-1, us, 0, VM_B_SINK, (int)nd[i].slot, 0);
        }
        units.push_back(u);
        ++depth;
    }

    /* The reverse handler of a node whose sweep continues into child kinds
       lk (left or only) and rk. */
    void emit(int n)
    {
        const Nd &N = nd[n];
        if (leaf(n)) {
            leaf_push(n);
            return;
        }
        if (!is_bin(N.op)) {
            const int ch = N.l;
            emit(ch);
            const bool need = nd[ch].on && !vo;
            const int us = need ? talloc() : usink;
            Unit u = blank();
            u.fx = VMF_UN | (need ? VMF_NEED : 0u) | un_bits(N.op);
            u.fy = yf(-1, us, 0, VM_B_SINK, -1, N.op);
            if (N.on) {
                if (kind(ch) == K_CONST) {
                    u.rkind = VMR_GEN;
                    u.ropt = VMR_UNARY | VMR_LCONST;
                    u.rop = N.op;
                    u.ut = us;
                    u.gl = (int)nd[ch].slot;
                    u.nodes_next = 0;
                } else {
                    u.rkind = VMR_UN;
                    u.ut = us;
                    u.nodes_next = 1;
                }
            }
            units.push_back(u);
            return;
        }
        int L = N.l, R = N.r;
        const Nd &NL = nd[L];
        const Nd &NR = nd[R];
        /* x * c */
        if (N.op == O_MUL && NR.op == O_CONST && NL.op == O_VAR) {
            const int us = depth > 0 ? ualloc() : usink;
            if (depth > 0) spill.push_back(us);
            ++depth;
            Unit u = blank();
            u.fx = VMF_VMC;
            u.fy = yf((int)NR.slot, us, 0, VM_B_SINK, (int)NL.slot, 0);
            u.rkind = VMR_VMC;
            u.xv = (int)NL.slot;
            u.gl = (int)NR.slot;
            u.nodes_next = 0;
            units.push_back(u);
            return;
        }
        /* f(g) * c */
        if (N.op == O_MUL && NR.op == O_CONST && !leaf(L) && !is_bin(NL.op) &&
            nd[NL.l].op != O_CONST) {
            const int g = NL.l;
            emit(g);
            const bool need = nd[g].on && !vo;
            const int us = talloc();
            const int bs = need ? balloc() : VM_B_SINK;
            Unit u = blank();
            u.fx = VMF_UNMC | (need ? VMF_NEED : 0u) | un_bits(NL.op);
            u.fy = yf((int)NR.slot, us, 0, bs, -1, NL.op);
            u.rkind = need ? VMR_UNMC_N : VMR_UNMC_P;
            u.ut = us;
            u.bt = bs;
            u.cslot = (int)NR.slot;
            u.gl = (int)NR.slot;
            u.nodes_next = need ? 1 : 0;
            units.push_back(u);
            return;
        }
        /* (.) * c */
        if (N.op == O_MUL && NR.op == O_CONST && !leaf(L)) {
            emit(L);
            const int us = talloc();
            Unit u = blank();
            u.fx = VMF_MULC;
            u.fy = yf((int)NR.slot, us, 0, VM_B_SINK, -1, 0);
            u.rkind = NL.on ? VMR_MULC_N : VMR_MULC_P;
            u.ut = us;
            u.cslot = (int)NR.slot;
            u.gl = (int)NR.slot;
            u.nodes_next = NL.on ? 1 : 0;
            units.push_back(u);
            return;
        }
        /* (.) + c */
        if (N.op == O_ADD && NR.op == O_CONST && NL.op != O_CONST) {
            emit(L);
            Unit u = blank();
            u.fx = VMF_ADDC;
            u.fy = yf((int)NR.slot, usink, 0, VM_B_SINK, -1, 0);
            u.rkind = NL.on ? VMR_ADDC_N : VMR_ADDC_P;
            u.gl = (int)NR.slot;
            u.nodes_next = NL.on ? 1 : 0;
            units.push_back(u);
            return;
        }
        /* (.) * x */
        if (N.op == O_MUL && NR.op == O_VAR && NL.op != O_CONST) {
            emit(L);
            Unit u = blank();
            u.fx = VMF_MULV;
            u.fy = yf(-1, usink, 0, VM_B_SINK, (int)NR.slot, 0);
            if (NL.on) {
                u.rkind = VMR_MULV;
                u.xv = (int)NR.slot;
                u.nodes_next = 1;
            }
            units.push_back(u);
            return;
        }
        /* General binary node. */
        const Kind lk = kind(L), rk = kind(R);
        const bool tape = !vo && (N.op == O_MUL || N.op == O_DIV || N.op == O_POW);
        const bool dl = tape && lk != K_OFF, dr = tape && rk != K_OFF;
        Unit u = blank();
        int ur = 0;
        uint32_t fx = VMF_BIN;
        int cs = -1, xv = -1;
        emit(L);
        if (NR.op == O_CONST) {
            fx |= VMF_RC;
            cs = (int)NR.slot;
        } else if (NR.op == O_VAR) {
            fx |= VMF_RV;
            xv = (int)NR.slot;
        } else {
            emit(R);
            if (spill.empty()) { overflow = true; return; }
            ur = spill.back();
            spill.pop_back();
            --depth;
        }
        const int us = dl ? ualloc() : usink;
        const int bs = dr ? balloc() : VM_B_SINK;
        if (!(fx & (VMF_RC | VMF_RV))) ufree[ur] = true;
        if (dl) fx |= VMF_D0;
        if (dr) fx |= VMF_D1;
        if (N.op == O_ADD && !(fx & (VMF_RC | VMF_RV))) {
            /* Both operands computed: the add folds into the word that
               finished the right operand when that word leaves ur unused,
               else the dedicated add. */
            Unit &prev = units.back();
            const uint32_t h = prev.fx;
            const bool ur_free = !(h & (VMF_ADD | VMF_ACC)) &&
                                 (!(h & VMF_BIN) || (h & (VMF_RC | VMF_RV)));
            if (ur_free) {
                prev.fx |= VMF_ACC;
                prev.fy = (prev.fy & ~(0xfu << 12)) | (((uint32_t)ur & 0xfu) << 12);
                u.ffused = true;
            }
            fx = VMF_ADD;
        }
        u.fx = fx;
        u.fy = yf(cs, us, ur, bs, xv, N.op);
        if (N.on) {
            if (N.op == O_ADD && lk == K_NODE && rk == K_NODE) {
                u.rkind = VMR_PUSH;
                u.nodes_next = 2;
            } else if (N.op == O_ADD && lk != K_CONST && rk != K_CONST) {
                u.rkind = 0; /* the adjoint passes through unchanged */
            } else {
                u.rkind = VMR_GEN;
                u.ropt = (lk == K_NODE ? VMR_LNODE : 0u) | (lk == K_CONST ? VMR_LCONST : 0u) |
                         (rk == K_NODE ? VMR_RNODE : 0u) | (rk == K_CONST ? VMR_RCONST : 0u);
                u.rop = N.op;
                u.ut = us;
                u.bt = bs;
                u.gl = lk == K_CONST ? (int)NL.slot : -1;
                u.gr = rk == K_CONST ? (int)NR.slot : -1;
                u.nodes_next = (lk == K_NODE ? 1 : 0) + (rk == K_NODE ? 1 : 0);
            }
        }
        units.push_back(u);
    }
};

} /* namespace */

extern "C" {

/* Compile one candidate for the fit VM.  rec: words/4 post-order records;
   p constants; nvars inputs; col_bytes: LDS bytes between gradient columns;
   flags: VMC_* options.
   fwd / rev receive word pairs (2 words each).  info[0..5] = forward pairs,
   reverse pairs, U slots, B slots, saved-adjoint slots, peak value depth.
   Returns 0, or -1 malformed records, -2 unsupported operator, -3 the
   program does not fit the VM, -4 an output buffer is too small. */
int rch_fitvm_compile(const unsigned int *rec, int words, int p, int nvars, int col_bytes,
                      unsigned int flags, unsigned int *fwd, int cap_f, unsigned int *rev,
                      int cap_r, int *info)
{
    if (!rec || words <= 0 || words % 4 != 0 || p < 0 || nvars < 0 || col_bytes <= 0) return -1;
    const int m = words / 4;
    Compiler C;
    C.nvars = nvars;
    C.vo = (flags & VMC_VALUE_ONLY) != 0;
    C.usink = (flags & VMC_SMALL_U) ? VM_U_SINK_SMALL : VM_U_SINK;
    C.nd.resize((size_t)m);
    C.units.reserve((size_t)m);
    C.spill.reserve((size_t)m);
    std::vector<int> st;
    st.reserve((size_t)m);
    bool unsupported = false;
    for (int i = 0; i < m; ++i) {
        Nd &N = C.nd[(size_t)i];
        N.op = rec[4 * i];
        N.l = N.r = -1;
        N.slot = rec[4 * i + 3];
        const uint32_t k0 = rec[4 * i + 1], k1 = rec[4 * i + 2];
        if (!op_ok(N.op)) unsupported = true;
        if (N.op == O_CONST) {
            if ((int)N.slot >= p) return -1;
            N.on = true;
            st.push_back(i);
        } else if (N.op == O_VAR) {
            if ((int)N.slot >= nvars) return -1;
            N.on = false;
            st.push_back(i);
        } else if (is_bin(N.op)) {
            if (st.size() < 2 || st.back() != (int)k1 || st[st.size() - 2] != (int)k0) return -1;
            st.pop_back();
            st.pop_back();
            N.l = (int)k0;
            N.r = (int)k1;
            N.on = C.nd[k0].on || C.nd[k1].on;
            st.push_back(i);
        } else {
            if (st.empty() || st.back() != (int)k0) return -1;
            st.pop_back();
            N.l = (int)k0;
            N.on = C.nd[k0].on;
            st.push_back(i);
        }
    }
    if (st.size() != 1 || st[0] != m - 1) return -1;
    if (unsupported) return -2;
    for (int i = 0; i < m; ++i)
        if (C.nd[(size_t)i].op == O_VAR && C.nd[(size_t)i].slot >= (uint32_t)VM_MAX_VARS) return -3;
    /* Leaves of ADD / MUL on the right: x*c and f*c fuse, and a leaf operand
       needs no spill. */
    for (int i = 0; i < m; ++i) {
        Nd &N = C.nd[(size_t)i];
        if (N.op != O_ADD && N.op != O_MUL) continue;
        const bool ll = C.leaf(N.l), rl = C.leaf(N.r);
        if ((ll && !rl) || (ll && rl && C.nd[N.l].op == O_CONST && C.nd[N.r].op == O_VAR))
            std::swap(N.l, N.r);
    }
    for (int s = 0; s < VM_U_SLOTS; ++s) C.ufree[s] = true;
    if (flags & VMC_INPUTS_IN_U) {
        /* the inputs occupy U slots 0..nvars-1 for the whole pass */
        if (nvars >= C.usink) return -3;
        for (int s = 0; s < nvars; ++s) C.ufree[s] = false;
        C.umax = nvars;
    }
    C.emit(m - 1);
    if (C.overflow) return -3;
    /* A lone constant: its gradient is the seed. */
    if (m == 1 && !C.vo) {
        Unit &u = C.units.back();
        u.rkind = VMR_ADDC_P;
        u.gl = (int)C.nd[0].slot;
        u.nodes_next = 0;
    }
    /* Reverse stream: units in reverse order with explicit saved-adjoint
       slots. */
    std::vector<uint32_t> rv;
    rv.reserve(C.vo ? 0 : 4 * C.units.size());
    int adepth = 0, apeak = 0;
    int popped = -1; /* the A slot the previous word ended by popping, else -1 */
    for (int i = C.vo ? -1 : (int)C.units.size() - 1; i >= 0; --i) {
        const Unit &u = C.units[(size_t)i];
        if (!u.rkind) continue;
        uint32_t x = u.rkind, y = 0;
        int as = VM_A_SINK, ar = VM_A_SINK;
        if (u.rkind == VMR_GEN) x |= u.ropt;
        if (u.nodes_next == 2) {
            if (adepth >= VM_A_SINK) return -3;
            as = adepth++;
            apeak = std::max(apeak, adepth);
            if (u.rkind == VMR_PUSH && popped == as) {
                /* The previous word popped the adjoint from this very slot:
                   it still holds the same bits, so the push is dropped (the
                   slot simply stays live). */
                popped = -1;
                continue;
            }
        } else if (u.nodes_next == 0) {
            ar = adepth > 0 ? --adepth : VM_A_SINK;
            if (u.rkind == VMR_GEN) x |= VMR_POP;
        }
        popped = (u.nodes_next == 0 && ar != VM_A_SINK) ? ar : -1;
        if (u.gl >= 0 && (u.gl + 1) * col_bytes >= (1 << (32 - VMR_GOFS_SHIFT))) return -3;
        if (u.gr >= 0 && (u.gr + 1) * col_bytes >= (1 << (32 - VMR_GOFS_SHIFT))) return -3;
        if (u.gl >= 0) x |= (uint32_t)(u.gl * col_bytes) << VMR_GOFS_SHIFT;
        if (u.cslot >= 0) y |= (uint32_t)(u.cslot * 4) & 0xffu;
        y |= ((uint32_t)(u.ut < 0 ? VM_U_SINK : u.ut) & 0xfu) << 8;
        y |= ((uint32_t)(u.bt < 0 ? VM_B_SINK : u.bt) & 0xfu) << 12;
        y |= ((uint32_t)as & 0x7u) << 16;
        y |= ((uint32_t)ar & 0x7u) << 19;
        y |= ((uint32_t)(u.xv < 0 ? 0 : u.xv) & 0x7u) << 22;
        y |= (u.rop & 0x1fu) << 25;
        rv.push_back(x);
        rv.push_back(y);
        if (u.rkind == VMR_GEN) {
            rv.push_back(u.gr >= 0 ? (uint32_t)(u.gr * col_bytes) << VMR_GOFS_SHIFT : 0u);
            rv.push_back(0u);
        }
    }
    int nf = 0;
    for (const Unit &u : C.units) nf += u.ffused ? 0 : 1;
    const int nr = (int)rv.size() / 2;
    if (!fwd || (!rev && nr > 0) || 2 * nf > cap_f || 2 * nr > cap_r) return -4;
    int w = 0;
    for (const Unit &u : C.units) {
        if (u.ffused) continue;
        fwd[2 * w] = u.fx;
        fwd[2 * w + 1] = u.fy;
        ++w;
    }
    if (!rv.empty()) memcpy(rev, rv.data(), rv.size() * 4);
    if (info) {
        info[0] = nf;
        info[1] = nr;
        info[2] = C.umax;
        info[3] = C.bnext;
        info[4] = apeak;
        info[5] = 0;
    }
    return 0;
}

} /* extern "C" */
