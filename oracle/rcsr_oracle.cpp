/*
 * CPU f64 oracle for the rcsr device arm — TEST INFRASTRUCTURE ONLY.
 *
 * A C++ restatement of the reference engine's exact CPU arithmetic.  Each
 * function cites the reference file and lines it restates (paths relative to
 * the reference repository's my-version/src).  Built with
 * -ffp-contract=off and no fast-math so every f64 operation is the IEEE
 * operation the reference performs; transcendental functions come from the
 * host libm, exactly as Rust's std does on this platform.
 *
 * Parity status of each function is recorded with its test:
 *   fit_lm / flatten / solve / sse / normal_equations: pinned bit-exactly by
 *     the six recorded golden cases (fit.rs:4654-4832).
 *   backward_impl: pinned by central differences and closed-form derivatives.
 *   irls_fit (L1): pinned by the weighted-median closed form, exact recovery
 *     of a clean law and the never-worse-than-start invariant.
 *   irls_fit (log-squashed): pinned by scipy's minimum of the same objective,
 *     by beating the squared fit on its own objective (fit.rs:4506) and by
 *     the base weight reaching the IRLS weight.
 *   metric_score: pinned by numpy means over the rows (squared, absolute,
 *     log-squashed; unweighted, weighted with zero weights, a weight of two
 *     as a repeated row) and the weight and finiteness rules.
 *   erf/erfc: pinned against the C library's erf/erfc (1e-14).
 *   implicit: pinned by the unit-circle closed form and scale invariance;
 *     implicit_score by a numpy weighted mean of the distances and the
 *     undefined-row rules.
 */
#include "rcsr_oracle.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <functional>
#include <limits>
#include <memory>
#include <thread>
#include <vector>

namespace {

const double NaN = std::numeric_limits<double>::quiet_NaN();
const double INF = std::numeric_limits<double>::infinity();

struct Node {
    int op;
    double val;
    std::vector<Node> kids;
};

int arity_of(int op)
{
    switch (op) {
        case OR_CONST: case OR_VAR: return 0;
        case OR_ADD: case OR_SUB: case OR_MUL: case OR_DIV: case OR_POW:
        case OR_MOD: case OR_ATAN2: case OR_MIN: case OR_MAX: case OR_EQUAL:
        case OR_LESS: case OR_LESS_OR_EQUAL: case OR_GREATER:
        case OR_GREATER_OR_EQUAL: case OR_AND: case OR_OR: case OR_XOR:
            return 2;
        case OR_IF: return 3;
        default:
            if (op >= 0 && op <= OR_ATANH) return 1;
            return -1;
    }
}

/* Parse a pre-order list; returns false on a malformed list. */
bool parse(const int *ops, const double *vals, int nn, int &pos, Node &out)
{
    if (pos >= nn) return false;
    out.op = ops[pos];
    out.val = vals[pos];
    ++pos;
    int a = arity_of(out.op);
    if (a < 0) return false;
    out.kids.resize(a);
    for (int i = 0; i < a; ++i)
        if (!parse(ops, vals, nn, pos, out.kids[i])) return false;
    return true;
}

bool build_tree(const int *ops, const double *vals, int nn, Node &root)
{
    int pos = 0;
    if (!parse(ops, vals, nn, pos, root)) return false;
    return pos == nn;
}

/* ---- expr.rs:930-1040: erf / erfc via the regularised incomplete gamma. */
const double ERF_EPS = 1e-15;
const double ERF_FPMIN = 1e-300;
const double PI = 3.14159265358979323846264338327950288;

double gcf_half(double x)
{
    const double a = 0.5;
    double b = x + 1.0 - a;
    double c = 1.0 / ERF_FPMIN;
    double d = 1.0 / b;
    double h = d;
    for (int i = 1; i < 300; ++i) {
        double an = -(double)i * ((double)i - a);
        b += 2.0;
        d = an * d + b;
        if (std::fabs(d) < ERF_FPMIN) d = ERF_FPMIN;
        c = b + an / c;
        if (std::fabs(c) < ERF_FPMIN) c = ERF_FPMIN;
        d = 1.0 / d;
        double del = d * c;
        h *= del;
        if (std::fabs(del - 1.0) < ERF_EPS) break;
    }
    return std::exp(-x + a * std::log(x) - 0.5 * std::log(PI)) * h;
}

double clamp01(double v)
{
    /* Rust f64::clamp(0, 1): NaN passes through. */
    if (v < 0.0) return 0.0;
    if (v > 1.0) return 1.0;
    return v;
}

double gammp_half(double x)
{
    if (std::isnan(x)) return NaN;
    if (x <= 0.0) return 0.0;
    if (!std::isfinite(x)) return 1.0;
    if (x < 1.5) {
        const double a = 0.5;
        double ap = a;
        double sum = 1.0 / a;
        double del = sum;
        for (int i = 0; i < 300; ++i) {
            ap += 1.0;
            del *= x / ap;
            sum += del;
            if (std::fabs(del) < std::fabs(sum) * ERF_EPS) break;
        }
        return clamp01(sum * std::exp(-x + a * std::log(x) - 0.5 * std::log(PI)));
    }
    return clamp01(1.0 - gcf_half(x));
}

double gammq_half(double x)
{
    if (std::isnan(x)) return NaN;
    if (x <= 0.0) return 1.0;
    if (!std::isfinite(x)) return 0.0;
    if (x < 1.5) return clamp01(1.0 - gammp_half(x));
    return clamp01(gcf_half(x));
}

double erf_ref(double x)
{
    if (std::isnan(x)) return NaN;
    double p = gammp_half(x * x);
    return x < 0.0 ? -p : p;
}

double erfc_ref(double x)
{
    if (std::isnan(x)) return NaN;
    double q = gammq_half(x * x);
    return x < 0.0 ? 2.0 - q : q;
}

/* Rust std's f64::asinh/acosh/atanh (library/std/src/f64.rs formulas). */
double rust_asinh(double x)
{
    double ax = std::fabs(x);
    double ix = 1.0 / ax;
    return std::copysign(std::log1p(ax + ax / (std::hypot(1.0, ix) + ix)), x);
}
double rust_acosh(double x)
{
    if (x < 1.0) return NaN;
    return std::log(x + std::sqrt(x - 1.0) * std::sqrt(x + 1.0));
}
double rust_atanh(double x) { return 0.5 * std::log1p((2.0 * x) / (1.0 - x)); }

double cmpf(double a, double b, int op)
{
    if (std::isnan(a) || std::isnan(b)) return NaN;
    bool r = false;
    switch (op) {
        case OR_EQUAL: r = a == b; break;
        case OR_LESS: r = a < b; break;
        case OR_LESS_OR_EQUAL: r = a <= b; break;
        case OR_GREATER: r = a > b; break;
        default: r = a >= b; break;
    }
    return r ? 1.0 : 0.0;
}

double logical(double a, double b, int op)
{
    if (std::isnan(a) || std::isnan(b)) return NaN;
    bool x = a > 0.0, y = b > 0.0;
    bool r = op == OR_AND ? (x && y) : op == OR_OR ? (x || y) : (x != y);
    return r ? 1.0 : 0.0;
}

/* ---- expr.rs:1054-1330: expr::eval with its domain guards. */
double eval(const Node &n, const double *x, int nx)
{
    auto k = [&](int i) { return eval(n.kids[i], x, nx); };
    switch (n.op) {
        case OR_CONST: return n.val;
        case OR_VAR: {
            size_t i = (size_t)n.val;
            return (n.val >= 0 && i < (size_t)nx) ? x[i] : NaN;
        }
        case OR_NEG: return -k(0);
        case OR_EXP: { double a = k(0); return a > 700.0 ? INF : std::exp(a); }
        case OR_LOG: { double a = k(0); return a > 0.0 ? std::log(a) : NaN; }
        case OR_ABS: return std::fabs(k(0));
        case OR_SQRT: { double a = k(0); return a >= 0.0 ? std::sqrt(a) : NaN; }
        case OR_SIN: return std::sin(k(0));
        case OR_COS: return std::cos(k(0));
        case OR_TAN: { double t = std::tan(k(0)); return std::isfinite(t) ? t : NaN; }
        case OR_SINH: return std::sinh(k(0));
        case OR_COSH: return std::cosh(k(0));
        case OR_TANH: return std::tanh(k(0));
        case OR_SIG: { double a = k(0); return 1.0 / (1.0 + std::exp(-a)); }
        case OR_STEP: return k(0) > 0.0 ? 1.0 : 0.0;
        case OR_GAUSS: { double a = k(0); return std::exp(-a * a); }
        case OR_SQUARE: { double a = k(0); return a * a; }
        case OR_CUBE: { double a = k(0); return a * a * a; }
        case OR_FLOOR: return std::floor(k(0));
        case OR_CEIL: return std::ceil(k(0));
        case OR_ROUND: return std::round(k(0));
        case OR_SIGN: {
            double a = k(0);
            if (a > 0.0) return 1.0;
            if (a < 0.0) return -1.0;
            if (a == 0.0) return 0.0;
            return NaN;
        }
        case OR_ERF: return erf_ref(k(0));
        case OR_ERFC: return erfc_ref(k(0));
        case OR_ASINH: return rust_asinh(k(0));
        case OR_ACOSH: { double a = k(0); return a >= 1.0 ? rust_acosh(a) : NaN; }
        case OR_ATANH: { double a = k(0); return std::fabs(a) < 1.0 ? rust_atanh(a) : NaN; }
        case OR_ADD: return k(0) + k(1);
        case OR_SUB: return k(0) - k(1);
        case OR_MUL: return k(0) * k(1);
        case OR_DIV: {
            double b = k(1);
            if (std::fabs(b) < 1e-12) return NaN;
            return k(0) / b;
        }
        case OR_POW: {
            double a = k(0), b = k(1);
            if (a < 0.0 && std::fabs(b - std::round(b)) > 1e-9) return NaN;
            return std::pow(a, b);
        }
        case OR_MOD: {
            double a = k(0), b = k(1);
            if (!std::isfinite(a) || !std::isfinite(b) || std::fabs(b) < 1e-12) return NaN;
            return a - b * std::floor(a / b);
        }
        case OR_MIN: {
            double a = k(0), b = k(1);
            if (std::isnan(a) || std::isnan(b)) return NaN;
            return std::fmin(a, b);
        }
        case OR_MAX: {
            double a = k(0), b = k(1);
            if (std::isnan(a) || std::isnan(b)) return NaN;
            return std::fmax(a, b);
        }
        case OR_ATAN2: {
            double y = k(0), xx = k(1);
            if (std::isnan(y) || std::isnan(xx)) return NaN;
            return std::atan2(y, xx);
        }
        case OR_IF: return k(0) > 0.0 ? k(1) : k(2);
        case OR_EQUAL: case OR_LESS: case OR_LESS_OR_EQUAL: case OR_GREATER:
        case OR_GREATER_OR_EQUAL: {
            double a = k(0), b = k(1);
            return cmpf(a, b, n.op);
        }
        case OR_AND: case OR_OR: case OR_XOR: {
            double a = k(0), b = k(1);
            return logical(a, b, n.op);
        }
        case OR_NOT: {
            double a = k(0);
            if (std::isnan(a)) return NaN;
            return a > 0.0 ? 0.0 : 1.0;
        }
        default: return NaN;
    }
}

/* ---- fit.rs:1101-1196: the flat post-order form (no block bodies here:
   the device arm's trees have no user-defined blocks). */
struct Flat {
    std::vector<int> op;
    std::vector<double> val;
    std::vector<int> k0, k1, k2, nk;
    std::vector<int> cslot;
    int root = 0;
    int nconst = 0;
};

int build_flat(const Node &n, Flat &f)
{
    int kid[3] = {0, 0, 0};
    for (size_t i = 0; i < n.kids.size(); ++i) kid[i] = build_flat(n.kids[i], f);
    int slot = -1;
    if (n.op == OR_CONST) slot = f.nconst++;
    f.op.push_back(n.op);
    f.val.push_back(n.val);
    f.k0.push_back(kid[0]);
    f.k1.push_back(kid[1]);
    f.k2.push_back(kid[2]);
    f.nk.push_back((int)n.kids.size());
    f.cslot.push_back(slot);
    return (int)f.op.size() - 1;
}

Flat flatten(const Node &root)
{
    Flat f;
    f.root = build_flat(root, f);
    return f;
}

/* Constants in slot order (= left-to-right leaf order = pre-order order). */
void collect_consts(const Node &n, std::vector<double> &out)
{
    if (n.op == OR_CONST) out.push_back(n.val);
    for (const Node &k : n.kids) collect_consts(k, out);
}

void write_consts(Node &n, const double *c, size_t &i)
{
    if (n.op == OR_CONST) n.val = c[i++];
    for (Node &k : n.kids) write_consts(k, c, i);
}

/* ---- fit.rs:1211-1365: value_of (the fitter's per-node arithmetic). */
double value_of(const Flat &f, int i, const std::vector<double> &v)
{
    const int op = f.op[i];
    auto a = [&](int j) { return v[j == 0 ? f.k0[i] : j == 1 ? f.k1[i] : f.k2[i]]; };
    switch (op) {
        case OR_CONST: case OR_VAR: return f.val[i];
        case OR_NEG: return -a(0);
        case OR_EXP: { double x = a(0); return x > 700.0 ? INF : std::exp(x); }
        case OR_LOG: return std::log(a(0));
        case OR_ABS: return std::fabs(a(0));
        case OR_SQRT: return std::sqrt(a(0));
        case OR_SIN: return std::sin(a(0));
        case OR_COS: return std::cos(a(0));
        case OR_TAN: return std::tan(a(0));
        case OR_SINH: return std::sinh(a(0));
        case OR_COSH: return std::cosh(a(0));
        case OR_TANH: return std::tanh(a(0));
        case OR_SIG: { double x = a(0); return 1.0 / (1.0 + std::exp(-x)); }
        case OR_STEP: return a(0) > 0.0 ? 1.0 : 0.0;
        case OR_GAUSS: { double x = a(0); return std::exp(-x * x); }
        case OR_SQUARE: return a(0) * a(0);
        case OR_CUBE: return a(0) * a(0) * a(0);
        case OR_FLOOR: return std::floor(a(0));
        case OR_CEIL: return std::ceil(a(0));
        case OR_ROUND: return std::round(a(0));
        case OR_SIGN: {
            double x = a(0);
            if (x > 0.0) return 1.0;
            if (x < 0.0) return -1.0;
            if (x == 0.0) return 0.0;
            return NaN;
        }
        case OR_ERF: return erf_ref(a(0));
        case OR_ERFC: return erfc_ref(a(0));
        case OR_ASINH: return rust_asinh(a(0));
        case OR_ACOSH: { double x = a(0); return x >= 1.0 ? rust_acosh(x) : NaN; }
        case OR_ATANH: { double x = a(0); return std::fabs(x) < 1.0 ? rust_atanh(x) : NaN; }
        case OR_ADD: return a(0) + a(1);
        case OR_SUB: return a(0) - a(1);
        case OR_MUL: return a(0) * a(1);
        case OR_DIV: { double b = a(1); return std::fabs(b) < 1e-12 ? NaN : a(0) / b; }
        case OR_POW: return std::pow(a(0), a(1));
        case OR_MOD: {
            double x = a(0), b = a(1);
            if (!std::isfinite(x) || !std::isfinite(b) || std::fabs(b) < 1e-12) return NaN;
            return x - b * std::floor(x / b);
        }
        case OR_MIN: {
            double x = a(0), y = a(1);
            if (std::isnan(x) || std::isnan(y)) return NaN;
            return std::fmin(x, y);
        }
        case OR_MAX: {
            double x = a(0), y = a(1);
            if (std::isnan(x) || std::isnan(y)) return NaN;
            return std::fmax(x, y);
        }
        case OR_ATAN2: return std::atan2(a(0), a(1));
        case OR_IF: return a(0) > 0.0 ? a(1) : a(2);
        case OR_EQUAL: case OR_LESS: case OR_LESS_OR_EQUAL: case OR_GREATER:
        case OR_GREATER_OR_EQUAL: return cmpf(a(0), a(1), op);
        case OR_AND: case OR_OR: case OR_XOR: return logical(a(0), a(1), op);
        case OR_NOT: {
            double x = a(0);
            if (std::isnan(x)) return NaN;
            return x > 0.0 ? 0.0 : 1.0;
        }
        default: return NaN;
    }
}

/* ---- fit.rs:1424-1440: eval_row. */
void eval_row(const Flat &f, const double *x, int nx, const double *c, std::vector<double> &v)
{
    const int m = (int)f.op.size();
    for (int i = 0; i < m; ++i) {
        if (f.cslot[i] >= 0) {
            v[i] = c[f.cslot[i]];
        } else if (f.op[i] == OR_VAR) {
            size_t idx = (size_t)f.val[i];
            v[i] = (f.val[i] >= 0 && idx < (size_t)nx) ? x[idx] : NaN;
        } else {
            v[i] = value_of(f, i, v);
        }
    }
}

double ind(double x)
{
    if (std::isnan(x)) return NaN;
    return x > 0.0 ? 1.0 : 0.0;
}

const double INV_SQRT_PI = 0.5641895835477563;

/* ---- fit.rs:1453-1671: backward_impl (adjoint sweep). */
void backward_impl(const Flat &f, const std::vector<double> &v, std::vector<double> &adj,
                   double *cg, int ncg, double *ig, int nig)
{
    const int m = (int)f.op.size();
    std::fill(adj.begin(), adj.end(), 0.0);
    if (cg) for (int i = 0; i < ncg; ++i) cg[i] = 0.0;
    if (ig) for (int i = 0; i < nig; ++i) ig[i] = 0.0;
    adj[f.root] = 1.0;
    for (int i = m - 1; i >= 0; --i) {
        const double a = adj[i];
        if (a == 0.0) continue;
        const int op = f.op[i];
        if (op == OR_CONST) {
            if (cg && f.cslot[i] >= 0 && f.cslot[i] < ncg) cg[f.cslot[i]] += a;
            continue;
        }
        if (op == OR_VAR) {
            size_t s = (size_t)f.val[i];
            if (ig && f.val[i] >= 0 && s < (size_t)nig) ig[s] += a;
            continue;
        }
        if (f.nk[i] == 0) continue;
        const int q0 = f.k0[i], q1 = f.k1[i], q2 = f.k2[i];
        auto get = [&](int j) { return v[j == 0 ? q0 : j == 1 ? q1 : q2]; };
        switch (op) {
            case OR_NEG: adj[q0] += -a; break;
            case OR_EXP: adj[q0] += a * std::exp(get(0)); break;
            case OR_LOG: adj[q0] += a / get(0); break;
            case OR_ABS: adj[q0] += a * (get(0) >= 0.0 ? 1.0 : -1.0); break;
            case OR_SQRT: adj[q0] += a / (2.0 * std::sqrt(get(0))); break;
            case OR_SIN: adj[q0] += a * std::cos(get(0)); break;
            case OR_COS: adj[q0] += -a * std::sin(get(0)); break;
            case OR_TAN: { double t = std::cos(get(0)); adj[q0] += a / (t * t); } break;
            case OR_SINH: adj[q0] += a * std::cosh(get(0)); break;
            case OR_COSH: adj[q0] += a * std::sinh(get(0)); break;
            case OR_TANH: { double t = std::tanh(get(0)); adj[q0] += a * (1.0 - t * t); } break;
            case OR_SIG: {
                double s = 1.0 / (1.0 + std::exp(-get(0)));
                adj[q0] += a * s * (1.0 - s);
            } break;
            case OR_STEP: break;
            case OR_GAUSS: adj[q0] += a * (-2.0 * get(0)) * std::exp(-get(0) * get(0)); break;
            case OR_SQUARE: adj[q0] += a * 2.0 * get(0); break;
            case OR_CUBE: adj[q0] += a * 3.0 * get(0) * get(0); break;
            case OR_FLOOR: case OR_CEIL: case OR_ROUND: case OR_SIGN: break;
            case OR_ERF: adj[q0] += a * 2.0 * INV_SQRT_PI * std::exp(-get(0) * get(0)); break;
            case OR_ERFC: adj[q0] += -a * 2.0 * INV_SQRT_PI * std::exp(-get(0) * get(0)); break;
            case OR_ASINH: adj[q0] += a / std::sqrt(get(0) * get(0) + 1.0); break;
            case OR_ACOSH: { double x = get(0); if (x > 1.0) adj[q0] += a / std::sqrt(x * x - 1.0); } break;
            case OR_ATANH: { double x = get(0); if (std::fabs(x) < 1.0) adj[q0] += a / (1.0 - x * x); } break;
            case OR_ADD: adj[q0] += a; adj[q1] += a; break;
            case OR_SUB: adj[q0] += a; adj[q1] += -a; break;
            case OR_MUL: adj[q0] += a * get(1); adj[q1] += a * get(0); break;
            case OR_DIV: {
                double d = get(1);
                adj[q0] += a / d;
                adj[q1] += -a * get(0) / (d * d);
            } break;
            case OR_POW: {
                double b = get(0), e = get(1);
                if (b > 0.0) {
                    adj[q0] += a * e * std::pow(b, e - 1.0);
                    adj[q1] += a * std::pow(b, e) * std::log(b);
                }
            } break;
            case OR_MOD: {
                double b = get(1);
                if (std::fabs(b) >= 1e-12) {
                    adj[q0] += a;
                    adj[q1] += -a * std::floor(get(0) / b);
                }
            } break;
            case OR_MIN: {
                double x = get(0), y = get(1);
                if (x < y) adj[q0] += a;
                else if (y < x) adj[q1] += a;
                else { adj[q0] += 0.5 * a; adj[q1] += 0.5 * a; }
            } break;
            case OR_MAX: {
                double x = get(0), y = get(1);
                if (x > y) adj[q0] += a;
                else if (y > x) adj[q1] += a;
                else { adj[q0] += 0.5 * a; adj[q1] += 0.5 * a; }
            } break;
            case OR_ATAN2: {
                double y = get(0), x = get(1);
                double d = x * x + y * y;
                if (d > 0.0) { adj[q0] += a * x / d; adj[q1] += -a * y / d; }
            } break;
            case OR_IF:
                if (get(0) > 0.0) adj[q1] += a; else adj[q2] += a;
                break;
            case OR_AND: adj[q0] += a * ind(get(1)); adj[q1] += a * ind(get(0)); break;
            case OR_OR:
                adj[q0] += a * (1.0 - ind(get(1)));
                adj[q1] += a * (1.0 - ind(get(0)));
                break;
            case OR_XOR:
                adj[q0] += a * (1.0 - 2.0 * ind(get(1)));
                adj[q1] += a * (1.0 - 2.0 * ind(get(0)));
                break;
            default: break;
        }
    }
}

/* Rust f64::clamp: NaN passes through unchanged. */
double clampr(double v, double lo, double hi)
{
    if (v < lo) return lo;
    if (v > hi) return hi;
    return v;
}

/* A row-major dataset view. */
struct Data {
    const double *x;
    int n;
    int nv;
    const double *y;
    const double *w; /* NULL = unweighted */
    const double *row(int i) const { return x + (size_t)i * (size_t)nv; }
};

/* ---- fit.rs:1686-1730: solve (plain Cholesky of A_ii(1+λ)+1e-12). */
bool solve(const std::vector<double> &A, const std::vector<double> &b, double lambda, int k,
           std::vector<double> &d)
{
    std::vector<double> m(A);
    for (int i = 0; i < k; ++i) m[i * k + i] = m[i * k + i] * (1.0 + lambda) + 1e-12;
    std::vector<double> L((size_t)k * k, 0.0);
    for (int i = 0; i < k; ++i) {
        for (int j = 0; j <= i; ++j) {
            double s = m[i * k + j];
            for (int p = 0; p < j; ++p) s -= L[i * k + p] * L[j * k + p];
            if (i == j) {
                if (s <= 1e-18) return false;
                L[i * k + j] = std::sqrt(s);
            } else {
                L[i * k + j] = s / L[j * k + j];
            }
        }
    }
    std::vector<double> y(k, 0.0);
    for (int i = 0; i < k; ++i) {
        double s = b[i];
        for (int p = 0; p < i; ++p) s -= L[i * k + p] * y[p];
        y[i] = s / L[i * k + i];
    }
    d.assign(k, 0.0);
    for (int i = k - 1; i >= 0; --i) {
        double s = y[i];
        for (int p = i + 1; p < k; ++p) s -= L[p * k + i] * d[p];
        d[i] = s / L[i * k + i];
    }
    return true;
}

/* ---- fit.rs:1732-1760: sse (weighted, zero-weight rows skipped). */
double sse(const Flat &f, const Data &D, const double *c, std::vector<double> &v)
{
    double s = 0.0;
    for (int i = 0; i < D.n; ++i) {
        double sw = D.w ? D.w[i] : 1.0;
        if (sw == 0.0) continue;
        eval_row(f, D.row(i), D.nv, c, v);
        double r = std::sqrt(sw) * (v[f.root] - D.y[i]);
        if (!std::isfinite(r)) return INF;
        s += r * r;
    }
    return s;
}

/* ---- fit.rs:1774-1827: normal_equations (JᵀWJ, JᵀWr). */
bool normal_equations(const Flat &f, const Data &D, const double *c, int k,
                      std::vector<double> &v, std::vector<double> &adj,
                      std::vector<double> &grad, std::vector<double> &jtj,
                      std::vector<double> &jtr)
{
    jtj.assign((size_t)k * k, 0.0);
    jtr.assign(k, 0.0);
    for (int row = 0; row < D.n; ++row) {
        double sw = D.w ? D.w[row] : 1.0;
        if (sw == 0.0) continue;
        eval_row(f, D.row(row), D.nv, c, v);
        double r = std::sqrt(sw) * (v[f.root] - D.y[row]);
        if (!std::isfinite(r)) return false;
        backward_impl(f, v, adj, grad.data(), k, nullptr, 0);
        double jw = std::sqrt(sw);
        for (int i = 0; i < k; ++i) {
            double gi = jw * grad[i];
            if (!std::isfinite(gi)) return false;
            jtr[i] += gi * r;
            for (int j = 0; j <= i; ++j) jtj[i * k + j] += gi * (jw * grad[j]);
        }
    }
    for (int i = 0; i < k; ++i)
        for (int j = 0; j < i; ++j) jtj[j * k + i] = jtj[i * k + j];
    return true;
}

/* ---- fit.rs:2318-2338: gn_trial. */
bool gn_trial(const Flat &f, const Data &D, const std::vector<double> &c, double lambda, int k,
              const std::vector<double> &jtj, const std::vector<double> &jtr,
              std::vector<double> &v, std::vector<double> &trial, double &tsse)
{
    std::vector<double> rhs(k);
    for (int i = 0; i < k; ++i) rhs[i] = -jtr[i];
    std::vector<double> step;
    if (!solve(jtj, rhs, lambda, k, step)) return false;
    trial.resize(k);
    for (int i = 0; i < k; ++i) trial[i] = clampr(c[i] + step[i], -1e6, 1e6);
    tsse = sse(f, D, trial.data(), v);
    return true;
}

/* ---- fit.rs:2534-2624: fit_lm. `c` holds the starting constants (slot
   order) and receives the best ones. Returns the best weighted SSE. */
double fit_lm(const Flat &f, const Data &D, std::vector<double> &c, int iters)
{
    const int k = f.nconst;
    const int m = (int)f.op.size();
    std::vector<double> v(std::max(m, 1), 0.0), adj(std::max(m, 1), 0.0), grad(std::max(k, 1), 0.0);
    if (k == 0 || D.n == 0) return sse(f, D, c.data(), v);
    if (iters == 0) return sse(f, D, c.data(), v);
    std::vector<double> best_c = c;
    double best_sse = sse(f, D, c.data(), v);
    double lambda = 1e-3;
    bool have_normal = false;
    std::vector<double> jtj, jtr;
    const int loops = std::max(iters, 1);
    for (int it = 0; it < loops; ++it) {
        if (!have_normal) {
            have_normal = normal_equations(f, D, c.data(), k, v, adj, grad, jtj, jtr);
        }
        if (!have_normal) break;
        std::vector<double> trial;
        double tsse;
        if (!gn_trial(f, D, c, lambda, k, jtj, jtr, v, trial, tsse)) break;
        if (tsse < best_sse) {
            c = trial;
            best_sse = tsse;
            best_c = c;
            lambda = std::fmax(lambda / 8.0, 1e-12);
            have_normal = false;
            if (best_sse == 0.0) break;
        } else {
            lambda *= 6.0;
            if (lambda > 1e9) break;
        }
    }
    for (int it = 0; it < 4; ++it) {
        if (best_sse == 0.0) break;
        std::vector<double> njtj, njtr;
        if (!normal_equations(f, D, best_c.data(), k, v, adj, grad, njtj, njtr)) break;
        std::vector<double> trial;
        double tsse;
        if (!gn_trial(f, D, best_c, 0.0, k, njtj, njtr, v, trial, tsse)) break;
        if (tsse < best_sse) {
            best_c = trial;
            best_sse = tsse;
        } else {
            break;
        }
    }
    c = best_c;
    return best_sse;
}

/* fit_weighted (fit.rs:2404-2416, Exact strategy): fit_lm on the tree. */
double fit_weighted(Node &tree, const Data &D, int iters)
{
    Flat f = flatten(tree);
    std::vector<double> c;
    collect_consts(tree, c);
    double s = fit_lm(f, D, c, iters);
    size_t i = 0;
    write_consts(tree, c.data(), i);
    return s;
}

std::vector<double> predictions(const Node &tree, const Data &D)
{
    std::vector<double> p(D.n);
    for (int i = 0; i < D.n; ++i) p[i] = eval(tree, D.row(i), D.nv);
    return p;
}

/* ---- metrics.rs:1467-1475 binary_labels: label 1 above the midpoint of
   the target's range, else 0. */
std::vector<double> binary_labels(const Data &D)
{
    double lo = INF, hi = -INF;
    for (int i = 0; i < D.n; ++i) {
        lo = std::fmin(lo, D.y[i]);
        hi = std::fmax(hi, D.y[i]);
    }
    const double mid = 0.5 * (lo + hi);
    std::vector<double> l(D.n);
    for (int i = 0; i < D.n; ++i) l[i] = D.y[i] > mid ? 1.0 : 0.0;
    return l;
}

/* ---- metrics.rs:1547-1550 softplus_neg: ln(1 + exp(-u)) without overflow. */
double softplus_neg(double u)
{
    const double a = std::fabs(u);
    return std::fmax(-u, 0.0) + std::log1p(std::exp(-a));
}

/* ---- metrics.rs:1610-1620 log_loss (unweighted): mean softplus_neg(t p). */
double log_loss_pred(const std::vector<double> &pred, const Data &D)
{
    if (pred.empty()) return INF;
    const std::vector<double> l = binary_labels(D);
    double acc = 0.0;
    for (size_t i = 0; i < pred.size(); ++i) acc += softplus_neg((2.0 * l[i] - 1.0) * pred[i]);
    return acc / (double)pred.size();
}

/* ---- metrics.rs:852-900 (unweighted) and 985-1060 (weighted) for the
   squared / log-squashed / absolute rows. */
double score_pred(const std::vector<double> &pred, const Data &D, int metric)
{
    const int n = (int)pred.size();
    if (n == 0) return INF;
    if (metric == OR_METRIC_LOG_LOSS) return log_loss_pred(pred, D);
    auto rho = [&](double e) {
        if (metric == OR_METRIC_ABSOLUTE) return std::fabs(e);
        if (metric == OR_METRIC_LOG_SQUASHED) return std::log(1.0 + e * e);
        return e * e;
    };
    if (!D.w) {
        for (double p : pred)
            if (!std::isfinite(p)) return INF;
        double s = 0.0;
        for (int i = 0; i < n; ++i) s += rho(pred[i] - D.y[i]);
        return s / (double)n;
    }
    double total = 0.0;
    for (int i = 0; i < n; ++i) {
        double wi = D.w[i];
        if (!std::isfinite(wi) || wi < 0.0) return INF;
        total += wi;
    }
    if (!(total > 0.0)) return INF;
    for (int i = 0; i < n; ++i)
        if (D.w[i] > 0.0 && !std::isfinite(pred[i])) return INF;
    double s = 0.0;
    for (int i = 0; i < n; ++i) {
        double e = D.w[i] > 0.0 ? pred[i] - D.y[i] : 0.0;
        s += D.w[i] * rho(e);
    }
    return s / total;
}

double metric_score(const Node &tree, const Data &D, int metric)
{
    return score_pred(predictions(tree, D), D, metric);
}

/* ---- fit.rs:3092-3170: irls_fit. */
double irls_fit(Node &tree, const Data &D, int iters, int metric)
{
    auto weight_of = [&](double r) {
        if (metric == OR_METRIC_ABSOLUTE) return 1.0 / std::fmax(std::fabs(r), 1e-6);
        return 1.0 / (1.0 + r * r);
    };
    double best = metric_score(tree, D, metric);
    if (!std::isfinite(best)) return best;
    const int rounds = std::min(std::max(iters, 1), 4);
    const int inner = std::max(iters / rounds, 1);
    std::vector<double> weights(D.n);
    for (int round = 0; round < rounds; ++round) {
        std::vector<double> pred = predictions(tree, D);
        bool finite = true;
        for (int i = 0; i < D.n; ++i) {
            double r = pred[i] - D.y[i];
            if (!std::isfinite(r)) { finite = false; break; }
            double base = D.w ? D.w[i] : 1.0;
            weights[i] = base > 0.0 ? base * std::fmax(weight_of(r), 0.0) : 0.0;
        }
        if (!finite) return best;
        Node trial = tree;
        Data Dw = D;
        Dw.w = weights.data();
        fit_weighted(trial, Dw, inner);
        double score = metric_score(trial, D, metric);
        if (std::isfinite(score) && score < best) {
            best = score;
            tree = trial;
        } else {
            break;
        }
    }
    std::vector<double> pred = predictions(tree, D);
    bool all_finite = true;
    for (double p : pred)
        if (!std::isfinite(p)) { all_finite = false; break; }
    if (all_finite) {
        for (int i = 0; i < D.n; ++i) {
            double r = pred[i] - D.y[i];
            double base = D.w ? D.w[i] : 1.0;
            weights[i] = base > 0.0 ? base * std::fmax(weight_of(r), 0.0) : 0.0;
        }
        Node trial = tree;
        Data Dw = D;
        Dw.w = weights.data();
        fit_weighted(trial, Dw, std::max(iters, 4));
        double score = metric_score(trial, D, metric);
        if (std::isfinite(score) && score < best) {
            best = score;
            tree = trial;
        }
    }
    return best;
}

/* ---- fit.rs:3271-3331: logistic_irls.  Each round solves the Newton
   step's weighted least squares min sum w q (1 - q) (p - z)^2 with
   q = sigmoid(t p) clamped to [1e-9, 1 - 1e-9] and z = p + t / q clamped to
   [-1e6, 1e6] (two inner fit iterations), and keeps the trial only when the
   log loss strictly improves; at most min(max(iters, 1), 16) rounds. */
double logistic_irls(Node &tree, const Data &D, int iters)
{
    double best = metric_score(tree, D, OR_METRIC_LOG_LOSS);
    if (!std::isfinite(best)) return best;
    const std::vector<double> l = binary_labels(D);
    const int rounds = std::min(std::max(iters, 1), 16);
    std::vector<double> target(D.n), weights(D.n);
    for (int round = 0; round < rounds; ++round) {
        const std::vector<double> pred = predictions(tree, D);
        bool finite = true;
        for (double p : pred)
            if (!std::isfinite(p)) { finite = false; break; }
        if (!finite) break;
        for (int i = 0; i < D.n; ++i) {
            const double t = 2.0 * l[i] - 1.0;
            const double u = t * pred[i];
            double q;
            if (u >= 0.0) {
                q = 1.0 / (1.0 + std::exp(-u));
            } else {
                const double e = std::exp(u);
                q = e / (1.0 + e);
            }
            q = std::fmin(std::fmax(q, 1e-9), 1.0 - 1e-9);
            const double base = D.w ? D.w[i] : 1.0;
            weights[i] = base > 0.0 ? base * q * (1.0 - q) : 0.0;
            target[i] = std::fmin(std::fmax(pred[i] + t / q, -1e6), 1e6);
        }
        Node trial = tree;
        Data Dw = D;
        Dw.y = target.data();
        Dw.w = weights.data();
        fit_weighted(trial, Dw, 2);
        const double score = metric_score(trial, D, OR_METRIC_LOG_LOSS);
        if (std::isfinite(score) && score < best) {
            best = score;
            tree = trial;
        } else {
            break;
        }
    }
    return best;
}

/* ---- fit.rs:2937-2990: fit_weighted_metric (squared / log / absolute /
   log loss). */
double fit_metric(Node &tree, const Data &D, int iters, int metric)
{
    if (metric == OR_METRIC_LOG_LOSS) return logistic_irls(tree, D, iters);
    if (iters == 0) {
        if (metric == OR_METRIC_SQUARED) return fit_weighted(tree, D, 0);
        return metric_score(tree, D, metric);
    }
    if (metric == OR_METRIC_SQUARED) return fit_weighted(tree, D, iters);
    return irls_fit(tree, D, iters, metric);
}

/* ---- implicit.rs:68-96: signed_distance. */
bool signed_distance(const Node &tree, const double *row, int nv, double &d)
{
    double fv = eval(tree, row, nv);
    if (!std::isfinite(fv)) return false;
    Flat f = flatten(tree);
    std::vector<double> c;
    collect_consts(tree, c);
    std::vector<double> v(f.op.size()), adj(f.op.size()), g(nv);
    eval_row(f, row, nv, c.data(), v);
    backward_impl(f, v, adj, nullptr, 0, g.data(), nv);
    double norm_sq = 0.0;
    for (int i = 0; i < nv; ++i) {
        if (!std::isfinite(g[i])) return false;
        norm_sq += g[i] * g[i];
    }
    double norm = std::sqrt(norm_sq);
    if (!(std::isfinite(norm) && norm > 0.0)) return false;
    d = fv / norm;
    return std::isfinite(d);
}

/* implicit.rs:122-155: score. */
double implicit_score(const Node &tree, const Data &D, int metric)
{
    if (D.n == 0) return INF;
    double total = 0.0, sum = 0.0;
    for (int i = 0; i < D.n; ++i) {
        double w = D.w ? D.w[i] : 1.0;
        if (!std::isfinite(w) || w < 0.0) return INF;
        if (w == 0.0) continue;
        double dist;
        if (metric == 0) {
            double sd;
            dist = signed_distance(tree, D.row(i), D.nv, sd) ? std::fabs(sd) : INF;
        } else {
            double fv = eval(tree, D.row(i), D.nv);
            dist = std::isfinite(fv) ? std::fabs(fv) : INF;
        }
        if (!std::isfinite(dist)) return INF;
        total += w;
        sum += w * dist;
    }
    if (!(std::isfinite(total) && total > 0.0)) return INF;
    return sum / total;
}

/* implicit.rs:159-180: residuals. */
bool implicit_residuals(const Node &tree, const Data &D, std::vector<double> &out)
{
    out.resize(D.n);
    for (int i = 0; i < D.n; ++i) {
        double wi = D.w ? D.w[i] : 1.0;
        if (!std::isfinite(wi) || wi < 0.0) return false;
        if (wi == 0.0) { out[i] = 0.0; continue; }
        double d;
        if (!signed_distance(tree, D.row(i), D.nv, d)) return false;
        out[i] = std::sqrt(wi) * d;
    }
    return true;
}

double cost(const std::vector<double> &r)
{
    double s = 0.0;
    for (double v : r) s += v * v;
    return s;
}

/* implicit.rs:215-253: Gaussian elimination with partial pivoting. */
bool implicit_solve(const std::vector<double> &A, const std::vector<double> &b, double lambda,
                    int k, std::vector<double> &d)
{
    std::vector<std::vector<double>> m(k, std::vector<double>(k + 1, 0.0));
    for (int i = 0; i < k; ++i) {
        for (int j = 0; j < k; ++j) m[i][j] = A[i * k + j];
        m[i][i] = m[i][i] * (1.0 + lambda) + 1e-12;
        m[i][k] = b[i];
    }
    for (int col = 0; col < k; ++col) {
        int piv = col;
        for (int r = col + 1; r < k; ++r)
            if (std::fabs(m[r][col]) > std::fabs(m[piv][col])) piv = r;
        if (std::fabs(m[piv][col]) < 1e-300) return false;
        std::swap(m[col], m[piv]);
        for (int r = col + 1; r < k; ++r) {
            double factor = m[r][col] / m[col][col];
            if (factor != 0.0)
                for (int c = col; c <= k; ++c) m[r][c] -= factor * m[col][c];
        }
    }
    d.assign(k, 0.0);
    for (int i = k - 1; i >= 0; --i) {
        double s = m[i][k];
        for (int j = i + 1; j < k; ++j) s -= m[i][j] * d[j];
        d[i] = s / m[i][i];
    }
    return true;
}

Node with_consts(const Node &base, const std::vector<double> &c)
{
    Node t = base;
    size_t i = 0;
    write_consts(t, c.data(), i);
    return t;
}

double rss_of(const Node &tree, const Data &D)
{
    std::vector<double> r;
    if (!implicit_residuals(tree, D, r)) return INF;
    return cost(r);
}

/* implicit.rs:277-409: fit (central-difference LM on f/||∇f||). */
double implicit_fit(Node &tree, const Data &D, int iters)
{
    if (D.n == 0) return INF;
    std::vector<double> initial;
    collect_consts(tree, initial);
    const int k = (int)initial.size();
    if (k == 0) return rss_of(tree, D);
    if (iters == 0) return rss_of(tree, D);
    const Node base = tree;
    std::vector<double> base_r;
    if (!implicit_residuals(base, D, base_r)) return INF;
    std::vector<double> best_c = initial;
    double best_sse = cost(base_r);
    std::vector<double> c = best_c;
    double lambda = 1e-3;
    const int n = D.n;
    for (int it = 0; it < std::max(iters, 1); ++it) {
        std::vector<double> r;
        if (!implicit_residuals(with_consts(base, c), D, r)) break;
        std::vector<double> jac((size_t)n * k, 0.0);
        bool ok = true;
        for (int j = 0; j < k && ok; ++j) {
            double h = 1e-6 * (1.0 + std::fabs(c[j]));
            std::vector<double> cp = c, cm = c;
            cp[j] += h;
            cm[j] -= h;
            std::vector<double> rp, rm;
            bool okp = implicit_residuals(with_consts(base, cp), D, rp);
            bool okm = okp && implicit_residuals(with_consts(base, cm), D, rm);
            if (!(okp && okm)) { ok = false; break; }
            for (int i = 0; i < n; ++i) jac[(size_t)i * k + j] = (rp[i] - rm[i]) / (2.0 * h);
        }
        if (!ok) break;
        std::vector<double> jtj((size_t)k * k, 0.0), jtr(k, 0.0);
        for (int i = 0; i < n && ok; ++i) {
            for (int a = 0; a < k; ++a) {
                double ga = jac[(size_t)i * k + a];
                if (!std::isfinite(ga)) { ok = false; break; }
                jtr[a] += ga * r[i];
                for (int b = 0; b <= a; ++b) jtj[a * k + b] += ga * jac[(size_t)i * k + b];
            }
        }
        if (!ok) break;
        for (int a = 0; a < k; ++a)
            for (int b = 0; b < a; ++b) jtj[b * k + a] = jtj[a * k + b];
        std::vector<double> rhs(k);
        for (int a = 0; a < k; ++a) rhs[a] = -jtr[a];
        std::vector<double> step;
        if (!implicit_solve(jtj, rhs, lambda, k, step)) {
            lambda *= 6.0;
            if (lambda > 1e9) break;
            continue;
        }
        std::vector<double> trial(k);
        for (int a = 0; a < k; ++a) trial[a] = clampr(c[a] + step[a], -1e6, 1e6);
        std::vector<double> rt;
        if (implicit_residuals(with_consts(base, trial), D, rt)) {
            double s = cost(rt);
            if (s < best_sse) {
                best_sse = s;
                best_c = trial;
                c = trial;
                lambda = std::fmax(lambda / 8.0, 1e-12);
                if (best_sse == 0.0) break;
            } else {
                lambda *= 6.0;
                if (lambda > 1e9) break;
            }
        } else {
            lambda *= 6.0;
            if (lambda > 1e9) break;
        }
    }
    if (best_sse < rss_of(tree, D)) {
        size_t i = 0;
        write_consts(tree, best_c.data(), i);
    }
    return best_sse;
}

void run_threads_impl(int ntree, int threads, const std::function<void(int)> &job)
{
    if (threads <= 1) {
        for (int t = 0; t < ntree; ++t) job(t);
        return;
    }
    std::atomic<int> next(0);
    std::vector<std::thread> pool;
    for (int i = 0; i < threads; ++i) {
        pool.emplace_back([&]() {
            for (;;) {
                int t = next.fetch_add(1);
                if (t >= ntree) break;
                job(t);
            }
        });
    }
    for (auto &th : pool) th.join();
}
} // namespace

extern "C" {

int or_arity(int op) { return arity_of(op); }

double or_eval(const int *ops, const double *vals, int nn, const double *x, int nx)
{
    Node t;
    if (!build_tree(ops, vals, nn, t)) return NaN;
    return eval(t, x, nx);
}

double or_erf(double x) { return erf_ref(x); }
double or_erfc(double x) { return erfc_ref(x); }

int or_nconst(const int *ops, const double *vals, int nn)
{
    Node t;
    if (!build_tree(ops, vals, nn, t)) return -1;
    return flatten(t).nconst;
}

double or_sse(const int *ops, const double *vals, int nn, const double *x, int n, int nv,
              const double *y, const double *w)
{
    Node t;
    if (!build_tree(ops, vals, nn, t)) return NaN;
    Flat f = flatten(t);
    std::vector<double> c;
    collect_consts(t, c);
    std::vector<double> v(std::max<size_t>(f.op.size(), 1));
    Data D{x, n, nv, y, w};
    return sse(f, D, c.data(), v);
}

void or_predict_flat(const int *ops, const double *vals, int nn, const double *x, int n, int nv,
                     double *out)
{
    Node t;
    if (!build_tree(ops, vals, nn, t)) {
        for (int i = 0; i < n; ++i) out[i] = NaN;
        return;
    }
    Flat f = flatten(t);
    std::vector<double> c;
    collect_consts(t, c);
    std::vector<double> v(std::max<size_t>(f.op.size(), 1));
    for (int i = 0; i < n; ++i) {
        eval_row(f, x + (size_t)i * (size_t)nv, nv, c.data(), v);
        out[i] = v[f.op.size() - 1];
    }
}

void or_gradients(const int *ops, const double *vals, int nn, const double *x, int nv,
                  double *g, double *gx)
{
    Node t;
    if (!build_tree(ops, vals, nn, t)) return;
    Flat f = flatten(t);
    std::vector<double> c;
    collect_consts(t, c);
    std::vector<double> v(f.op.size()), adj(f.op.size());
    eval_row(f, x, nv, c.data(), v);
    backward_impl(f, v, adj, g, f.nconst, gx, nv);
}

double or_fit_weighted(const int *ops, const double *vals, int nn, const double *x, int n,
                       int nv, const double *y, const double *w, int iters, double *out_c)
{
    Node t;
    if (!build_tree(ops, vals, nn, t)) return NaN;
    Data D{x, n, nv, y, w};
    double s = fit_weighted(t, D, iters);
    std::vector<double> c;
    collect_consts(t, c);
    if (out_c) std::memcpy(out_c, c.data(), c.size() * sizeof(double));
    return s;
}

double or_fit_metric(const int *ops, const double *vals, int nn, const double *x, int n,
                     int nv, const double *y, const double *w, int iters, int metric,
                     double *out_c)
{
    Node t;
    if (!build_tree(ops, vals, nn, t)) return NaN;
    Data D{x, n, nv, y, w};
    double s = fit_metric(t, D, iters, metric);
    std::vector<double> c;
    collect_consts(t, c);
    if (out_c) std::memcpy(out_c, c.data(), c.size() * sizeof(double));
    return s;
}

double or_metric_score(const int *ops, const double *vals, int nn, const double *x, int n,
                       int nv, const double *y, const double *w, int metric)
{
    Node t;
    if (!build_tree(ops, vals, nn, t)) return NaN;
    Data D{x, n, nv, y, w};
    return metric_score(t, D, metric);
}

double or_implicit_score(const int *ops, const double *vals, int nn, const double *x, int n,
                         int nv, const double *w, int metric)
{
    Node t;
    if (!build_tree(ops, vals, nn, t)) return NaN;
    Data D{x, n, nv, nullptr, w};
    return implicit_score(t, D, metric);
}

int or_signed_distance(const int *ops, const double *vals, int nn, const double *x, int nv,
                       double *d)
{
    Node t;
    if (!build_tree(ops, vals, nn, t)) return -1;
    double v;
    if (!signed_distance(t, x, nv, v)) return -1;
    *d = v;
    return 0;
}

double or_implicit_fit(const int *ops, const double *vals, int nn, const double *x, int n,
                       int nv, const double *w, int iters, double *out_c)
{
    Node t;
    if (!build_tree(ops, vals, nn, t)) return NaN;
    Data D{x, n, nv, nullptr, w};
    double s = implicit_fit(t, D, iters);
    std::vector<double> c;
    collect_consts(t, c);
    if (out_c) std::memcpy(out_c, c.data(), c.size() * sizeof(double));
    return s;
}

void or_fit_batch(const int *ops, const double *vals, const int *tree_off, int ntree,
                  const int *metric, const double *x, int n, int nv, const double *y,
                  const double *w, int iters, int kmax, int threads, double *out_c,
                  double *out_loss)
{
    run_threads_impl(ntree, threads, [&](int t) {
        const int a = tree_off[t], b = tree_off[t + 1];
        double *c = out_c + (size_t)t * (size_t)kmax;
        for (int i = 0; i < kmax; ++i) c[i] = 0.0;
        Node tree;
        if (!build_tree(ops + a, vals + a, b - a, tree)) { out_loss[t] = NaN; return; }
        Data D{x, n, nv, y, w};
        int mt = metric ? metric[t] : OR_METRIC_SQUARED;
        out_loss[t] = fit_metric(tree, D, iters, mt);
        std::vector<double> cc;
        collect_consts(tree, cc);
        for (size_t i = 0; i < cc.size() && (int)i < kmax; ++i) c[i] = cc[i];
    });
}

void or_implicit_fit_batch(const int *ops, const double *vals, const int *tree_off, int ntree,
                           const double *x, int n, int nv, const double *w, int iters, int kmax,
                           int threads, double *out_c, double *out_loss)
{
    run_threads_impl(ntree, threads, [&](int t) {
        const int a = tree_off[t], b = tree_off[t + 1];
        double *c = out_c + (size_t)t * (size_t)kmax;
        for (int i = 0; i < kmax; ++i) c[i] = 0.0;
        Node tree;
        if (!build_tree(ops + a, vals + a, b - a, tree)) { out_loss[t] = NaN; return; }
        Data D{x, n, nv, nullptr, w};
        out_loss[t] = implicit_fit(tree, D, iters);
        std::vector<double> cc;
        collect_consts(tree, cc);
        for (size_t i = 0; i < cc.size() && (int)i < kmax; ++i) c[i] = cc[i];
    });
}

} // extern "C"

namespace {

/* ---- the correlation fit arm (metrics 2 and 3) ------------------------- */

/* The equilibrated, damped Cholesky step of the device fitter (eval.cu:
   1874-1949): jtj packed lower triangle (p x p), delta = -(D^-1/2 (A + lambda
   diag A) D^-1/2 + 1e-12)^-1 jtr in scaled form.  Returns 1 unusable, 4
   solved but rank-deficient (ln det below the floor), 0 solved. */
int eq_solve(const std::vector<double> &jtj, const std::vector<double> &jtr, double lambda, int p,
             std::vector<double> &delta)
{
    const double LN_DET_FLOOR = -27.631021115928547;
    std::vector<double> dsc(p), L((size_t)p * p, 0.0);
    for (int i = 0; i < p; ++i) {
        const double aii = jtj[i * (i + 1) / 2 + i];
        if (!(aii > 0.0) || !std::isfinite(aii)) return 1;
        dsc[i] = std::sqrt(aii);
    }
    bool ok = true;
    double sum_ln = 0.0;
    for (int i = 0; i < p && ok; ++i) {
        for (int j = 0; j <= i; ++j) {
            double s = jtj[i * (i + 1) / 2 + j] / (dsc[i] * dsc[j]);
            if (!std::isfinite(s)) return 1;
            for (int q = 0; q < j; ++q) s = std::fma(-L[i * p + q], L[j * p + q], s);
            if (i == j) {
                if (!(s > 0.0) || !std::isfinite(s)) { ok = false; break; }
                L[i * p + i] = std::sqrt(s);
                sum_ln = std::fma(0.5, std::log(s), sum_ln);
            } else {
                L[i * p + j] = s / L[j * p + j];
            }
        }
    }
    bool rank = !ok;
    if (ok) {
        const double ln_det = 2.0 * sum_ln;
        if (!std::isfinite(ln_det) || ln_det < LN_DET_FLOOR) rank = true;
    }
    for (int i = 0; i < p; ++i) {
        for (int j = 0; j <= i; ++j) {
            double s = jtj[i * (i + 1) / 2 + j] / (dsc[i] * dsc[j]);
            if (i == j) s = std::fma(s, 1.0 + lambda, 1e-12);
            for (int q = 0; q < j; ++q) s = std::fma(-L[i * p + q], L[j * p + q], s);
            if (i == j) {
                if (!(s > 1e-18) || !std::isfinite(s)) return 1;
                L[i * p + i] = std::sqrt(s);
            } else {
                L[i * p + j] = s / L[j * p + j];
            }
        }
    }
    std::vector<double> rhs(p);
    for (int i = 0; i < p; ++i) rhs[i] = -jtr[i] / dsc[i];
    for (int i = 0; i < p; ++i) {
        double s = rhs[i];
        for (int q = 0; q < i; ++q) s = std::fma(-L[i * p + q], rhs[q], s);
        rhs[i] = s / L[i * p + i];
    }
    delta.assign(p, 0.0);
    for (int i = p - 1; i >= 0; --i) {
        double s = rhs[i];
        for (int q = i + 1; q < p; ++q) s = std::fma(-L[q * p + i], delta[q], s);
        delta[i] = s / L[i * p + i];
    }
    for (int i = 0; i < p; ++i) delta[i] /= dsc[i];
    return rank ? 4 : 0;
}

struct CorrState {
    std::vector<double> jtj, jtr;
    double sff = 0.0, syy = 0.0, rho = 0.0, score = 1.0;
    bool bad = false;
};

/* correlation.rs moments / ft_corr_assemble (eval.cu:2009-2177): weighted
   moments over the rows with positive weight (a non-finite prediction,
   target or derivative marks the pass bad), the rho gradient and its
   Gauss-Newton curvature. */
void corr_assemble(const Flat &f, const Data &D, const std::vector<double> &c, int metric,
                   CorrState &S)
{
    const int p = f.nconst;
    const int m = (int)f.op.size();
    std::vector<double> v(m), adj(m), g(p);
    std::vector<double> C((size_t)p * (p + 1) / 2, 0.0), sg(p, 0.0), sgf(p, 0.0), sgy(p, 0.0);
    double W = 0.0, sf = 0.0, sf2 = 0.0, sft = 0.0, st = 0.0, st2 = 0.0;
    S.bad = false;
    for (int i = 0; i < D.n; ++i) {
        const double w = D.w ? D.w[i] : 1.0;
        if (!(w > 0.0)) continue;
        eval_row(f, D.row(i), D.nv, c.data(), v);
        const double fv = v[f.root];
        const double t = D.y[i];
        bool ok = std::isfinite(fv) && std::fabs(fv) <= 1e30 && std::isfinite(t);
        backward_impl(f, v, adj, g.data(), p, nullptr, 0);
        for (int j = 0; j < p; ++j)
            if (!std::isfinite(g[j])) ok = false;
        if (!ok) {
            S.bad = true;
            continue;
        }
        int k = 0;
        for (int a = 0; a < p; ++a)
            for (int b = 0; b <= a; ++b, ++k) C[k] += (w * g[a]) * g[b];
        for (int j = 0; j < p; ++j) {
            sg[j] += w * g[j];
            sgf[j] += (w * fv) * g[j];
            sgy[j] += (w * t) * g[j];
        }
        W += w;
        sf += w * fv;
        sf2 += (w * fv) * fv;
        st += w * t;
        sft += (w * t) * fv;
        st2 += (w * t) * t;
    }
    S.jtj = C;
    S.jtr.assign(p, 0.0);
    double sff = 0.0, syy = 0.0, rho = 0.0;
    if (W > 0.0) {
        const double mf = sf / W, mt = st / W;
        sff = sf2 - W * mf * mf;
        syy = st2 - W * mt * mt;
        const double sfy = sft - W * mf * mt;
        if (sff > 0.0 && syy > 0.0) rho = sfy / std::sqrt(sff * syy);
        std::vector<double> A(p), B(p), gbar(p);
        for (int j = 0; j < p; ++j) {
            gbar[j] = sg[j] / W;
            A[j] = sgf[j] - W * gbar[j] * mf;
            B[j] = sgy[j] - W * gbar[j] * mt;
        }
        if (sff > 0.0 && syy > 0.0) {
            const double root = std::sqrt(sff * syy);
            for (int j = 0; j < p; ++j) S.jtr[j] = B[j] / root - rho * A[j] / sff;
            int k = 0;
            for (int a = 0; a < p; ++a)
                for (int b = 0; b <= a; ++b, ++k)
                    S.jtj[k] = ((C[k] - W * gbar[a] * gbar[b]) - A[a] * A[b] / sff) / sff;
        } else {
            int k = 0;
            for (int a = 0; a < p; ++a)
                for (int b = 0; b <= a; ++b, ++k) S.jtj[k] = (a == b) ? 1.0 : 0.0;
        }
    }
    S.sff = sff;
    S.syy = syy;
    S.rho = rho;
    S.score = metric == 3 ? 1.0 - rho : 1.0 - std::fabs(rho);
    if (!std::isfinite(S.score) || !std::isfinite(sff) || !std::isfinite(syy)) S.bad = true;
}

/* The objective at c (correlation.rs objective of the weighted moments); a
   non-finite row scores 1e30. */
double corr_score(const Flat &f, const Data &D, const std::vector<double> &c, int metric)
{
    std::vector<double> v(f.op.size());
    double W = 0.0, sf = 0.0, sf2 = 0.0, sft = 0.0, st = 0.0, st2 = 0.0;
    for (int i = 0; i < D.n; ++i) {
        const double w = D.w ? D.w[i] : 1.0;
        if (!(w > 0.0)) continue;
        eval_row(f, D.row(i), D.nv, c.data(), v);
        const double fv = v[f.root];
        const double t = D.y[i];
        if (!std::isfinite(fv) || std::fabs(fv) > 1e30 || !std::isfinite(t)) return 1e30;
        W += w;
        sf += w * fv;
        sf2 += w * fv * fv;
        sft += w * fv * t;
        st += w * t;
        st2 += w * t * t;
    }
    if (!(W > 0.0)) return 1e30;
    const double mf = sf / W, mt = st / W;
    const double sff = sf2 - W * mf * mf, syy = st2 - W * mt * mt, sfy = sft - W * mf * mt;
    if (!(sff > 0.0) || !(syy > 0.0)) return 1.0;
    const double rho = sfy / std::sqrt(sff * syy);
    return metric == 3 ? 1.0 - rho : 1.0 - std::fabs(rho);
}

/* One step: the assembled system, objective gradient -sigma * d rho, gauge
   rows pinned (eval.cu:2244-2262), the equilibrated solve, the clamp. */
bool corr_step(const CorrState &S, const std::vector<double> &c, int p, double lambda, int metric,
               int gs, int go, std::vector<double> &trial, bool &rank)
{
    std::vector<double> jtj = S.jtj, jtr = S.jtr;
    const double sigma = metric == 3 ? 1.0 : (S.rho >= 0.0 ? 1.0 : -1.0);
    for (int j = 0; j < p; ++j) jtr[j] = -sigma * jtr[j];
    for (int i = 0; i < p; ++i) {
        const bool gi = gs == i || go == i;
        for (int j = 0; j <= i; ++j) {
            const bool gj = gs == j || go == j;
            if (gi || gj) jtj[i * (i + 1) / 2 + j] = (i == j) ? 1.0 : 0.0;
        }
        if (gi) jtr[i] = 0.0;
    }
    std::vector<double> delta;
    const int rc = eq_solve(jtj, jtr, lambda, p, delta);
    if (rc == 1) return false;
    if (rc == 4) rank = true;
    trial.resize(p);
    for (int i = 0; i < p; ++i) trial[i] = std::fmin(1e6, std::fmax(-1e6, c[i] + delta[i]));
    return true;
}

/* The device correlation arm (eval.cu:2320-2448, each step on a freshly
   masked system): gauge renormalisation, LM with the damping schedule of
   the other arms, four undamped polish steps, status labels. */
double corr_fit(const Flat &f, const Data &D, std::vector<double> &c, int iters, int metric,
                int gs, int go, int &status)
{
    const int p = f.nconst;
    if (gs >= 0 && gs < p && std::isfinite(c[gs]) && c[gs] != 0.0) c[gs] = c[gs] > 0.0 ? 1.0 : -1.0;
    if (go >= 0 && go < p) c[go] = 0.0;
    status = 0;
    double best = 1.0;
    bool start = false, improved = false, rank = false;
    CorrState S;
    if (iters > 0) {
        corr_assemble(f, D, c, metric, S);
        start = !S.bad;
        if (S.bad) {
            status = 2;
        } else if (!(S.sff > 0.0) || !(S.syy > 0.0)) {
            best = 1.0;
        } else if (S.sff < 1e-24 * S.syy) {
            status = 1;
        } else {
            best = S.score;
            double lambda = 1e-3;
            bool assembled = true;
            for (int it = 0; it < iters && status == 0; ++it) {
                if (!assembled) {
                    corr_assemble(f, D, c, metric, S);
                    assembled = true;
                    if (S.bad) { status = 2; break; }
                    if (S.sff < 1e-24 * S.syy) { status = 1; break; }
                }
                std::vector<double> trial;
                if (!corr_step(S, c, p, lambda, metric, gs, go, trial, rank)) { status = 1; break; }
                const double t = corr_score(f, D, trial, metric);
                if (t < best) {
                    c = trial;
                    best = t;
                    lambda = std::fmax(lambda / 8.0, 1e-12);
                    assembled = false;
                    improved = true;
                    if (best <= 0.0) break;
                } else {
                    lambda *= 6.0;
                    if (lambda > 1e9) break;
                }
            }
            for (int it = 0; it < 4 && status == 0 && best > 0.0; ++it) {
                corr_assemble(f, D, c, metric, S);
                if (S.bad) { status = 2; break; }
                std::vector<double> trial;
                if (!corr_step(S, c, p, 0.0, metric, gs, go, trial, rank)) break;
                const double t = corr_score(f, D, trial, metric);
                if (t < best) {
                    c = trial;
                    best = t;
                    improved = true;
                } else {
                    break;
                }
            }
        }
    } else {
        best = corr_score(f, D, c, metric);
    }
    if (status == 1 || status == 2) status = !start ? 7 : (improved ? 5 : 6);
    else if (status == 0 && rank) status = 4;
    return best;
}

} // namespace (correlation)

extern "C" {

double or_corr_fit(const int *ops, const double *vals, int nn, const double *x, int n, int nv,
                   const double *y, const double *w, int iters, int metric, int gauge_scale,
                   int gauge_offset, double *out_c, int *status)
{
    Node t;
    *status = -1;
    if (!build_tree(ops, vals, nn, t)) return NaN;
    Flat f = flatten(t);
    std::vector<double> c;
    collect_consts(t, c);
    Data D{x, n, nv, y, w};
    const double best = corr_fit(f, D, c, iters, metric, gauge_scale, gauge_offset, *status);
    for (size_t i = 0; i < c.size(); ++i) out_c[i] = c[i];
    return best;
}

double or_corr_score(const int *ops, const double *vals, int nn, const double *x, int n, int nv,
                     const double *y, const double *w, int metric)
{
    Node t;
    if (!build_tree(ops, vals, nn, t)) return NaN;
    Flat f = flatten(t);
    std::vector<double> c;
    collect_consts(t, c);
    Data D{x, n, nv, y, w};
    return corr_score(f, D, c, metric);
}

/* ---- order, rank and maximum statistics of residual vectors ----------- */

/* metrics::median_abs (metrics.rs:1193-1205): the middle |r| of the sorted
   absolute errors, 0.5 * (a + b) of the two middle values for even n. */
double or_median_abs(const double *r, int n)
{
    if (n <= 0) return 0.0;
    std::vector<double> s((size_t)n);
    for (int i = 0; i < n; ++i) s[i] = std::fabs(r[i]);
    std::sort(s.begin(), s.end());
    return (n % 2 == 1) ? s[n / 2] : 0.5 * (s[n / 2 - 1] + s[n / 2]);
}

/* metrics::interquartile_mean_abs (metrics.rs:1209-1223): the mean, summed
   left to right, of the sorted |r| at ranks [n/4, 3n/4) (all rows below
   four). */
double or_iqr_mean_abs(const double *r, int n)
{
    if (n <= 0) return 0.0;
    std::vector<double> s((size_t)n);
    for (int i = 0; i < n; ++i) s[i] = std::fabs(r[i]);
    std::sort(s.begin(), s.end());
    int lo = 0, hi = n;
    if (n >= 4) {
        lo = n / 4;
        hi = std::max((3 * n) / 4, lo + 1);
    }
    double sum = 0.0;
    for (int i = lo; i < hi; ++i) sum += s[i];
    return sum / (double)(hi - lo);
}

/* metrics::centred_ranks_doubled (metrics.rs:1258-1286): A = 2 * average rank
   - (n + 1) = i + j - n for the tie run at sorted positions [i, j); ties are
   exact f64 equality (-0.0 == +0.0).  Returns -1 when a value is not
   finite. */
int or_centred_ranks_doubled(const double *v, int n, long long *out)
{
    for (int i = 0; i < n; ++i)
        if (!std::isfinite(v[i])) return -1;
    std::vector<int> idx((size_t)n);
    for (int i = 0; i < n; ++i) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return v[a] < v[b]; });
    int i = 0;
    while (i < n) {
        int j = i + 1;
        while (j < n && v[idx[j]] == v[idx[i]]) ++j;
        const long long a = (long long)i + (long long)j - (long long)n;
        for (int q = i; q < j; ++q) out[idx[q]] = a;
        i = j;
    }
    return 0;
}

/* RankSums::between (metrics.rs:1307-1323): exact sum(a*b), sum(a*a),
   sum(b*b) as two's-complement 128-bit (hi, lo) pairs in out[0..5]. */
void or_rank_sums(const long long *a, const long long *b, int n, unsigned long long *out)
{
    __int128 ab = 0, aa = 0, bb = 0;
    for (int i = 0; i < n; ++i) {
        ab += (__int128)a[i] * (__int128)b[i];
        aa += (__int128)a[i] * (__int128)a[i];
        bb += (__int128)b[i] * (__int128)b[i];
    }
    const __int128 v[3] = {ab, aa, bb};
    for (int k = 0; k < 3; ++k) {
        const unsigned __int128 u = (unsigned __int128)v[k];
        out[2 * k] = (unsigned long long)(u >> 64);
        out[2 * k + 1] = (unsigned long long)u;
    }
}

/* The exact maximum-error scan of one residual vector (the device stage
   eval.cu:4505-4600): max |r| and its lowest row; with a finite level t >= 0
   and the residuals' derivatives b along one constant, the intersection
   [max lo_i, min hi_i] of {delta : |r_i + b_i delta| <= t}, where a row with
   b_i == 0 only vetoes (|r_i| > t makes the level infeasible: lo = +inf,
   hi = -inf).  A non-finite residual or derivative is status 2 (vmax NaN,
   row -1).  b may be NULL (maximum only). */
void or_max_scan(const double *r, const double *b, int n, double level, double *vmax, int *row,
                 double *lo, double *hi, int *status)
{
    const bool want = b != nullptr && std::isfinite(level) && level >= 0.0;
    double best = -1.0;
    int brow = -1;
    double L = -INF, H = INF;
    bool bad = false, veto = false;
    for (int i = 0; i < n; ++i) {
        if (!std::isfinite(r[i]) || (want && !std::isfinite(b[i]))) {
            bad = true;
            break;
        }
        const double a = std::fabs(r[i]);
        if (a > best) {
            best = a;
            brow = i;
        }
        if (!want) continue;
        if (b[i] == 0.0) {
            if (a > level) veto = true;
        } else {
            const double e1 = (-level - r[i]) / b[i];
            const double e2 = (level - r[i]) / b[i];
            L = std::max(L, std::fmin(e1, e2));
            H = std::min(H, std::fmax(e1, e2));
        }
    }
    *status = bad ? 2 : 0;
    *vmax = bad ? NaN : (brow < 0 ? 0.0 : best);
    *row = bad ? -1 : (brow < 0 ? 0 : brow);
    *lo = veto ? INF : L;
    *hi = veto ? -INF : H;
}

} // extern "C"
