// flop_count.cpp -- MEASUREMENT INFRASTRUCTURE (not part of the product): counts the
// floating-point operations of the restated algorithm per interior cell-stage update.
//
// The CPU restatement (wenofv_oracle.c) performs exactly the arithmetic of the all-FP64 GPU
// path: both reproduce the reference solver bit for bit, operation by operation.  Compiling
// it here as C++ with every `double` replaced by a counting type therefore gives the
// algorithm's operation count as executed, without retyping any formula.  Convention:
// add, subtract, multiply, divide, square root, fabs and each comparison (the min/max of the
// WENO-Z normalisation, Riemann speeds and limiter, and the few branch tests of the equation
// of state and HLLC) count 1; pow(x, n) for integer n counts n - 1 multiplications; exp, log
// and tanh (THINC only) count 1 each; negation and NaN tests count 0.  The restatement is
// built with contraction off, so there are no fused multiply-adds to count as 2.
//
// Two counts per configuration:
//  - "literal": every operation the reference's code performs, including what it recomputes
//    (the volume fraction is solved anew inside every pressure and sound-speed call, the
//    squared sound speeds inside every equation-of-state call);
//  - "distinct": each value computed once, by value numbering: an operation counts only the
//    first time its operator and operands (identified by their bits at the leaves, by their
//    own numbers above; + and * commutative) occur.  This is what a kernel computing each
//    quantity once, as the GPU path does, must execute, and it is the count used as the
//    algorithm's FLOPs.
//
// Each configuration evaluates one right-hand side and one SSP-RK3 step with its CFL time
// step on four periodic grid shapes of a seeded random two-phase state (random, so that no
// two cells share operand bits), and the per-cell coefficient is separated from the
// per-row/column (boundary ring) and constant terms by solving
// count = a nx ny + b nx + c ny + d.  Output: FLOPs per cell-stage per configuration (the
// configurations named on the command line, or all).
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <cmath>
#include <vector>

namespace fc {
struct Ops {
    uint64_t n[2];     // literal, distinct
    uint64_t kind[8];  // distinct, by kind: add/sub, mul, div, sqrt, fabs, compare, transcendental
};
static Ops g;

// open-addressing set of value numbers
struct Seen {
    std::vector<uint64_t> slot;
    uint64_t mask = 0;
    size_t used = 0;
    void reset(int log2) {
        slot.assign((size_t)1 << log2, 0);
        mask = ((uint64_t)1 << log2) - 1;
        used = 0;
    }
    bool insert(uint64_t h) {  // true if new
        if (h == 0) h = 1;
        for (uint64_t i = h & mask;; i = (i + 1) & mask) {
            if (slot[i] == h) return false;
            if (slot[i] == 0) {
                slot[i] = h;
                if (++used * 2 > slot.size()) { fprintf(stderr, "value-number table full\n"); exit(3); }
                return true;
            }
        }
    }
};
static Seen seen;

constexpr uint64_t mix(uint64_t x) {
    x ^= x >> 33; x *= 0xff51afd7ed558ccdULL; x ^= x >> 33; x *= 0xc4ceb9fe1a85ec53ULL; x ^= x >> 33;
    return x;
}
constexpr uint64_t bits(double x) { return __builtin_bit_cast(uint64_t, x); }
constexpr uint64_t node(uint64_t op, uint64_t a, uint64_t b) { return mix(mix(op * 0x9e3779b97f4a7c15ULL ^ a) + b); }

enum Kind { K_ADD, K_MUL, K_DIV, K_SQRT, K_FABS, K_CMP, K_TRANS };
inline void count(uint64_t id, Kind k) {
    ++g.n[0];
    if (seen.insert(id)) { ++g.n[1]; ++g.kind[k]; }
}

struct CD {
    double v;
    uint64_t id;
    CD() = default;
    constexpr CD(double x) : v(x), id(mix(bits(x) ^ 0x5bd1e995ULL)) {}
    constexpr CD(double x, uint64_t i) : v(x), id(i) {}
    explicit operator int() const { return (int)v; }
    explicit operator bool() const { return v != 0.0; }
};
enum : uint64_t { ADD = 1, SUB, MUL, DIV, SQRT, FABS, LT, LE, EQ, EXP, LOG, TANH, POW };
inline Kind kind_of(uint64_t op) {
    switch (op) {
    case ADD: case SUB: return K_ADD;
    case MUL: return K_MUL;
    case DIV: return K_DIV;
    case SQRT: return K_SQRT;
    case FABS: return K_FABS;
    case LT: case LE: case EQ: return K_CMP;
    default: return K_TRANS;
    }
}
inline CD op2(uint64_t op, CD a, CD b, double r, bool commutative) {
    uint64_t x = a.id, y = b.id;
    if (commutative && x > y) { const uint64_t t = x; x = y; y = t; }
    const uint64_t id = node(op, x, y);
    count(id, kind_of(op));
    return CD(r, id);
}
inline CD op1(uint64_t op, CD a, double r) {
    const uint64_t id = node(op, a.id, 0);
    count(id, kind_of(op));
    return CD(r, id);
}
inline CD operator+(CD a, CD b) { return op2(ADD, a, b, a.v + b.v, true); }
inline CD operator-(CD a, CD b) { return op2(SUB, a, b, a.v - b.v, false); }
inline CD operator*(CD a, CD b) { return op2(MUL, a, b, a.v * b.v, true); }
inline CD operator/(CD a, CD b) { return op2(DIV, a, b, a.v / b.v, false); }
inline CD &operator+=(CD &a, CD b) { return a = a + b; }
inline CD &operator-=(CD &a, CD b) { return a = a - b; }
inline CD &operator*=(CD &a, CD b) { return a = a * b; }
inline CD &operator/=(CD &a, CD b) { return a = a / b; }
inline CD operator-(CD a) { return CD(-a.v, mix(a.id ^ 0x2545f4914f6cdd1dULL)); }
inline CD operator+(CD a) { return a; }
// a comparison counts once per distinct (ordered) operand pair; b > a is a < b
inline bool cmp(uint64_t op, CD a, CD b, bool r) { count(node(op, a.id, b.id), K_CMP); return r; }
inline bool operator<(CD a, CD b) { return cmp(LT, a, b, a.v < b.v); }
inline bool operator>(CD a, CD b) { return cmp(LT, b, a, a.v > b.v); }
inline bool operator<=(CD a, CD b) { return cmp(LE, a, b, a.v <= b.v); }
inline bool operator>=(CD a, CD b) { return cmp(LE, b, a, a.v >= b.v); }
inline bool operator==(CD a, CD b) { return cmp(EQ, a, b, a.v == b.v); }
inline bool operator!=(CD a, CD b) { return !(a == b); }
inline CD sqrt(CD a) { return op1(SQRT, a, std::sqrt(a.v)); }
inline CD fabs(CD a) { return op1(FABS, a, std::fabs(a.v)); }
inline CD exp(CD a) { return op1(EXP, a, std::exp(a.v)); }
inline CD log(CD a) { return op1(LOG, a, std::log(a.v)); }
inline CD tanh(CD a) { return op1(TANH, a, std::tanh(a.v)); }
inline bool isnan(CD a) { return std::isnan(a.v); }
// integer powers as repeated products (x^2 = x*x, x^3 = x^2*x, ...), others one operation
inline CD pow(CD a, CD n) {
    const double k = n.v;
    if (k == std::floor(k) && k >= 1 && k <= 8) {
        CD r = a;
        for (int i = 1; i < (int)k; ++i) r = r * a;
        return CD(std::pow(a.v, n.v), r.id);
    }
    return op2(POW, a, n, std::pow(a.v, n.v), false);
}
}  // namespace fc

using fc::CD;
using fc::sqrt; using fc::fabs; using fc::exp; using fc::log; using fc::tanh; using fc::isnan; using fc::pow;

#define double CD
#include "wenofv_oracle.c"
#undef double

namespace {
struct Cfg {
    const char *name;
    int scheme, recon, flux, bp;
};

// seeded random two-phase state around a smooth mean (no limiting, no supersonic faces)
void random_state(int nx, int ny, CD *U, uint64_t seed) {
    const size_t n = (size_t)nx * ny;
    uint64_t s = seed;
    auto uni = [&]() { s = fc::mix(s + 0x9e3779b97f4a7c15ULL); return (double)(s >> 11) / 9007199254740992.0; };
    for (size_t c = 0; c < n; ++c) {
        const double z1 = 0.3 + 0.4 * uni(), p = 1.0 + 0.05 * uni();
        const double u = 0.3 + 0.1 * uni(), v = -0.2 + 0.1 * uni();
        const double r1 = 1.0 + p / 100.0, r2 = 0.1 + p / 100.0;
        const double a = z1 * r1, b = (1 - z1) * r2;
        U[c] = a; U[n + c] = b; U[2 * n + c] = (a + b) * u; U[3 * n + c] = (a + b) * v;
    }
}

struct Count { double rhs[2], step[2], kind[7]; };

Count run(const Cfg &cf, int nx, int ny) {
    CD cx[256], cw[256], cy[256];
    fc::seen.reset(25);
    orc_geometry(nx, ny, 0.0, 1.0, 0.0, 1.0, 1, cx, cw, cy, nullptr, nullptr);
    orc_params p;
    memset(&p, 0, sizeof p);
    p.nx = nx; p.ny = ny; p.scheme = cf.scheme; p.recon = cf.recon; p.flux = cf.flux; p.bp = cf.bp;
    p.bc = 0; p.rk = 2; p.cfl = 0.4;
    p.dX = 1.0 / nx; p.dY = 1.0 / ny;
    p.rho10 = 1.0; p.rho20 = 0.1; p.sound1 = 10.0; p.sound2 = 10.0; p.preb = 0.0; p.gravity = 0.0;
    p.dxC = cx; p.dxW = cw; p.dyC = cy; p.ghost = nullptr;
    const size_t n = (size_t)4 * nx * ny;
    CD *U = (CD *)malloc(n * sizeof(CD)), *rhs = (CD *)malloc(n * sizeof(CD));
    random_state(nx, ny, U, 1000003ULL * nx + ny);
    Count r;
    fc::g = fc::Ops{};
    fc::seen.reset(25);
    orc_assemble_rhs(&p, U, rhs, nullptr, nullptr);
    for (int k = 0; k < 2; ++k) r.rhs[k] = (double)fc::g.n[k];
    fc::g = fc::Ops{};
    fc::seen.reset(25);
    CD dt;
    orc_compute_dt(&p, U, &dt);
    orc_rk_step(&p, U, dt);
    for (int k = 0; k < 2; ++k) r.step[k] = (double)fc::g.n[k];
    for (int k = 0; k < 7; ++k) r.kind[k] = (double)fc::g.kind[k];
    free(U);
    free(rhs);
    return r;
}

// solve count = a nx ny + b nx + c ny + d from the four shapes (N|2N) x (N|2N):
// a = (C(2N,2N) - C(2N,N) - C(N,2N) + C(N,N)) / N^2
constexpr int N = 12;
double per_cell(double n11, double n21, double n12, double n22) { return (n22 - n21 - n12 + n11) / double(N * N); }
}  // namespace

// known answers for the counting itself: literal and distinct counts of small expressions
int selftest() {
    int bad = 0;
    auto expect = [&](const char *what, uint64_t lit, uint64_t dis) {
        if (fc::g.n[0] != lit || fc::g.n[1] != dis) {
            printf("selftest %s: literal %llu distinct %llu, expected %llu %llu\n", what,
                   (unsigned long long)fc::g.n[0], (unsigned long long)fc::g.n[1], (unsigned long long)lit,
                   (unsigned long long)dis);
            ++bad;
        }
        fc::g = fc::Ops{};
        fc::seen.reset(10);
    };
    fc::seen.reset(10);
    const CD a = 1.25, b = -3.5, c = 7.0;
    CD r = (a + b) * (b + a);  // + is commutative: one add, one multiply
    expect("(a+b)*(b+a)", 3, 2);
    r = (a - b) * (b - a);  // - is not
    expect("(a-b)*(b-a)", 3, 3);
    r = pow(a, 3.0) + pow(a, 2.0);  // a*a shared by both powers
    expect("a^3+a^2", 4, 3);
    r = sqrt(fabs(c)) / sqrt(fabs(c));
    expect("sqrt|c|/sqrt|c|", 5, 3);
    const bool m = (a < b) || (b > a);  // b > a is a < b
    expect("a<b, b>a", 2, 1);
    r = -a * b + a * b;  // negation is free and makes a new value
    expect("-a*b+a*b", 3, 3);
    (void)r; (void)m;
    printf("selftest %s\n", bad ? "FAILED" : "ok");
    return bad ? 1 : 0;
}

// argv: configuration names to count (default: all), or --selftest
int main(int argc, char **argv) {
    if (argc > 1 && strcmp(argv[1], "--selftest") == 0) return selftest();
    const Cfg cfgs[] = {
        {"js_cons_llf", 0, 0, 0, 0},      // tanh_interface 128^2 options
        {"z_cons_hllc_bp", 1, 0, 1, 1},   // liquid-column workloads and batch256
        {"z_cons_llf", 1, 0, 0, 0},
        {"z_char_llf", 1, 1, 0, 0},       // shipped settings.cfg / smooth_1024 options
        {"z_char_hllc_bp", 1, 1, 1, 1},
        {"z_prim_hllc_bp", 1, 2, 1, 1},
        {"z_prim2_hllc_bp", 1, 3, 1, 1},
        {"z_prim3_hllc_bp", 1, 4, 1, 1},
        {"exp_cons_hllc_bp", 2, 0, 1, 1},
    };
    printf("# config distinct_flops_per_cell_stage literal_flops_per_cell_stage distinct_rhs_per_cell literal_rhs_per_cell"
           " | distinct per cell-stage: add mul div sqrt fabs cmp trans\n");
    for (const Cfg &c : cfgs) {
        bool want = argc < 2;
        for (int k = 1; k < argc; ++k) want = want || strcmp(argv[k], c.name) == 0;
        if (!want) continue;
        const Count a = run(c, N, N), b = run(c, 2 * N, N), d = run(c, N, 2 * N), e = run(c, 2 * N, 2 * N);
        double st[2], rh[2];
        for (int k = 0; k < 2; ++k) {
            st[k] = per_cell(a.step[k], b.step[k], d.step[k], e.step[k]) / 3.0;  // one RK3 step = 3 stages
            rh[k] = per_cell(a.rhs[k], b.rhs[k], d.rhs[k], e.rhs[k]);
        }
        printf("%s %.1f %.1f %.1f %.1f |", c.name, st[1], st[0], rh[1], rh[0]);
        for (int k = 0; k < 7; ++k) printf(" %.1f", per_cell(a.kind[k], b.kind[k], d.kind[k], e.kind[k]) / 3.0);
        printf("\n");
    }
    return 0;
}
