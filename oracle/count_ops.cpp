// count_ops.cpp -- arithmetic operation count of the reference's Eulerian step.
//
// Test infrastructure (like the rest of oracle/): compiles the oracle's C
// restatement, which follows the reference operation for operation
// (GKMflux.cpp:11-229, update.cpp:214-680), as C++ with `double` replaced by a
// counting number type, and counts +, -, *, / and each sqrt / exp / erfc / pow
// as one operation (comparisons are listed, not counted).
//
// Prints two figures per cell-update:
//   * algorithmic: one WENO reconstruction per interface, i.e. per cell
//     3 x (WENO + collision time + GKS flux + interface slopes) + the update;
//   * as written: a whole oracle step on a 16^3 cube divided by 16^3 (the
//     reference evaluates five WENO reconstructions per interface).
//
//   make -C oracle count && oracle/count_ops
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <omp.h>

namespace {
struct Ops {
    long add = 0, mul = 0, div = 0, sqrt = 0, exp = 0, erfc = 0, pow = 0, cmp = 0;
    long total() const { return add + mul + div + sqrt + exp + erfc + pow; }
};
Ops g_ops;

struct Num {
    double v;
    Num() : v(0) {}
    Num(double x) : v(x) {}
    explicit operator int() const { return (int)v; }
    explicit operator long() const { return (long)v; }
};
inline Num operator+(Num a, Num b) { g_ops.add++; return a.v + b.v; }
inline Num operator-(Num a, Num b) { g_ops.add++; return a.v - b.v; }
inline Num operator*(Num a, Num b) { g_ops.mul++; return a.v * b.v; }
inline Num operator/(Num a, Num b) { g_ops.div++; return a.v / b.v; }
inline Num operator-(Num a) { return -a.v; }
inline Num& operator+=(Num& a, Num b) { return a = a + b; }
inline Num& operator-=(Num& a, Num b) { return a = a - b; }
inline Num& operator*=(Num& a, Num b) { return a = a * b; }
inline Num& operator/=(Num& a, Num b) { return a = a / b; }
inline bool operator<(Num a, Num b) { g_ops.cmp++; return a.v < b.v; }
inline bool operator>(Num a, Num b) { g_ops.cmp++; return a.v > b.v; }
inline bool operator<=(Num a, Num b) { g_ops.cmp++; return a.v <= b.v; }
inline bool operator>=(Num a, Num b) { g_ops.cmp++; return a.v >= b.v; }
inline bool operator==(Num a, Num b) { return a.v == b.v; }
inline bool operator!=(Num a, Num b) { return a.v != b.v; }
inline Num sqrt(Num a) { g_ops.sqrt++; return std::sqrt(a.v); }
inline Num exp(Num a) { g_ops.exp++; return std::exp(a.v); }
inline Num erfc(Num a) { g_ops.erfc++; return std::erfc(a.v); }
inline Num pow(Num a, Num b) { g_ops.pow++; return std::pow(a.v, b.v); }
inline Num fabs(Num a) { return std::fabs(a.v); }
inline Num fmax(Num a, Num b) { g_ops.cmp++; return std::fmax(a.v, b.v); }
inline Num fmin(Num a, Num b) { g_ops.cmp++; return std::fmin(a.v, b.v); }
inline Num floor(Num a) { return std::floor(a.v); }
inline Num fmod(Num a, Num b) { return std::fmod(a.v, b.v); }
inline Num modf(Num a, Num* ip)
{
    double i;
    const double f = std::modf(a.v, &i);
    *ip = i;
    return f;
}
}  // namespace

#define double Num
#include "gkm_oracle.c"
#undef double

static void print(const char* what, const Ops& o)
{
    std::printf("%-22s add %6ld mul %6ld div %5ld sqrt %3ld exp %3ld erfc %3ld pow %2ld (cmp %3ld) total %6ld\n", what,
                o.add, o.mul, o.div, o.sqrt, o.exp, o.erfc, o.pow, o.cmp, o.total());
}

int main()
{
    orc_gas g;
    g.nc = 16;
    g.K = 2;
    g.gam = 1.4;
    g.R = 287.0;
    g.prn = 0.7;
    g.mu0 = 0.5;
    g.T0 = 300.0;
    g.dx = 2 * M_PI / 16;
    g.eps = 1e-10;

    // one interface: 6 cells of a smooth state (the count does not depend on the values)
    Num S[30];
    for (int s = 0; s < 6; s++) {
        const double rho = 1 + 0.01 * s, u = 30 + s, v = -20 + 2 * s, w = 10 - s, p = 86100 * (1 + 0.01 * s);
        S[5 * s + 0] = rho;
        S[5 * s + 1] = rho * u;
        S[5 * s + 2] = rho * v;
        S[5 * s + 3] = rho * w;
        S[5 * s + 4] = p / 0.4 + 0.5 * rho * (u * u + v * v + w * w);
    }
    g_ops = Ops();
    Num Wl[5], Wr[5];
    orc_weno(S, &g, Wl, Wr);
    const Ops weno = g_ops;
    g_ops = Ops();
    const Num tau = orc_tau(S + 10, &g);
    const Ops ctau = g_ops;
    Num D[6][5];
    for (int a = 0; a < 6; a++)
        for (int q = 0; q < 5; q++) D[a][q] = Num(0.01 * (a + 1) * (q + 1));
    g_ops = Ops();
    Num F[5];
    orc_flux(S + 10, S + 15, Wl, Wr, D[0], D[1], D[2], D[3], D[4], D[5], tau, Num(2e-4), Num(g.dx), &g, F);
    const Ops flux = g_ops;
    print("WENO (both sides)", weno);
    print("collision time", ctau);
    print("GKS flux", flux);

    // interface slopes (update.cpp:258-265): 2 normal 2(Wl-W)/dx, 4 tangential (W+ - W-)/(2dx),
    // 3 operations per value; update (update.cpp:505-516): 3 flux differences, 2 sums, 1 scale, 1 subtract
    const long slopes = 30 * 3, update = 5 * 7;
    const long per_face = weno.total() + ctau.total() + flux.total() + slopes;
    const long algorithmic = 3 * per_face + update;
    std::printf("per interface (one WENO) %ld; algorithmic per cell-update %ld\n", per_face, algorithmic);

    // the oracle's whole step on 16^3, as written
    const int nc = 16, nt = nc + 6;
    const long Nt = (long)nt * nt * nt;
    Num* W = new Num[5 * Nt];
    for (long i = 0; i < Nt; i++) {
        const int z = i % nt, x = (i / nt) % nt, y = i / ((long)nt * nt);
        const double u = 40 * std::sin(2 * M_PI * x / nc), v = 30 * std::cos(2 * M_PI * z / nc), w = 20 * std::sin(2 * M_PI * y / nc);
        const double rho = 1.0, p = 86100.0;
        W[0 * Nt + i] = rho;
        W[1 * Nt + i] = rho * u;
        W[2 * Nt + i] = rho * v;
        W[3 * Nt + i] = rho * w;
        W[4 * Nt + i] = p / 0.4 + 0.5 * rho * (u * u + v * v + w * w);
    }
    g_ops = Ops();
    orc_eulerian_step(nc, W, Num(1e-4), &g, 1);
    const Ops step = g_ops;
    delete[] W;
    print("step 16^3 (as written)", step);
    std::printf("as written per cell-update %.0f\n", (double)step.total() / (nc * nc * nc));
    return 0;
}
