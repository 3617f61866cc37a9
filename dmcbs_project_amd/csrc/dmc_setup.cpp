// Host-side set-up of one DMC system: box, units, k shells, beta(r) table, Jastrow and
// potential tables from the cosh two-body equation, initial walkers and RNG jump tables.
// These run once per system on the CPU; the tables are then uploaded to the device.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>

#include "dmc_internal.h"

namespace dmc {
namespace {

constexpr uint64_t kMask48 = (1ULL << 48) - 1ULL;
constexpr uint64_t kM1 = 44485709377909ULL, kA1 = 96309754297ULL;  // ran.f90:15,18
constexpr uint64_t kM2 = 34522712143931ULL, kA2 = 55789347517ULL;  // ran.f90:36,39

inline double lcg(uint64_t &s, uint64_t m, uint64_t a) {
    s = (s * m + a) & kMask48;
    return double(s | 1ULL) * 0x1p-48;
}

inline int isq(const std::array<int, 3> &v) { return v[0] * v[0] + v[1] * v[1] + v[2] * v[2]; }

}  // namespace

// Half-space integer vectors grouped in shells of equal |k|^2, in the order of
// kstates3d (kshell.f90:20-156): enumeration over the positive octant with the sign
// variants that complete a half space, then an exchange sort on |k|^2.

int enumerate_kshells(int nsh0, std::vector<std::array<int, 3>> &ik, std::vector<int> &numsh,
                      std::string &err) {
    const int i2max = 2 * nsh0;
    const int imax = int(std::lround(std::sqrt(float(i2max))));
    std::vector<std::array<int, 3>> v;
    for (int ix = 0; ix <= imax; ++ix)
        for (int iy = 0; iy <= imax; ++iy)
            for (int iz = 0; iz <= imax; ++iz) {
                if (ix * ix + iy * iy + iz * iz > i2max) continue;
                v.push_back({ix, iy, iz});
                if (ix == 0 && iy != 0 && iz != 0) v.push_back({ix, iy, -iz});
                if (iy == 0 && ix != 0 && iz != 0) v.push_back({ix, iy, -iz});
                if (iz == 0 && ix != 0 && iy != 0) v.push_back({ix, -iy, iz});
                if (ix != 0 && iy != 0 && iz != 0) {
                    v.push_back({ix, iy, -iz});
                    v.push_back({ix, -iy, iz});
                    v.push_back({ix, -iy, -iz});
                }
            }
    const int nk = int(v.size());
    for (int i = 0; i < nk - 1; ++i)
        for (int j = i + 1; j < nk; ++j)
            if (isq(v[i]) >= isq(v[j])) std::swap(v[i], v[j]);
    std::vector<int> count(nk + 2, 0);
    int last = 0, nsh = 1;
    for (int i = 0; i < nk; ++i) {
        if (isq(v[i]) > last) { ++nsh; last = isq(v[i]); }
        ++count[nsh];
    }
    if (nsh0 >= nsh) {
        err = "Too many states: the shell count exceeds the enumerated shells";
        return DMC_ERR_SETUP;
    }
    numsh.assign(nsh0, 0);
    ik.clear();
    int k = 0;
    for (int sh = 0; sh < nsh0; ++sh) {
        numsh[sh] = count[sh + 1];
        const int m = std::max(1, count[sh + 1]);
        for (int j = 0; j < m; ++j, ++k) ik.push_back(v[k]);
    }
    return DMC_OK;
}

// setupenk (momentum3d.f90:13-79): enumeration over the octant up to int(sqrt(ntab)) + 1 per
// axis, then the reference's exchange pass (each later vector against every earlier one)
// into |m|^2 order, then shells of equal |m|^2
void enk_shells(int ntab, std::vector<std::array<int, 3>> &ik, std::vector<int> &first,
                std::vector<int> &count) {
    const int nnk = int(std::sqrt(double(ntab))) + 1;
    ik.clear();
    for (int ix = 0; ix <= nnk; ++ix)
        for (int iy = 0; iy <= nnk; ++iy)
            for (int iz = 0; iz <= nnk; ++iz)
                if (ix * ix + iy * iy + iz * iz <= ntab) ik.push_back({ix, iy, iz});
    const int ntot = int(ik.size());
    for (int i = 1; i < ntot; ++i)
        for (int j = 0; j < i; ++j)
            if (isq(ik[i]) < isq(ik[j])) std::swap(ik[i], ik[j]);
    first.assign(1, 0);
    count.assign(1, 1);
    for (int i = 1; i < ntot; ++i) {
        if (isq(ik[i]) != isq(ik[i - 1])) {
            first.push_back(i);
            count.push_back(0);
        }
        ++count.back();
    }
}

namespace {

int kshells(System &s, std::string &err) {
    int st = enumerate_kshells(s.p.norb, s.ik, s.numsh, err);
    if (st) return st;
    s.kshell.clear();
    for (int sh = 0; sh < s.p.norb; ++sh)
        for (int j = 0; j < std::max(1, s.numsh[sh]); ++j) s.kshell.push_back(sh);
    s.nk = int(s.ik.size());
    if (s.nk - 1 > kMaxQ) {
        err = "too many plane waves for the device kernels";
        return DMC_ERR_ARG;
    }
    return DMC_OK;
}

// value with its derivatives with respect to (b, d): forward-mode differentiation of the
// beta table (its d/db and d/dd tables feed the parameter adjoint)
struct Dual {
    double v = 0, b = 0, d = 0;
    Dual() = default;
    Dual(double x) : v(x) {}  // NOLINT: constants promote
    Dual(double x, double xb, double xd) : v(x), b(xb), d(xd) {}
};
inline Dual operator+(Dual x, Dual y) { return {x.v + y.v, x.b + y.b, x.d + y.d}; }
inline Dual operator-(Dual x, Dual y) { return {x.v - y.v, x.b - y.b, x.d - y.d}; }
inline Dual operator-(Dual x) { return {-x.v, -x.b, -x.d}; }
inline Dual operator*(Dual x, Dual y) {
    return {x.v * y.v, x.b * y.v + x.v * y.b, x.d * y.v + x.v * y.d};
}
inline Dual operator/(Dual x, Dual y) {
    const double q = x.v / y.v;
    return {q, (x.b - q * y.b) / y.v, (x.d - q * y.d) / y.v};
}
inline Dual &operator+=(Dual &x, Dual y) { return x = x + y; }
inline Dual &operator-=(Dual &x, Dual y) { return x = x - y; }
inline Dual &operator*=(Dual &x, Dual y) { return x = x * y; }
inline Dual &operator/=(Dual &x, Dual y) { return x = x / y; }
inline Dual exp(Dual x) {
    const double e = std::exp(x.v);
    return {e, e * x.b, e * x.d};
}
inline double val(double x) { return x; }
inline double val(const Dual &x) { return x.v; }

// beta0 and beta (betafun.f90:127-171) for the current (b, d, c); T = double or Dual
template <class T>
struct Beta {
    T b, d, c, c0, c1, c2, c3, c4;
    double small, el;
    void at0(double r, T &f, T &fp, T &fpp) const {
        using std::exp;
        const T x = b * r;
        if (r < small) {
            f = c0 + x * (c1 + x * (c2 + x * (c3 + x * c4)));
            fp = c1 + x * (2.0 * c2 + x * (3.0 * c3 + 4.0 * c4 * x));
            fpp = 2.0 * c2 + x * (6.0 * c3 + 12.0 * c4 * x);
        } else {
            const T ex = exp(-x), ed = exp(-(d * x));
            f = (1.0 + c * x) * (1.0 - ed) * ex / x;
            fp = ex * (-1.0 - x * (1.0 + c * x) + ed * (1.0 + x * (1.0 + d) * (1.0 + c * x))) / (x * x);
            fpp = ex * (2.0 + x * (2.0 + x * (1.0 + c * x)) -
                        ed * (2.0 + x * (1.0 + d) * (2.0 + x * (1.0 + d) * (1.0 + x * c)))) /
                  (x * x * x);
        }
        fp *= b;
        fpp *= b * b;
        f /= d;
        fp /= d;
        fpp /= d;
    }
    void at(double r, T &f, T &fp, T &fpp) const {
        if (r > 0.5 * el) { f = fp = fpp = 0.0; return; }
        T h0, t1, t2, f1, f1p, f1pp, f2, f2p, f2pp;
        at0(0.5 * el, h0, t1, t2);
        at0(r, f1, f1p, f1pp);
        at0(el - r, f2, f2p, f2pp);
        f = f1 + f2 - 2.0 * h0;
        fp = r < 1.0e-14 ? 0.0 : f1p - f2p;
        fpp = f1pp + f2pp;
    }
};

double vcosh(const System &s, double v0, double r) {
    const double ch = std::cosh(s.p.amu * r);
    return -2.0 * v0 * s.p.amu * s.p.amu * s.tab_hbar / (ch * ch);
}

// RK4 integration of -hb u'' + (V - E) u = 0 from u(0) = 0, u'(0) = 1 (rungekutta.f90:127)
int integrate(const System &s, double v0, double e, double hb, double dr, int nt,
              std::vector<double> &u, std::vector<double> &up) {
    u[0] = 0.0;
    up[0] = 1.0;
    int ncross = 0;
    for (int i = 1; i <= nt; ++i) {
        const double r = (i - 1) * dr;
        const double va = vcosh(s, v0, r), vh = vcosh(s, v0, r + 0.5 * dr), vb = vcosh(s, v0, r + dr);
        const double k1a = -(e - va) * u[i - 1] / hb, k1b = up[i - 1];
        const double k2a = -(e - vh) * (u[i - 1] + 0.5 * dr * k1b) / hb, k2b = up[i - 1] + 0.5 * dr * k1a;
        const double k3a = -(e - vh) * (u[i - 1] + 0.5 * dr * k2b) / hb, k3b = up[i - 1] + 0.5 * dr * k2a;
        const double k4a = -(e - vb) * (u[i - 1] + dr * k3b) / hb, k4b = up[i - 1] + dr * k3a;
        up[i] = up[i - 1] + (k1a + 2.0 * k2a + 2.0 * k3a + k4a) * dr / 6.0;
        u[i] = u[i - 1] + (k1b + 2.0 * k2b + 2.0 * k3b + k4b) * dr / 6.0;
        if (u[i] != std::copysign(u[i], u[i - 1])) ++ncross;
    }
    return ncross;
}

// Zero-node solution with u'(R)/u(R) = 1/R at R = nt*dr, by bisection on E
// (rungekutta.f90:8-100); returns f = u/r, f', f'' on [0, nt].
int two_body(const System &s, double v0, int nt, double dr, std::vector<double> &f,
             std::vector<double> &fp, std::vector<double> &fpp, std::string &err) {
    const double hb = s.tab_hbar;
    double emin = vcosh(s, v0, 0.0);
    for (int i = 1; i <= nt; ++i) emin = std::min(emin, vcosh(s, v0, i * dr));
    double emax = 0.001;
    while (integrate(s, v0, emax, hb, dr, nt, f, fp) <= 0) emax *= 2.0;
    double e = 0.5 * (emin + emax);
    for (int iter = 0;; ++iter) {
        integrate(s, v0, e, hb, dr, nt, f, fp);
        const double rr = dr * nt;
        const double der = ((fp[nt] - f[nt] / rr) / rr) / f[nt];
        if (der < 0.0) emax = e; else emin = e;
        if (std::fabs(der) < 1.0e-8) break;
        e = 0.5 * (emin + emax);
        if (iter + 1 > 100) {
            err = "orbital: no convergence";
            return DMC_ERR_SETUP;
        }
    }
    f[0] = fp[0];
    fpp[0] = (vcosh(s, v0, 0.0) - e) * fp[0] / (3.0 * hb);
    fp[0] = 0.0;
    for (int i = 1; i <= nt; ++i) {
        const double r = i * dr;
        fp[i] = (fp[i] - f[i] / r) / r;
        f[i] = f[i] / r;
        fpp[i] = (vcosh(s, v0, r) - e) * f[i] / hb - 2.0 * fp[i] / r;
    }
    return DMC_OK;
}

// Jastrow (cjas * ln f, its derivative and (ln f)'') and potential tables
// (coshpot.f90:71-133), opposite-spin channel.
int jastrow_tables(System &s, std::string &err) {
    const int nt = s.ntab;
    s.range = 0.5 * s.el;
    const double dr = s.range / nt;
    s.scale = 1.0 / dr;
    s.vt.assign(nt + 1, 0.0);
    for (int i = 0; i <= nt; ++i) s.vt[i] = vcosh(s, s.p.v0, i * dr);
    s.ntab0 = int(nt * s.p.djas / (0.5 * s.el));
    s.ut.assign(nt + 1, 0.0);
    s.dut.assign(nt + 1, 0.0);
    s.d2ut.assign(nt + 1, 0.0);
    const double cj = s.p.cjas1;
    const double v0op = s.p.v0op != 0.0 ? s.p.v0op : s.p.v0;
    if (cj != 0.0) {
        const int n0 = s.ntab0;
        std::vector<double> f(n0 + 1), fp(n0 + 1), fpp(n0 + 1);
        int st = two_body(s, v0op, n0, dr, f, fp, fpp, err);
        if (st) return st;
        for (int i = 0; i <= std::min(n0, nt); ++i) {
            s.ut[i] = f[i] / f[n0];
            s.dut[i] = fp[i] / f[n0];
            s.d2ut[i] = fpp[i] / f[n0];
        }
    }
    for (int i = s.ntab0 + 1; i <= nt; ++i) { s.ut[i] = 1.0; s.dut[i] = 0.0; s.d2ut[i] = 0.0; }
    for (int i = 0; i <= nt; ++i) {
        if (cj != 0.0) {
            s.dut[i] = s.dut[i] / s.ut[i];
            s.d2ut[i] = s.d2ut[i] / s.ut[i] - s.dut[i] * s.dut[i];
            s.ut[i] = std::log(s.ut[i]);
        }
        s.ut[i] *= cj;
        s.dut[i] *= cj;
        s.d2ut[i] *= cj;
    }
    return DMC_OK;
}

}  // namespace

namespace {

// beta table (initbeta, betafun.f90:74-125) for parameters (b, d) of type T
template <class T>
void beta_table(const System &s, T b, T d, std::vector<T> &bt, std::vector<T> &dbt,
                std::vector<T> &d2bt) {
    using std::exp;
    const double el = s.el;
    const int nt = s.ntab;
    Beta<T> B{};
    B.b = b;
    B.d = d;
    B.el = el;
    const T e1 = exp(b * el * (1.0 + d)), e2 = exp(d * b * el);
    B.c = 0.5 * (2.0 + 2.0 * d * b * el + (d * b * el) * (d * b * el) * e1 +
                 2.0 * d * b * b * el * el * e1 + 2.0 * b * el - 2.0 * e2 * b * el - 2.0 * e2) /
          (el * el * (e2 - 1.0 - d + d * e1) * b * b);
    const T c = B.c;
    B.c0 = d;
    B.small = 1.0e-3 / val(b);
    B.c1 = -d + c * d - d * d / 2.0;
    B.c2 = d / 2.0 + d * d / 2.0 + d * d * d / 6.0 + c * (-d - d * d / 2.0);
    B.c3 = -d / 6.0 - d * d / 4.0 - d * d * d / 6.0 - d * d * d * d / 24.0 +
           c * (d / 2.0 + d * d / 2.0 + d * d * d / 6.0);
    B.c4 = d / 24.0 + d * d / 12.0 + d * d * d / 12.0 + d * d * d * d / 24.0 +
           d * d * d * d * d / 120.0 - c * (d / 6.0 + d * d / 4.0 + d * d * d / 6.0 + d * d * d * d / 24.0);
    // number of periodic images needed (betafun.f90:94-99)
    int nj = 0;
    for (double bn = 10.0; std::fabs(bn) > 1.0e-10;) {
        ++nj;
        T v, t1, t2;
        B.at(el * nj, v, t1, t2);
        bn = val(v);
    }
    bt.assign(nt + 1, T(0.0));
    dbt.assign(nt + 1, T(0.0));
    d2bt.assign(nt + 1, T(0.0));
    const double dr = 0.5 * el / nt;
    for (int i = 0; i <= nt; ++i) {
        const double r = i * dr;
        T f, fp, fpp;
        B.at(r, f, fp, fpp);
        for (int j = 1; j <= nj; ++j) {
            T a0, a1, a2, c0_, c1_, c2_;
            B.at(j * el - r, a0, a1, a2);
            B.at(j * el + r, c0_, c1_, c2_);
            f += a0 + c0_;
            fp += -a1 + c1_;
            fpp += a2 + c2_;
        }
        bt[i] = f;
        dbt[i] = fp;
        d2bt[i] = fpp;
    }
    const T top = bt[nt];
    const double fac = std::fabs(s.p.bfact);
    for (int i = 0; i <= nt; ++i) {
        bt[i] -= top;
        // store the radial laplacian f'' + 2 f'/r (3 f'' at the origin)
        d2bt[i] = i ? d2bt[i] + 2.0 * dbt[i] / (i * dr) : 3.0 * d2bt[i];
        bt[i] *= fac;
        dbt[i] *= fac;
        d2bt[i] *= fac;
    }
}

}  // namespace

void build_beta(System &s, double b, double d) {
    s.b = b;
    s.d = d;
    beta_table<double>(s, b, d, s.bt, s.dbt, s.d2bt);
    // derivative tables with respect to b and d
    std::vector<Dual> t0, t1, t2;
    beta_table<Dual>(s, Dual(b, 1.0, 0.0), Dual(d, 0.0, 1.0), t0, t1, t2);
    const int n1 = s.ntab + 1;
    s.bt_b.resize(n1);
    s.dbt_b.resize(n1);
    s.d2bt_b.resize(n1);
    s.bt_d.resize(n1);
    s.dbt_d.resize(n1);
    s.d2bt_d.resize(n1);
    for (int i = 0; i < n1; ++i) {
        s.bt_b[i] = t0[i].b;
        s.dbt_b[i] = t1[i].b;
        s.d2bt_b[i] = t2[i].b;
        s.bt_d[i] = t0[i].d;
        s.dbt_d[i] = t1[i].d;
        s.d2bt_d[i] = t2[i].d;
    }
}

int build_system(const dmc_params &p, System &s, std::string &err) {
    s = System{};
    s.p = p;
    if (p.npart < 2 || p.npart % 2 || p.npart / 2 > 64) {
        err = "npart must be even with npart/2 <= 64";
        return DMC_ERR_ARG;
    }
    if (p.norb < 1 || p.norb > DMC_MAXORB || p.ntab < 8 || p.rho <= 0 || p.mass <= 0) {
        err = "invalid norb/ntab/rho/mass";
        return DMC_ERR_ARG;
    }
    if (p.cjas2 != 0.0) {
        err = "equal-spin Jastrow (cjas(2) != 0) is not supported on the device path";
        return DMC_ERR_ARG;
    }
    s.npart = p.npart;
    s.n = p.npart / 2;
    s.ntab = p.ntab;
    s.el = std::pow(p.npart / p.rho, 1.0 / 3.0);
    s.eli = 1.0 / s.el;
    s.mass = p.mass;
    s.sqrtm = std::sqrt(p.mass);
    s.hbar = p.hbar;
    const double mu = p.mass * p.mass / (p.mass + p.mass);
    s.tab_hbar = 2.0 * (p.hbar / (2.0 * mu));
    const double pi = 4.0 * std::atan(1.0);
    if (p.units) {
        const double rho = p.npart / (s.el * s.el * s.el);
        const double kf = std::pow(3.0 * pi * pi * rho, 1.0 / 3.0);
        s.eunit = 3.0 / 5.0 * p.hbar / (2.0 * mu) * kf * kf;
    }
    s.etrial = p.etrial * s.eunit * p.npart;
    s.nfac = 1.0 / s.eunit / p.npart;
    s.sigma = std::sqrt(2.0 * p.hbar * p.dt) / std::sqrt(p.mass);
    s.sig2 = s.sigma * s.sigma;
    s.cutdrift = p.dcut >= 0.0;
    s.dcut_abs = s.cutdrift ? s.sigma * p.dcut : 0.0;
    s.vorb.assign(p.vorb, p.vorb + p.norb);
    int st = kshells(s, err);
    if (st) return st;
    s.nobeta = p.bfact == 0.0;
    s.rangeb = 0.5 * s.el;
    s.scaleb = p.ntab / s.rangeb;
    if (p.bfact < 0.0) {
        err = "beta table from file (bfact < 0) is not supported; use dmc_set_tables";
        return DMC_ERR_ARG;
    }
    if (!s.nobeta) {
        build_beta(s, p.bbeta, p.dbeta);
    } else {
        s.bt.assign(p.ntab + 1, 0.0);
        s.dbt = s.d2bt = s.bt_b = s.dbt_b = s.d2bt_b = s.bt_d = s.dbt_d = s.d2bt_d = s.bt;
    }
    return jastrow_tables(s, err);
}

void initial_walkers(const System &s, int64_t seed, int nwalk, std::vector<double> &x,
                     std::vector<int64_t> &irn) {
    const int np = s.npart;
    uint64_t stream = uint64_t(seed) & kMask48;  // setrn
    uint64_t own = uint64_t(seed);
    std::vector<double> cfg(3 * np);
    for (int i = 0; i < 3 * np; ++i) cfg[i] = s.el * (lcg(stream, kM1, kA1) - 0.5);
    x.resize(size_t(nwalk) * 3 * np);
    irn.resize(nwalk);
    for (int k = 0; k < nwalk; ++k) {
        for (int i = 0; i < np; ++i) {
            const double u = lcg(stream, kM1, kA1);
            const int j = int(np * u + 1.0) - 1;
            for (int d = 0; d < 3; ++d) std::swap(cfg[3 * i + d], cfg[3 * j + d]);
        }
        std::memcpy(&x[size_t(k) * 3 * np], cfg.data(), sizeof(double) * 3 * np);
        own = (own * kM2 + kA2) & kMask48;
        irn[k] = int64_t(own);
    }
}

static void jump_table(int n, uint64_t m, uint64_t add, std::vector<uint64_t> &a,
                       std::vector<uint64_t> &c) {
    a.resize(n);
    c.resize(n);
    uint64_t am = 1, cm = 0;  // identity
    for (int t = 0; t < n; ++t) {
        am = (am * m) & kMask48;
        cm = (cm * m + add) & kMask48;
        a[t] = am;  // after t+1 draws
        c[t] = cm;
    }
}
void ran1_jump_table(int n, std::vector<uint64_t> &a, std::vector<uint64_t> &c) {
    jump_table(n, kM1, kA1, a, c);
}
void ran2_jump_table(int n, std::vector<uint64_t> &a, std::vector<uint64_t> &c) {
    jump_table(n, kM2, kA2, a, c);
}

}  // namespace dmc
