// CPU FP64 oracle for the covariant-tensor partial-wave likelihood path.
//
// TEST INFRASTRUCTURE ONLY.  This file restates the reference algorithm in plain
// C++17 + OpenMP so that tests (and the CPU baseline leg of the benchmark) can check
// the HIP product path against it.  Nothing in the product imports, links or runs it.
//
// Every function cites the reference file:line it follows.  It deliberately keeps the
// reference's formulation (angles via acos/atan2, Wigner D by the explicit sum,
// generic tensor contraction, lineshapes evaluated per event) rather than the
// factorised form the GPU path uses, so that the two are independent computations of
// the same quantities.
//
// Parity pins (see tests/): CG against sympy, Y_lm against scipy, Wigner d against the
// matrix exponential of J_y, lineshapes against closed forms, the massive spin transform
// against the Thomas-Wigner rotation composed from explicit Lorentz boosts, the massless
// one against its m -> 0 limit, the node tensor against the canonical amplitude in a
// moving frame, gradient and Hessian against central finite differences of this file's
// own NLL, and likelihood invariances (global phase, single-amplitude independence of v).

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef std::complex<double> cd;

// ===========================================================================
// Special functions (src/Amplitude.cu:5-230)
// ===========================================================================

// factorial(n) with the reference's semantics (src/Amplitude.cu:5-91): exact for
// integers and half-integers up to 11.5, Stirling with two correction terms for
// integers >= 12, and the Gamma-function branch (which evaluates Gamma(n)) otherwise.
static double ref_factorial(double n) {
    if (n < 0.0 || std::isnan(n)) return 0.0;
    double ri = std::round(n);
    bool is_int = std::fabs(n - ri) < 1e-10;
    double hr = std::round(2.0 * n) / 2.0;
    bool is_half = !is_int && std::fabs(n - hr) < 1e-10;
    if (n <= 11.5) {
        if (is_int) {
            double f = 1.0;
            for (int i = 2; i <= (int)ri; ++i) f *= i;
            return f;
        }
        if (is_half) {
            // Gamma(k + 3/2) = (2k+1)!! / 2^(k+1) * sqrt(pi), n = k + 1/2
            int k = (int)std::round(n - 0.5);
            double df = 1.0;
            for (int i = 1; i <= 2 * k + 1; i += 2) df *= i;
            return df / std::pow(2.0, k + 1) * std::sqrt(M_PI);
        }
    }
    if (is_int && (int)ri >= 12) {
        int m = (int)ri;
        double st = std::sqrt(2.0 * M_PI * m) * std::pow(m / M_E, m);
        return st * (1.0 + 1.0 / (12.0 * m) + 1.0 / (288.0 * m * m));
    }
    return std::tgamma(n);
}

// Clebsch-Gordan <j1 m1; j2 m2 | J M> by the Racah sum (src/Amplitude.cu:93-138)
static double ref_cg(int dim_j1, int m1_x2, int dim_j2, int m2_x2, int dim_J, int M_x2) {
    double j1 = (dim_j1 - 1) / 2.0, j2 = (dim_j2 - 1) / 2.0, J = (dim_J - 1) / 2.0;
    double m1 = m1_x2 / 2.0, m2 = m2_x2 / 2.0, M = M_x2 / 2.0;
    if (std::fabs(M - (m1 + m2)) > 1e-10) return 0.0;
    if (J < std::fabs(j1 - j2) || J > j1 + j2) return 0.0;
    if (std::fabs(m1) > j1 || std::fabs(m2) > j2 || std::fabs(M) > J) return 0.0;
    double t1 = std::sqrt((2.0 * J + 1.0) * ref_factorial(j1 + j2 - J) * ref_factorial(j1 - j2 + J) *
                          ref_factorial(-j1 + j2 + J) / ref_factorial(j1 + j2 + J + 1.0));
    double t2 = std::sqrt(ref_factorial(J + M) * ref_factorial(J - M) * ref_factorial(j1 - m1) *
                          ref_factorial(j1 + m1) * ref_factorial(j2 - m2) * ref_factorial(j2 + m2));
    int kmin = (int)std::ceil(std::max(0.0, std::max(j2 - J - m1, j1 - J + m2)));
    int kmax = (int)std::floor(std::min(j1 + j2 - J, std::min(j1 - m1, j2 + m2)));
    double sum = 0.0;
    for (int k = kmin; k <= kmax; ++k) {
        double den = ref_factorial(k) * ref_factorial(j1 + j2 - J - k) * ref_factorial(j1 - m1 - k) *
                     ref_factorial(j2 + m2 - k) * ref_factorial(J - j2 + m1 + k) *
                     ref_factorial(J - j1 - m2 + k);
        if (std::fabs(den) > 1e-15) sum += ((k % 2 == 0) ? 1.0 : -1.0) / den;
    }
    return t1 * t2 * sum;
}

// associated Legendre P_l^m(x), m >= 0, Condon-Shortley phase (src/Amplitude.cu:140-190)
static double ref_legendre(int l, int m, double x) {
    if (l < 0 || m < 0 || m > l) return 0.0;
    if (m == 0) {
        if (l == 0) return 1.0;
        if (l == 1) return x;
        if (l == 2) return 0.5 * (3.0 * x * x - 1.0);
        if (l == 3) return 0.5 * (5.0 * x * x * x - 3.0 * x);
        double a = 1.0, b = x, c = 0.0;
        for (int n = 2; n <= l; ++n) {
            c = ((2.0 * n - 1.0) * x * b - (n - 1.0) * a) / n;
            a = b;
            b = c;
        }
        return c;
    }
    double pmm = 1.0;
    for (int i = 1; i <= m; ++i) pmm *= (2 * i - 1) * std::sqrt(1.0 - x * x);
    pmm *= std::pow(-1.0, m);
    if (l == m) return pmm;
    double pm1 = x * (2.0 * m + 1.0) * pmm;
    if (l == m + 1) return pm1;
    double a = pmm, b = pm1, c = 0.0;
    for (int n = m + 2; n <= l; ++n) {
        c = ((2.0 * n - 1.0) * x * b - (n + m - 1.0) * a) / (n - m);
        a = b;
        b = c;
    }
    return c;
}

// Y_l^m(theta, phi) (src/Amplitude.cu:192-230)
static cd ref_ylm(int l, int m, double theta, double phi) {
    if (l < 0 || std::abs(m) > l) return 0.0;
    if (l == 0 && m == 0) return cd(0.5 * std::sqrt(1.0 / M_PI), 0.0);
    double x = std::cos(theta);
    int am = std::abs(m);
    double P = ref_legendre(l, am, x);
    double norm = std::sqrt((2.0 * l + 1.0) * ref_factorial(l - am) / (4.0 * M_PI * ref_factorial(l + am)));
    double phase = (m < 0 && (am % 2) != 0) ? -1.0 : 1.0;
    double a = phase * norm * P;
    return cd(a * std::cos(m * phi), a * std::sin(m * phi));
}

// reference small-d sum, DFunc(dim, m1, m2, beta) (src/Amplitude.cu:272-318)
static cd ref_dfunc(int dim, int m1_x2, int m2_x2, double beta) {
    double j = (dim - 1) / 2.0, m1 = m1_x2 / 2.0, m2 = m2_x2 / 2.0;
    if (std::fabs(m1) > j || std::fabs(m2) > j) return 0.0;
    int kmin = (m1 < m2) ? 0 : (int)(m1 - m2);
    int kmax = (m1 > -m2) ? (int)(j - m2) : (int)(j + m1);
    double sum = 0.0;
    double ch = std::cos(beta / 2.0), sh = std::sin(beta / 2.0);
    for (int k = kmin; k <= kmax; ++k) {
        double num = ref_factorial(j + m2) * ref_factorial(j - m2) * ref_factorial(j + m1) * ref_factorial(j - m1);
        double den = ref_factorial(k) * ref_factorial(j - m2 - k) * ref_factorial(j + m1 - k) *
                     ref_factorial(k - m1 + m2);
        double comb = std::sqrt(num) / den;
        double ct = std::pow(ch, 2 * j + m1 - m2 - 2 * k);
        double st = std::pow(sh, 2 * k - m1 + m2);
        double ph = std::pow(-1.0, k - m1 + m2);
        sum += ph * comb * ct * st;
    }
    return cd(sum, 0.0);
}

static cd ref_wignerD(int dim, int m1_x2, int m2_x2, double a, double b, double g) {
    cd d = ref_dfunc(dim, m1_x2, m2_x2, b);
    double m1 = m1_x2 / 2.0, m2 = m2_x2 / 2.0;
    return std::exp(cd(0.0, m1 * a)) * d * std::exp(cd(0.0, m2 * g));
}

// ===========================================================================
// Four-vectors and the two-body node tensor (include/Amplitude.cuh:9-119,
// src/Amplitude.cu:232-644)
// ===========================================================================
struct LV {
    double E = 0, x = 0, y = 0, z = 0;
    LV() {}
    LV(double e, double a, double b, double c) : E(e), x(a), y(b), z(c) {}
    double P() const { return std::sqrt(x * x + y * y + z * z); }
    double M() const {
        double m2 = E * E - x * x - y * y - z * z;
        return m2 < 0 ? 0.0 : std::sqrt(m2);
    }
    double Dot(const LV& o) const { return E * o.E - (x * o.x + y * o.y + z * o.z); }
    LV operator+(const LV& o) const { return LV(E + o.E, x + o.x, y + o.y, z + o.z); }
};

struct Angles { double theta, phi, qs; };

// pair-rest-frame direction of daughter 1 (src/Amplitude.cu:237-257)
static Angles cm_angles(const LV& p1, const LV& p2) {
    double m1 = p1.M(), m2 = p2.M();
    double p1p2 = p1.Dot(p2);
    double m = std::sqrt(m1 * m1 + m2 * m2 + 2 * p1p2);
    double lamb = (2 * m * p1.E + m * m + m1 * m1 - m2 * m2) / (2 * m * (m + p1.E + p2.E));
    double qs = std::sqrt(std::pow(m, 4) + std::pow(m1 * m1 - m2 * m2, 2) - 2 * m * m * (m1 * m1 + m2 * m2)) / (2 * m);
    double nx = (p1.x - lamb * (p1.x + p2.x)) / qs;
    double ny = (p1.y - lamb * (p1.y + p2.y)) / qs;
    double nz = (p1.z - lamb * (p1.z + p2.z)) / qs;
    nz = std::max(-1.0, std::min(1.0, nz));
    return Angles{std::acos(nz), std::atan2(ny, nx), qs};
}

// MassiveTrans (src/Amplitude.cu:331-404)
static void massive_trans(cd* T, const LV& p, const LV& q, int dim) {
    for (int i = 0; i < dim; i++)
        for (int j = 0; j < dim; j++) T[i * dim + j] = (i == j) ? 1.0 : 0.0;
    double Pp = p.P(), Pq = q.P();
    double hx = p.y * q.z - p.z * q.y, hy = p.z * q.x - p.x * q.z, hz = p.x * q.y - p.y * q.x;
    if (std::fabs(hx) < 1e-10 && std::fabs(hy) < 1e-10 && std::fabs(hz) < 1e-10) return;
    double m1 = p.M(), g1 = p.E / m1, m2 = q.M(), g2 = q.E / m2;
    double ca = (p.x * q.x + p.y * q.y + p.z * q.z) / (-1.0 * Pp * Pq);
    double sa = std::sqrt(1 - ca * ca);
    double nrm = std::sqrt(hx * hx + hy * hy + hz * hz);
    double th = std::acos(hz / nrm), ph = std::atan2(hy / nrm, hx / nrm);
    double num = sa * std::sqrt((g1 - 1.0) * (g2 - 1.0));
    double den = std::sqrt((g1 + 1.0) * (g2 + 1.0)) + ca * std::sqrt((g1 - 1.0) * (g2 - 1.0));
    double psi = 2.0 * std::atan2(num, den);
    std::vector<cd> w1(dim * dim), w2(dim * dim);
    for (int i = 0; i < dim; i++)
        for (int j = 0; j < dim; j++) {
            int a = dim - 1 - 2 * i, b = dim - 1 - 2 * j;
            w1[i * dim + j] = ref_wignerD(dim, a, b, -ph, -th, -psi);
            w2[i * dim + j] = ref_wignerD(dim, a, b, 0.0, th, ph);
        }
    for (int i = 0; i < dim; i++)
        for (int j = 0; j < dim; j++) {
            cd s = 0.0;
            for (int k = 0; k < dim; k++) s += w1[i * dim + k] * w2[k * dim + j];
            T[i * dim + j] = s;
        }
}

// MasslessTrans (src/Amplitude.cu:409-470)
static void massless_trans(cd* T, const LV& P, const LV& p, int dim) {
    for (int i = 0; i < dim; i++)
        for (int j = 0; j < dim; j++) T[i * dim + j] = (i == j) ? 1.0 : 0.0;
    double Pm = P.P();
    if (Pm < 1e-10) return;
    double pm = p.P();
    if (pm < 1e-10) return;
    double n1x = -P.x / Pm, n1y = -P.y / Pm, n1z = -P.z / Pm;
    double n2x = p.x / pm, n2y = p.y / pm, n2z = p.z / pm;
    double b1 = Pm / P.E, g1 = 1.0 / std::sqrt(1.0 - b1 * b1);
    double sh = std::sqrt((g1 - 1.0) / 2.0), ch = std::sqrt((g1 + 1.0) / 2.0);
    double d12 = n1x * n2x + n1y * n2y + n1z * n2z;
    double cz = n1x * n2y - n1y * n2x;
    double x12 = sh * (d12 - n1z) + ch * (1.0 - n2z), y12 = sh * cz;
    double x22 = sh * (d12 + n1z) + ch * (1.0 + n2z), y22 = -y12;
    double tot = -(std::atan2(y12, x12) + std::atan2(y22, x22));
    for (int i = 0; i < dim; i++) {
        double m2 = (dim - 1 - 2 * i) / 2.0;
        T[i * dim + i] = cd(std::cos(m2 * tot), std::sin(m2 * tot));
    }
}

// multiplyDFuncLeft (src/Amplitude.cu:474-499)
static void dfunc_left(cd* mat, int dim, double a, double b, double g) {
    std::vector<cd> D(dim * dim), tmp(dim * dim);
    for (int i = 0; i < dim; i++)
        for (int j = 0; j < dim; j++) D[i * dim + j] = ref_wignerD(dim, dim - 1 - 2 * i, dim - 1 - 2 * j, a, b, g);
    for (int i = 0; i < dim; i++)
        for (int j = 0; j < dim; j++) {
            cd s = 0.0;
            for (int k = 0; k < dim; k++) s += D[i * dim + k] * mat[k * dim + j];
            tmp[i * dim + j] = s;
        }
    for (int i = 0; i < dim * dim; i++) mat[i] = tmp[i];
}

// pwa_amp: node tensor amp[sigma][sigma1][sigma2] (src/Amplitude.cu:500-644)
static void pwa_amp(cd* amp, const LV& p1, int dj1, const LV& p2, int dj2, int dj, int dS, int L) {
    int kron = dj1 * dj2, total = dj * kron;
    std::vector<cd> a0(total);
    Angles an = cm_angles(p1, p2);
    for (int si = 0; si < dj; si++) {
        int s = dj - 1 - 2 * si;
        for (int s1i = 0; s1i < dj1; s1i++) {
            int s1 = dj1 - 1 - 2 * s1i;
            for (int s2i = 0; s2i < dj2; s2i++) {
                int s2 = dj2 - 1 - 2 * s2i;
                cd Y = ref_ylm(L, (s - s1 - s2) / 2, an.theta, -1.0 * an.phi);
                double cg1 = ref_cg(dj1, s1, dj2, s2, dS, s1 + s2);
                double cg2 = ref_cg(dS, s1 + s2, 2 * L + 1, s - s1 - s2, dj, s);
                a0[si * kron + s1i * dj2 + s2i] = (std::pow(an.qs, L) / std::sqrt((double)dj)) * cg1 * cg2 * Y;
            }
        }
    }
    LV Pt = p1 + p2;
    std::vector<cd> T1(dj1 * dj1), T2(dj2 * dj2);
    double m1 = p1.M(), m2 = p2.M();
    if (m1 > 1e-10) massive_trans(T1.data(), Pt, p1, dj1);
    else {
        massless_trans(T1.data(), Pt, p1, dj1);
        dfunc_left(T1.data(), dj1, -an.phi, -an.theta, 0.0);
    }
    if (m2 > 1e-10) massive_trans(T2.data(), Pt, p2, dj2);
    else {
        massless_trans(T2.data(), Pt, p2, dj2);
        dfunc_left(T2.data(), dj2, -(M_PI + an.phi), -(M_PI - an.theta), 0.0);
    }
    // transposed application of the daughter transforms (src/Amplitude.cu:594-635)
    std::vector<cd> tmp(kron);
    for (int si = 0; si < dj; si++) {
        int base = si * kron;
        for (int n1 = 0; n1 < dj1; n1++)
            for (int o2 = 0; o2 < dj2; o2++) {
                cd s = 0.0;
                for (int o1 = 0; o1 < dj1; o1++) s += T1[n1 + dj1 * o1] * a0[base + o1 * dj2 + o2];
                tmp[n1 * dj2 + o2] = s;
            }
        for (int n1 = 0; n1 < dj1; n1++)
            for (int n2 = 0; n2 < dj2; n2++) {
                cd s = 0.0;
                for (int o2 = 0; o2 < dj2; o2++) s += tmp[n1 * dj2 + o2] * T2[n2 + dj2 * o2];
                amp[base + n1 * dj2 + n2] = s * 10.0;
            }
    }
}

// ===========================================================================
// Forward-mode derivatives (value, gradient, optional Hessian) over a block's
// local resonance parameters.
// ===========================================================================
static const int MAXV = 8;
struct Jet {
    double v = 0;
    double g[MAXV];
    double h[MAXV][MAXV];
    Jet() { std::memset(g, 0, sizeof g); std::memset(h, 0, sizeof h); }
    explicit Jet(double c) : v(c) { std::memset(g, 0, sizeof g); std::memset(h, 0, sizeof h); }
};
static int g_nv = 0;                 // active variables (thread-local below)
#pragma omp threadprivate(g_nv)

static Jet jvar(double val, int i) { Jet r(val); r.g[i] = 1.0; return r; }
static Jet operator+(const Jet& a, const Jet& b) {
    Jet r(a.v + b.v);
    for (int i = 0; i < g_nv; i++) { r.g[i] = a.g[i] + b.g[i]; for (int k = 0; k < g_nv; k++) r.h[i][k] = a.h[i][k] + b.h[i][k]; }
    return r;
}
static Jet operator-(const Jet& a, const Jet& b) {
    Jet r(a.v - b.v);
    for (int i = 0; i < g_nv; i++) { r.g[i] = a.g[i] - b.g[i]; for (int k = 0; k < g_nv; k++) r.h[i][k] = a.h[i][k] - b.h[i][k]; }
    return r;
}
static Jet operator-(const Jet& a) { return Jet(0.0) - a; }
static Jet operator*(const Jet& a, const Jet& b) {
    Jet r(a.v * b.v);
    for (int i = 0; i < g_nv; i++) r.g[i] = a.g[i] * b.v + a.v * b.g[i];
    for (int i = 0; i < g_nv; i++)
        for (int k = 0; k < g_nv; k++)
            r.h[i][k] = a.h[i][k] * b.v + a.v * b.h[i][k] + a.g[i] * b.g[k] + a.g[k] * b.g[i];
    return r;
}
// unary function with f, f', f''
static Jet jfun(const Jet& a, double f, double d1, double d2) {
    Jet r(f);
    for (int i = 0; i < g_nv; i++) r.g[i] = d1 * a.g[i];
    for (int i = 0; i < g_nv; i++)
        for (int k = 0; k < g_nv; k++) r.h[i][k] = d1 * a.h[i][k] + d2 * a.g[i] * a.g[k];
    return r;
}
static Jet inv(const Jet& a) { double x = a.v; return jfun(a, 1.0 / x, -1.0 / (x * x), 2.0 / (x * x * x)); }
static Jet operator/(const Jet& a, const Jet& b) { return a * inv(b); }
static Jet jsqrt(const Jet& a) { double s = std::sqrt(a.v); return jfun(a, s, 0.5 / s, -0.25 / (s * a.v)); }
static Jet jlog(const Jet& a) { return jfun(a, std::log(a.v), 1.0 / a.v, -1.0 / (a.v * a.v)); }
static Jet jpowi(const Jet& a, int n) {
    if (n == 0) return Jet(1.0);
    double x = a.v;
    return jfun(a, std::pow(x, n), n * std::pow(x, n - 1), n * (n - 1.0) * std::pow(x, n - 2));
}
static Jet operator*(double c, const Jet& a) { return Jet(c) * a; }
static Jet operator+(double c, const Jet& a) { return Jet(c) + a; }
static Jet operator-(double c, const Jet& a) { return Jet(c) - a; }
static Jet operator+(const Jet& a, double c) { return a + Jet(c); }
static Jet operator-(const Jet& a, double c) { return a - Jet(c); }
static Jet operator*(const Jet& a, double c) { return a * Jet(c); }
static Jet operator/(const Jet& a, double c) { return a * Jet(1.0 / c); }

struct CJ { Jet re, im; };
static CJ cmul(const CJ& a, const CJ& b) { return CJ{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
static CJ cscale(const CJ& a, const Jet& s) { return CJ{a.re * s, a.im * s}; }
static CJ cinv_xy(const Jet& x, const Jet& y) {          // 1 / (x - i y) = (x + i y)/(x^2+y^2)
    Jet s = x * x + y * y;
    Jet is = inv(s);
    return CJ{x * is, y * is};
}

// ---------------------------------------------------------------------------
// Lineshapes (include/ResModel.cuh:88-451, src/ResModel.cu:484-561)
// ---------------------------------------------------------------------------
static double breakup_q(double m, double m1, double m2) {          // src/AmpGen.cu:2733-2737
    double q2 = (m * m - (m1 + m2) * (m1 + m2)) * (m * m - (m1 - m2) * (m1 - m2));
    if (q2 <= 0.0) return 0.0;
    return std::sqrt(q2) / (2.0 * m);
}
static Jet breakup_jet(const Jet& m, const Jet& m1, const Jet& m2) {
    Jet s = m1 + m2, d = m1 - m2;
    Jet q2 = (m * m - s * s) * (m * m - d * d);
    if (q2.v <= 0.0) return Jet(0.0);
    return jsqrt(q2) / (2.0 * m);
}
static Jet q0ad_jet(const Jet& m0, const Jet& m1, const Jet& m2) {   // computeQ0AD, ResModel.cuh:88-97
    Jet s = m1 + m2, d = m1 - m2;
    Jet m0s = m0 * m0;
    Jet q2 = (m0s - s * s) * (m0s - d * d) / (4.0 * m0s);
    if (q2.v < 0.0) return Jet(0.0);
    return jsqrt(q2);
}
static Jet bf_poly(int L, const Jet& z) {
    Jet z2 = z * z;
    switch (L) {
    case 0: return Jet(1.0);
    case 1: return 1.0 + z2;
    case 2: return 9.0 + 3.0 * z2 + z2 * z2;
    case 3: return jpowi(z, 6) + 6.0 * jpowi(z, 4) + 45.0 * z2 + 225.0;
    case 4: return jpowi(z, 8) + 10.0 * jpowi(z, 6) + 135.0 * jpowi(z, 4) + 1575.0 * z2 + 11025.0;
    case 5: return jpowi(z, 10) + 15.0 * jpowi(z, 8) + 315.0 * jpowi(z, 6) + 6300.0 * jpowi(z, 4) + 99225.0 * z2 + 893025.0;
    case 6: return jpowi(z, 12) + 21.0 * jpowi(z, 10) + 630.0 * jpowi(z, 8) + 17325.0 * jpowi(z, 6) + 363825.0 * jpowi(z, 4) + 6185025.0 * z2 + 540326025.0;
    default: return Jet(1.0);
    }
}
// Bf(L, q, q0, d) (include/ResModel.cuh:184-211)
static Jet bf_jet(int L, const Jet& q, const Jet& q0, double d) {
    if (L == 0 || L > 6) return Jet(1.0);
    return jsqrt(bf_poly(L, q0 * Jet(d)) / bf_poly(L, q * Jet(d)));
}
// BWR (include/ResModel.cuh:213-223)
static CJ bwr_jet(const Jet& m, const Jet& m0, const Jet& g0, int L, const Jet& q, const Jet& q0, double d) {
    Jet bf = bf_jet(L, q, q0, d);
    Jet gam = g0 * jpowi(q / q0, 2 * L + 1) * (m0 / m) * (bf * bf);
    Jet x = m0 * m0 - m * m, y = m0 * gam;
    return cinv_xy(x, y);
}
static CJ bw_jet(const Jet& m, const Jet& m0, const Jet& g0) {      // ResModel.cuh:397-405
    return cinv_xy(m0 * m0 - m * m, m0 * g0);
}
// Flatte (include/ResModel.cuh:418-451)
static CJ flatte_jet(double m, const Jet& m0, int nch, const Jet* g, const double* ch) {
    double s = m * m;
    Jet ir(0.0), ii(0.0);
    for (int i = 0; i < nch; ++i) {
        double ma = ch[2 * i], mb = ch[2 * i + 1];
        double f1sq = 1.0 - (ma + mb) * (ma + mb) / s;
        double f1r = f1sq >= 0 ? std::sqrt(f1sq) : 0.0, f1i = f1sq >= 0 ? 0.0 : std::sqrt(-f1sq);
        double f2sq = 1.0 - (ma - mb) * (ma - mb) / s;
        double f2 = f2sq > 0 ? std::sqrt(f2sq) : 0.0;
        Jet gw = (m0 / Jet(2.0)) * g[i];
        ir = ir + gw * Jet(f1r * f2);
        ii = ii + gw * Jet(f1i * f2);
    }
    Jet dre = m0 * m0 - Jet(s) + ii;
    Jet dim = Jet(0.0) - ir;
    Jet dsq = dre * dre + dim * dim;
    return CJ{dre / dsq, (Jet(0.0) - dim) / dsq};
}
// Gounaris-Sakurai (src/ResModel.cu:484-561)
static CJ gs_jet(double m, const Jet& m0, const Jet& g0, double q, const Jet& q0, double mpi) {
    const double pi = 3.14159265358979323846;
    double hm = (2.0 / pi) * (q / m) * std::log((m + 2.0 * q) / (2.0 * mpi));
    Jet q02 = q0 * q0, q03 = q02 * q0;
    Jet logA = jlog((m0 + 2.0 * q0) / Jet(2.0 * mpi));
    Jet dval = (Jet(3.0 / pi) * (Jet(mpi * mpi) / q02)) * logA + m0 / (Jet(2.0 * pi) * q0)
               - Jet(mpi * mpi) * m0 / (Jet(pi) * q03);
    Jet h0 = Jet(2.0 / pi) * (q0 / m0) * logA;
    Jet dh0 = h0 * (inv(Jet(8.0) * q02) - inv(Jet(2.0) * m0 * m0)) + inv(Jet(2.0 * pi) * m0 * m0);
    Jet qr = Jet(q) / q0;
    Jet gam = g0 * (qr * qr * qr) * (m0 / Jet(m));
    Jet f = g0 * (m0 * m0 / q03) * (Jet(q * q) * (Jet(hm) - h0) + (m0 * m0 - Jet(m * m)) * q02 * dh0);
    Jet num = 1.0 + dval * g0 / m0;
    Jet dre = m0 * m0 - Jet(m * m) + f;
    Jet dimv = Jet(0.0) - m0 * gam;
    Jet dsq = dre * dre + dimv * dimv;
    return CJ{num * dre / dsq, Jet(0.0) - num * dimv / dsq};
}

// ===========================================================================
// Plan (the model description written by ctpwa_amd.model.Plan.to_text)
// ===========================================================================
struct ONode { int dj, dj1, dj2, d1ref, d2ref, has_bf, lmin; double bf_d, mm, md1, md2; };
struct OTopo { std::vector<ONode> nodes; std::vector<std::vector<std::pair<int, int>>> slc; };
struct ORes { int node, model; std::vector<double> params; std::vector<int> slots; std::vector<double> ch;
              std::vector<std::string> expr; std::vector<double> tab; };
// c0: first SL combination of the topology used by the block (runs of one L)
struct OBlock { int topo, site, nsl, c0 = 0; std::vector<ORes> res; };
struct OPlan {
    int nfinal = 0, npol = 0;
    std::vector<int> polmap;
    std::vector<OTopo> topos;
    std::vector<OBlock> blocks;
    int ntheta = 0, namp = 0, nfree = 0;
    std::vector<double> th_init;
    std::vector<int> amap;
    std::vector<double> aratio;
    std::vector<int> g_slot;
    std::vector<double> g_mu, g_sig;
};

static OPlan parse_plan(const char* text) {
    OPlan P;
    std::istringstream in(text);
    std::string line;
    std::getline(in, line);
    if (line.rfind("CTPWA_PLAN", 0) != 0) throw std::runtime_error("not a plan");
    while (std::getline(in, line)) {
        std::istringstream ls(line);
        std::string key;
        ls >> key;
        if (key == "nfinal") ls >> P.nfinal;
        else if (key == "npol") { ls >> P.npol; P.polmap.resize(P.npol); for (auto& x : P.polmap) ls >> x; }
        else if (key == "topo") { int t, nn, nc; ls >> t >> nn >> nc; P.topos.push_back(OTopo()); }
        else if (key == "node") {
            ONode n; int k;
            ls >> k >> n.dj >> n.dj1 >> n.dj2 >> n.d1ref >> n.d2ref >> n.has_bf >> n.bf_d >> n.lmin >> n.mm >> n.md1 >> n.md2;
            P.topos.back().nodes.push_back(n);
        } else if (key == "slc") {
            std::vector<std::pair<int, int>> c;
            int S, L;
            while (ls >> S >> L) c.push_back({S, L});
            P.topos.back().slc.push_back(c);
        } else if (key == "block") {
            OBlock b; int idx, nres;
            ls >> idx >> b.topo >> b.site >> b.nsl >> nres;
            P.blocks.push_back(b);
        } else if (key == "res") {
            ORes r; int np;
            ls >> r.node >> r.model >> np;
            r.params.resize(np); for (auto& x : r.params) ls >> x;
            r.slots.resize(np); for (auto& x : r.slots) ls >> x;
            int nch; ls >> nch; r.ch.resize(2 * nch); for (auto& x : r.ch) ls >> x;
            P.blocks.back().res.push_back(r);
        } else if (key == "cexpr") {
            int n; ls >> n;
            std::vector<std::string> toks(n);
            for (auto& t : toks) ls >> t;
            P.blocks.back().res.back().expr = toks;
        } else if (key == "rtab") {
            long long n; ls >> n;
            std::vector<double> t((size_t)n);
            for (auto& x : t) ls >> x;
            P.blocks.back().res.back().tab = t;
        } else if (key == "bslc0") {
            for (auto& B : P.blocks) ls >> B.c0;
        } else if (key == "theta") { int j; double a, lo, hi; ls >> j >> a >> lo >> hi; P.th_init.push_back(a); }
        else if (key == "ntheta") ls >> P.ntheta;
        else if (key == "namp") ls >> P.namp >> P.nfree;
        else if (key == "amap") { P.amap.resize(P.namp); for (auto& x : P.amap) ls >> x; }
        else if (key == "aratio") { P.aratio.resize(P.namp); for (auto& x : P.aratio) ls >> x; }
        else if (key == "gauss") { int s; double mu, sg; ls >> s >> mu >> sg; P.g_slot.push_back(s); P.g_mu.push_back(mu); P.g_sig.push_back(sg); }
    }
    return P;
}

// ===========================================================================
// Oracle state
// ===========================================================================
struct Sample {
    int n = 0;
    std::vector<double> mom;                 // [n][nfinal][4]
    std::vector<double> w;                   // per-event weight (empty = 1)
    // per topology: slamp[slc][e][pol]
    std::vector<std::vector<cd>> slamp;
    // per topology: per node kinematics [e][node][KS] (mm, qq, md1, md2, then |p|, E,
    // cos(theta), phi of daughter 1 and of daughter 2 in the event frame)
    std::vector<std::vector<double>> kin;
};

struct Oracle {
    OPlan P;
    Sample S[3];                             // 0 = phsp, 1 = data, 2 = bkg
    double Wd = 0, Wb = 0;
    // per block: local variable -> global theta slot
    std::vector<std::vector<int>> bvars;
};

static const int KS = 12;   // kinematic values per node

// momenta of a topology's particles: finals, root (sum of finals in data order),
// intermediates (d1 + d2 of their node) (src/AmpGen.cu:531-653, src/Analysis.cu:148-160)
static LV final_lv(const double* m, int i) { return LV(m[4 * i], m[4 * i + 1], m[4 * i + 2], m[4 * i + 3]); }
static LV node_mother(const OTopo& T, const double* m, int nfinal, int k, std::vector<int>& done, std::vector<LV>& cache);
static LV ref_lv(const OTopo& T, const double* m, int nfinal, int ref, std::vector<int>& done, std::vector<LV>& cache) {
    if (ref >= 0) return final_lv(m, ref);
    return node_mother(T, m, nfinal, -ref - 1, done, cache);
}
static LV node_mother(const OTopo& T, const double* m, int nfinal, int k, std::vector<int>& done, std::vector<LV>& cache) {
    if (done[k]) return cache[k];
    LV r;
    if (k == 0) {
        for (int i = 0; i < nfinal; ++i) r = r + final_lv(m, i);
    } else {
        LV a = ref_lv(T, m, nfinal, T.nodes[k].d1ref, done, cache);
        LV b = ref_lv(T, m, nfinal, T.nodes[k].d2ref, done, cache);
        r = a + b;
    }
    done[k] = 1; cache[k] = r;
    return r;
}

// SL amplitudes of one event for one topology (src/AmpGen.cu:924-1219)
static void event_slamp(const OPlan& P, const OTopo& T, const double* m, cd* out /*[nslc][npol]*/, double* kin) {
    int nn = (int)T.nodes.size();
    std::vector<int> done(nn, 0);
    std::vector<LV> cache(nn);
    std::vector<LV> pd1(nn), pd2(nn), pm(nn);
    for (int k = 0; k < nn; ++k) {
        pm[k] = node_mother(T, m, P.nfinal, k, done, cache);
        pd1[k] = ref_lv(T, m, P.nfinal, T.nodes[k].d1ref, done, cache);
        pd2[k] = ref_lv(T, m, P.nfinal, T.nodes[k].d2ref, done, cache);
        double mm = pm[k].M(), m1 = pd1[k].M(), m2 = pd2[k].M();
        double* kk = kin + KS * k;
        kk[0] = mm;
        kk[1] = breakup_q(mm, m1, m2);
        kk[2] = m1;
        kk[3] = m2;
        // daughter momenta for Custom expressions: |p|, E, p_z/|p| (0 at rest),
        // atan2(p_y, p_x) (src/AmpGen.cu:1707-1712)
        const LV* dd[2] = {&pd1[k], &pd2[k]};
        for (int j = 0; j < 2; ++j) {
            double P = dd[j]->P();
            kk[4 + 4 * j] = P;
            kk[5 + 4 * j] = dd[j]->E;
            kk[6 + 4 * j] = (P > 0) ? dd[j]->z / P : 0.0;
            kk[7 + 4 * j] = std::atan2(dd[j]->y, dd[j]->x);
        }
    }
    for (size_t c = 0; c < T.slc.size(); ++c) {
        std::vector<cd> cur;
        std::vector<int> shape, labels;
        for (int k = 0; k < nn; ++k) {
            const ONode& n = T.nodes[k];
            std::vector<cd> amp(n.dj * n.dj1 * n.dj2);
            pwa_amp(amp.data(), pd1[k], n.dj1, pd2[k], n.dj2, n.dj, T.slc[c][k].first, T.slc[c][k].second);
            int lab1 = n.d1ref >= 0 ? n.d1ref : 1000 + (-n.d1ref - 1);
            int lab2 = n.d2ref >= 0 ? n.d2ref : 1000 + (-n.d2ref - 1);
            if (k == 0) {
                cur = amp; shape = {n.dj, n.dj1, n.dj2}; labels = {-1, lab1, lab2};
                continue;
            }
            int ci = -1;
            for (size_t j = 0; j < labels.size(); ++j) if (labels[j] == 1000 + k) { ci = (int)j; break; }
            if (ci < 0) throw std::runtime_error("contraction label not found");
            // out[...A without ci..., dj1, dj2] = sum_x A[..x..] * B[x, s1, s2]
            std::vector<int> oshape;
            for (size_t j = 0; j < shape.size(); ++j) if ((int)j != ci) oshape.push_back(shape[j]);
            oshape.push_back(n.dj1); oshape.push_back(n.dj2);
            int osz = 1; for (int d : oshape) osz *= d;
            std::vector<int> astr(shape.size());
            astr.back() = 1;
            for (int j = (int)shape.size() - 2; j >= 0; --j) astr[j] = astr[j + 1] * shape[j + 1];
            std::vector<cd> o(osz);
            int ra = (int)shape.size();
            std::vector<int> idx(oshape.size());
            for (int lin = 0; lin < osz; ++lin) {
                int t = lin;
                for (int j = (int)oshape.size() - 1; j >= 0; --j) { idx[j] = t % oshape[j]; t /= oshape[j]; }
                int abase = 0, pos = 0;
                for (int j = 0; j < ra; ++j) if (j != ci) abase += idx[pos++] * astr[j];
                int s1 = idx[oshape.size() - 2], s2 = idx[oshape.size() - 1];
                cd s = 0.0;
                for (int x = 0; x < shape[ci]; ++x)
                    s += cur[abase + x * astr[ci]] * amp[x * n.dj1 * n.dj2 + s1 * n.dj2 + s2];
                o[lin] = s;
            }
            cur.swap(o);
            shape = oshape;
            labels.erase(labels.begin() + ci);
            labels.push_back(lab1); labels.push_back(lab2);
        }
        for (int p = 0; p < P.npol; ++p) out[c * P.npol + p] = cur[P.polmap[p]];
    }
}

static void build_sample(Oracle& O, int cls) {
    Sample& s = O.S[cls];
    s.slamp.assign(O.P.topos.size(), {});
    s.kin.assign(O.P.topos.size(), {});
    for (size_t t = 0; t < O.P.topos.size(); ++t) {
        const OTopo& T = O.P.topos[t];
        size_t nslc = T.slc.size();
        s.slamp[t].assign(nslc * (size_t)s.n * O.P.npol, 0.0);
        s.kin[t].assign((size_t)s.n * T.nodes.size() * KS, 0.0);
        int npol = O.P.npol;
        int nn = (int)T.nodes.size();
#pragma omp parallel
        {
            std::vector<cd> buf(nslc * npol);
#pragma omp for schedule(static)
            for (int e = 0; e < s.n; ++e) {
                event_slamp(O.P, T, &s.mom[(size_t)e * O.P.nfinal * 4], buf.data(), &s.kin[t][(size_t)e * nn * KS]);
                for (size_t c = 0; c < nslc; ++c)
                    for (int p = 0; p < npol; ++p) s.slamp[t][(c * s.n + e) * npol + p] = buf[c * npol + p];
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Lineshape factor F of one (block, SL combination) for one event, with
// derivatives over the block's local variables (src/AmpGen.cu:1593-1932, the
// per-evaluation path for blocks with free parameters).
// ---------------------------------------------------------------------------
static Jet res_param(const ORes& r, int p, const std::vector<int>& vars, const std::vector<double>& theta) {
    if (p >= (int)r.params.size()) return Jet(0.0);
    int s = r.slots[p];
    if (s < 0) return Jet(r.params[p]);
    for (size_t i = 0; i < vars.size(); ++i) if (vars[i] == s) return jvar(theta[s], (int)i);
    return Jet(theta[s]);
}

// ---------------------------------------------------------------------------
// Custom lineshape expressions (src/CustomExpr.cu:95-228 semantics), evaluated on
// complex jets from the plan's prefix token list
// ---------------------------------------------------------------------------
// real function of two jets, with its first and second partial derivatives
static Jet jfun2(const Jet& x, const Jet& y, double f, double fx, double fy, double fxx, double fyy, double fxy) {
    Jet r(f);
    for (int i = 0; i < g_nv; i++) {
        r.g[i] = fx * x.g[i] + fy * y.g[i];
        for (int k = 0; k < g_nv; k++)
            r.h[i][k] = fx * x.h[i][k] + fy * y.h[i][k] + fxx * x.g[i] * x.g[k] + fyy * y.g[i] * y.g[k]
                        + fxy * (x.g[i] * y.g[k] + y.g[i] * x.g[k]);
    }
    return r;
}
static Jet jexp(const Jet& a) { double e = std::exp(a.v); return jfun(a, e, e, e); }
static Jet jsin(const Jet& a) { return jfun(a, std::sin(a.v), std::cos(a.v), -std::sin(a.v)); }
static Jet jcos(const Jet& a) { return jfun(a, std::cos(a.v), -std::sin(a.v), -std::cos(a.v)); }
static Jet jatan2(const Jet& y, const Jet& x) {
    double r2 = x.v * x.v + y.v * y.v, r4 = r2 * r2;
    return jfun2(x, y, std::atan2(y.v, x.v), -y.v / r2, x.v / r2, 2 * x.v * y.v / r4, -2 * x.v * y.v / r4,
                 (y.v * y.v - x.v * x.v) / r4);
}
static bool jzero(const Jet& a) {
    if (a.v != 0.0) return false;
    for (int i = 0; i < g_nv; i++) if (a.g[i] != 0.0) return false;
    return true;
}
static CJ cadd(const CJ& a, const CJ& b) { return CJ{a.re + b.re, a.im + b.im}; }
static CJ csub(const CJ& a, const CJ& b) { return CJ{a.re - b.re, a.im - b.im}; }
static CJ cdiv(const CJ& a, const CJ& b) {
    Jet d = b.re * b.re + b.im * b.im;
    return CJ{(a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d};
}
static CJ cexpj(const CJ& a) { Jet e = jexp(a.re); return CJ{e * jcos(a.im), e * jsin(a.im)}; }
static CJ clogj(const CJ& a) { return CJ{0.5 * jlog(a.re * a.re + a.im * a.im), jatan2(a.im, a.re)}; }
static CJ csqrtj(const CJ& a) {
    // principal square root; a real argument keeps a real (or purely imaginary) root
    if (jzero(a.im)) {
        if (a.re.v >= 0) return CJ{jsqrt(a.re), Jet(0.0)};
        return CJ{Jet(0.0), jsqrt(Jet(0.0) - a.re)};
    }
    Jet r = jsqrt(a.re * a.re + a.im * a.im);
    Jet re = jsqrt((r + a.re) * 0.5), im = jsqrt((r - a.re) * 0.5);
    if (a.im.v < 0) im = Jet(0.0) - im;
    return CJ{re, im};
}

struct CExpr {
    const std::vector<std::string>& t;
    size_t i = 0;
    const ORes& R;
    const std::vector<int>& vars;
    const std::vector<double>& theta;
    double m, q;
    const double* dv;      // |p|, E, cos(theta), phi of daughter 1, then of daughter 2
    CJ eval() {
        const std::string& k = t.at(i++);
        auto real = [&](const Jet& x) { return CJ{x, Jet(0.0)}; };
        if (k.rfind("n:", 0) == 0) return real(Jet(std::stod(k.substr(2))));
        if (k == "I") return CJ{Jet(0.0), Jet(1.0)};
        if (k == "v:m") return real(Jet(m));
        {
            static const char* dn[8] = {"v:p1p", "v:p1e", "v:p1costheta", "v:p1phi",
                                        "v:p2p", "v:p2e", "v:p2costheta", "v:p2phi"};
            for (int j = 0; j < 8; ++j) if (k == dn[j]) return real(Jet(dv[j]));
        }
        // Flatte channel phase-space factor of the event mass (src/CustomExpr.cu:844-870,
        // include/ResModel.cuh:418-451): rho = sqrt(1 - (ma+mb)^2/s) sqrt(1 - (ma-mb)^2/s)
        // above threshold (real part), its analytic continuation below (imaginary part);
        // no parameter dependence
        if (k == "rr" || k == "ri") {
            CJ a = eval(), b = eval();
            double ma = a.re.v, mb = b.re.v, s2 = m * m;
            double f1 = 1.0 - (ma + mb) * (ma + mb) / s2, f2 = 1.0 - (ma - mb) * (ma - mb) / s2;
            double g2 = f2 > 0.0 ? std::sqrt(f2) : 0.0;
            double v = (k == "rr") ? ((f1 >= 0.0 ? std::sqrt(f1) : 0.0) * g2) : ((f1 >= 0.0 ? 0.0 : std::sqrt(-f1)) * g2);
            return real(Jet(v));
        }
        if (k == "v:q") return real(Jet(q));
        if (k.rfind("p:", 0) == 0) return real(res_param(R, std::stoi(k.substr(2)), vars, theta));
        if (k == "+") { CJ a = eval(), b = eval(); return cadd(a, b); }
        if (k == "-") { CJ a = eval(), b = eval(); return csub(a, b); }
        if (k == "*") { CJ a = eval(), b = eval(); return cmul(a, b); }
        if (k == "/") { CJ a = eval(), b = eval(); return cdiv(a, b); }
        if (k == "^") { CJ a = eval(), b = eval(); return cexpj(cmul(b, clogj(a))); }
        if (k == "neg") { CJ a = eval(); return CJ{Jet(0.0) - a.re, Jet(0.0) - a.im}; }
        if (k.rfind("f:", 0) == 0) {
            std::string f = k.substr(2);
            CJ a = eval();
            if (f == "exp") return cexpj(a);
            if (f == "log") return clogj(a);
            if (f == "sin") {   // sin(x + iy) = sin x cosh y + i cos x sinh y
                Jet ep = jexp(a.im), em = jexp(Jet(0.0) - a.im);
                return CJ{jsin(a.re) * (ep + em) * 0.5, jcos(a.re) * (ep - em) * 0.5};
            }
            if (f == "cos") {
                Jet ep = jexp(a.im), em = jexp(Jet(0.0) - a.im);
                return CJ{jcos(a.re) * (ep + em) * 0.5, Jet(0.0) - jsin(a.re) * (ep - em) * 0.5};
            }
            if (f == "sqrt") return csqrtj(a);
            if (f == "abs") return real(jsqrt(a.re * a.re + a.im * a.im));
            if (f == "re") return real(a.re);
            if (f == "im") return real(a.im);
            if (f == "conj") return CJ{a.re, Jet(0.0) - a.im};
            throw std::runtime_error("custom expression: unknown function " + f);
        }
        if (k == "q0") { CJ x = eval(), a = eval(), b = eval(); return real(q0ad_jet(x.re, a.re, b.re)); }
        if (k.rfind("bf:", 0) == 0) {
            int L = std::stoi(k.substr(3));
            CJ qq = eval(), qz = eval(), d = eval();
            return real(bf_jet(L, qq.re, qz.re, d.re.v));
        }
        if (k == "bw") { CJ mm = eval(), m0 = eval(), g0 = eval(); return bw_jet(mm.re, m0.re, g0.re); }
        if (k.rfind("bwr:", 0) == 0) {
            int L = std::stoi(k.substr(4));
            CJ mm = eval(), m0 = eval(), g0 = eval(), qq = eval(), qz = eval(), d = eval();
            return bwr_jet(mm.re, m0.re, g0.re, L, qq.re, qz.re, d.re.v);
        }
        throw std::runtime_error("custom expression: unknown token " + k);
    }
};

// Hist lineshape (include/ResModel.cuh:67-85, 153-157): table [m_min, m_max, n, v...];
// for m_min < m < m_max the value interpolates linearly between v[i] and v[i+1] at
// position x = n (m - m_min) / (m_max - m_min), with x held to [0, n - 2]; v[0] at and
// below m_min, v[n-1] at and above m_max, 0 for a negative or NaN mass.
static double hist_lookup(const std::vector<double>& t, double m) {
    if (!(m >= 0.0)) return 0.0;
    int n = (int)t[2];
    if (n < 2 || !(t[1] > t[0])) return 0.0;
    if (m <= t[0]) return t[3];
    if (m >= t[1]) return t[3 + n - 1];
    double x = std::min(std::max(n * (m - t[0]) / (t[1] - t[0]), 0.0), (double)(n - 2));
    int i = (int)x;
    double w = x - i;
    return (1.0 - w) * t[3 + i] + w * t[3 + i + 1];
}

// Interp lineshape, linear (include/ResModel.cuh:226-290, src/Resonance.cu:76-128):
// equal spacing [1, N, x0, dx, re, im]: segment i = trunc((x - x0) / dx), below 0 ->
// segment 0 at its start, at or above N-1 -> segment N-2 at its end, else the fraction
// within it; unequal spacing [-2, N, xs, re, im]: the segment by bisection, with the
// fraction unclamped (linear extrapolation past the ends).
static cd interp_lookup(const std::vector<double>& t, double x) {
    int N = (int)t[1];
    const double *re, *im;
    int i;
    double w;
    if (t[0] >= 0) {
        re = &t[4];
        im = &t[4 + N];
        double pos = (x - t[2]) / t[3];
        i = (int)pos;                       // truncation toward zero
        if (i < 0) { i = 0; w = 0.0; }
        else if (i >= N - 1) { i = N - 2; w = 1.0; }
        else w = pos - i;
    } else {
        const double* xs = &t[2];
        re = &t[2 + N];
        im = &t[2 + 2 * N];
        int a = 0, b = N - 1;
        while (b - a > 1) {
            int c = (a + b) / 2;
            if (x < xs[c]) b = c; else a = c;
        }
        i = a;
        w = (x - xs[i]) / (xs[i + 1] - xs[i]);
    }
    return cd(re[i] + w * (re[i + 1] - re[i]), im[i] + w * (im[i + 1] - im[i]));
}

static CJ block_factor(const OPlan& P, const OBlock& B, int c, const double* kin,
                       const std::vector<int>& vars, const std::vector<double>& theta) {
    const OTopo& T = P.topos[B.topo];
    int nn = (int)T.nodes.size();
    // resonance occupying each node's mother
    std::vector<int> res_at(nn, -1);
    for (size_t r = 0; r < B.res.size(); ++r) res_at[B.res[r].node] = (int)r;
    CJ F{Jet(1.0), Jet(0.0)};
    for (int k = 0; k < nn; ++k) {
        const ONode& n = T.nodes[k];
        int L = T.slc[B.c0 + c][k].second;
        double mm = kin[KS * k], qq = kin[KS * k + 1], md1 = kin[KS * k + 2], md2 = kin[KS * k + 3];
        // q0 daughter masses: config mass, else the mass parameter of the block resonance
        // occupying the daughter (1 when it has no parameters, src/AmpGen.cu:1504-1527),
        // else the event mass
        auto dmass = [&](double cfg, int ref, double evm) -> Jet {
            if (cfg > 0) return Jet(cfg);
            if (ref < 0) {
                int kk = -ref - 1;
                if (res_at[kk] >= 0)
                    return B.res[res_at[kk]].params.empty() ? Jet(1.0) : res_param(B.res[res_at[kk]], 0, vars, theta);
            }
            return Jet(evm);
        };
        Jet m1q = dmass(n.md1, n.d1ref, md1), m2q = dmass(n.md2, n.d2ref, md2);
        int r = res_at[k];
        if (k == 0 || r < 0) {
            double m0c = (n.mm > 0) ? n.mm : 1.0;
            if (n.has_bf) {
                Jet q0 = breakup_jet(Jet(m0c), m1q, m2q);
                Jet bf = bf_jet(L, Jet(qq), q0, n.bf_d);
                F = cscale(F, bf);
            }
            continue;
        }
        const ORes& R = B.res[r];
        Jet m0 = R.params.empty() ? Jet(1.0) : res_param(R, 0, vars, theta);
        Jet q0 = q0ad_jet(m0, m1q, m2q);
        Jet m(mm), q(qq);
        CJ f;
        switch (R.model) {
        case 0: {   // BWR x Bf  (src/ResModel.cu:417-444)
            Jet g0 = res_param(R, 1, vars, theta);
            f = bwr_jet(m, m0, g0, n.lmin, q, q0, n.has_bf ? n.bf_d : 0.0);
            if (n.has_bf) f = cscale(f, bf_jet(L, q, q0, n.bf_d));
            break;
        }
        case 1: {   // BW
            Jet g0 = res_param(R, 1, vars, theta);
            f = bw_jet(m, m0, g0);
            break;
        }
        case 2: {   // ONE
            f = CJ{n.has_bf ? bf_jet(L, q, q0, n.bf_d) : Jet(1.0), Jet(0.0)};
            break;
        }
        case 3: {   // Flatte x Bf
            int nch = (int)R.ch.size() / 2;
            std::vector<Jet> g(nch);
            for (int i = 0; i < nch; ++i) g[i] = res_param(R, 1 + i, vars, theta);
            f = flatte_jet(mm, m0, nch, g.data(), R.ch.data());
            if (n.has_bf) f = cscale(f, bf_jet(L, q, q0, n.bf_d));
            break;
        }
        case 4: {   // GS x Bf
            Jet g0 = res_param(R, 1, vars, theta);
            double mpi = R.ch.empty() ? 0.1396 : R.ch[0];
            f = gs_jet(mm, m0, g0, qq, q0, mpi);
            if (n.has_bf) f = cscale(f, bf_jet(L, q, q0, n.bf_d));
            break;
        }
        case 5: {   // Custom expression: no vertex barrier besides what it contains
            CExpr ce{R.expr, 0, R, vars, theta, mm, qq, kin + KS * k + 4};
            f = ce.eval();
            if (ce.i != R.expr.size()) throw std::runtime_error("custom expression: trailing tokens");
            break;
        }
        case 6: {   // Hist: the table value, real, no vertex barrier
            f = CJ{Jet(hist_lookup(R.tab, mm)), Jet(0.0)};
            break;
        }
        case 7: {   // Interp: the table value, no vertex barrier
            cd z = interp_lookup(R.tab, mm);
            f = CJ{Jet(z.real()), Jet(z.imag())};
            break;
        }
        default: throw std::runtime_error("model not supported by the oracle");
        }
        F = cmul(F, f);
    }
    return F;
}

// ===========================================================================
// Likelihood, gradient and Hessian (src/Analysis.cu:625-1262, 2689-3378;
// src/ComputeNLL.cu:21-84; src/AmpGen.cu:2762-2822)
// ===========================================================================
struct EvalOut {
    double nll = 0;
    std::vector<double> grad;      // [2 nfree + ntheta]
    std::vector<double> hess;      // optional
};

static void extend_v(const OPlan& P, const double* params, std::vector<cd>& v) {
    v.resize(P.namp);
    for (int a = 0; a < P.namp; ++a) {
        int f = P.amap[a];
        v[a] = P.aratio[a] * cd(params[f], params[P.nfree + f]);
    }
}

// per-event pieces: S_p, dS_p/dx (x = extended params Re v, Im v, theta), d2S
struct Acc {
    int nx;
    double val = 0;                        // data/bkg: sum w ln I (with I=0 rule); phsp: sum I
    std::vector<double> gI;                // sum coef * dI/dx  (x over extended params)
    std::vector<double> hI;                // sum c1 * d2I + c2 * dI dI^T
    Acc(int n, bool hess) : nx(n), gI(n, 0.0), hI(hess ? (size_t)n * n : 0, 0.0) {}
};

static void eval_class(const Oracle& O, int cls, const std::vector<cd>& v, const std::vector<double>& theta,
                       bool hess, Acc& tot) {
    const OPlan& P = O.P;
    const Sample& s = O.S[cls];
    int na = P.namp, nt = P.ntheta, npol = P.npol;
    int nx = 2 * na + nt;
    int nthr = 1;
#ifdef _OPENMP
    nthr = omp_get_max_threads();
#endif
    std::vector<Acc> accs(nthr, Acc(nx, hess));
#pragma omp parallel
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        Acc& acc = accs[tid];
        std::vector<cd> A(na * npol), dA((size_t)na * npol * nt, 0.0);
        std::vector<cd> d2S(hess ? (size_t)npol * nt * nt : 0);
        std::vector<cd> S(npol), dS((size_t)npol * nt);
        std::vector<double> dI(nx), ddI(hess ? (size_t)nx * nx : 0);
#pragma omp for schedule(static)
        for (int e = 0; e < s.n; ++e) {
            std::fill(dA.begin(), dA.end(), 0.0);
            if (hess) std::fill(d2S.begin(), d2S.end(), 0.0);
            for (const OBlock& B : P.blocks) {
                if (B.nsl == 0) continue;
                const std::vector<int>& vars = O.bvars[&B - &P.blocks[0]];
                g_nv = (int)vars.size();
                const OTopo& T = P.topos[B.topo];
                const double* kin = &s.kin[B.topo][(size_t)e * T.nodes.size() * KS];
                for (int c = 0; c < B.nsl; ++c) {
                    CJ F = block_factor(P, B, c, kin, vars, theta);
                    int a = B.site + c;
                    for (int p = 0; p < npol; ++p) {
                        cd sl = s.slamp[B.topo][((size_t)(B.c0 + c) * s.n + e) * npol + p];
                        A[a * npol + p] = cd(F.re.v, F.im.v) * sl;
                        for (int i = 0; i < g_nv; ++i)
                            dA[((size_t)a * npol + p) * nt + vars[i]] += cd(F.re.g[i], F.im.g[i]) * sl;
                        if (hess)
                            for (int i = 0; i < g_nv; ++i)
                                for (int k = 0; k < g_nv; ++k)
                                    d2S[((size_t)p * nt + vars[i]) * nt + vars[k]] += v[a] * cd(F.re.h[i][k], F.im.h[i][k]) * sl;
                    }
                }
            }
            double I = 0;
            for (int p = 0; p < npol; ++p) {
                cd sp = 0.0;
                for (int a = 0; a < na; ++a) sp += v[a] * A[a * npol + p];
                S[p] = sp;
                I += std::norm(sp);
                for (int j = 0; j < nt; ++j) {
                    cd d = 0.0;
                    for (int a = 0; a < na; ++a) d += v[a] * dA[((size_t)a * npol + p) * nt + j];
                    dS[p * nt + j] = d;
                }
            }
            double w = s.w.empty() ? 1.0 : s.w[e];
            // dI/dx, x = (Re v_a, Im v_a, theta_j)
            if (cls == 0 || hess) {
                for (int a = 0; a < na; ++a) {
                    cd g = 0.0;
                    for (int p = 0; p < npol; ++p) g += std::conj(A[a * npol + p]) * S[p];
                    dI[a] = 2.0 * g.real();
                    dI[na + a] = 2.0 * g.imag();
                }
                for (int j = 0; j < nt; ++j) {
                    double t = 0;
                    for (int p = 0; p < npol; ++p) t += 2.0 * (std::conj(S[p]) * dS[p * nt + j]).real();
                    dI[2 * na + j] = t;
                }
            }
            if (!hess) {
                if (cls == 0) {
                    acc.val += I;
                    for (int x = 0; x < nx; ++x) acc.gI[x] += dI[x];
                } else {
                    // computeFactorsAndWeightsKernel rules (src/ComputeNLL.cu:44-66)
                    std::vector<cd> wp(npol, 0.0);
                    if (I == 0.0) {
                        acc.val += std::log(1e-10) * w;
                    } else {
                        acc.val += std::log(I) * w;
                        double inv_f = w / I;
                        for (int p = 0; p < npol; ++p) {
                            double wx = std::min(std::max(S[p].real() * inv_f, -1e10), 1e10);
                            double wy = std::min(std::max(S[p].imag() * inv_f, -1e10), 1e10);
                            wp[p] = cd(wx, wy);
                        }
                    }
                    // sum_p conj(A) w  and  2 Re sum_p conj(w) dS
                    for (int a = 0; a < na; ++a) {
                        cd g = 0.0;
                        for (int p = 0; p < npol; ++p) g += std::conj(A[a * npol + p]) * wp[p];
                        acc.gI[a] += 2.0 * g.real();
                        acc.gI[na + a] += 2.0 * g.imag();
                    }
                    for (int j = 0; j < nt; ++j) {
                        double t = 0;
                        for (int p = 0; p < npol; ++p) t += 2.0 * (std::conj(wp[p]) * dS[p * nt + j]).real();
                        acc.gI[2 * na + j] += t;
                    }
                }
                continue;
            }
            // Hessian pieces: d2I_xy = 2 Re sum_p [conj(dS_y) dS_x + conj(S) d2S_xy]
            auto Jcol = [&](int x, int p) -> cd {
                if (x < na) return A[x * npol + p];
                if (x < 2 * na) return cd(0, 1) * A[(x - na) * npol + p];
                return dS[p * nt + (x - 2 * na)];
            };
            for (int x = 0; x < nx; ++x)
                for (int y = 0; y < nx; ++y) {
                    double t = 0;
                    for (int p = 0; p < npol; ++p) t += 2.0 * (std::conj(Jcol(y, p)) * Jcol(x, p)).real();
                    ddI[(size_t)x * nx + y] = t;
                }
            for (int p = 0; p < npol; ++p) {
                cd cs = std::conj(S[p]);
                for (int a = 0; a < na; ++a)
                    for (int j = 0; j < nt; ++j) {
                        cd d = dA[((size_t)a * npol + p) * nt + j];
                        double re = 2.0 * (cs * d).real();
                        double im = 2.0 * (cs * cd(0, 1) * d).real();
                        ddI[(size_t)a * nx + 2 * na + j] += re;
                        ddI[(size_t)(2 * na + j) * nx + a] += re;
                        ddI[(size_t)(na + a) * nx + 2 * na + j] += im;
                        ddI[(size_t)(2 * na + j) * nx + na + a] += im;
                    }
                for (int j = 0; j < nt; ++j)
                    for (int k = 0; k < nt; ++k)
                        ddI[(size_t)(2 * na + j) * nx + 2 * na + k] += 2.0 * (cs * d2S[((size_t)p * nt + j) * nt + k]).real();
            }
            if (cls == 0) {
                acc.val += I;
                for (int x = 0; x < nx; ++x) acc.gI[x] += dI[x];
                for (size_t q = 0; q < ddI.size(); ++q) acc.hI[q] += ddI[q];
            } else {
                // w ln I: gradient w/I dI, Hessian w (d2I/I - dI dI^T / I^2)
                acc.val += (I == 0.0 ? std::log(1e-10) : std::log(I)) * w;
                if (I != 0.0) {
                    double c1 = w / I, c2 = -w / (I * I);
                    for (int x = 0; x < nx; ++x) acc.gI[x] += c1 * dI[x];
                    for (int x = 0; x < nx; ++x)
                        for (int y = 0; y < nx; ++y)
                            acc.hI[(size_t)x * nx + y] += c1 * ddI[(size_t)x * nx + y] + c2 * dI[x] * dI[y];
                }
            }
        }
    }
    for (int t = 0; t < nthr; ++t) {
        tot.val += accs[t].val;
        for (int x = 0; x < nx; ++x) tot.gI[x] += accs[t].gI[x];
        for (size_t q = 0; q < tot.hI.size(); ++q) tot.hI[q] += accs[t].hI[q];
    }
}

static int evaluate(Oracle& O, const double* params, EvalOut& out, bool hess) {
    const OPlan& P = O.P;
    int na = P.namp, nt = P.ntheta, nf = P.nfree;
    int nx = 2 * na + nt;
    std::vector<cd> v;
    extend_v(P, params, v);
    std::vector<double> theta(nt);
    for (int j = 0; j < nt; ++j) theta[j] = params[2 * nf + j];
    Acc ph(nx, hess), da(nx, hess), bk(nx, hess);
    if (O.S[0].n > 0) eval_class(O, 0, v, theta, hess, ph);
    if (O.S[1].n > 0) eval_class(O, 1, v, theta, hess, da);
    if (O.S[2].n > 0) eval_class(O, 2, v, theta, hess, bk);
    double Np = O.S[0].n;
    double pf = (Np > 0) ? ph.val / Np : 0.0;
    double Neff = O.Wd - O.Wb;
    bool reset = false;
    if (pf <= 1e-30 || std::isnan(pf) || std::isinf(pf)) pf = 1e-30;
    double loss = -da.val + bk.val + Neff * std::log(pf);
    for (size_t g = 0; g < P.g_slot.size(); ++g) {
        double d = theta[P.g_slot[g]] - P.g_mu[g];
        loss += d * d / (2.0 * P.g_sig[g] * P.g_sig[g]);
    }
    if (std::isnan(loss) || std::isinf(loss)) { loss = 1e30; reset = true; }
    // gradient in extended params x
    std::vector<double> gx(nx, 0.0);
    if (!reset) {
        for (int x = 0; x < nx; ++x) gx[x] = -da.gI[x] + bk.gI[x] + (Np > 0 ? Neff / (pf * Np) * ph.gI[x] : 0.0);
        for (size_t g = 0; g < P.g_slot.size(); ++g)
            gx[2 * na + P.g_slot[g]] += (theta[P.g_slot[g]] - P.g_mu[g]) / (P.g_sig[g] * P.g_sig[g]);
    }
    // map extended (Re v, Im v, theta) -> fit params (Re p, Im p, theta); v_a = r_a p_f(a)
    int np = 2 * nf + nt;
    std::vector<double> J((size_t)nx * np, 0.0);
    for (int a = 0; a < na; ++a) {
        int f = P.amap[a];
        J[(size_t)a * np + f] = P.aratio[a];
        J[(size_t)(na + a) * np + nf + f] = P.aratio[a];
    }
    for (int j = 0; j < nt; ++j) J[(size_t)(2 * na + j) * np + 2 * nf + j] = 1.0;
    out.nll = loss;
    out.grad.assign(np, 0.0);
    for (int q = 0; q < np; ++q)
        for (int x = 0; x < nx; ++x) out.grad[q] += J[(size_t)x * np + q] * gx[x];
    if (hess) {
        std::vector<double> hx((size_t)nx * nx, 0.0);
        double c = (Np > 0) ? Neff / (pf * Np) : 0.0;
        double c2 = (Np > 0) ? -Neff / (pf * pf * Np * Np) : 0.0;
        for (int x = 0; x < nx; ++x)
            for (int y = 0; y < nx; ++y)
                hx[(size_t)x * nx + y] = -da.hI[(size_t)x * nx + y] + bk.hI[(size_t)x * nx + y]
                                       + c * ph.hI[(size_t)x * nx + y] + c2 * ph.gI[x] * ph.gI[y];
        for (size_t g = 0; g < P.g_slot.size(); ++g)
            hx[(size_t)(2 * na + P.g_slot[g]) * nx + 2 * na + P.g_slot[g]] += 1.0 / (P.g_sig[g] * P.g_sig[g]);
        out.hess.assign((size_t)np * np, 0.0);
        for (int q = 0; q < np; ++q)
            for (int r = 0; r < np; ++r) {
                double s = 0;
                for (int x = 0; x < nx; ++x) {
                    double jx = J[(size_t)x * np + q];
                    if (jx == 0) continue;
                    for (int y = 0; y < nx; ++y) s += jx * hx[(size_t)x * nx + y] * J[(size_t)y * np + r];
                }
                out.hess[(size_t)q * np + r] = s;
            }
    }
    return reset ? 1 : 0;
}

// ===========================================================================
// C interface used by tests and the CPU-baseline leg of the benchmark
// ===========================================================================
extern "C" {

void* orc_create(const char* plan, int nfinal,
                 const double* mom_p, int n_p, const double* w_p,
                 const double* mom_d, int n_d, const double* w_d,
                 const double* mom_b, int n_b, const double* w_b) {
    try {
        Oracle* O = new Oracle();
        O->P = parse_plan(plan);
        if (O->P.nfinal != nfinal) throw std::runtime_error("nfinal mismatch");
        const double* moms[3] = {mom_p, mom_d, mom_b};
        const double* ws[3] = {w_p, w_d, w_b};
        int ns[3] = {n_p, n_d, n_b};
        for (int c = 0; c < 3; ++c) {
            O->S[c].n = ns[c];
            if (ns[c] > 0) O->S[c].mom.assign(moms[c], moms[c] + (size_t)ns[c] * nfinal * 4);
            if (ws[c] && ns[c] > 0) O->S[c].w.assign(ws[c], ws[c] + ns[c]);
            build_sample(*O, c);
        }
        O->Wd = 0;
        for (int e = 0; e < n_d; ++e) O->Wd += O->S[1].w.empty() ? 1.0 : O->S[1].w[e];
        O->Wb = 0;
        for (int e = 0; e < n_b; ++e) O->Wb += O->S[2].w.empty() ? 1.0 : O->S[2].w[e];
        for (const OBlock& B : O->P.blocks) {
            std::vector<int> vars;
            for (const ORes& r : B.res)
                for (int s : r.slots)
                    if (s >= 0 && std::find(vars.begin(), vars.end(), s) == vars.end()) vars.push_back(s);
            if ((int)vars.size() > MAXV) throw std::runtime_error("too many parameters in one block");
            O->bvars.push_back(vars);
        }
        return O;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "oracle: %s\n", e.what());
        return nullptr;
    }
}

void orc_destroy(void* h) { delete (Oracle*)h; }

void orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#endif
}

int orc_max_threads() {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

int orc_dims(void* h, int* namp, int* nfree, int* ntheta, int* npol) {
    Oracle* O = (Oracle*)h;
    *namp = O->P.namp; *nfree = O->P.nfree; *ntheta = O->P.ntheta; *npol = O->P.npol;
    return 0;
}

int orc_nll_grad(void* h, const double* params, double* nll, double* grad) {
    EvalOut o;
    int r = evaluate(*(Oracle*)h, params, o, false);
    *nll = o.nll;
    std::copy(o.grad.begin(), o.grad.end(), grad);
    return r;
}

int orc_hessian(void* h, const double* params, double* nll, double* grad, double* H) {
    EvalOut o;
    int r = evaluate(*(Oracle*)h, params, o, true);
    *nll = o.nll;
    std::copy(o.grad.begin(), o.grad.end(), grad);
    std::copy(o.hess.begin(), o.hess.end(), H);
    return r;
}

// per-event amplitudes A[e][p][a] (interleaved complex) of class cls
int orc_amplitudes(void* h, const double* params, int cls, double* out) {
    Oracle& O = *(Oracle*)h;
    const OPlan& P = O.P;
    const Sample& s = O.S[cls];
    int nt = P.ntheta;
    std::vector<double> theta(nt);
    for (int j = 0; j < nt; ++j) theta[j] = params[2 * P.nfree + j];
#pragma omp parallel for schedule(static)
    for (int e = 0; e < s.n; ++e) {
        for (size_t b = 0; b < P.blocks.size(); ++b) {
            const OBlock& B = P.blocks[b];
            g_nv = (int)O.bvars[b].size();
            const OTopo& T = P.topos[B.topo];
            const double* kin = &s.kin[B.topo][(size_t)e * T.nodes.size() * KS];
            for (int c = 0; c < B.nsl; ++c) {
                CJ F = block_factor(P, B, c, kin, O.bvars[b], theta);
                for (int p = 0; p < P.npol; ++p) {
                    cd a = cd(F.re.v, F.im.v) * s.slamp[B.topo][((size_t)(B.c0 + c) * s.n + e) * P.npol + p];
                    size_t i = ((size_t)e * P.npol + p) * P.namp + B.site + c;
                    out[2 * i] = a.real();
                    out[2 * i + 1] = a.imag();
                }
            }
        }
    }
    return 0;
}

// contracted node tensors of topology `topo` for class cls, before any lineshape or
// barrier factor: out[c][e][p] (interleaved complex) for SL combination c, event e and
// polarisation p, exactly as event_slamp leaves them (src/AmpGen.cu:924-1217)
int orc_slamp(void* h, int cls, int topo, double* out) {
    Oracle& O = *(Oracle*)h;
    if (topo < 0 || topo >= (int)O.P.topos.size()) return -1;
    const Sample& s = O.S[cls];
    const size_t nslc = O.P.topos[topo].slc.size(), npol = O.P.npol;
    for (size_t c = 0; c < nslc; ++c)
        for (int e = 0; e < s.n; ++e)
            for (size_t p = 0; p < npol; ++p) {
                const cd a = s.slamp[topo][((size_t)c * s.n + e) * npol + p];
                const size_t i = ((size_t)c * s.n + e) * npol + p;
                out[2 * i] = a.real();
                out[2 * i + 1] = a.imag();
            }
    return 0;
}

// per-event intensity I = sum_p |sum_a v_a A|^2 of class cls
int orc_intensity(void* h, const double* params, int cls, double* out) {
    Oracle& O = *(Oracle*)h;
    const OPlan& P = O.P;
    int n = O.S[cls].n;
    std::vector<double> amps((size_t)n * P.npol * P.namp * 2);
    orc_amplitudes(h, params, cls, amps.data());
    std::vector<cd> v;
    extend_v(P, params, v);
    for (int e = 0; e < n; ++e) {
        double I = 0;
        for (int p = 0; p < P.npol; ++p) {
            cd s = 0.0;
            for (int a = 0; a < P.namp; ++a) {
                size_t i = ((size_t)e * P.npol + p) * P.namp + a;
                s += v[a] * cd(amps[2 * i], amps[2 * i + 1]);
            }
            I += std::norm(s);
        }
        out[e] = I;
    }
    return 0;
}

// ---- component functions exported for the pinning tests ----
double orc_cg(int dj1, int m1, int dj2, int m2, int dJ, int M) { return ref_cg(dj1, m1, dj2, m2, dJ, M); }
void orc_ylm(int l, int m, double th, double ph, double* out) { cd y = ref_ylm(l, m, th, ph); out[0] = y.real(); out[1] = y.imag(); }
double orc_dfunc(int dim, int m1, int m2, double beta) { return ref_dfunc(dim, m1, m2, beta).real(); }
double orc_factorial(double n) { return ref_factorial(n); }
double orc_bf(int L, double q, double q0, double d) { g_nv = 0; return bf_jet(L, Jet(q), Jet(q0), d).v; }
void orc_bwr(double m, double m0, double g0, int L, double q, double q0, double d, double* out) {
    g_nv = 0;
    CJ z = bwr_jet(Jet(m), Jet(m0), Jet(g0), L, Jet(q), Jet(q0), d);
    out[0] = z.re.v; out[1] = z.im.v;
}
void orc_flatte(double m, double m0, int nch, const double* g, const double* ch, double* out) {
    g_nv = 0;
    std::vector<Jet> gj(nch);
    for (int i = 0; i < nch; ++i) gj[i] = Jet(g[i]);
    CJ z = flatte_jet(m, Jet(m0), nch, gj.data(), ch);
    out[0] = z.re.v; out[1] = z.im.v;
}
void orc_gs(double m, double m0, double g0, double q, double q0, double mpi, double* out) {
    g_nv = 0;
    CJ z = gs_jet(m, Jet(m0), Jet(g0), q, Jet(q0), mpi);
    out[0] = z.re.v; out[1] = z.im.v;
}
double orc_hist(const double* tab, int n, double m) { return hist_lookup(std::vector<double>(tab, tab + n), m); }
void orc_interp(const double* tab, int n, double x, double* out) {
    cd z = interp_lookup(std::vector<double>(tab, tab + n), x);
    out[0] = z.real(); out[1] = z.imag();
}
void orc_node_tensor(const double* p1, int dj1, const double* p2, int dj2, int dj, int dS, int L, double* out) {
    std::vector<cd> amp(dj * dj1 * dj2);
    pwa_amp(amp.data(), LV(p1[0], p1[1], p1[2], p1[3]), dj1, LV(p2[0], p2[1], p2[2], p2[3]), dj2, dj, dS, L);
    for (size_t i = 0; i < amp.size(); ++i) { out[2 * i] = amp[i].real(); out[2 * i + 1] = amp[i].imag(); }
}

void orc_massive_trans(const double* P, const double* p, int dim, double* out) {
    std::vector<cd> T(dim * dim);
    massive_trans(T.data(), LV(P[0], P[1], P[2], P[3]), LV(p[0], p[1], p[2], p[3]), dim);
    for (int i = 0; i < dim * dim; ++i) { out[2 * i] = T[i].real(); out[2 * i + 1] = T[i].imag(); }
}
void orc_massless_trans(const double* P, const double* p, int dim, double* out) {
    std::vector<cd> T(dim * dim);
    massless_trans(T.data(), LV(P[0], P[1], P[2], P[3]), LV(p[0], p[1], p[2], p[3]), dim);
    for (int i = 0; i < dim * dim; ++i) { out[2 * i] = T[i].real(); out[2 * i + 1] = T[i].imag(); }
}

}  // extern "C"
