// One-time per-event work: spin amplitudes (SL tensors) and event invariants.
//
// For every event this kernel
//  1. builds the four-momenta of each topology's particles (intermediates are the sum
//     of their daughters, the initial state is the sum of the final state in data
//     order; src/AmpGen.cu:531-653, src/Analysis.cu:148-160);
//  2. evaluates each two-body node tensor amp[s][s1][s2] = qs^L/sqrt(2j+1) CG CG
//     Y_L(theta, -phi) with the daughters' Wigner transforms applied transposed and
//     the factor 10 (src/Amplitude.cu:232-644), contracts the nodes in chain order
//     (src/AmpGen.cu:1034-1219) and keeps the polarisation-mapped entries;
//  3. multiplies in the event-only part 1/sqrt(N_L(q d)) of every Blatt-Weisskopf
//     factor of the tensor's barrier mask;
//  4. writes the event invariants used by the lineshape denominators.
//
// Angles are handled algebraically (direction cosines and unit phasors) instead of
// acos/atan2 round trips; Clebsch-Gordan products and Y_lm normalisations come from a
// host-built table.
#include <algorithm>

#include "device.hpp"

namespace ctpwa {

struct LV { double E, x, y, z; };
struct Cx { double r, i; };

__device__ inline LV lv_add(const LV& a, const LV& b) { return {a.E + b.E, a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ inline double lv_mass(const LV& p) {
    double m2 = p.E * p.E - p.x * p.x - p.y * p.y - p.z * p.z;
    return m2 < 0 ? 0.0 : sqrt(m2);
}
__device__ inline double lv_dot(const LV& a, const LV& b) { return a.E * b.E - (a.x * b.x + a.y * b.y + a.z * b.z); }
__device__ inline Cx cmul(Cx a, Cx b) { return {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; }
__device__ inline Cx cconj(Cx a) { return {a.r, -a.i}; }
// unit phasor raised to an integer power (negative powers conjugate)
__device__ inline Cx cpowu(Cx u, int n) {
    if (n < 0) { u = cconj(u); n = -n; }
    Cx r{1.0, 0.0};
    for (int i = 0; i < n; ++i) r = cmul(r, u);
    return r;
}
__device__ inline double ipow(double x, int n) {
    double r = 1.0;
    for (int i = 0; i < n; ++i) r *= x;
    return r;
}
// e^{i a/2} from e^{i a} with a in (-pi, pi]
__device__ inline Cx half_phasor(Cx u) {
    double c = sqrt(fmax(0.0, 0.5 * (1.0 + u.r)));
    double s = sqrt(fmax(0.0, 0.5 * (1.0 - u.r)));
    return {c, (u.i < 0.0) ? -s : s};
}

__constant__ double c_fact[21] = {
    1.0, 1.0, 2.0, 6.0, 24.0, 120.0, 720.0, 5040.0, 40320.0, 362880.0, 3628800.0,
    39916800.0, 479001600.0, 6227020800.0, 87178291200.0, 1307674368000.0,
    20922789888000.0, 355687428096000.0, 6402373705728000.0, 121645100408832000.0,
    2432902008176640000.0};

// The reference's small-d sum DFunc(dim, m1, m2, beta) (src/Amplitude.cu:272-318) with
// c = cos(beta/2), s = sin(beta/2).  All factorial arguments are integers.
__device__ double dsmall(int dim, int m1x2, int m2x2, double c, double s) {
    int j2 = dim - 1;
    int dm = (m1x2 - m2x2) / 2;                 // m1 - m2
    int a = (j2 + m2x2) / 2, b = (j2 - m2x2) / 2, cc = (j2 + m1x2) / 2, dd = (j2 - m1x2) / 2;
    int kmin = (m1x2 < m2x2) ? 0 : dm;
    int kmax = (m1x2 > -m2x2) ? b : cc;
    double num = sqrt(c_fact[a] * c_fact[b] * c_fact[cc] * c_fact[dd]);
    double sum = 0.0;
    for (int k = kmin; k <= kmax; ++k) {
        double den = c_fact[k] * c_fact[b - k] * c_fact[cc - k] * c_fact[k - dm];
        double ph = ((k - dm) & 1) ? -1.0 : 1.0;
        sum += ph * num / den * ipow(c, j2 + dm - 2 * k) * ipow(s, 2 * k - dm);
    }
    return sum;
}

// D(alpha, beta, gamma) matrix [i][j], index i <-> 2m = dim-1-2i, given the half-angle
// phasors e^{i alpha/2}, e^{i gamma/2} and cos/sin(beta/2).
__device__ void dmatrix(Cx* D, int dim, Cx ha, double cb, double sb, Cx hg) {
    for (int i = 0; i < dim; ++i)
        for (int j = 0; j < dim; ++j) {
            int m1 = dim - 1 - 2 * i, m2 = dim - 1 - 2 * j;
            double d = dsmall(dim, m1, m2, cb, sb);
            Cx e = cmul(cpowu(ha, m1), cpowu(hg, m2));
            D[i * dim + j] = {d * e.r, d * e.i};
        }
}

__device__ void matmul(Cx* C, const Cx* A, const Cx* B, int dim) {
    for (int i = 0; i < dim; ++i)
        for (int j = 0; j < dim; ++j) {
            Cx s{0.0, 0.0};
            for (int k = 0; k < dim; ++k) {
                Cx t = cmul(A[i * dim + k], B[k * dim + j]);
                s.r += t.r; s.i += t.i;
            }
            C[i * dim + j] = s;
        }
}

__device__ void identity(Cx* T, int dim) {
    for (int i = 0; i < dim * dim; ++i) T[i] = {0.0, 0.0};
    for (int i = 0; i < dim; ++i) T[i * dim + i] = {1.0, 0.0};
}

// MassiveTrans(p = pair momentum, q = daughter) (src/Amplitude.cu:331-404):
// T = D(-phi, -theta, -psi) D(0, theta, phi)
__device__ void massive_trans(Cx* T, const LV& p, const LV& q, int dim) {
    identity(T, dim);
    if (dim == 1) return;
    double hx = p.y * q.z - p.z * q.y, hy = p.z * q.x - p.x * q.z, hz = p.x * q.y - p.y * q.x;
    if (fabs(hx) < 1e-10 && fabs(hy) < 1e-10 && fabs(hz) < 1e-10) return;
    double Pp = sqrt(p.x * p.x + p.y * p.y + p.z * p.z), Pq = sqrt(q.x * q.x + q.y * q.y + q.z * q.z);
    double m1 = lv_mass(p), g1 = p.E / m1, m2 = lv_mass(q), g2 = q.E / m2;
    double ca = (p.x * q.x + p.y * q.y + p.z * q.z) / (-1.0 * Pp * Pq);
    double sa = sqrt(1 - ca * ca);
    double nrm = sqrt(hx * hx + hy * hy + hz * hz);
    double nx = hx / nrm, ny = hy / nrm, nz = fmin(1.0, fmax(-1.0, hz / nrm));
    double rho = sqrt(nx * nx + ny * ny);
    Cx ephi = (rho > 0.0) ? Cx{nx / rho, ny / rho} : Cx{1.0, 0.0};
    Cx hphi = half_phasor(ephi);
    double ct = sqrt(fmax(0.0, 0.5 * (1.0 + nz))), st = sqrt(fmax(0.0, 0.5 * (1.0 - nz)));
    double num = sa * sqrt((g1 - 1.0) * (g2 - 1.0));
    double den = sqrt((g1 + 1.0) * (g2 + 1.0)) + ca * sqrt((g1 - 1.0) * (g2 - 1.0));
    double rr = sqrt(num * num + den * den);
    Cx hpsi = (rr > 0.0) ? Cx{den / rr, num / rr} : Cx{1.0, 0.0};   // e^{i psi/2}
    Cx W1[25], W2[25];
    dmatrix(W1, dim, cconj(hphi), ct, -st, cconj(hpsi));
    dmatrix(W2, dim, Cx{1.0, 0.0}, ct, st, hphi);
    matmul(T, W1, W2, dim);
}

// MasslessTrans followed by the left D multiplication (src/Amplitude.cu:409-499,570-586)
__device__ void massless_trans(Cx* T, const LV& P, const LV& p, int dim, Cx hA, double cB, double sB) {
    identity(T, dim);
    double Pm = sqrt(P.x * P.x + P.y * P.y + P.z * P.z);
    double pm = sqrt(p.x * p.x + p.y * p.y + p.z * p.z);
    if (Pm >= 1e-10 && pm >= 1e-10) {
        double n1x = -P.x / Pm, n1y = -P.y / Pm, n1z = -P.z / Pm;
        double n2x = p.x / pm, n2y = p.y / pm, n2z = p.z / pm;
        double b1 = Pm / P.E, g1 = 1.0 / sqrt(1.0 - b1 * b1);
        double sh = sqrt((g1 - 1.0) / 2.0), ch = sqrt((g1 + 1.0) / 2.0);
        double d12 = n1x * n2x + n1y * n2y + n1z * n2z;
        double cz = n1x * n2y - n1y * n2x;
        double x12 = sh * (d12 - n1z) + ch * (1.0 - n2z), y12 = sh * cz;
        double x22 = sh * (d12 + n1z) + ch * (1.0 + n2z), y22 = -y12;
        double r1 = sqrt(x12 * x12 + y12 * y12), r2 = sqrt(x22 * x22 + y22 * y22);
        Cx e1 = (r1 > 0) ? Cx{x12 / r1, y12 / r1} : Cx{1.0, 0.0};
        Cx e2 = (r2 > 0) ? Cx{x22 / r2, y22 / r2} : Cx{1.0, 0.0};
        Cx etot = cconj(cmul(e1, e2));               // e^{i tot}, tot = -(psi1 + psi2)
        Cx htot = half_phasor(etot);
        for (int i = 0; i < dim; ++i) T[i * dim + i] = cpowu(htot, dim - 1 - 2 * i);
    }
    Cx D[25], tmp[25];
    dmatrix(D, dim, hA, cB, sB, Cx{1.0, 0.0});
    matmul(tmp, D, T, dim);
    for (int i = 0; i < dim * dim; ++i) T[i] = tmp[i];
}

// associated Legendre P_l^m(x), m >= 0, Condon-Shortley phase
__device__ double plm(int l, int m, double x) {
    double s = sqrt(fmax(0.0, 1.0 - x * x));
    double pmm = 1.0;
    for (int i = 1; i <= m; ++i) pmm *= -(2 * i - 1) * s;
    if (l == m) return pmm;
    double pm1 = x * (2 * m + 1) * pmm;
    if (l == m + 1) return pm1;
    double a = pmm, b = pm1, c = 0.0;
    for (int n = m + 2; n <= l; ++n) {
        c = ((2.0 * n - 1.0) * x * b - (n + m - 1.0) * a) / (n - m);
        a = b;
        b = c;
    }
    return c;
}

__device__ LV load_final(const double* m, int i) { return {m[4 * i], m[4 * i + 1], m[4 * i + 2], m[4 * i + 3]}; }

__device__ inline LV ref_mom(const LV* pm, int nfinal, int ref) {
    return (ref >= 0) ? pm[ref] : pm[nfinal + (-ref - 1)];
}

// node tensor amp[s][s1][s2] (src/Amplitude.cu:500-644)
__device__ void node_tensor(Cx* amp, const LV& p1, const LV& p2, const DNode& n, int L, const double* cg) {
    int dj = n.dj, dj1 = n.dj1, dj2 = n.dj2, kron = dj1 * dj2;
    double m1 = lv_mass(p1), m2 = lv_mass(p2);
    double m = sqrt(m1 * m1 + m2 * m2 + 2 * lv_dot(p1, p2));
    double lamb = (2 * m * p1.E + m * m + m1 * m1 - m2 * m2) / (2 * m * (m + p1.E + p2.E));
    double d12 = m1 * m1 - m2 * m2;
    double qs = sqrt(m * m * m * m + d12 * d12 - 2 * m * m * (m1 * m1 + m2 * m2)) / (2 * m);
    double nx = (p1.x - lamb * (p1.x + p2.x)) / qs;
    double ny = (p1.y - lamb * (p1.y + p2.y)) / qs;
    double nz = fmax(-1.0, fmin(1.0, (p1.z - lamb * (p1.z + p2.z)) / qs));
    double rho = sqrt(nx * nx + ny * ny);
    Cx ephi = (rho > 0.0) ? Cx{nx / rho, ny / rho} : Cx{1.0, 0.0};
    double qsL = ipow(qs, L);
    double P[8];
    for (int mm = 0; mm <= L && mm < 8; ++mm) P[mm] = plm(L, mm, nz);
    for (int si = 0; si < dj; ++si)
        for (int s1i = 0; s1i < dj1; ++s1i)
            for (int s2i = 0; s2i < dj2; ++s2i) {
                int idx = si * kron + s1i * dj2 + s2i;
                int M = ((dj - 1 - 2 * si) - (dj1 - 1 - 2 * s1i) - (dj2 - 1 - 2 * s2i)) / 2;
                int am = M < 0 ? -M : M;
                double c = (am <= L) ? cg[idx] * qsL * P[am] : 0.0;
                Cx e = cpowu(cconj(ephi), M);        // e^{-i M phi}
                amp[idx] = {c * e.r, c * e.i};
            }
    LV Pt = lv_add(p1, p2);
    Cx T1[25], T2[25];
    if (dj1 > 1) {
        if (m1 > 1e-10) massive_trans(T1, Pt, p1, dj1);
        else {
            Cx hphi = half_phasor(ephi);
            massless_trans(T1, Pt, p1, dj1, cconj(hphi), sqrt(fmax(0.0, 0.5 * (1 + nz))), -sqrt(fmax(0.0, 0.5 * (1 - nz))));
        }
    }
    if (dj2 > 1) {
        if (m2 > 1e-10) massive_trans(T2, Pt, p2, dj2);
        else {
            // alpha = -(pi + phi): e^{i alpha/2} = -i e^{-i phi/2}; beta = -(pi - theta)
            Cx hphi = half_phasor(ephi);
            Cx ha = cmul(Cx{0.0, -1.0}, cconj(hphi));
            massless_trans(T2, Pt, p2, dj2, ha, sqrt(fmax(0.0, 0.5 * (1 - nz))), -sqrt(fmax(0.0, 0.5 * (1 + nz))));
        }
    }
    if (dj1 == 1 && dj2 == 1) {
        for (int i = 0; i < dj; ++i) { amp[i].r *= 10.0; amp[i].i *= 10.0; }
        return;
    }
    // transposed application: new[n1][n2] = sum_o1,o2 T1[o1][n1] old[o1][o2] T2[o2][n2]
    Cx tmp[MAXT];
    for (int si = 0; si < dj; ++si) {
        int base = si * kron;
        for (int n1 = 0; n1 < dj1; ++n1)
            for (int o2 = 0; o2 < dj2; ++o2) {
                Cx s{0.0, 0.0};
                for (int o1 = 0; o1 < dj1; ++o1) {
                    Cx t = (dj1 > 1) ? cmul(T1[o1 * dj1 + n1], amp[base + o1 * dj2 + o2]) : amp[base + o1 * dj2 + o2];
                    s.r += t.r; s.i += t.i;
                }
                tmp[n1 * dj2 + o2] = s;
            }
        for (int n1 = 0; n1 < dj1; ++n1)
            for (int n2 = 0; n2 < dj2; ++n2) {
                Cx s{0.0, 0.0};
                for (int o2 = 0; o2 < dj2; ++o2) {
                    Cx t = (dj2 > 1) ? cmul(tmp[n1 * dj2 + o2], T2[o2 * dj2 + n2]) : tmp[n1 * dj2 + o2];
                    s.r += t.r; s.i += t.i;
                }
                amp[base + n1 * dj2 + n2] = {s.r * 10.0, s.i * 10.0};
            }
    }
}

// Hist shape table t = [m_min, m_max, n, v_0..v_{n-1}] at mass m: linear between the bin
// values on x = (m - m_min) / (m_max - m_min) * n clamped to [0, n - 2], the first/last
// value outside the range, 0 for a negative or NaN mass (include/ResModel.cuh:67-85)
__device__ double hist_value(const double* t, double m) {
    if (!(m >= 0.0)) return 0.0;
    const double lo = t[0], hi = t[1];
    const int n = (int)t[2];
    const double* v = t + 3;
    if (n < 2 || hi <= lo) return 0.0;
    if (m <= lo) return v[0];
    if (m >= hi) return v[n - 1];
    double x = (m - lo) / (hi - lo) * n;
    if (x < 0.0) x = 0.0;
    if (x >= (double)(n - 1)) x = (double)(n - 2);
    const int i = (int)x;
    const double f = x - i;
    return v[i] * (1.0 - f) + v[i + 1] * f;
}

// Interp table (linear): t = [1, N, x_min, dx, re..., im...] for equally spaced points or
// [-2, N, x..., re..., im...] otherwise; the segment index is clamped to [0, N - 2] and
// the fraction is not, so the end segments extrapolate linearly
// (include/ResModel.cuh:226-290)
__device__ void interp_value(const double* t, double x, double& fr, double& fi) {
    const int hdr = (int)t[0];
    const int N = (int)t[1];
    const double *re, *im;
    int i;
    double frac;
    if (hdr >= 0) {
        const double xmin = t[2], dx = t[3];
        re = t + 4;
        im = t + 4 + N;
        const double pos = (x - xmin) / dx;
        i = (int)pos;
        if (i < 0) { i = 0; frac = 0.0; }
        else if (i >= N - 1) { i = N - 2; frac = 1.0; }
        else frac = pos - (double)i;
    } else {
        const double* xv = t + 2;
        re = t + 2 + N;
        im = t + 2 + 2 * N;
        int lo = 0, hi = N - 1;
        while (lo < hi - 1) {
            const int mid = (lo + hi) / 2;
            if (x < xv[mid]) hi = mid; else lo = mid;
        }
        i = lo;
        frac = (x - xv[i]) / (xv[i + 1] - xv[i]);
    }
    fr = re[i] + (re[i + 1] - re[i]) * frac;
    fi = im[i] + (im[i + 1] - im[i]) * frac;
}

__global__ void __launch_bounds__(128)
k_init_events(const double* __restrict__ mom, long long N, int nfinal,
              const DTopo* __restrict__ topos, int ntopo, const DNode* __restrict__ nodes,
              const DKin* __restrict__ kins, int nkin,
              const DU* __restrict__ us, int nu, const double* __restrict__ cgtab,
              const int* __restrict__ polmap, int npol,
              const DInv* __restrict__ invs, int ninv, const double* __restrict__ tabs,
              double2* __restrict__ sl, double* __restrict__ inv, long long ld, long long e_off) {
    long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= N) return;
    const double* m = mom + (size_t)e * nfinal * 4;
    long long eo = e_off + e;
    double kmm[MAXKIN], kq[MAXKIN];
    LV pm[MAXPART];
    Cx cur[MAXT], nxt[MAXT], amp[MAXT];
    for (int t = 0; t < ntopo; ++t) {
        const DTopo& T = topos[t];
        for (int i = 0; i < nfinal; ++i) pm[i] = load_final(m, i);
        LV root = {0.0, 0.0, 0.0, 0.0};
        for (int i = 0; i < nfinal; ++i) root = lv_add(root, pm[i]);
        pm[nfinal] = root;
        for (int oi = 0; oi < T.nn - 1; ++oi) {
            int k = T.order[oi];
            const DNode& n = nodes[T.node0 + k];
            pm[nfinal + k] = lv_add(ref_mom(pm, nfinal, n.d1ref), ref_mom(pm, nfinal, n.d2ref));
        }
        for (int k = 0; k < T.nn; ++k) {
            const DNode& n = nodes[T.node0 + k];
            const DKin& K = kins[n.kin];
            if (K.owner_topo != t || K.owner_node != k) continue;
            double mm = lv_mass(pm[nfinal + k]);
            kmm[n.kin] = mm;
            kq[n.kin] = breakup(mm, lv_mass(ref_mom(pm, nfinal, n.d1ref)), lv_mass(ref_mom(pm, nfinal, n.d2ref)));
        }
        for (int u = 0; u < nu; ++u) {
            const DU& U = us[u];
            if (U.topo != t) continue;
            int rank = 0;
            int shape[2 * MAXNODE + 2], labels[2 * MAXNODE + 2];
            for (int k = 0; k < T.nn; ++k) {
                const DNode& n = nodes[T.node0 + k];
                LV p1 = ref_mom(pm, nfinal, n.d1ref), p2 = ref_mom(pm, nfinal, n.d2ref);
                int l1 = n.d1ref >= 0 ? n.d1ref : 1000 + (-n.d1ref - 1);
                int l2 = n.d2ref >= 0 ? n.d2ref : 1000 + (-n.d2ref - 1);
                if (k == 0) {
                    node_tensor(cur, p1, p2, n, U.L[k], cgtab + U.cg[k]);
                    rank = 3;
                    shape[0] = n.dj; shape[1] = n.dj1; shape[2] = n.dj2;
                    labels[0] = -1; labels[1] = l1; labels[2] = l2;
                    continue;
                }
                node_tensor(amp, p1, p2, n, U.L[k], cgtab + U.cg[k]);
                int ci = 0;
                for (int j = 0; j < rank; ++j) if (labels[j] == 1000 + k) { ci = j; break; }
                int astr[2 * MAXNODE + 2];
                astr[rank - 1] = 1;
                for (int j = rank - 2; j >= 0; --j) astr[j] = astr[j + 1] * shape[j + 1];
                int oshape[2 * MAXNODE + 2], orank = 0;
                for (int j = 0; j < rank; ++j) if (j != ci) oshape[orank++] = shape[j];
                oshape[orank++] = n.dj1;
                oshape[orank++] = n.dj2;
                int osz = 1;
                for (int j = 0; j < orank; ++j) osz *= oshape[j];
                int kb = n.dj1 * n.dj2;
                for (int lin = 0; lin < osz; ++lin) {
                    int tt = lin, idx[2 * MAXNODE + 2];
                    for (int j = orank - 1; j >= 0; --j) { idx[j] = tt % oshape[j]; tt /= oshape[j]; }
                    int abase = 0, pos = 0;
                    for (int j = 0; j < rank; ++j) if (j != ci) abase += idx[pos++] * astr[j];
                    int s12 = idx[orank - 2] * n.dj2 + idx[orank - 1];
                    Cx s{0.0, 0.0};
                    for (int x = 0; x < shape[ci]; ++x) {
                        Cx tv = cmul(cur[abase + x * astr[ci]], amp[x * kb + s12]);
                        s.r += tv.r; s.i += tv.i;
                    }
                    nxt[lin] = s;
                }
                for (int i = 0; i < osz; ++i) cur[i] = nxt[i];
                int nr = 0;
                for (int j = 0; j < rank; ++j) if (j != ci) { labels[nr] = labels[j]; shape[nr] = shape[j]; ++nr; }
                labels[nr] = l1; shape[nr] = n.dj1; ++nr;
                labels[nr] = l2; shape[nr] = n.dj2; ++nr;
                rank = nr;
            }
            // event-only barrier parts of this tensor's mask
            double X = 1.0;
            for (int k = 0; k < T.nn; ++k) {
                if (!((U.mask >> k) & 1)) continue;
                const DNode& n = nodes[T.node0 + k];
                int L = U.L[k];
                if (L >= 1 && L <= 6) X /= sqrt(bf_poly(L, kq[n.kin] * n.bf_d));
            }
            for (int p = 0; p < npol; ++p) {
                Cx v = cur[polmap[p]];
                sl[((size_t)u * npol + p) * ld + eo] = make_double2(v.r * X, v.i * X);
            }
        }
    }
    // invariants
    for (int i = 0; i < ninv; ++i) {
        const DInv& I = invs[i];
        double mm = kmm[I.kin], q = kq[I.kin];
        double val = 0.0;
        switch (I.kind) {
        case 1: val = mm * mm; break;                                     // s
        case 2: {                                                          // q^(2L+1) / (m N_L(q d))
            double nq = (I.d != 0.0) ? bf_poly(I.L, q * I.d) : 1.0;
            val = ipow(q, 2 * I.L + 1) / (mm * nq);
            break;
        }
        case 3:
        case 4: {                                                          // Flatte rho (ResModel.cuh:418-451)
            double s = mm * mm;
            double f1 = 1.0 - (I.ma + I.mb) * (I.ma + I.mb) / s;
            double f2 = 1.0 - (I.ma - I.mb) * (I.ma - I.mb) / s;
            double g2 = f2 > 0.0 ? sqrt(f2) : 0.0;
            if (I.kind == 3) val = (f1 >= 0.0 ? sqrt(f1) : 0.0) * g2;
            else val = (f1 >= 0.0 ? 0.0 : sqrt(-f1)) * g2;
            break;
        }
        case 5: {                                                          // GS q^2 h(m)
            const double pi = 3.14159265358979323846;
            double h = (2.0 / pi) * (q / mm) * log((mm + 2.0 * q) / (2.0 * I.ma));
            val = q * q * h;
            break;
        }
        case 6: val = q * q; break;
        case 7: val = q * q * q / mm; break;
        case 8: val = mm; break;                                           // m (Custom)
        case 9: val = q; break;                                            // q (Custom)
        case 12: case 13: case 14: case 15:                                // daughter momenta in the
        case 16: case 17: case 18: case 19: {                              // event frame (Custom)
            const int j = I.kind - 12;
            const uint32_t set = (j < 4) ? kins[I.kin].d1set : kins[I.kin].d2set;
            LV d = {0.0, 0.0, 0.0, 0.0};
            for (int f = 0; f < nfinal; ++f)
                if ((set >> f) & 1u) d = lv_add(d, load_final(m, f));
            const double P = sqrt(d.x * d.x + d.y * d.y + d.z * d.z);
            switch (j & 3) {
            case 0: val = P; break;
            case 1: val = d.E; break;
            case 2: val = (P > 0.0) ? d.z / P : 0.0; break;
            default: val = atan2(d.y, d.x);
            }
            break;
        }
        case 10:
        case 11: {                                                         // table lineshapes
            const double* t = tabs + I.L;
            if ((int)I.d == 6) val = (I.kind == 10) ? hist_value(t, mm) : 0.0;
            else {
                double fr, fi;
                interp_value(t, mm, fr, fi);
                val = (I.kind == 10) ? fr : fi;
            }
            break;
        }
        default: val = 1.0;
        }
        inv[(size_t)i * ld + eo] = val;
    }
}

// mixed precision: a batch of FP64 SL tensors and invariants (leading dimension n) is
// rounded into the FP32 stores; invariants with imap >= 0 are copied to the FP64 store
__global__ void __launch_bounds__(256)
k_to_mixed(const double2* __restrict__ sl64, const double* __restrict__ inv64, long long n, int nrow, int ninv,
           const int* __restrict__ imap, float2* __restrict__ slf, double* __restrict__ invd, float* __restrict__ invf,
           long long ld, long long e_off) {
    long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    for (int r = 0; r < nrow; ++r) {
        double2 v = sl64[(long long)r * n + e];
        slf[(long long)r * ld + e_off + e] = make_float2((float)v.x, (float)v.y);
    }
    for (int i = 0; i < ninv; ++i) {
        double v = inv64[(long long)i * n + e];
        int m = imap[i];
        if (m >= 0) invd[(long long)m * ld + e_off + e] = v;
        else invf[(long long)(-m - 1) * ld + e_off + e] = (float)v;
    }
}

}  // namespace ctpwa

extern "C" __attribute__((visibility("hidden"))) hipError_t ctpwa_launch_to_mixed(
    const double2* sl64, const double* inv64, long long n, int nrow, int ninv, const int* imap, float2* slf,
    double* invd, float* invf, long long ld, long long e_off, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    long long nb = (n + 255) / 256;
    hipLaunchKernelGGL(ctpwa::k_to_mixed, dim3((unsigned)nb), dim3(256), 0, stream, sl64, inv64, n, nrow, ninv, imap,
                       slf, invd, invf, ld, e_off);
    return hipGetLastError();
}

// host launcher (called from capi.hip)
extern "C" __attribute__((visibility("hidden"))) hipError_t ctpwa_launch_init(const double* mom, long long N, int nfinal,
                                         const ctpwa::DTopo* topos, int ntopo, const ctpwa::DNode* nodes,
                                         const ctpwa::DKin* kins, int nkin, const ctpwa::DU* us, int nu,
                                         const double* cgtab, const int* polmap, int npol,
                                         const ctpwa::DInv* invs, int ninv, const double* tabs, double2* sl,
                                         double* inv, long long ld, long long e_off, hipStream_t stream) {
    if (N <= 0) return hipSuccess;
    int bs = 128;
    long long nb = (N + bs - 1) / bs;
    hipLaunchKernelGGL(ctpwa::k_init_events, dim3((unsigned)nb), dim3(bs), 0, stream, mom, N, nfinal, topos, ntopo,
                       nodes, kins, nkin, us, nu, cgtab, polmap, npol, invs, ninv, tabs, sl, inv, ld, e_off);
    return hipGetLastError();
}

namespace ctpwa {

// mirror-image check of the two polarisations of every SL tensor over events
// [e_off, e_off + n) of a [nu][2][ld] FP64 store: acc[4 u + 0] = max |sl1 - conj(sl0)|,
// acc[4 u + 1] = max |sl1 + conj(sl0)|, acc[4 u + 2] = max |sl0| (componentwise maxima,
// non-negative doubles kept as their bit patterns, so an atomic integer maximum orders
// them; any order of the atomics gives the same maxima)
__global__ void __launch_bounds__(256) k_polsym(const double2* __restrict__ sl, long long ld, long long e_off, long long n,
                                                int nu, unsigned long long* __restrict__ acc) {
    const int u = blockIdx.y;
    double dp = 0.0, dm = 0.0, m0 = 0.0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const double2 a = sl[((long long)u * 2 + 0) * ld + e_off + i];
        const double2 b = sl[((long long)u * 2 + 1) * ld + e_off + i];
        dp = fmax(dp, fmax(fabs(b.x - a.x), fabs(b.y + a.y)));
        dm = fmax(dm, fmax(fabs(b.x + a.x), fabs(b.y - a.y)));
        m0 = fmax(m0, fmax(fabs(a.x), fabs(a.y)));
    }
    for (int off = 32; off > 0; off >>= 1) {
        dp = fmax(dp, __shfl_xor(dp, off, 64));
        dm = fmax(dm, __shfl_xor(dm, off, 64));
        m0 = fmax(m0, __shfl_xor(m0, off, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMax(&acc[4 * u + 0], (unsigned long long)__double_as_longlong(dp));
        atomicMax(&acc[4 * u + 1], (unsigned long long)__double_as_longlong(dm));
        atomicMax(&acc[4 * u + 2], (unsigned long long)__double_as_longlong(m0));
    }
}

}  // namespace ctpwa

extern "C" __attribute__((visibility("hidden"))) hipError_t ctpwa_launch_polsym(const double2* sl, long long ld,
                                                                                 long long e_off, long long n, int nu,
                                                                                 unsigned long long* acc,
                                                                                 hipStream_t stream) {
    if (n <= 0 || nu <= 0) return hipSuccess;
    const unsigned nb = (unsigned)std::min<long long>(256, (n + 255) / 256);
    hipLaunchKernelGGL(ctpwa::k_polsym, dim3(nb, (unsigned)nu), dim3(256), 0, stream, sl, ld, e_off, n, nu, acc);
    return hipGetLastError();
}
