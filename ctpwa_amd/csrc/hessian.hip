// Hessian of the likelihood (replaces analysis::getHessian, src/Analysis.cu:2689-3378).
//
// With S_p = sum_a v_a A_a(p), I = sum_p |S_p|^2 and real parameters x = (Re v, Im v,
// theta), every event needs
//   dI/dx           = 2 Re <S, dS/dx>
//   d2I/dx dy       = 2 Re sum_p [conj(dS/dy) dS/dx + conj(S) d2S/dx dy]
// The first part of d2I is the real embedding of the complex Gram of the Jacobian
// columns J = [A_1..A_n, dS/dtheta_1..dS/dtheta_P] (rows: event x polarisation), so
// the sum over events is a weighted Gram  sum_e alpha_e J^H J  (a K = N*nPol tall-skinny
// contraction).  The second part only couples (v, theta) and (theta, theta) and reduces
// to per-event scalars: <S, dA_a/dtheta> = dF_a/dtheta conj(T_u) and
// <S, d2S> = sum_a v_a d2F_a conj(T_u), T_u = sum_p conj(sl'_u,p) S_p.
//
// Per class the device forms
//   G   = sum_e alpha_e Jhat^T Jhat      (Jhat = [Re J | Im J])
//   U   = sum_e beta_e u u^T,  u = dI/dx
//   Z   = sum_e alpha_e z,     z = second-part scalars and I
//   g   = sum_e u
// with alpha = w/I, beta = -w/I^2 (data, background: second derivative of w ln I) and
// alpha = 1, beta = 0 (phase space), then assembles
//   H = -H_data + H_bkg + N_eff [ A2 / (N_p pf) - g g^T / (N_p pf)^2 ] (+ Gaussian terms)
// and maps it to the fit parameters (v_a = r_a p_f(a)).
//
// Events are processed in chunks; every sum is a fixed-order reduction.
#include <algorithm>

#include "device.hpp"

namespace ctpwa {



struct Cd { double r, i; };
__device__ inline Cd cm(Cd a, Cd b) { return {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; }
__device__ inline Cd ca(Cd a, Cd b) { return {a.r + b.r, a.i + b.i}; }
__device__ inline Cd cs(Cd a, Cd b) { return {a.r - b.r, a.i - b.i}; }
__device__ inline Cd csc(Cd a, double s) { return {a.r * s, a.i * s}; }
__device__ inline Cd cinv(Cd a) { double q = 1.0 / (a.r * a.r + a.i * a.i); return {a.r * q, -a.i * q}; }
__device__ inline Cd ccj(Cd a) { return {a.r, -a.i}; }

__device__ inline double rec_v(const HessArgs& a, int rec) { return a.table[a.off_rec + (long long)rec * REC]; }
__device__ inline double rec_g(const HessArgs& a, int rec, int i) { return a.table[a.off_rec + (long long)rec * REC + 1 + i]; }
__device__ inline double rec_h(const HessArgs& a, int rec, int i, int k) {
    return a.table[a.off_rec + (long long)rec * REC + 1 + NLOC + i * NLOC + k];
}

// the factor product of a block and its derivatives over the block's local variables:
//   psi, P1_i = -d psi / d theta_i, P2_ik = d2 psi / d theta_i d theta_k.
// The propagators enter through their log-derivatives (psi = prod 1/D_d, L_i = sum
// D_i / D, M_ik = sum (D_ik / D - D_i D_k / D^2): P1 = psi L, P2 = psi (L L - M)); each
// Custom factor F (value, F_i, F_ik per event, second-order rows) then multiplies in by
// the product rule, so a factor that vanishes needs no division.
// NL bounds the block's local variables (B.nloc <= NL): every array is indexed with
// unrolled constant indices and lives in registers.
// H2 = false: value and first derivatives only (Mm is left unset)
template <int NL, bool H2 = true>
__device__ inline void block_terms(const HessArgs& a, int b, long long e, Cd& psi, Cd* L, Cd* Mm) {
    const DBlock& B = a.blocks[b];
    const int nl = B.nloc;
    psi = {1.0, 0.0};
#pragma unroll
    for (int i = 0; i < NL; ++i) {
        L[i] = {0.0, 0.0};
#pragma unroll
        for (int k = 0; k < NL; ++k) Mm[i * NL + k] = {0.0, 0.0};
    }
    for (int dd = 0; dd < B.nden; ++dd) {
        const DDen& D = a.dens[B.den0 + dd];
        const DDenTerms& T = a.dterms[B.den0 + dd];
        Cd Dv{0.0, 0.0}, dD[NL], d2D[NL * NL];
#pragma unroll
        for (int i = 0; i < NL; ++i) {
            dD[i] = {0.0, 0.0};
#pragma unroll
            for (int k = 0; k < NL; ++k) d2D[i * NL + k] = {0.0, 0.0};
        }
        for (int t = 0; t < T.nre + T.nim; ++t) {
            const DTerm& tm = a.terms[T.t0 + t];
            double x = (tm.inv < 0) ? 1.0 : ev_inv(a.st, tm.inv, e);
            bool im = t >= T.nre;           // D = sum re - i sum im
            double c = (tm.coef < 0) ? tm.cval : rec_v(a, D.rec + tm.coef);
            if (im) Dv.i -= c * x; else Dv.r += c * x;
            if (tm.coef < 0) continue;
#pragma unroll
            for (int i = 0; i < NL; ++i) {
                if (i >= nl) break;
                double gi = rec_g(a, D.rec + tm.coef, i) * x;
                if (im) dD[i].i -= gi; else dD[i].r += gi;
                if constexpr (!H2) continue;
#pragma unroll
                for (int k = 0; k < NL; ++k) {
                    if (k >= nl) break;
                    double hik = rec_h(a, D.rec + tm.coef, i, k) * x;
                    if (im) d2D[i * NL + k].i -= hik; else d2D[i * NL + k].r += hik;
                }
            }
        }
        // a table lineshape is the factor sum re + i sum im itself (no theta dependence)
        Cd pd = T.stat ? Cd{Dv.r, -Dv.i} : cinv(Dv);
        psi = cm(psi, pd);
#pragma unroll
        for (int i = 0; i < NL; ++i) {
            if (i >= nl) break;
            Cd li = cm(pd, dD[i]);
            L[i] = ca(L[i], li);
            if constexpr (!H2) continue;
#pragma unroll
            for (int k = 0; k < NL; ++k) {
                if (k >= nl) break;
                Mm[i * NL + k] = ca(Mm[i * NL + k], cs(cm(pd, d2D[i * NL + k]), cm(li, cm(pd, dD[k]))));
            }
        }
    }
    // log-derivatives -> derivatives of psi (P2 first: it reads L)
#pragma unroll
    for (int i = 0; i < NL; ++i) {
        if (!H2 || i >= nl) break;
#pragma unroll
        for (int k = 0; k < NL; ++k) {
            if (k >= nl) break;
            Mm[i * NL + k] = cm(psi, cs(cm(L[i], L[k]), Mm[i * NL + k]));
        }
    }
#pragma unroll
    for (int i = 0; i < NL; ++i) L[i] = (i < nl) ? cm(psi, L[i]) : Cd{0.0, 0.0};
    // Custom factors of the block
    for (int q = a.bcp[b]; q < a.bcp[b + 1]; ++q) {
        const double* f = a.cfb2 + (long long)a.bcl[q] * a.ncf2 * a.cld + e;
        const long long ld = a.cld;
        const int ncv = a.ncv, nh = ncv * (ncv + 1) / 2;
        const Cd F{f[0], f[ld]};
        Cd Fi[NL];
#pragma unroll
        for (int i = 0; i < NL; ++i)
            Fi[i] = (i < nl) ? Cd{f[(2 + i) * ld], f[(2 + ncv + i) * ld]} : Cd{0.0, 0.0};
#pragma unroll
        for (int i = 0; i < NL; ++i) {
            if (!H2 || i >= nl) break;
#pragma unroll
            for (int k = 0; k < NL; ++k) {
                if (k >= nl) break;
                const int lo = i < k ? i : k, hi = i < k ? k : i;
                const int h = lo * ncv - (lo * (lo - 1)) / 2 + (hi - lo);
                const Cd Fik{f[(2 + 2 * ncv + h) * ld], f[(2 + 2 * ncv + nh + h) * ld]};
                // d2(psi F) = psi_ik F + psi_i F_k + psi_k F_i + psi F_ik, psi_i = -P1_i
                Mm[i * NL + k] = ca(cs(cs(cm(Mm[i * NL + k], F), cm(L[i], Fi[k])), cm(L[k], Fi[i])), cm(psi, Fik));
            }
        }
#pragma unroll
        for (int i = 0; i < NL; ++i)
            if (i < nl) L[i] = cs(cm(L[i], F), cm(psi, Fi[i]));   // -(psi F)_i = P1_i F - psi F_i
        psi = cm(psi, F);
    }
}

// F = K psi, dF_i = K_i psi - K P1_i, d2F_ik = K_ik psi - K_i P1_k - K_k P1_i + K P2_ik
__device__ inline Cd amp_F(double K, Cd psi) { return csc(psi, K); }
__device__ inline Cd amp_dF(const HessArgs& a, int rec, double K, Cd psi, const Cd* P1, int i) {
    return cs(csc(psi, rec_g(a, rec, i)), csc(P1[i], K));
}
template <int NL>
__device__ inline Cd amp_d2F(const HessArgs& a, int rec, double K, Cd psi, const Cd* P1, const Cd* P2, int i, int k) {
    Cd t = csc(psi, rec_h(a, rec, i, k));
    t = cs(t, csc(P1[k], rec_g(a, rec, i)));
    t = cs(t, csc(P1[i], rec_g(a, rec, k)));
    return ca(t, csc(P2[i * NL + k], K));
}

// column sums over a wave's 64 events, taken inside k_hess_rows.  Every lane hands in the
// same sequence of columns; the wave stages SR_K values per lane in its LDS slab (rows
// padded to SR_K + 1 doubles against bank conflicts) and then lane 4k + q adds column k of
// lanes 16q .. 16q + 15 in lane order, and the four quarters combine by two exchanges: a
// fixed order.  The wave's sum of column c is stored at out[c * ld].
constexpr int HR_THREADS = 128;
constexpr int SR_K = 16;
struct WaveColSums {
    double (*slab)[SR_K + 1];
    double* out;
    long long ld;
    int lane, n, col;
    __device__ void put(double v) {
        slab[lane][n] = v;
        if (++n == SR_K) flush();
    }
    __device__ void flush() {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const int k = lane >> 2, q = lane & 3;
        double s = 0.0;
#pragma unroll
        for (int t = 0; t < 16; ++t) s += slab[16 * q + t][k];
        s += __shfl_xor(s, 1, 64);
        s += __shfl_xor(s, 2, 64);
        if (q == 0 && k < n) out[(long long)(col + k) * ld] = s;
        col += n;
        n = 0;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
};

// one event per thread.  NL: block-size class (max local variables, 1/2/4/8); NPT: the
// number of polarisations when it is 1..4 (S in registers), 0 for any other (S in
// private memory, runtime bounds).  Lanes past the chunk's end repeat its last event,
// store nothing and hand zeros to the column sums, so every lane of a wave reaches them.
template <int NL, int NPT>
__global__ void __launch_bounds__(HR_THREADS) k_hess_rows(HessArgs a, long long e0, long long n) {
    __shared__ double slab[HR_THREADS / 64][64][SR_K + 1];
    const long long r0 = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = r0 < n;
    const long long r = live ? r0 : n - 1;
    const long long e = e0 + r;
    const int wv = threadIdx.x >> 6;
    WaveColSums cs{slab[wv], a.cs + (long long)blockIdx.x * (HR_THREADS / 64) + wv, a.cs_ld, (int)(threadIdx.x & 63), 0,
                   0};
    constexpr int SP = NPT > 0 ? NPT : MAXPOL;
    const int NP = NPT > 0 ? NPT : a.npol;
    const int C = a.C, na = a.namp;
    const long long ldJ = a.ld * NP;
    // column c, polarisation p of this event: Jc[c * ldJ + p * n]
    double* Jc = a.J + r;
    Cd S[SP];
#pragma unroll
    for (int p = 0; p < SP; ++p) S[p] = {0.0, 0.0};
    if (live)
        for (int p = 0; p < NP; ++p)
            for (int c = na; c < C; ++c) {
                Jc[c * ldJ + p * n] = 0.0;
                Jc[(C + c) * ldJ + p * n] = 0.0;
            }
    Cd psi, L[NL], Mm[NL * NL];
    // pass 1: A columns, S, dS/dtheta columns.  For small blocks (NL * SP <= 4) a block's
    // dS/dtheta entries are loaded once, added over its amplitudes in order and stored once
    // (the same additions as adding into memory amplitude by amplitude)
    constexpr bool JREG = NL * SP <= 4;
    for (int b = 0; b < a.nblock; ++b) {
        const DBlock& B = a.blocks[b];
        if (B.nsl == 0) continue;
        block_terms<NL, false>(a, b, e, psi, L, Mm);
        Cd dJ[JREG ? NL * SP : 1];
        if constexpr (JREG) {
#pragma unroll
            for (int i = 0; i < NL; ++i)
#pragma unroll
                for (int p = 0; p < SP; ++p)
                    dJ[i * SP + p] = (live && i < B.nloc && p < NP)
                                         ? Cd{Jc[(na + B.vars[i]) * ldJ + p * n], Jc[(C + na + B.vars[i]) * ldJ + p * n]}
                                         : Cd{0.0, 0.0};
        }
        for (int c = 0; c < B.nsl; ++c) {
            int am = B.site + c;
            const DAmp& A = a.amps[am];
            double K = rec_v(a, A.rec);
            Cd F = amp_F(K, psi);
            Cd v{a.table[a.off_v + 2 * am], a.table[a.off_v + 2 * am + 1]};
            Cd dF[NL];
#pragma unroll
            for (int i = 0; i < NL; ++i) dF[i] = (i < B.nloc) ? cm(v, amp_dF(a, A.rec, K, psi, L, i)) : Cd{0.0, 0.0};
#pragma unroll
            for (int p = 0; p < SP; ++p) {
                if (p >= NP) break;
                double2 s = ev_sl(a.st, (long long)A.u * NP + p, e);
                Cd sp{s.x, s.y};
                Cd Aap = cm(F, sp);
                S[p] = ca(S[p], cm(v, Aap));
                if (live) {
                    Jc[am * ldJ + p * n] = Aap.r;
                    Jc[(C + am) * ldJ + p * n] = Aap.i;
                }
#pragma unroll
                for (int i = 0; i < NL; ++i) {
                    if (i >= B.nloc) break;
                    Cd d = cm(dF[i], sp);
                    if constexpr (JREG) {
                        dJ[i * SP + p] = ca(dJ[i * SP + p], d);
                    } else if (live) {
                        int j = B.vars[i];
                        Jc[(na + j) * ldJ + p * n] += d.r;
                        Jc[(C + na + j) * ldJ + p * n] += d.i;
                    }
                }
            }
        }
        if constexpr (JREG) {
            if (live)
#pragma unroll
                for (int i = 0; i < NL; ++i)
#pragma unroll
                    for (int p = 0; p < SP; ++p)
                        if (i < B.nloc && p < NP) {
                            Jc[(na + B.vars[i]) * ldJ + p * n] = dJ[i * SP + p].r;
                            Jc[(C + na + B.vars[i]) * ldJ + p * n] = dJ[i * SP + p].i;
                        }
        }
    }
    double I = 0.0;
#pragma unroll
    for (int p = 0; p < SP; ++p) {
        if (p >= NP) break;
        I += S[p].r * S[p].r + S[p].i * S[p].i;
    }
    double w = (a.cls == 0) ? 1.0 : (a.wts ? a.wts[e] : a.wconst);
    double alpha, beta;
    if (a.cls == 0) { alpha = 1.0; beta = 0.0; }
    else if (I == 0.0) { alpha = 0.0; beta = 0.0; }
    else { alpha = w / I; beta = -w / (I * I); }
    if (live) {
        a.W[r] = alpha;
        a.W[a.ld + r] = beta;
    }
    // pass 2: u = dI/dx (column k of this event: Ur[k * ld]) and the second-part scalars,
    // which only enter through their alpha-weighted sums over the events
    double* Ur = a.U + r;
    const long long ld = a.ld;
    if (live)
        for (int j = 0; j < a.ntheta; ++j) Ur[(2 * na + j) * ld] = 0.0;
    for (int b = 0; b < a.nblock; ++b) {
        const DBlock& B = a.blocks[b];
        if (B.nsl == 0) continue;
        block_terms<NL>(a, b, e, psi, L, Mm);
        double Rb[NL * NL];
#pragma unroll
        for (int q = 0; q < NL * NL; ++q) Rb[q] = 0.0;
        // this block's dI/dtheta columns: loaded once, added over its amplitudes in order,
        // stored once (the same additions as adding into memory amplitude by amplitude)
        double ut[NL];
#pragma unroll
        for (int i = 0; i < NL; ++i) ut[i] = (live && i < B.nloc) ? Ur[(2 * na + B.vars[i]) * ld] : 0.0;
        for (int c = 0; c < B.nsl; ++c) {
            int am = B.site + c;
            const DAmp& A = a.amps[am];
            double K = rec_v(a, A.rec);
            Cd F = amp_F(K, psi);
            Cd v{a.table[a.off_v + 2 * am], a.table[a.off_v + 2 * am + 1]};
            // T_u = sum_p conj(sl) S ; conj(T_u) = <S, sl>
            Cd T{0.0, 0.0};
#pragma unroll
            for (int p = 0; p < SP; ++p) {
                if (p >= NP) break;
                double2 s = ev_sl(a.st, (long long)A.u * NP + p, e);
                T = ca(T, cm(Cd{s.x, -s.y}, S[p]));
            }
            Cd cT = ccj(T);
            Cd g = cm(ccj(F), T);                     // dI/dv_a* = <A_a, S>
            if (live) {
                Ur[am * ld] = 2.0 * g.r;
                Ur[(na + am) * ld] = 2.0 * g.i;
            }
#pragma unroll
            for (int i = 0; i < NL; ++i) {
                if (i >= B.nloc) break;
                Cd dF = amp_dF(a, A.rec, K, psi, L, i);
                Cd q = cm(dF, cT);                    // <S, dA_a/dtheta_i>
                cs.put(live ? alpha * q.r : 0.0);
                cs.put(live ? alpha * q.i : 0.0);
                Cd vq = cm(v, q);                     // contributes to dI/dtheta
                ut[i] += 2.0 * vq.r;
#pragma unroll
                for (int k = 0; k < NL; ++k) {
                    if (k < i) continue;
                    if (k >= B.nloc) break;
                    Cd d2 = cm(v, cm(amp_d2F<NL>(a, A.rec, K, psi, L, Mm, i, k), cT));
                    Rb[i * NL + k] += d2.r;
                }
            }
        }
#pragma unroll
        for (int i = 0; i < NL; ++i) {
            if (i >= B.nloc) break;
#pragma unroll
            for (int k = 0; k < NL; ++k) {
                if (k < i) continue;
                if (k >= B.nloc) break;
                cs.put(live ? alpha * Rb[i * NL + k] : 0.0);
            }
        }
        if (live)
#pragma unroll
            for (int i = 0; i < NL; ++i)
                if (i < B.nloc) Ur[(2 * na + B.vars[i]) * ld] = ut[i];
    }
    cs.put(live ? alpha * I : 0.0);
    // then the columns of u, unweighted, read back eight at a time (one load latency each)
    for (int k0 = 0; k0 < a.NU; k0 += 8) {
        double u[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) u[j] = (live && k0 + j < a.NU) ? Ur[(k0 + j) * ld] : 0.0;
#pragma unroll
        for (int j = 0; j < 8; ++j)
            if (k0 + j < a.NU) cs.put(u[j]);
    }
    if (cs.n > 0) cs.flush();
}

// weighted Gram of a column-major matrix X[col][row] (column stride ldx, rows
// [0, nrows)): part[wg][entry] = sum_rows w(row) X[k][row] X[l][row] for the entries
// [ent0, ent0 + nent) of the upper-triangle list kl[]; the weight of row R is
// Wt[R mod nev] (rows p * nev + r share event r's weight)
constexpr int GR_THREADS = 256;
constexpr int GR_EPT = 8;
constexpr int GR_TILE = 16;

__global__ void __launch_bounds__(GR_THREADS)
k_wgram(const double* __restrict__ X, int ncol, long long ldx, long long nrows, const double* __restrict__ Wt,
        long long nev, const int2* __restrict__ kl, int ent0, int nent, double* __restrict__ part) {
    extern __shared__ double tile[];        // [GR_TILE][ncol] + [GR_TILE] weights
    double* wt = tile + GR_TILE * ncol;
    int nwg = gridDim.x, wg = blockIdx.x;
    long long per = (nrows + nwg - 1) / nwg;
    long long r0 = (long long)wg * per, r1 = r0 + per < nrows ? r0 + per : nrows;
    double acc[GR_EPT];
    int2 ids[GR_EPT];
    for (int q = 0; q < GR_EPT; ++q) {
        acc[q] = 0.0;
        int ent = threadIdx.x + q * GR_THREADS;
        ids[q] = (ent < nent) ? kl[ent0 + ent] : make_int2(-1, -1);
    }
    for (long long t0 = r0; t0 < r1; t0 += GR_TILE) {
        int nt = (int)((r1 - t0) < GR_TILE ? (r1 - t0) : GR_TILE);
        __syncthreads();
        // consecutive threads: consecutive rows of one column
        for (int q = threadIdx.x; q < nt * ncol; q += GR_THREADS) {
            int col = q / nt, rr = q - col * nt;
            tile[rr * ncol + col] = X[(long long)col * ldx + t0 + rr];
        }
        for (int q = threadIdx.x; q < nt; q += GR_THREADS) wt[q] = Wt[(t0 + q) % nev];
        __syncthreads();
        for (int q = 0; q < GR_EPT; ++q) {
            if (ids[q].x < 0) continue;
            double s = 0.0;
            for (int rr = 0; rr < nt; ++rr) s += wt[rr] * tile[rr * ncol + ids[q].x] * tile[rr * ncol + ids[q].y];
            acc[q] += s;
        }
    }
    for (int q = 0; q < GR_EPT; ++q) {
        int ent = threadIdx.x + q * GR_THREADS;
        if (ent < nent) part[(long long)wg * nent + ent] = acc[q];
    }
}

// The same weighted Gram on the matrix cores, for matrices of at most MG_MAXCOL
// columns: v_mfma_f64_16x16x4 accumulates the 16x16 tiles (I <= J) of the padded Gram,
// D[i][j] += sum_k (w_k X[k][16 I + i]) X[k][16 J + j] over 4 rows k per instruction.
// Every wave owns a contiguous range of rows, stages 32-row slabs of all columns in its
// own LDS region (column stride padded by one double against bank conflicts) and keeps
// its tiles in registers; it writes one partial per tile, and k_accum_mfma sums the
// partials in a fixed order.  Lane layout of the instruction (16x16x4, f64; measured on
// the device): A[i][k] in lane 16 k + i, B[k][j] in lane 16 k + j, D[4 r + q][j] in
// component r of lane 16 q + j.
constexpr int MG_WAVES = 4;                  // waves per workgroup
constexpr int MG_ROWS = 32;                  // rows per LDS slab
constexpr int MG_MAXCOL = 64;                // padded columns (4 tiles of 16)
constexpr int MG_MAXPAIR = 10;               // tile pairs I <= J of a 4 x 4 tiling
typedef double mg_acc_t __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(64 * MG_WAVES)
k_wgram_mfma(const double* __restrict__ X, int ncol, long long ldx, long long nrows, const double* __restrict__ Wt,
             long long nev, double* __restrict__ part) {
    extern __shared__ double smem[];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int T = (ncol + 15) >> 4, ncp = T * 16, stride = MG_ROWS + 1;
    const int npair = T * (T + 1) / 2;
    double* xs = smem + (size_t)wave * (ncp * stride + MG_ROWS);
    double* ws = xs + ncp * stride;
    // tile pairs in row-major upper-triangle order: LDS offsets of their A and B columns
    int aoff[MG_MAXPAIR], boff[MG_MAXPAIR];
    {
        int q = 0;
        for (int I = 0; I < 4; ++I)
            for (int J = I; J < 4; ++J) {
                if (q < MG_MAXPAIR) {
                    aoff[q] = (I * 16 + (lane & 15)) * stride;
                    boff[q] = (J * 16 + (lane & 15)) * stride;
                }
                if (I < T && J < T) ++q;
            }
    }
    // this wave's rows: the grid's waves split [0, nrows) into ranges of whole slabs
    const long long nw = (long long)gridDim.x * MG_WAVES;
    const long long gw = (long long)blockIdx.x * MG_WAVES + wave;
    const long long per = ((nrows + nw - 1) / nw + MG_ROWS - 1) / MG_ROWS * MG_ROWS;
    const long long r0 = gw * per < nrows ? gw * per : nrows;
    const long long r1 = r0 + per < nrows ? r0 + per : nrows;
    mg_acc_t acc[MG_MAXPAIR];
#pragma unroll
    for (int q = 0; q < MG_MAXPAIR; ++q) acc[q] = (mg_acc_t){0.0, 0.0, 0.0, 0.0};
    // software pipeline: the next slab is loaded into registers while the current one is
    // multiplied from LDS.  Lane l always carries row (l & 31) of columns 2 it + (l >> 5).
    const int rr = lane & 31, ch = lane >> 5, nit = ncp >> 1;
    double pre[MG_MAXCOL / 2];
    double wpre = 0.0;
    auto load = [&](long long t0) {
        const int nt = (int)((r1 - t0) < MG_ROWS ? (r1 - t0) : MG_ROWS);
#pragma unroll
        for (int it = 0; it < MG_MAXCOL / 2; ++it) {
            const int c = 2 * it + ch;
            pre[it] = (it < nit && c < ncol && rr < nt) ? X[(long long)c * ldx + t0 + rr] : 0.0;
        }
        wpre = (ch == 0 && rr < nt) ? Wt[(t0 + rr) % nev] : 0.0;
    };
    if (r0 < r1) load(r0);
    for (long long t0 = r0; t0 < r1; t0 += MG_ROWS) {
#pragma unroll
        for (int it = 0; it < MG_MAXCOL / 2; ++it)
            if (it < nit) xs[(2 * it + ch) * stride + rr] = pre[it];
        if (ch == 0) ws[rr] = wpre;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        if (t0 + MG_ROWS < r1) load(t0 + MG_ROWS);
#pragma unroll
        for (int k0 = 0; k0 < MG_ROWS; k0 += 4) {
            const int kr = k0 + (lane >> 4);
            const double w = ws[kr];
#pragma unroll
            for (int q = 0; q < MG_MAXPAIR; ++q) {
                if (q < npair) {
                    const double av = w * xs[aoff[q] + kr];
                    const double bv = xs[boff[q] + kr];
                    acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc[q], 0, 0, 0);
                }
            }
        }
        // the slab is consumed before the next one overwrites it
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    }
    // the workgroup's partial: waves 1..3 leave their tiles in LDS (the slab regions are
    // free once every wave has finished), wave 0 adds them in wave order and writes
    // part[(blockIdx.x * npair + q) * 256 + 16 i + j]
    __syncthreads();
    if (wave > 0) {
        double* o = smem + (size_t)(wave - 1) * npair * 256;
#pragma unroll
        for (int q = 0; q < MG_MAXPAIR; ++q)
            if (q < npair) {
#pragma unroll
                for (int r = 0; r < 4; ++r) o[q * 256 + (4 * r + (lane >> 4)) * 16 + (lane & 15)] = acc[q][r];
            }
    }
    __syncthreads();
    if (wave == 0) {
#pragma unroll
        for (int q = 0; q < MG_MAXPAIR; ++q)
            if (q < npair) {
                double* o = part + ((size_t)blockIdx.x * npair + q) * 256;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int e = q * 256 + (4 * r + (lane >> 4)) * 16 + (lane & 15);
                    double v = acc[q][r];
                    for (int w = 1; w < MG_WAVES; ++w) v += smem[(size_t)(w - 1) * npair * 256 + e];
                    o[e - q * 256] = v;
                }
            }
    }
}

// tot[idx(a, b)] += sum over the nwg workgroup partials of k_wgram_mfma, for every
// entry a <= b < ncol of the Gram (packed upper triangle, row-major).  Each workgroup
// takes 64 consecutive tile elements (one per lane, coalesced reads); its four waves add
// the partials g = w, w + 4, ... with eight interleaved sums combined by a fixed tree,
// and wave 0 adds the waves' sums in wave order: a fixed order throughout.
__global__ void __launch_bounds__(256) k_accum_mfma(const double* __restrict__ part, int nwg, int ncol,
                                                    double* __restrict__ tot) {
    __shared__ double sh[4][64];
    const int T = (ncol + 15) >> 4, npair = T * (T + 1) / 2, E = npair * 256;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int e = blockIdx.x * 64 + lane;
    double s[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (e < E) {
        int g = wave;
        for (; g + 28 < nwg; g += 32) {
#pragma unroll
            for (int k = 0; k < 8; ++k) s[k] += part[(size_t)(g + 4 * k) * E + e];
        }
        for (; g < nwg; g += 4) s[0] += part[(size_t)g * E + e];
    }
    sh[wave][lane] = ((s[0] + s[1]) + (s[2] + s[3])) + ((s[4] + s[5]) + (s[6] + s[7]));
    __syncthreads();
    if (wave == 0 && e < E) {
        const double v = ((sh[0][lane] + sh[1][lane]) + sh[2][lane]) + sh[3][lane];
        // element e of tile pair q = (I, J) (row-major upper-triangle order)
        int q = e >> 8, I = 0;
        while (q >= T - I) { q -= T - I; ++I; }
        const int J = I + q;
        const int a = I * 16 + ((e >> 4) & 15), b = J * 16 + (e & 15);
        if (a < ncol && b < ncol && a <= b) tot[(long long)a * ncol - (long long)a * (a - 1) / 2 + (b - a)] += v;
    }
}

// tot[c] += the sum over the nw wave sums of column c (cs[c * ld + w]) left by
// k_hess_rows: one workgroup per column; each thread adds its strided share in two
// interleaved sums, the waves combine by exchanges and thread 0 adds the four waves in
// order: a fixed order
__global__ void __launch_bounds__(256) k_accum_cols(const double* __restrict__ cs, long long ld, int nw,
                                                    double* __restrict__ tot) {
    __shared__ double sh[4];
    const double* p = cs + (long long)blockIdx.x * ld;
    double s0 = 0.0, s1 = 0.0;
    int w = threadIdx.x;
    for (; w + 256 < nw; w += 512) {
        s0 += p[w];
        s1 += p[w + 256];
    }
    if (w < nw) s0 += p[w];
    double s = s0 + s1;
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) tot[blockIdx.x] += (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

// tot[c] += sum_wg part[wg][c]  (fixed order)
__global__ void k_accum_parts(const double* __restrict__ part, int nwg, int ncol, double* __restrict__ tot) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncol) return;
    double s[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    int g = 0;
    for (; g + 8 <= nwg; g += 8) {
#pragma unroll
        for (int k = 0; k < 8; ++k) s[k] += part[(long long)(g + k) * ncol + c];
    }
    for (int k = 0; g < nwg; ++g, ++k) s[k] += part[(long long)g * ncol + c];
    tot[c] += ((s[0] + s[1]) + (s[2] + s[3])) + ((s[4] + s[5]) + (s[6] + s[7]));
}



__device__ inline double gram_at(const double* G, int n2, int k, int l) {
    // index of (min, max) in the row-major upper-triangle list of an n2 x n2 matrix
    int a = k < l ? k : l, b = k < l ? l : k;
    long long idx = (long long)a * n2 - (long long)a * (a - 1) / 2 + (b - a);
    return G[idx];
}

// sum_e alpha d2I[x][y] for one class from its Gram and second-part sums
__device__ double d2I_sum(const HessAsm& h, int c, int x, int y) {
    const int n = h.namp, C = h.C, C2 = 2 * C;
    const double* G = h.G[c];
    // complex Gram entry G[l][k] = sum conj(J_l) J_k
    auto Gre = [&](int l, int k) { return gram_at(G, C2, l, k) + gram_at(G, C2, C + l, C + k); };
    auto Gim = [&](int l, int k) { return gram_at(G, C2, l, C + k) - gram_at(G, C2, C + l, k); };
    // column index and factor (1 or i) of a real parameter
    auto col = [&](int p, int& k, bool& im) {
        if (p < n) { k = p; im = false; }
        else if (p < 2 * n) { k = p - n; im = true; }
        else { k = n + (p - 2 * n); im = false; }
    };
    int kx, ly;
    bool ix, iy;
    col(x, kx, ix);
    col(y, ly, iy);
    // 2 Re(conj(c_y) c_x G[ly][kx])
    double gr = Gre(ly, kx), gi = Gim(ly, kx);
    double v;
    if (!ix && !iy) v = 2.0 * gr;
    else if (!ix && iy) v = 2.0 * gi;
    else if (ix && !iy) v = -2.0 * gi;
    else v = 2.0 * gr;
    return v;
}

// complex Gram Q[l][k] = sum conj(A_l) A_k (n x n, interleaved) from the packed real Gram G
// of the 2n columns [Re A | Im A]
__global__ void k_gram_to_complex(const double* __restrict__ G, int n, double* __restrict__ Q) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n * n) return;
    const int l = q / n, k = q - l * n, n2 = 2 * n;
    Q[2 * q] = gram_at(G, n2, l, k) + gram_at(G, n2, n + l, n + k);
    Q[2 * q + 1] = gram_at(G, n2, l, n + k) - gram_at(G, n2, n + l, k);
}

__global__ void __launch_bounds__(256) k_hess_assemble(HessAsm h) {
    const int n = h.namp, NU = h.NU, P = h.ntheta;
    // phase-space mean
    double Np = h.Np, Neff = h.Wd - h.Wb;
    double pf = (Np > 0) ? h.Zs[0][h.NZ - 1] / Np : 0.0;
    if (!(pf > 1e-30) || isinf(pf)) pf = 1e-30;
    double cp = (Np > 0) ? Neff / (Np * pf) : 0.0;
    double cq = (Np > 0) ? -Neff / (Np * pf * Np * pf) : 0.0;
    for (int q = threadIdx.x; q < NU * NU; q += blockDim.x) {
        int x = q / NU, y = q - x * NU;
        double H = 0.0;
        for (int c = 0; c < 3; ++c) {
            double s = (c == 1) ? -1.0 : 1.0;        // -data + bkg; phase space scaled below
            double v = d2I_sum(h, c, x, y);
            if (c != 0) v += gram_at(h.Uu[c], NU, x, y);
            if (c == 0) H += cp * v + cq * h.gs[0][x] * h.gs[0][y];
            else H += s * v;
        }
        h.Hext[q] = H;
    }
    __syncthreads();
    // second-part terms: (v_a, theta_j) and (theta_j, theta_k)
    if (threadIdx.x == 0) {
        for (int c = 0; c < 3; ++c) {
            double s = (c == 0) ? cp : ((c == 1) ? -1.0 : 1.0);
            const double* Z = h.Zs[c];
            int zq = 0;
            for (int b = 0; b < h.nblock; ++b) {
                const DBlock& B = h.blocks[b];
                if (B.nsl == 0) continue;
                for (int cc = 0; cc < B.nsl; ++cc) {
                    int am = B.site + cc;
                    double vr = h.table[h.off_v + 2 * am], vi = h.table[h.off_v + 2 * am + 1];
                    for (int i = 0; i < B.nloc; ++i) {
                        double qr = Z[zq++], qi = Z[zq++];
                        int j = 2 * n + B.vars[i];
                        // d2S/dRe v_a dtheta = dA_a/dtheta: 2 Re Q ; d2S/dIm v_a dtheta = i dA: -2 Im Q
                        double re = 2.0 * s * qr, im = -2.0 * s * qi;
                        h.Hext[(long long)am * NU + j] += re;
                        h.Hext[(long long)j * NU + am] += re;
                        h.Hext[(long long)(n + am) * NU + j] += im;
                        h.Hext[(long long)j * NU + n + am] += im;
                        (void)vr; (void)vi;
                    }
                }
                for (int i = 0; i < B.nloc; ++i)
                    for (int k = i; k < B.nloc; ++k) {
                        double rv = 2.0 * s * Z[zq++];
                        int j1 = 2 * n + B.vars[i], j2 = 2 * n + B.vars[k];
                        h.Hext[(long long)j1 * NU + j2] += rv;
                        if (j1 != j2) h.Hext[(long long)j2 * NU + j1] += rv;
                    }
            }
        }
        for (int g = 0; g < h.ngauss; ++g) {
            int j = 2 * n + h.g_slot[g];
            h.Hext[(long long)j * NU + j] += 1.0 / (h.g_sig[g] * h.g_sig[g]);
        }
    }
    __syncthreads();
    // fit-parameter layout [Re p, Im p, theta]; v_a = r_a p_f(a)
    const int nf = h.nfree, np = 2 * nf + P;
    for (int q = threadIdx.x; q < np * np; q += blockDim.x) {
        int r = q / np, s = q - r * np;
        double H = 0.0;
        auto sel = [&](int pidx, int* xs, double* rs) -> int {
            int cnt = 0;
            if (pidx >= 2 * nf) { xs[0] = 2 * n + (pidx - 2 * nf); rs[0] = 1.0; return 1; }
            int f = pidx < nf ? pidx : pidx - nf;
            int off = pidx < nf ? 0 : n;
            for (int a = 0; a < n && cnt < 64; ++a)
                if (h.amps[a].freeidx == f) { xs[cnt] = off + a; rs[cnt] = h.amps[a].ratio; ++cnt; }
            return cnt;
        };
        int xr[64], xs[64];
        double rr[64], rs[64];
        int nr = sel(r, xr, rr), ns = sel(s, xs, rs);
        for (int u = 0; u < nr; ++u)
            for (int w = 0; w < ns; ++w) H += rr[u] * rs[w] * h.Hext[(long long)xr[u] * NU + xs[w]];
        h.out[q] = H;
    }
}

}  // namespace ctpwa

using namespace ctpwa;

extern "C" __attribute__((visibility("hidden"))) long long ctpwa_hess_rows_waves(long long n);

template <int NL>
static void launch_rows_np(const HessArgs* a, long long e0, long long n, dim3 g, dim3 b, hipStream_t s) {
    switch (a->npol) {
    case 1: hipLaunchKernelGGL((k_hess_rows<NL, 1>), g, b, 0, s, *a, e0, n); break;
    case 2: hipLaunchKernelGGL((k_hess_rows<NL, 2>), g, b, 0, s, *a, e0, n); break;
    case 3: hipLaunchKernelGGL((k_hess_rows<NL, 3>), g, b, 0, s, *a, e0, n); break;
    case 4: hipLaunchKernelGGL((k_hess_rows<NL, 4>), g, b, 0, s, *a, e0, n); break;
    default: hipLaunchKernelGGL((k_hess_rows<NL, 0>), g, b, 0, s, *a, e0, n); break;
    }
}

extern "C" __attribute__((visibility("hidden"))) hipError_t ctpwa_launch_hess_rows(const HessArgs* a, long long e0,
                                                                                   long long n, int maxloc, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    if (ctpwa_hess_rows_waves(n) > a->cs_ld) return hipErrorInvalidValue;
    dim3 g((unsigned)((n + HR_THREADS - 1) / HR_THREADS)), b(HR_THREADS);
    if (maxloc <= 1) launch_rows_np<1>(a, e0, n, g, b, s);
    else if (maxloc <= 2) launch_rows_np<2>(a, e0, n, g, b, s);
    else if (maxloc <= 4) launch_rows_np<4>(a, e0, n, g, b, s);
    else launch_rows_np<8>(a, e0, n, g, b, s);
    return hipGetLastError();
}

__global__ void k_zero(double* __restrict__ p, long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = 0.0;
}

// zero fill by a kernel (not hipMemsetAsync: the call may be captured in a graph)
extern "C" __attribute__((visibility("hidden"))) hipError_t ctpwa_launch_zero(double* p, long long n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    unsigned nb = (unsigned)std::min<long long>(1024, (n + 255) / 256);
    hipLaunchKernelGGL(k_zero, dim3(nb), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

extern "C" __attribute__((visibility("hidden"))) hipError_t ctpwa_launch_wgram(const double* X, int ncol, long long ldx,
                                                                               long long nrows, const double* Wt,
                                                                               long long nev, const int2* kl, int ent0,
                                                                               int nent, double* part, int nwg,
                                                                               hipStream_t s) {
    size_t sh = (size_t)(GR_TILE * ncol + GR_TILE) * sizeof(double);
    hipLaunchKernelGGL(k_wgram, dim3(nwg), dim3(GR_THREADS), sh, s, X, ncol, ldx, nrows, Wt, nev, kl, ent0, nent, part);
    return hipGetLastError();
}

// weighted Gram on the matrix cores (ncol <= MG_MAXCOL): nwg workgroups of MG_WAVES waves,
// partials [nwg * MG_WAVES][pairs][256] in part, summed into tot (packed upper triangle)
extern "C" __attribute__((visibility("hidden"))) int ctpwa_mfma_gram_ok(int ncol) { return ncol <= MG_MAXCOL; }
extern "C" __attribute__((visibility("hidden"))) long long ctpwa_mfma_gram_part_doubles(int ncol, int nwg) {
    const int T = (ncol + 15) / 16;
    return (long long)nwg * (T * (T + 1) / 2) * 256;
}
extern "C" __attribute__((visibility("hidden"))) hipError_t ctpwa_launch_wgram_mfma(const double* X, int ncol, long long ldx,
                                                                                    long long nrows, const double* Wt,
                                                                                    long long nev, const int2* kl, int nent,
                                                                                    double* part, int nwg, double* tot,
                                                                                    hipStream_t s) {
    if (ncol > MG_MAXCOL || ncol <= 0) return hipErrorInvalidValue;
    const int ncp = (ncol + 15) / 16 * 16, T = ncp / 16;
    // slab regions of the waves, or the tiles of waves 1..3 when the workgroup combines them
    size_t sh = std::max((size_t)MG_WAVES * (ncp * (MG_ROWS + 1) + MG_ROWS), (size_t)(MG_WAVES - 1) * (T * (T + 1) / 2) * 256) *
                sizeof(double);
    hipLaunchKernelGGL(k_wgram_mfma, dim3(nwg), dim3(64 * MG_WAVES), sh, s, X, ncol, ldx, nrows, Wt, nev, part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int E = (T * (T + 1) / 2) * 256;
    hipLaunchKernelGGL(k_accum_mfma, dim3((E + 63) / 64), dim3(256), 0, s, part, nwg, ncol, tot);
    return hipGetLastError();
}

// waves of k_hess_rows over n events: the row count of its column-sum partials
extern "C" __attribute__((visibility("hidden"))) long long ctpwa_hess_rows_waves(long long n) {
    return (n + HR_THREADS - 1) / HR_THREADS * (HR_THREADS / 64);
}

// tot[c] += the column sums left by k_hess_rows over n events, for the ncol columns
extern "C" __attribute__((visibility("hidden"))) hipError_t ctpwa_launch_accum_cols(const double* cs, long long ld,
                                                                                    long long n, int ncol, double* tot,
                                                                                    hipStream_t s) {
    const long long nw = ctpwa_hess_rows_waves(n);
    if (ncol <= 0 || nw <= 0) return hipSuccess;
    if (nw > ld) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_accum_cols, dim3(ncol), dim3(256), 0, s, cs, ld, (int)nw, tot);
    return hipGetLastError();
}

extern "C" __attribute__((visibility("hidden"))) hipError_t ctpwa_launch_accum(const double* part, int nwg, int ncol,
                                                                               double* tot, hipStream_t s) {
    hipLaunchKernelGGL(k_accum_parts, dim3((ncol + 255) / 256), dim3(256), 0, s, part, nwg, ncol, tot);
    return hipGetLastError();
}

extern "C" __attribute__((visibility("hidden"))) hipError_t ctpwa_launch_gram_to_complex(const double* G, int n, double* Q,
                                                                                        hipStream_t s) {
    hipLaunchKernelGGL(k_gram_to_complex, dim3((n * n + 255) / 256), dim3(256), 0, s, G, n, Q);
    return hipGetLastError();
}

extern "C" __attribute__((visibility("hidden"))) hipError_t ctpwa_launch_hess_assemble(const HessAsm* h, hipStream_t s) {
    hipLaunchKernelGGL(k_hess_assemble, dim3(1), dim3(256), 0, s, *h);
    return hipGetLastError();
}
