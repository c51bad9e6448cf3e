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
#include <cstdlib>

#include "device.hpp"

namespace ctpwa {



// Weighted Gram of a column-major matrix X[col][row] (column stride ldx, rows
// [0, nrows)) on the matrix cores, for any number of columns:
//   G[a][b] = sum_rows w(row) X[a][row] X[b][row],  w(row) = Wt[row mod nev]
// (rows p * nev + r share event r's weight).  Columns are grouped in panels of 64 (four
// 16-column tiles); workgroup (x, y) takes panel pair y = (P, Q), P <= Q, and share x of
// the rows.  Each of its four waves owns a contiguous range of the share, stages R-row
// slabs of the pair's columns in its own LDS region (register prefetch of the next slab,
// column stride padded by one element against bank conflicts) and accumulates the pair's
// 16x16 tile products (I <= J on a diagonal pair) in registers with
// v_mfma_f64_16x16x4 (T = double) or v_mfma_f32_16x16x4 (T = float: the mixed form's
// FP32 Jacobian, products summed in FP32 over one slab, then added into FP64
// accumulators).  The waves combine their tiles through LDS in wave order and the
// workgroup writes (or, for a later chunk of rows, adds) one partial per tile pair;
// k_accum_panels sums the partials in a fixed order.
// Lane layouts (16x16x4, one block): A[i][k] in lane 16 k + i, B[k][j] in lane 16 k + j;
// D[4 r + q][j] (f64) or D[4 q + r][j] (f32) in component r of lane 16 q + j.
constexpr int GP_WAVES = 4;
constexpr int GP_PAIRS = 16;                 // tile pairs of two different panels
template <typename T> struct GP;
template <> struct GP<double> {
    static constexpr int R = 16;
    typedef double acc_t __attribute__((ext_vector_type(4)));
    __device__ static inline int row(int lane, int r) { return 4 * r + (lane >> 4); }
};
template <> struct GP<float> {
    static constexpr int R = 32;
    typedef float acc_t __attribute__((ext_vector_type(4)));
    __device__ static inline int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};
typedef double gp_acc64 __attribute__((ext_vector_type(4)));

// LDS bytes of k_gram_panels<T, NT, FULL>: per wave a slab of the pair's columns, or the
// combined tile pairs
inline size_t gram_panels_lds(bool f32, int NT, bool full) {
    const int R = f32 ? 32 : 16;
    const size_t es = f32 ? 4 : 8;
    const int cols = full ? 128 : NT * 16;
    const int npair = full ? NT * NT : NT * (NT + 1) / 2;
    size_t slabs = (size_t)GP_WAVES * cols * (R + 1) * es;
    size_t comb = (size_t)npair * 256 * sizeof(double);
    return std::max(slabs, comb);
}

// NT: tiles per panel; FULL: all NT x NT tile pairs of the panel pair (a Gram of several
// panels: a diagonal pair then also computes its tiles below the diagonal, which the
// accumulation skips), else the NT (NT + 1) / 2 pairs I <= J (a Gram of one panel).  With
// the pair count fixed at compile time every pair's product is unconditional.
template <typename T, int NT, bool FULL>
__global__ void __launch_bounds__(64 * GP_WAVES)
k_gram_panels(const T* __restrict__ X, int ncol, long long ldx, long long nrows, const double* __restrict__ Wt,
              long long nev, const int4* __restrict__ pp, int ntp, double* __restrict__ part, int accumulate) {
    constexpr int R = GP<T>::R, ST = R + 1;
    constexpr int NPAIR = FULL ? NT * NT : NT * (NT + 1) / 2;
    typedef typename GP<T>::acc_t acc_t;
    extern __shared__ __attribute__((aligned(16))) unsigned char gp_smem[];
    __shared__ double wsh[GP_WAVES][R];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int4 pq = pp[blockIdx.y];                 // panels P, Q; first tile pair; tile pairs
    const int P = pq.x, Q = pq.y, tp0 = pq.z;
    const bool diag = P == Q;
    const int wP = min(64, ncol - 64 * P), wQ = min(64, ncol - 64 * Q);
    const int ncl = diag ? 64 : 128;
    constexpr int NCOLS = FULL ? 128 : NT * 16;
    T* xs = (T*)gp_smem + (size_t)wave * NCOLS * ST;
    double* ws = wsh[wave];
    const int boff0 = diag ? 0 : 64;
    // this wave's rows: the grid's waves split [0, nrows) into ranges of whole slabs
    const long long nw = (long long)gridDim.x * GP_WAVES;
    const long long gw = (long long)blockIdx.x * GP_WAVES + wave;
    const long long per = ((nrows + nw - 1) / nw + R - 1) / R * R;
    const long long r0 = gw * per < nrows ? gw * per : nrows;
    const long long r1 = r0 + per < nrows ? r0 + per : nrows;
    gp_acc64 acc[NPAIR];
#pragma unroll
    for (int q = 0; q < NPAIR; ++q) acc[q] = (gp_acc64){0.0, 0.0, 0.0, 0.0};
    // 16-byte loads: lane l carries rows VW (l mod LPC) .. + VW - 1 of column CPI it + l / LPC
    // (the chunk's leading dimension is a multiple of 4, so every such run is aligned)
    constexpr int VW = 16 / (int)sizeof(T), LPC_ = R / VW, CPI = 64 / LPC_;
    constexpr int NLD = (FULL ? 128 : NT * 16) / CPI;
    typedef T vec_t __attribute__((ext_vector_type(VW)));
    const int rr0 = (lane % LPC_) * VW, ch = lane / LPC_;
    const int rw = lane % R;
    // event of row t + rw is (t + rw) mod nev: one division for the wave's first slab, then
    // the base of the next slab to load advances by R with subtractions (no 64-bit division
    // per slab)
    long long eload = r0 < r1 ? r0 % nev : 0;
    auto gcol = [&](int c) -> int {           // global column of LDS column c, or -1
        if (c < 64) return c < wP ? 64 * P + c : -1;
        return (c - 64) < wQ ? 64 * Q + (c - 64) : -1;
    };
    auto load = [&](vec_t (&pre)[NLD], double& wpre, long long t0) {
        const int nt = (int)((r1 - t0) < R ? (r1 - t0) : R);
#pragma unroll
        for (int it = 0; it < NLD; ++it) {
            const int c = CPI * it + ch;
            const int g = c < ncl ? gcol(c) : -1;
            const T* src = X + (long long)(g < 0 ? 0 : g) * ldx + t0 + rr0;
            if (g >= 0 && rr0 + VW <= nt) {
                pre[it] = *(const vec_t*)src;
            } else {
#pragma unroll
                for (int v = 0; v < VW; ++v) pre[it][v] = (g >= 0 && rr0 + v < nt) ? src[v] : (T)0;
            }
        }
        long long ev = eload + rw;
        while (ev >= nev) ev -= nev;              // rows that wrap to the next polarisation
        wpre = (lane < R && rw < nt) ? Wt[ev] : 0.0;
        eload += R;
        while (eload >= nev) eload -= nev;        // once at most unless nev < R
    };
    // one slab: its registers to the wave's LDS region, the refill of those registers with
    // the slab at tnext (loads in flight during this slab's products), then the products
    auto slab = [&](vec_t (&pre)[NLD], double& wpre, long long tnext) {
#pragma unroll
        for (int it = 0; it < NLD; ++it) {
            const int c = CPI * it + ch;
            if (c < ncl) {
#pragma unroll
                for (int v = 0; v < VW; ++v) xs[c * ST + rr0 + v] = pre[it][v];
            }
        }
        if (lane < R) ws[rw] = wpre;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        if (tnext < r1) load(pre, wpre, tnext);
        const int li = lane & 15;
        if constexpr (sizeof(T) == 8) {
#pragma unroll
            for (int k0 = 0; k0 < R; k0 += 4) {
                const int kr = k0 + (lane >> 4);
                const double w = ws[kr];
                int q = 0;
#pragma unroll
                for (int I = 0; I < NT; ++I) {
                    const double av = w * (double)xs[(I * 16 + li) * ST + kr];
#pragma unroll
                    for (int J = FULL ? 0 : I; J < NT; ++J, ++q)
                        acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, (double)xs[(boff0 + J * 16 + li) * ST + kr], acc[q],
                                                                      0, 0, 0);
                }
            }
        } else {
            acc_t t[NPAIR];
#pragma unroll
            for (int q = 0; q < NPAIR; ++q) t[q] = (acc_t){0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int k0 = 0; k0 < R; k0 += 4) {
                const int kr = k0 + (lane >> 4);
                const float w = (float)ws[kr];
                int q = 0;
#pragma unroll
                for (int I = 0; I < NT; ++I) {
                    const float av = w * xs[(I * 16 + li) * ST + kr];
#pragma unroll
                    for (int J = FULL ? 0 : I; J < NT; ++J, ++q)
                        t[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, (float)xs[(boff0 + J * 16 + li) * ST + kr], t[q], 0, 0, 0);
                }
            }
#pragma unroll
            for (int q = 0; q < NPAIR; ++q) {
                acc[q][0] += (double)t[q][0];
                acc[q][1] += (double)t[q][1];
                acc[q][2] += (double)t[q][2];
                acc[q][3] += (double)t[q][3];
            }
        }
        // the slab is consumed before the next one overwrites it
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    };
    // one register buffer: the next slab's loads are in flight while this one is consumed
    // (a second buffer, two slabs ahead, measured no faster at two waves per SIMD)
    vec_t pre[NLD];
    double wpre = 0.0;
    if (r0 < r1) load(pre, wpre, r0);
    for (long long t0 = r0; t0 < r1; t0 += R) slab(pre, wpre, t0 + R);
    // combine the waves in wave order in LDS (the slab regions are free once every wave
    // has finished), element e = 16 i + j of tile pair q at comb[q * 256 + e]
    double* comb = (double*)gp_smem;
    __syncthreads();
    for (int w = 0; w < GP_WAVES; ++w) {
        if (wave == w) {
#pragma unroll
            for (int q = 0; q < NPAIR; ++q) {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int e = q * 256 + GP<T>::row(lane, r) * 16 + (lane & 15);
                    comb[e] = (w == 0) ? acc[q][r] : comb[e] + acc[q][r];
                }
            }
        }
        __syncthreads();
    }
    double* o = part + ((size_t)blockIdx.x * ntp + tp0) * 256;
    for (int e = threadIdx.x; e < NPAIR * 256; e += 64 * GP_WAVES) o[e] = accumulate ? o[e] + comb[e] : comb[e];
}

// tot[idx(a, b)] += the sum over the nwg workgroup partials of k_gram_panels, for every
// entry a <= b < ncol of the Gram (packed upper triangle, row-major).  tiles[q] = (I, J):
// the 16-column tiles of tile pair q.  Each workgroup takes 64 consecutive tile elements
// (one per lane, coalesced reads); its four waves add the partials g = w, w + 4, ...
// with eight interleaved sums combined by a fixed tree, and wave 0 adds the waves' sums
// in wave order: a fixed order throughout.
__global__ void __launch_bounds__(256) k_accum_panels(const double* __restrict__ part, int nwg, int ntp,
                                                      const int2* __restrict__ tiles, int ncol, double* __restrict__ tot) {
    __shared__ double sh[4][64];
    const long long E = (long long)ntp * 256;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const long long e = (long long)blockIdx.x * 64 + lane;
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
        const int2 t = tiles[e >> 8];
        const int a = t.x * 16 + (int)((e >> 4) & 15), b = t.y * 16 + (int)(e & 15);
        if (a < ncol && b < ncol && a <= b) tot[(long long)a * ncol - (long long)a * (a - 1) / 2 + (b - a)] += v;
    }
}

// tot[c] += the sum over the nw wave sums of column c (cs[c * ld + w]) left by
// k_hess_rows: one workgroup per column; each thread adds its strided share in two
// interleaved sums, the waves combine by exchanges and thread 0 adds the four waves in
// order: a fixed order
__global__ void __launch_bounds__(256) k_accum_cols(const double* __restrict__ cs, long long ld, int nw,
                                                    const int* __restrict__ perm, double* __restrict__ tot) {
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
    if (threadIdx.x == 0) tot[perm ? perm[blockIdx.x] : (int)blockIdx.x] += (sh[0] + sh[1]) + (sh[2] + sh[3]);
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

// Hessian assembly, three kernels:
//  k_hess_ext     Hext[x][y] over the extended parameters (Re v, Im v, theta) from the
//                 class Grams, the u Grams and the phase-space normalisation terms,
//                 H = -H_data + H_bkg + N_eff [A2 / (N_p pf) - g g^T / (N_p pf)^2];
//  k_hess_second  adds the second-part terms of the (v, theta) and (theta, theta) blocks
//                 and the Gaussian constraints (one thread, a fixed order);
//  k_hess_map     maps to the fit parameters [Re p, Im p, theta] (v_a = r_a p_f(a)).
__device__ inline void hess_pf(const HessAsm& h, double& cp, double& cq) {
    double Np = h.Np, Neff = h.Wd - h.Wb;
    double pf = (Np > 0) ? h.Zs[0][h.NZ - 1] / Np : 0.0;
    if (!(pf > 1e-30) || isinf(pf)) pf = 1e-30;
    cp = (Np > 0) ? Neff / (Np * pf) : 0.0;
    cq = (Np > 0) ? -Neff / (Np * pf * Np * pf) : 0.0;
}

__global__ void __launch_bounds__(256) k_hess_ext(HessAsm h) {
    const int NU = h.NU;
    double cp, cq;
    hess_pf(h, cp, cq);
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (long long)NU * NU) return;
    const int x = (int)(q / NU), y = (int)(q - (long long)x * NU);
    double H = 0.0;
    for (int c = 0; c < 3; ++c) {
        double v = d2I_sum(h, c, x, y);
        if (c != 0) v += gram_at(h.Uu[c], NU, x, y);
        if (c == 0) H += cp * v + cq * h.gs[0][x] * h.gs[0][y];
        else H += (c == 1 ? -1.0 : 1.0) * v;
    }
    h.Hext[q] = H;
}

// z offset of block b in the class sums: the blocks before it with waves
__device__ inline int hess_zoff(const HessAsm& h, int b) {
    int z = 0;
    for (int q = 0; q < b; ++q) {
        const DBlock& B = h.blocks[q];
        if (B.nsl > 0) z += B.nsl * 2 * B.nloc + B.nloc * (B.nloc + 1) / 2;
    }
    return z;
}

// second-part terms: every (amplitude, local parameter) entry of the (v, theta) blocks
// belongs to one thread (the classes added in a fixed order); the (theta, theta) terms
// of all blocks, which may share parameters, and the Gaussian constraints are gathered
// in LDS by one thread in block order and then added in parallel
constexpr int HS_MAXTH = 64;
__global__ void __launch_bounds__(256) k_hess_second(HessAsm h) {
    __shared__ double tt[HS_MAXTH * HS_MAXTH];
    const int n = h.namp, NU = h.NU, P = h.ntheta;
    double cp, cq;
    hess_pf(h, cp, cq);
    const double sc[3] = {cp, -1.0, 1.0};
    for (int a = threadIdx.x; a < n; a += blockDim.x) {
        const int b = h.amps[a].block;
        const DBlock& B = h.blocks[b];
        if (B.nsl == 0) continue;
        const int z0 = hess_zoff(h, b) + (a - B.site) * 2 * B.nloc;
        for (int i = 0; i < B.nloc; ++i) {
            double re = 0.0, im = 0.0;
            for (int c = 0; c < 3; ++c) {
                // d2S/dRe v_a dtheta = dA_a/dtheta: 2 Re Q ; d2S/dIm v_a dtheta = i dA: -2 Im Q
                re += 2.0 * sc[c] * h.Zs[c][z0 + 2 * i];
                im += -2.0 * sc[c] * h.Zs[c][z0 + 2 * i + 1];
            }
            const int j = 2 * n + B.vars[i];
            h.Hext[(long long)a * NU + j] += re;
            h.Hext[(long long)j * NU + a] += re;
            h.Hext[(long long)(n + a) * NU + j] += im;
            h.Hext[(long long)j * NU + n + a] += im;
        }
    }
    if (P > HS_MAXTH) return;     // rejected on the host
    for (int q = threadIdx.x; q < P * P; q += blockDim.x) tt[q] = 0.0;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int c = 0; c < 3; ++c) {
            int zq = 0;
            for (int b = 0; b < h.nblock; ++b) {
                const DBlock& B = h.blocks[b];
                if (B.nsl == 0) continue;
                zq += B.nsl * 2 * B.nloc;
                for (int i = 0; i < B.nloc; ++i)
                    for (int k = i; k < B.nloc; ++k) {
                        const double rv = 2.0 * sc[c] * h.Zs[c][zq++];
                        const int j1 = B.vars[i], j2 = B.vars[k];
                        tt[j1 * P + j2] += rv;
                        if (j1 != j2) tt[j2 * P + j1] += rv;
                    }
            }
        }
        for (int g = 0; g < h.ngauss; ++g) {
            const int j = h.g_slot[g];
            tt[j * P + j] += 1.0 / (h.g_sig[g] * h.g_sig[g]);
        }
    }
    __syncthreads();
    for (int q = threadIdx.x; q < P * P; q += blockDim.x) {
        const int j1 = q / P, j2 = q - j1 * P;
        h.Hext[(long long)(2 * n + j1) * NU + 2 * n + j2] += tt[q];
    }
}

__global__ void __launch_bounds__(256) k_hess_map(HessAsm h) {
    const int np = 2 * h.nfree + h.ntheta, NU = h.NU;
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (long long)np * np) return;
    const int r = (int)(q / np), c = (int)(q - (long long)r * np);
    double H = 0.0;
    for (int u = h.map_ptr[r]; u < h.map_ptr[r + 1]; ++u)
        for (int w = h.map_ptr[c]; w < h.map_ptr[c + 1]; ++w)
            H += h.map_r[u] * h.map_r[w] * h.Hext[(long long)h.map_x[u] * NU + h.map_x[w]];
    h.out[q] = H;
}

}  // namespace ctpwa

using namespace ctpwa;

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

// panel pairs (P, Q, first tile pair, tile pairs) and the tile-pair list (I, J) of the
// matrix-core Gram of ncol columns; returns the number of tile pairs
extern "C" __attribute__((visibility("hidden"))) int ctpwa_gram_tables(int ncol, int4* pp, int* npp, int2* tiles) {
    const int npan = (ncol + 63) / 64;
    const int T1 = (ncol + 15) / 16;        // tiles of a one-panel Gram
    int q = 0, y = 0;
    for (int P = 0; P < npan; ++P)
        for (int Q = P; Q < npan; ++Q) {
            const int q0 = q;
            // one panel: the pairs I <= J of its tiles; several: all 4 x 4 pairs of a panel
            // pair (tiles past the last column give entries the accumulation skips)
            for (int I = 0; I < 4; ++I)
                for (int J = 0; J < 4; ++J) {
                    const bool use = (npan == 1) ? (I < T1 && J < T1 && I <= J) : true;
                    if (!use) continue;
                    if (tiles) tiles[q] = make_int2(4 * P + I, 4 * Q + J);
                    ++q;
                }
            if (pp) pp[y] = make_int4(P, Q, q0, q - q0);
            ++y;
        }
    if (npp) *npp = y;
    return q;
}

// workgroups along the rows for a Gram of npp panel pairs (about 2048 in all)
extern "C" __attribute__((visibility("hidden"))) int ctpwa_gram_nwg(int npp) {
    // environment CTPWA_GRAM_WGS overrides the per-launch workgroup budget (measurements)
    static const int budget = [] {
        const char* e = std::getenv("CTPWA_GRAM_WGS");
        return e ? std::max(32, std::atoi(e)) : 2048;
    }();
    const int cap = std::max(512, budget / 4);
    return std::max(32, std::min(cap, budget / std::max(1, npp)));
}

// weighted Gram on the matrix cores: partials [nwg][ntp][256] written (accumulate = 0)
// or added (accumulate = 1) in part; FP64 input, or FP32 input (f32 = 1)
extern "C" __attribute__((visibility("hidden"))) hipError_t ctpwa_launch_gram_panels(const void* X, int f32, int ncol,
                                                                                     long long ldx, long long nrows,
                                                                                     const double* Wt, long long nev,
                                                                                     const int4* pp, int npp, int ntp,
                                                                                     double* part, int nwg, int accumulate,
                                                                                     hipStream_t s) {
    if (ncol <= 0 || nrows <= 0) return hipSuccess;
    const dim3 g(nwg, npp), b(64 * GP_WAVES);
#define GP_LAUNCH(TT, NTT, FF)                                                                                   \
    hipLaunchKernelGGL((k_gram_panels<TT, NTT, FF>), g, b, gram_panels_lds(f32 != 0, NTT, FF), s, (const TT*)X, ncol, \
                       ldx, nrows, Wt, nev, pp, ntp, part, accumulate)
    const int T1 = (ncol + 15) / 16;
    if (ncol > 64) {
        if (f32) GP_LAUNCH(float, 4, true); else GP_LAUNCH(double, 4, true);
    } else if (f32) {
        switch (T1) {
        case 1: GP_LAUNCH(float, 1, false); break;
        case 2: GP_LAUNCH(float, 2, false); break;
        case 3: GP_LAUNCH(float, 3, false); break;
        default: GP_LAUNCH(float, 4, false); break;
        }
    } else {
        switch (T1) {
        case 1: GP_LAUNCH(double, 1, false); break;
        case 2: GP_LAUNCH(double, 2, false); break;
        case 3: GP_LAUNCH(double, 3, false); break;
        default: GP_LAUNCH(double, 4, false); break;
        }
    }
#undef GP_LAUNCH
    return hipGetLastError();
}

// tot += the Gram from the partials of k_gram_panels (packed upper triangle)
extern "C" __attribute__((visibility("hidden"))) hipError_t ctpwa_launch_accum_panels(const double* part, int nwg, int ntp,
                                                                                      const int2* tiles, int ncol,
                                                                                      double* tot, hipStream_t s) {
    const long long E = (long long)ntp * 256;
    hipLaunchKernelGGL(k_accum_panels, dim3((unsigned)((E + 63) / 64)), dim3(256), 0, s, part, nwg, ntp, tiles, ncol, tot);
    return hipGetLastError();
}

// waves of the Hessian rows kernel (256 threads per workgroup) over n events: the row
// count of its column-sum partials
extern "C" __attribute__((visibility("hidden"))) long long ctpwa_hess_rows_waves(long long n) {
    return (n + 255) / 256 * 4;
}

// tot[perm[c]] += the column sums left by the rows kernel over n events, for its ncol
// summed columns (perm: the rows kernel's put order -> the assembly's order)
extern "C" __attribute__((visibility("hidden"))) hipError_t ctpwa_launch_accum_cols(const double* cs, long long ld,
                                                                                    long long n, int ncol, const int* perm,
                                                                                    double* tot, hipStream_t s) {
    const long long nw = ctpwa_hess_rows_waves(n);
    if (ncol <= 0 || nw <= 0) return hipSuccess;
    if (nw > ld) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_accum_cols, dim3(ncol), dim3(256), 0, s, cs, ld, (int)nw, perm, tot);
    return hipGetLastError();
}

extern "C" __attribute__((visibility("hidden"))) hipError_t ctpwa_launch_gram_to_complex(const double* G, int n, double* Q,
                                                                                        hipStream_t s) {
    hipLaunchKernelGGL(k_gram_to_complex, dim3((n * n + 255) / 256), dim3(256), 0, s, G, n, Q);
    return hipGetLastError();
}

extern "C" __attribute__((visibility("hidden"))) hipError_t ctpwa_launch_hess_assemble(const HessAsm* h, hipStream_t s) {
    const long long ne = (long long)h->NU * h->NU;
    hipLaunchKernelGGL(k_hess_ext, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, s, *h);
    if (h->ntheta > HS_MAXTH) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_hess_second, dim3(1), dim3(256), 0, s, *h);
    const int np = 2 * h->nfree + h->ntheta;
    const long long nm = (long long)np * np;
    hipLaunchKernelGGL(k_hess_map, dim3((unsigned)((nm + 255) / 256)), dim3(256), 0, s, *h);
    return hipGetLastError();
}
