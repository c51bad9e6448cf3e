// Per-evaluation theta-only work and the final assembly.
//
//  k_prep      params -> coefficient table: couplings v_a, c_a = v_a K_a, amplitude
//              numerators K_a and every propagator-denominator coefficient, each with
//              its first and second derivatives over the block's free parameters
//              (forward-mode jets).  One thread per item; runs in microseconds.
//  k_reduce, k_assemble  per-workgroup partial sums -> NLL and gradient in the fit-parameter
//              layout [Re p, Im p, theta], following the assembly of the reference's
//              NLLFunction::forward/backward (src/Analysis.cu:1134-1261): phase-space
//              mean, N_eff log term, NaN/Inf -> 1e30 with zero gradient, Gaussian
//              constraints (src/Parameters.cu:769-806).
//  k_amplitudes / k_intensity   per-event amplitudes and intensities (diagnostics,
//              event generation).
#include "device.hpp"

namespace ctpwa {

// ---------------------------------------------------------------------------
// forward-mode jet over NV block-local variables: value, gradient and (when H)
// Hessian.  NV is the smallest of 1/2/4/8 that holds the model's largest block, so
// the jets stay in registers.
// ---------------------------------------------------------------------------
template <int NV, bool H>
struct DJ {
    double v, g[NV], h[H ? NV : 1][H ? NV : 1];
};
template <int NV, bool H>
__device__ inline DJ<NV, H> jc(double c) {
    DJ<NV, H> r;
    r.v = c;
#pragma unroll
    for (int i = 0; i < NV; ++i) r.g[i] = 0.0;
    if constexpr (H) {
#pragma unroll
        for (int i = 0; i < NV; ++i)
#pragma unroll
            for (int k = 0; k < NV; ++k) r.h[i][k] = 0.0;
    }
    return r;
}
template <int NV, bool H>
__device__ inline DJ<NV, H> jv(double c, int i) {
    DJ<NV, H> r = jc<NV, H>(c);
#pragma unroll
    for (int k = 0; k < NV; ++k) if (k == i) r.g[k] = 1.0;
    return r;
}
template <int NV, bool H>
__device__ inline DJ<NV, H> jadd(const DJ<NV, H>& a, const DJ<NV, H>& b) {
    DJ<NV, H> r;
    r.v = a.v + b.v;
#pragma unroll
    for (int i = 0; i < NV; ++i) r.g[i] = a.g[i] + b.g[i];
    if constexpr (H) {
#pragma unroll
        for (int i = 0; i < NV; ++i)
#pragma unroll
            for (int k = 0; k < NV; ++k) r.h[i][k] = a.h[i][k] + b.h[i][k];
    }
    return r;
}
template <int NV, bool H>
__device__ inline DJ<NV, H> jscale(const DJ<NV, H>& a, double c) {
    DJ<NV, H> r;
    r.v = a.v * c;
#pragma unroll
    for (int i = 0; i < NV; ++i) r.g[i] = a.g[i] * c;
    if constexpr (H) {
#pragma unroll
        for (int i = 0; i < NV; ++i)
#pragma unroll
            for (int k = 0; k < NV; ++k) r.h[i][k] = a.h[i][k] * c;
    }
    return r;
}
template <int NV, bool H>
__device__ inline DJ<NV, H> jsub(const DJ<NV, H>& a, const DJ<NV, H>& b) { return jadd(a, jscale(b, -1.0)); }
template <int NV, bool H>
__device__ inline DJ<NV, H> jmul(const DJ<NV, H>& a, const DJ<NV, H>& b) {
    DJ<NV, H> r;
    r.v = a.v * b.v;
#pragma unroll
    for (int i = 0; i < NV; ++i) r.g[i] = a.g[i] * b.v + a.v * b.g[i];
    if constexpr (H) {
#pragma unroll
        for (int i = 0; i < NV; ++i)
#pragma unroll
            for (int k = 0; k < NV; ++k)
                r.h[i][k] = a.h[i][k] * b.v + a.v * b.h[i][k] + a.g[i] * b.g[k] + a.g[k] * b.g[i];
    }
    return r;
}
// f(a) given f, f', f''
template <int NV, bool H>
__device__ inline DJ<NV, H> jfun(const DJ<NV, H>& a, double f, double d1, double d2) {
    DJ<NV, H> r;
    r.v = f;
#pragma unroll
    for (int i = 0; i < NV; ++i) r.g[i] = d1 * a.g[i];
    if constexpr (H) {
#pragma unroll
        for (int i = 0; i < NV; ++i)
#pragma unroll
            for (int k = 0; k < NV; ++k) r.h[i][k] = d1 * a.h[i][k] + d2 * a.g[i] * a.g[k];
    }
    return r;
}
template <int NV, bool H>
__device__ inline DJ<NV, H> jinv(const DJ<NV, H>& a) { double x = a.v; return jfun(a, 1.0 / x, -1.0 / (x * x), 2.0 / (x * x * x)); }
template <int NV, bool H>
__device__ inline DJ<NV, H> jdiv(const DJ<NV, H>& a, const DJ<NV, H>& b) { return jmul(a, jinv(b)); }
template <int NV, bool H>
__device__ inline DJ<NV, H> jsqrt(const DJ<NV, H>& a) { double s = sqrt(a.v); return jfun(a, s, 0.5 / s, -0.25 / (s * a.v)); }
template <int NV, bool H>
__device__ inline DJ<NV, H> jlog(const DJ<NV, H>& a) { return jfun(a, log(a.v), 1.0 / a.v, -1.0 / (a.v * a.v)); }
template <int NV, bool H>
__device__ inline DJ<NV, H> jpowi(const DJ<NV, H>& a, int n) {
    if (n == 0) return jc<NV, H>(1.0);
    if (n == 1) return a;
    double x = a.v, xn2 = 1.0;
    for (int i = 0; i < n - 2; ++i) xn2 *= x;
    double xn1 = xn2 * x, xn = xn1 * x;
    return jfun(a, xn, n * xn1, n * (n - 1.0) * xn2);
}
// N_L(z) with z = q0 d as a jet in q0
template <int NV, bool H>
__device__ inline DJ<NV, H> jbfpoly(int L, const DJ<NV, H>& q0, double d) {
    double c[7];
    int deg = 0;
    switch (L) {
    case 1: c[0] = 1; c[1] = 1; deg = 1; break;
    case 2: c[0] = 9; c[1] = 3; c[2] = 1; deg = 2; break;
    case 3: c[0] = 225; c[1] = 45; c[2] = 6; c[3] = 1; deg = 3; break;
    case 4: c[0] = 11025; c[1] = 1575; c[2] = 135; c[3] = 10; c[4] = 1; deg = 4; break;
    case 5: c[0] = 893025; c[1] = 99225; c[2] = 6300; c[3] = 315; c[4] = 15; c[5] = 1; deg = 5; break;
    case 6: c[0] = 540326025; c[1] = 6185025; c[2] = 363825; c[3] = 17325; c[4] = 630; c[5] = 21; c[6] = 1; deg = 6; break;
    default: return jc<NV, H>(1.0);
    }
    DJ<NV, H> z = jscale(q0, d);
    DJ<NV, H> z2 = jmul(z, z);
    DJ<NV, H> r = jc<NV, H>(c[deg]);          // Horner in z^2
    for (int i = deg - 1; i >= 0; --i) r = jadd(jmul(r, z2), jc<NV, H>(c[i]));
    return r;
}
// breakup momentum (src/AmpGen.cu:2733-2737) and computeQ0AD (include/ResModel.cuh:88-97)
template <int NV, bool H>
__device__ inline DJ<NV, H> jbreakup(const DJ<NV, H>& m, const DJ<NV, H>& m1, const DJ<NV, H>& m2) {
    DJ<NV, H> s = jadd(m1, m2), d = jsub(m1, m2);
    DJ<NV, H> mm = jmul(m, m);
    DJ<NV, H> q2 = jmul(jsub(mm, jmul(s, s)), jsub(mm, jmul(d, d)));
    if (q2.v <= 0.0) return jc<NV, H>(0.0);
    return jdiv(jsqrt(q2), jscale(m, 2.0));
}
template <int NV, bool H>
__device__ inline DJ<NV, H> jq0ad(const DJ<NV, H>& m0, const DJ<NV, H>& m1, const DJ<NV, H>& m2) {
    DJ<NV, H> s = jadd(m1, m2), d = jsub(m1, m2);
    DJ<NV, H> m0s = jmul(m0, m0);
    DJ<NV, H> q2 = jdiv(jmul(jsub(m0s, jmul(s, s)), jsub(m0s, jmul(d, d))), jscale(m0s, 4.0));
    if (q2.v < 0.0) return jc<NV, H>(0.0);
    return jsqrt(q2);
}

__device__ inline int local_index(const DBlock& B, int slot) {
    for (int i = 0; i < B.nloc; ++i) if (B.vars[i] == slot) return i;
    return -1;
}
template <int NV, bool H>
__device__ inline DJ<NV, H> res_param(const DRes& R, int p, const DBlock& B, const double* theta) {
    if (p >= R.np) return jc<NV, H>(0.0);
    int s = R.slot[p];
    if (s < 0) return jc<NV, H>(R.p[p]);
    return jv<NV, H>(theta[s], local_index(B, s));
}
// mass used in q0 for a daughter: configured mass, else the mass parameter of the
// block's resonance that occupies it (src/AmpGen.cu:1664-1731)
template <int NV, bool H>
__device__ inline DJ<NV, H> q0_dmass(double cfg, int ref, const DBlock& B, const DRes* res, const double* theta) {
    if (cfg > 0) return jc<NV, H>(cfg);
    // a final-state particle configured with mass 0: the reference falls back to the
    // event mass here, which is 0 for a massless particle (up to rounding)
    if (ref >= 0) return jc<NV, H>(0.0);
    int kk = -ref - 1;
    for (int r = 0; r < B.nres; ++r) {
        const DRes& R = res[B.res0 + r];
        if (R.node == kk && R.np > 0) return res_param<NV, H>(R, 0, B, theta);
    }
    return jc<NV, H>(1.0);   // unreachable: the host rejects event-dependent q0
}
template <int NV, bool H>
__device__ inline void store(double* t, const DJ<NV, H>& x) {
    t[0] = x.v;
#pragma unroll
    for (int i = 0; i < NV; ++i) t[1 + i] = x.g[i];
    if constexpr (H) {
#pragma unroll
        for (int i = 0; i < NV; ++i)
#pragma unroll
            for (int k = 0; k < NV; ++k) t[1 + NLOC + i * NLOC + k] = x.h[i][k];
    }
}

template <int NV, bool H>
__device__ __forceinline__ void prep_body(const double* __restrict__ gparams, double* __restrict__ table, PrepArgs a) {
    using J = DJ<NV, H>;
    extern __shared__ uint4 smeta[];
    // the parameter vector goes to LDS behind the model tables (meta_bytes is a multiple of
    // 16), loaded in the same phase as the tables, so the item's chain starts with both
    // in LDS instead of waiting for a dependent global load of its parameters
    double* params = (double*)((char*)smeta + a.meta_bytes);
    for (int i = threadIdx.x; i < 2 * a.nfree + a.ntheta; i += blockDim.x) params[i] = gparams[i];
    if (a.meta_bytes > 0) {
        // the model tables are read many times with dependent indices: stage them in LDS
        const uint4* src = (const uint4*)a.meta;
        for (int i = threadIdx.x; i < a.meta_bytes / 16; i += blockDim.x) smeta[i] = src[i];
        const char* base = (const char*)smeta;
        a.blocks = (const DBlock*)(base + a.o_blocks);
        a.res = (const DRes*)(base + a.o_res);
        a.nodes = (const DNode*)(base + a.o_nodes);
        a.topos = (const DTopo*)(base + a.o_topos);
        a.slcL = (const int*)(base + a.o_slcL);
        a.slc0 = (const int*)(base + a.o_slc0);
        a.dens = (const DDen*)(base + a.o_dens);
        a.amps = (const DAmp*)(base + a.o_amps);
    }
    __syncthreads();
    // one item (denominator or amplitude record) per workgroup: the items' chains of
    // dependent jet arithmetic then run on different compute units at the same time instead
    // of one after another in a single divergent wave
    int item = blockIdx.x;
    const double* theta = params + 2 * a.nfree;
    if (blockIdx.x == 0)
        for (int j = threadIdx.x; j < a.ntheta; j += blockDim.x) {
            table[a.off_th + j] = theta[j];
            if (a.remember) table[a.off_lth + j] = theta[j];
        }
    if (threadIdx.x != 0) return;
    if (item < a.nden) {
        const DDen& D = a.dens[item];
        const DBlock& B = a.blocks[D.block];
        const DRes& R = a.res[B.res0 + D.res];
        const DTopo& T = a.topos[B.topo];
        const DNode& n = a.nodes[T.node0 + R.node];
        J m0 = (R.np > 0) ? res_param<NV, H>(R, 0, B, theta) : jc<NV, H>(1.0);
        J q0 = jq0ad(m0, q0_dmass<NV, H>(n.md1, n.d1ref, B, a.res, theta), q0_dmass<NV, H>(n.md2, n.d2ref, B, a.res, theta));
        double* t = table + a.off_rec + (long long)D.rec * REC;
        J m02 = jmul(m0, m0);
        switch (R.model) {
        case 0: {   // BWR: D = m0^2 - s - i g0 m0^2 N_L(q0 d')/q0^(2L+1) * E
            J g0 = res_param<NV, H>(R, 1, B, theta);
            double dp = n.has_bf ? n.bf_d : 0.0;
            J N0 = (dp != 0.0) ? jbfpoly(n.lmin, q0, dp) : jc<NV, H>(1.0);
            J b = jdiv(jmul(jmul(g0, m02), N0), jpowi(q0, 2 * n.lmin + 1));
            store(t, m02);
            store(t + REC, b);
            break;
        }
        case 1: {   // BW: D = m0^2 - s - i m0 g0
            J g0 = res_param<NV, H>(R, 1, B, theta);
            store(t, m02);
            store(t + REC, jmul(m0, g0));
            break;
        }
        case 3: {   // Flatte: D = m0^2 - s + sum b_i rho_im,i - i sum b_i rho_re,i, b_i = m0 g_i / 2
            store(t, m02);
            for (int i = 0; i < R.nch; ++i) store(t + (1 + i) * REC, jscale(jmul(m0, res_param<NV, H>(R, 1 + i, B, theta)), 0.5));
            break;
        }
        case 4: {   // Gounaris-Sakurai (src/ResModel.cu:484-561)
            const double pi = 3.14159265358979323846;
            J g0 = res_param<NV, H>(R, 1, B, theta);
            double mpi = R.nch > 0 ? R.ch[0] : 0.1396;
            J q02 = jmul(q0, q0), q03 = jmul(q02, q0);
            J logA = jlog(jscale(jadd(m0, jscale(q0, 2.0)), 1.0 / (2.0 * mpi)));
            J h0 = jmul(jscale(jdiv(q0, m0), 2.0 / pi), logA);
            J dh0 = jadd(jmul(h0, jsub(jinv(jscale(q02, 8.0)), jinv(jscale(m02, 2.0)))), jinv(jscale(m02, 2.0 * pi)));
            J c1 = jdiv(jmul(g0, m02), q03);
            J c3 = jdiv(jmul(jmul(g0, m02), dh0), q0);
            J one_c3 = jadd(jc<NV, H>(1.0), c3);
            store(t, jmul(one_c3, m02));
            store(t + REC, jscale(one_c3, -1.0));
            store(t + 2 * REC, c1);
            store(t + 3 * REC, jscale(jmul(c1, h0), -1.0));
            break;
        }
        default: break;
        }
        return;
    }
    item -= a.nden;
    if (item >= a.namp) return;
    const DAmp& A = a.amps[item];
    const DBlock& B = a.blocks[A.block];
    const DTopo& T = a.topos[B.topo];
    const int* Ls = a.slcL + (long long)(a.slc0[B.topo] + A.slc) * MAXNODE;
    J K = jc<NV, H>(1.0);
    for (int k = 0; k < T.nn; ++k) {
        const DNode& n = a.nodes[T.node0 + k];
        int L = Ls[k];
        int r = -1;
        for (int i = 0; i < B.nres; ++i) if (a.res[B.res0 + i].node == k) { r = i; break; }
        J m1 = q0_dmass<NV, H>(n.md1, n.d1ref, B, a.res, theta), m2 = q0_dmass<NV, H>(n.md2, n.d2ref, B, a.res, theta);
        if (k == 0 || r < 0) {
            if (n.has_bf && L >= 1 && L <= 6) {
                J q0 = jbreakup(jc<NV, H>(n.mm > 0 ? n.mm : 1.0), m1, m2);
                K = jmul(K, jsqrt(jbfpoly(L, q0, n.bf_d)));
            }
            continue;
        }
        const DRes& R = a.res[B.res0 + r];
        // BW, Custom and the table lineshapes (Hist, Interp) carry no automatic barrier
        if (R.model == 1 || R.model == 5 || R.model == 6 || R.model == 7) continue;
        J m0 = (R.np > 0) ? res_param<NV, H>(R, 0, B, theta) : jc<NV, H>(1.0);
        J q0 = jq0ad(m0, m1, m2);
        if (n.has_bf && L >= 1 && L <= 6) K = jmul(K, jsqrt(jbfpoly(L, q0, n.bf_d)));
        if (R.model == 4) {                                // GS numerator 1 + d(m0) g0 / m0
            const double pi = 3.14159265358979323846;
            J g0 = res_param<NV, H>(R, 1, B, theta);
            double mpi = R.nch > 0 ? R.ch[0] : 0.1396;
            J q02 = jmul(q0, q0), q03 = jmul(q02, q0);
            J logA = jlog(jscale(jadd(m0, jscale(q0, 2.0)), 1.0 / (2.0 * mpi)));
            J dv = jadd(jsub(jmul(jscale(jdiv(jc<NV, H>(mpi * mpi), q02), 3.0 / pi), logA), jmul(jc<NV, H>(mpi * mpi / pi), jdiv(m0, q03))),
                         jdiv(m0, jscale(q0, 2.0 * pi)));
            K = jmul(K, jadd(jc<NV, H>(1.0), jdiv(jmul(dv, g0), m0)));
        }
    }
    store(table + a.off_rec + (long long)A.rec * REC, K);
    double pr = params[A.freeidx] * A.ratio, pi_ = params[a.nfree + A.freeidx] * A.ratio;
    table[a.off_v + 2 * item] = pr;
    table[a.off_v + 2 * item + 1] = pi_;
    table[a.off_c + 2 * item] = pr * K.v;
    table[a.off_c + 2 * item + 1] = pi_ * K.v;
}

template <int NV, bool H>
__global__ void __launch_bounds__(64) k_prep(const double* __restrict__ params, double* __restrict__ table, PrepArgs a) {
    prep_body<NV, H>(params, table, a);
}

// many likelihoods of one model: blockIdx.y selects the likelihood (its parameters and
// coefficient table); the model tables in `a` are the same for all of them
template <int NV, bool H>
// one workgroup row per job.  hdr (graph replays; NULL otherwise): hdr[0] is the number
// of jobs, rows at or beyond it leave at once, and each job's prep arguments come from
// args[row] (its own handle's model tables) instead of `a`
__global__ void __launch_bounds__(64) k_prep_batch(const double* const* __restrict__ params, double* const* __restrict__ tables,
                                                   PrepArgs a, const PrepArgs* __restrict__ args, const int* __restrict__ hdr) {
    if (hdr && (int)blockIdx.y >= hdr[0]) return;
    if (args) a = args[blockIdx.y];
    a.remember = 1;                 // NLL evaluations: each table keeps its theta as the last evaluated one
    prep_body<NV, H>(params[blockIdx.y], tables[blockIdx.y], a);
}

// ---------------------------------------------------------------------------
// final assembly
// ---------------------------------------------------------------------------

constexpr int RED_THREADS = 256;
constexpr int ASM_THREADS = 256;

// this thread's share of class c's total of accumulator column col: the class's
// workgroup partials in a fixed order, or, for the phase space with a stored Gram (no free
// resonance parameter), the column from G and c_a = v_a K_a:
//   col 0: sum_a Re(conj(c_a) (G c)_a) = sum_e I(e);  col 1 + 2a / 2 + 2a: Re / Im (G c)_a
// (sum_e conj(Psi_a sl_a) S, the fused kernel's accX_a); a tree sum over the workgroup
// follows, so the result does not depend on timing
__device__ __forceinline__ double class_share(const FinArgs& f, int c, int col) {
    double s = 0.0;
    if (c == 0 && f.gram) {
        const int n = f.namp;
        const double2* cv = (const double2*)(f.table + f.off_c);
        if (col == 0) {
            for (int a = threadIdx.x; a < n; a += RED_THREADS) {
                double gr = 0.0, gi = 0.0;
                for (int b = 0; b < n; ++b) {
                    const double2 G = f.gram[(long long)a * n + b], cb = cv[b];
                    gr += G.x * cb.x - G.y * cb.y;
                    gi += G.x * cb.y + G.y * cb.x;
                }
                const double2 ca = cv[a];
                s += ca.x * gr + ca.y * gi;
            }
        } else {
            const int a = (col - 1) >> 1, im = (col - 1) & 1;
            for (int b = threadIdx.x; b < n; b += RED_THREADS) {
                const double2 G = f.gram[(long long)a * n + b], cb = cv[b];
                s += im ? (G.x * cb.y + G.y * cb.x) : (G.x * cb.x - G.y * cb.y);
            }
        }
        return s;
    }
    const double* src = f.part + (long long)col * f.nwg;
    for (int g = f.g0[c] + threadIdx.x; g < f.g1[c]; g += RED_THREADS) s += src[g];
    return s;
}

// fixed-order sum over a workgroup of RED_THREADS threads, valid in thread 0: a butterfly
// within each wave (lane 0 always adds in the same order), then the wave sums in wave order
// through LDS (one barrier instead of one per tree level)
__device__ __forceinline__ double wg_sum(double v, double* sh) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0) {
        t = sh[0];
        for (int k = 1; k < RED_THREADS / 64; ++k) t += sh[k];
    }
    return t;
}

// stage 1: one workgroup per (class, accumulator column); fixed summation order
__global__ void __launch_bounds__(RED_THREADS) k_reduce(FinArgs f) {
    __shared__ double sh[RED_THREADS / 64];
    int pair = blockIdx.x;
    int c = pair / f.nacc, col = pair - c * f.nacc;
    const double tot = wg_sum(class_share(f, c, col), sh);
    if (threadIdx.x == 0) f.totals[pair] = tot;
}

__global__ void __launch_bounds__(RED_THREADS) k_reduce_batch(const FinArgs* __restrict__ fins, const int* __restrict__ hdr) {
    if (hdr && (int)blockIdx.y >= hdr[0]) return;
    const FinArgs& f = fins[blockIdx.y];
    if ((int)blockIdx.x >= 3 * f.nacc) return;      // the whole workgroup leaves together
    __shared__ double sh[RED_THREADS / 64];
    int pair = blockIdx.x;
    int c = pair / f.nacc, col = pair - c * f.nacc;
    const double tot = wg_sum(class_share(f, c, col), sh);
    if (threadIdx.x == 0) f.totals[pair] = tot;
}

// stage 2: NLL and gradient from the class totals.  Every global read happens first, in
// parallel over the threads: the gradient entries' weights (scale times their table
// coefficient), accumulator indices and row pointers (assemble_prefill, which needs no
// totals and so runs in k_finalize before a workgroup's reduction, off the last
// workgroup's path), then the class-combined accumulators.  The gradient rows are then
// fixed-order sums over LDS (one wave per theta with a butterfly sum), so no thread waits
// on a chain of dependent global loads.
// LDS layout: assemble_lds_bytes (device.hpp)
struct AsmLds {
    double *C, *wcp, *wth, *vre, *vim;
    int *icp, *ith, *cpp, *thp;
};
__device__ __forceinline__ AsmLds asm_lds(const FinArgs& f) {
    extern __shared__ double smem_asm[];
    AsmLds L;
    L.C = smem_asm;                          // [nacc] = -data + bkg + cp * phsp
    L.wcp = L.C + f.nacc;
    L.wth = L.wcp + f.ncpe;
    L.vre = L.wth + f.nthe;
    L.vim = L.vre + f.nthe;
    L.icp = (int*)(L.vim + f.nthe);
    L.ith = L.icp + f.ncpe;
    L.cpp = L.ith + f.nthe;
    L.thp = L.cpp + f.nfree + 1;
    return L;
}
__device__ __forceinline__ void assemble_prefill(const FinArgs& f) {
    const AsmLds L = asm_lds(f);
    const int tid = threadIdx.x;
    // coupling entries: weight r_a K_a, accumulator pair (1 + 2a, 2 + 2a)
    for (int k = tid; k < f.ncpe; k += ASM_THREADS) {
        DGEnt e = f.cpe[k];
        L.wcp[k] = e.scale * f.table[f.off_rec + e.rec];
        L.icp[k] = e.idx;
    }
    // theta entries: idx >= 0 -> w C[idx]; idx < 0 -> w (vr C[1 + 2a] + vi C[2 + 2a])
    for (int k = tid; k < f.nthe; k += ASM_THREADS) {
        DGEnt e = f.the[k];
        double d = (e.rec < 0) ? 1.0 : f.table[f.off_rec + e.rec];
        L.wth[k] = e.scale * d;
        L.ith[k] = e.idx;
        if (e.idx < 0) {
            int a = -e.idx - 1;
            L.vre[k] = f.table[f.off_v + 2 * a];
            L.vim[k] = f.table[f.off_v + 2 * a + 1];
        }
    }
    for (int k = tid; k <= f.nfree; k += ASM_THREADS) L.cpp[k] = f.cpp[k];
    for (int k = tid; k <= f.ntheta; k += ASM_THREADS) L.thp[k] = f.thp[k];
}
// the assembly proper, after assemble_prefill and a barrier
__device__ __forceinline__ void assemble_main(const FinArgs& f) {
    const AsmLds L = asm_lds(f);
    double* C = L.C;
    const double* wcp = L.wcp;
    const double* wth = L.wth;
    const double* vre = L.vre;
    const double* vim = L.vim;
    const int* icp = L.icp;
    const int* ith = L.ith;
    const int* cpp = L.cpp;
    const int* thp = L.thp;
    int tid = threadIdx.x;
    const double* tot = f.totals;            // tot[c * nacc + col]
    double Np = f.Np, Neff = f.Wd - f.Wb;
    double pf = (Np > 0) ? tot[0] / Np : 0.0;
    if (!(pf > 1e-30) || isinf(pf)) pf = 1e-30;
    double cp = (Np > 0) ? Neff / (pf * Np) : 0.0;
    const double* theta = f.params + 2 * f.nfree;
    for (int col = tid; col < f.nacc; col += ASM_THREADS)
        C[col] = -tot[f.nacc + col] + tot[2 * f.nacc + col] + cp * tot[col];
    double loss = -tot[f.nacc] + tot[2 * f.nacc] + Neff * log(pf);
    for (int g = 0; g < f.ngauss; ++g) {
        double d = theta[f.g_slot[g]] - f.g_mu[g];
        loss += d * d / (2.0 * f.g_sig[g] * f.g_sig[g]);
    }
    bool reset = isnan(loss) || isinf(loss);
    __syncthreads();
    double* out = f.out;
    if (tid == 0) out[0] = reset ? 1e30 : loss;
    // coupling gradient: dL/dp_f* = sum_a r_a K_a X_a, X_a = (C[1 + 2a], C[2 + 2a]);
    // output [2 Re, 2 Im]
    for (int fi = tid; fi < f.nfree; fi += ASM_THREADS) {
        double gr = 0.0, gi = 0.0;
        for (int k = cpp[fi]; k < cpp[fi + 1]; ++k) {
            double w = wcp[k];
            gr += w * C[1 + 2 * icp[k]];
            gi += w * C[2 + 2 * icp[k]];
        }
        out[1 + fi] = reset ? 0.0 : 2.0 * gr;
        out[1 + f.nfree + fi] = reset ? 0.0 : 2.0 * gi;
    }
    // theta gradient: propagator part -2 sum_t dcoef_t/dtheta_j C_t plus numerator part
    // 2 Re sum_a v_a dK_a/dtheta_j conj(X_a)
    const int lane = tid & 63, nw = ASM_THREADS / 64;
    for (int j = tid >> 6; j < f.ntheta; j += nw) {
        double g = 0.0;
        for (int k = thp[j] + lane; k < thp[j + 1]; k += 64) {
            int idx = ith[k];
            if (idx >= 0) {
                g += wth[k] * C[idx];
            } else {
                int a = -idx - 1;
                g += wth[k] * (vre[k] * C[1 + 2 * a] + vim[k] * C[2 + 2 * a]);
            }
        }
        for (int o = 32; o > 0; o >>= 1) g += __shfl_xor(g, o);
        if (lane == 0) {
            for (int q = 0; q < f.ngauss; ++q)
                if (f.g_slot[q] == j) g += (theta[j] - f.g_mu[q]) / (f.g_sig[q] * f.g_sig[q]);
            out[1 + 2 * f.nfree + j] = reset ? 0.0 : g;
        }
    }
}

__device__ __forceinline__ void assemble_body(const FinArgs& f) {
    assemble_prefill(f);
    __syncthreads();
    assemble_main(f);
}

__global__ void __launch_bounds__(ASM_THREADS) k_assemble(FinArgs f) { assemble_body(f); }

// k_reduce and k_assemble in one launch: every workgroup sums one (class, accumulator)
// over the class's workgroup partials (the order of k_reduce), and the workgroup that
// finishes last, found by a device-scope counter, runs the assembly on the totals and
// resets the counter for the next evaluation.  Results are those of the two kernels.
static_assert(RED_THREADS == ASM_THREADS, "one workgroup size for both parts");
__global__ void __launch_bounds__(ASM_THREADS) k_finalize(FinArgs f, unsigned int* __restrict__ done) {
    __shared__ double sh[RED_THREADS / 64];
    __shared__ int last;
    const int pair = blockIdx.x;
    const int c = pair / f.nacc, col = pair - c * f.nacc;
    // the assembly's inputs that need no totals, loaded by every workgroup while its
    // reduction loads are in flight (any workgroup may turn out to be the last)
    assemble_prefill(f);
    const double tot = wg_sum(class_share(f, c, col), sh);
    if (threadIdx.x == 0) {
        f.totals[pair] = tot;
        __threadfence();                                   // the total is visible before the count
        last = atomicAdd(done, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return;
    __threadfence();                                       // every total is visible to this workgroup
    assemble_main(f);
    if (threadIdx.x == 0) *done = 0u;
}

__global__ void __launch_bounds__(ASM_THREADS) k_assemble_batch(const FinArgs* __restrict__ fins, const int* __restrict__ hdr) {
    if (hdr && (int)blockIdx.x >= hdr[0]) return;
    assemble_body(fins[blockIdx.x]);
}

// ---------------------------------------------------------------------------
// generic per-event amplitudes and intensities
// ---------------------------------------------------------------------------

__device__ inline double2 block_psi(const AuxArgs& a, int b, long long e) {
    const DBlock& B = a.blocks[b];
    double pr = 1.0, pi = 0.0;
    for (int d = 0; d < B.nden; ++d) {
        const DDen& D = a.dens[B.den0 + d];
        const DDenTerms& DT = a.dterms[B.den0 + d];
        double dr = 0.0, di = 0.0;
        for (int t = 0; t < DT.nre + DT.nim; ++t) {
            const DTerm& T = a.terms[DT.t0 + t];
            double c = (T.coef < 0) ? T.cval : a.table[a.off_rec + (long long)(D.rec + T.coef) * REC];
            double x = (T.inv < 0) ? 1.0 : ev_inv(a.st, T.inv, e);
            if (t < DT.nre) dr += c * x; else di += c * x;
        }
        double xr = dr, xi = di;                          // table lineshape: the factor itself
        if (!DT.stat) {
            const double q = 1.0 / (dr * dr + di * di);
            xr = dr * q;                                  // 1 / (dr - i di)
            xi = di * q;
        }
        double nr = pr * xr - pi * xi, ni = pr * xi + pi * xr;
        pr = nr; pi = ni;
    }
    // Custom lineshape factors (written for this parameter point by the Custom kernel)
    for (int q = a.bcp[b]; q < a.bcp[b + 1]; ++q) {
        const double* f = a.cfb + (long long)a.bcl[q] * a.ncf * a.cld + e;
        const double fr = f[0], fi = f[a.cld];
        const double nr = pr * fr - pi * fi, ni = pr * fi + pi * fr;
        pr = nr; pi = ni;
    }
    return make_double2(pr, pi);
}

// A[e][p][a] = K_a Psi_b(e) sl_u,p(e)  (interleaved complex), events [e0, e0 + n)
__global__ void k_amplitudes(AuxArgs a, long long e0, long long n, double* __restrict__ out) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    long long e = e0 + i;
    for (int am = 0; am < a.namp; ++am) {
        const DAmp& A = a.amps[am];
        double2 psi = block_psi(a, A.block, e);
        double K = a.table[a.off_rec + (long long)A.rec * REC];
        for (int p = 0; p < a.npol; ++p) {
            double2 s = ev_sl(a.st, A.u, p, a.npol, e);
            double r = K * (psi.x * s.x - psi.y * s.y), im = K * (psi.x * s.y + psi.y * s.x);
            long long o = ((i * a.npol + p) * a.namp + am) * 2;
            out[o] = r;
            out[o + 1] = im;
        }
    }
}

// amplitude rows for Grams of the amplitudes (fit fractions, interference): column-major
// X[col][R], R = p * n + i (polarisation p of event e0 + i), columns Re A_a (a < namp)
// then Im A_a; A_a = K_a Psi_b sl_u,p without couplings (withK == 0: Psi_b sl_u,p)
__global__ void k_amp_rows(AuxArgs a, long long e0, long long n, double* __restrict__ X, long long ldx, int withK) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    long long e = e0 + i;
    for (int am = 0; am < a.namp; ++am) {
        const DAmp& A = a.amps[am];
        double2 psi = block_psi(a, A.block, e);
        double K = withK ? a.table[a.off_rec + (long long)A.rec * REC] : 1.0;
        for (int p = 0; p < a.npol; ++p) {
            double2 s = ev_sl(a.st, A.u, p, a.npol, e);
            X[(long long)am * ldx + p * n + i] = K * (psi.x * s.x - psi.y * s.y);
            X[(long long)(a.namp + am) * ldx + p * n + i] = K * (psi.x * s.y + psi.y * s.x);
        }
    }
}

// per-event partial-wave intensities: out[b][i] = sum_p |P_b(e, p)|^2 for the nblock
// partial waves (blocks) and out[nblock][i] = I(e) = sum_p |sum_b P_b|^2, with
// P_b = Psi_b sum_{a in b} c_a sl_u(a),p and c_a = v_a K_a (src/ComputeResults.cu:57-190)
// partial wave w = blocks wblk[wptr[w] .. wptr[w+1]) (a block and its identical-particle
// permuted copies): out[w][i] = sum_pol |sum over the wave's blocks P_b|^2,
// out[nwave][i] = the total intensity
__global__ void k_partial_int(AuxArgs a, const int* __restrict__ wptr, const int* __restrict__ wblk, int nwave,
                              long long e0, long long n, double* __restrict__ out) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    long long e = e0 + i;
    double Sr[MAXPOL], Si[MAXPOL], Wr[MAXPOL], Wi[MAXPOL];
    for (int p = 0; p < a.npol; ++p) { Sr[p] = 0.0; Si[p] = 0.0; }
    for (int w = 0; w < nwave; ++w) {
        for (int p = 0; p < a.npol; ++p) { Wr[p] = 0.0; Wi[p] = 0.0; }
        for (int k = wptr[w]; k < wptr[w + 1]; ++k) {
            const DBlock& B = a.blocks[wblk[k]];
            if (B.nsl <= 0) continue;
            double2 psi = block_psi(a, wblk[k], e);
            for (int p = 0; p < a.npol; ++p) {
                double yr = 0.0, yi = 0.0;
                for (int c = 0; c < B.nsl; ++c) {
                    const int am = B.site + c;
                    const double cr = a.table[a.off_c + 2 * am], ci = a.table[a.off_c + 2 * am + 1];
                    double2 s = ev_sl(a.st, a.amps[am].u, p, a.npol, e);
                    yr += cr * s.x - ci * s.y;
                    yi += cr * s.y + ci * s.x;
                }
                Wr[p] += psi.x * yr - psi.y * yi;
                Wi[p] += psi.x * yi + psi.y * yr;
            }
        }
        double v = 0.0;
        for (int p = 0; p < a.npol; ++p) {
            v += Wr[p] * Wr[p] + Wi[p] * Wi[p];
            Sr[p] += Wr[p];
            Si[p] += Wi[p];
        }
        out[(long long)w * n + i] = v;
    }
    double I = 0.0;
    for (int p = 0; p < a.npol; ++p) I += Sr[p] * Sr[p] + Si[p] * Si[p];
    out[(long long)nwave * n + i] = I;
}

// coherent partial-wave amplitudes: out[w][i][p] = sum over the wave's blocks of P_b
// (couplings applied), interleaved complex, events [e0, e0 + n)
__global__ void k_partial_amp(AuxArgs a, const int* __restrict__ wptr, const int* __restrict__ wblk, int nwave,
                              long long e0, long long n, double2* __restrict__ out) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    long long e = e0 + i;
    for (int w = 0; w < nwave; ++w) {
        for (int p = 0; p < a.npol; ++p) {
            double wr = 0.0, wi = 0.0;
            for (int k = wptr[w]; k < wptr[w + 1]; ++k) {
                const DBlock& B = a.blocks[wblk[k]];
                if (B.nsl <= 0) continue;
                double2 psi = block_psi(a, wblk[k], e);
                double yr = 0.0, yi = 0.0;
                for (int c = 0; c < B.nsl; ++c) {
                    const int am = B.site + c;
                    const double cr = a.table[a.off_c + 2 * am], ci = a.table[a.off_c + 2 * am + 1];
                    double2 s = ev_sl(a.st, a.amps[am].u, p, a.npol, e);
                    yr += cr * s.x - ci * s.y;
                    yi += cr * s.y + ci * s.x;
                }
                wr += psi.x * yr - psi.y * yi;
                wi += psi.x * yi + psi.y * yr;
            }
            out[((long long)w * n + i) * a.npol + p] = make_double2(wr, wi);
        }
    }
}

__global__ void k_intensity(AuxArgs a, long long e0, long long n, double* __restrict__ out) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    long long e = e0 + i;
    double Sr[MAXPOL], Si[MAXPOL];
    for (int p = 0; p < a.npol; ++p) { Sr[p] = 0.0; Si[p] = 0.0; }
    for (int am = 0; am < a.namp; ++am) {
        const DAmp& A = a.amps[am];
        double2 psi = block_psi(a, A.block, e);
        double cr = a.table[a.off_c + 2 * am], ci = a.table[a.off_c + 2 * am + 1];
        double yr = cr * psi.x - ci * psi.y, yi = cr * psi.y + ci * psi.x;
        for (int p = 0; p < a.npol; ++p) {
            double2 s = ev_sl(a.st, A.u, p, a.npol, e);
            Sr[p] += yr * s.x - yi * s.y;
            Si[p] += yr * s.y + yi * s.x;
        }
    }
    double I = 0.0;
    for (int p = 0; p < a.npol; ++p) I += Sr[p] * Sr[p] + Si[p] * Si[p];
    out[i] = I;
}

}  // namespace ctpwa

template <int NV, bool H>
static hipError_t launch_prep_t(const double* params, double* table, const ctpwa::PrepArgs* a, int remember,
                                hipStream_t stream) {
    int items = a->nden + a->namp;
    ctpwa::PrepArgs b = *a;
    b.remember = remember;
    hipLaunchKernelGGL((ctpwa::k_prep<NV, H>), dim3(items), dim3(64), ctpwa::prep_lds_bytes(*a), stream, params, table, b);
    return hipGetLastError();
}

// nv: largest number of free parameters in one block; flags: bit 0 also second
// derivatives (Hessian), bit 1 theta is also stored as the last evaluated theta (the
// NLL and Hessian entry points; the per-event outputs leave it)
extern "C" __attribute__((visibility("hidden"))) hipError_t ctpwa_launch_prep(const double* params, double* table,
                                                                              const ctpwa::PrepArgs* a, int nv,
                                                                              int flags, hipStream_t stream) {
    const int hess = flags & 1, rem = (flags >> 1) & 1;
    if (nv <= 1) return hess ? launch_prep_t<1, true>(params, table, a, rem, stream) : launch_prep_t<1, false>(params, table, a, rem, stream);
    if (nv <= 2) return hess ? launch_prep_t<2, true>(params, table, a, rem, stream) : launch_prep_t<2, false>(params, table, a, rem, stream);
    if (nv <= 4) return hess ? launch_prep_t<4, true>(params, table, a, rem, stream) : launch_prep_t<4, false>(params, table, a, rem, stream);
    return hess ? launch_prep_t<8, true>(params, table, a, rem, stream) : launch_prep_t<8, false>(params, table, a, rem, stream);
}

template <int NV, bool H>
static hipError_t launch_prep_batch_t(const double* const* params, double* const* tables, const ctpwa::PrepArgs* a,
                                      const ctpwa::PrepArgs* args, int rows, const int* hdr, hipStream_t stream) {
    int items = a->nden + a->namp;
    hipLaunchKernelGGL((ctpwa::k_prep_batch<NV, H>), dim3(items, rows), dim3(64), ctpwa::prep_lds_bytes(*a), stream, params,
                       tables, *a, args, hdr);
    return hipGetLastError();
}

// rows: workgroup rows launched (the job count, or with hdr a bound on the job count hdr[0]
// on the device, extra rows leave); args: NULL, or the jobs' own prep arguments on the
// device (a then only gives the launch shape)
extern "C" __attribute__((visibility("hidden"))) hipError_t ctpwa_launch_prep_batch(const double* const* params,
                                                                                    double* const* tables,
                                                                                    const ctpwa::PrepArgs* a, int nv,
                                                                                    const ctpwa::PrepArgs* args, int rows,
                                                                                    const int* hdr, hipStream_t stream) {
    if (rows <= 0) return hipSuccess;
    if (nv <= 1) return launch_prep_batch_t<1, false>(params, tables, a, args, rows, hdr, stream);
    if (nv <= 2) return launch_prep_batch_t<2, false>(params, tables, a, args, rows, hdr, stream);
    if (nv <= 4) return launch_prep_batch_t<4, false>(params, tables, a, args, rows, hdr, stream);
    return launch_prep_batch_t<8, false>(params, tables, a, args, rows, hdr, stream);
}

// fins: device array of the jobs' arguments; rows and hdr as above; max_nacc: the largest
// accumulator count among the jobs
extern "C" __attribute__((visibility("hidden"))) hipError_t ctpwa_launch_finalize_batch(const ctpwa::FinArgs* fins, int rows,
                                                                                        int max_nacc, size_t lds_bytes,
                                                                                        const int* hdr, hipStream_t stream) {
    if (rows <= 0) return hipSuccess;
    hipLaunchKernelGGL(ctpwa::k_reduce_batch, dim3(3 * max_nacc, rows), dim3(ctpwa::RED_THREADS), 0, stream, fins, hdr);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(ctpwa::k_assemble_batch, dim3(rows), dim3(ctpwa::ASM_THREADS), lds_bytes, stream, fins, hdr);
    return hipGetLastError();
}

// done: a device counter that is zero between evaluations (reset by the kernel itself)
extern "C" __attribute__((visibility("hidden"))) hipError_t ctpwa_launch_finalize(const ctpwa::FinArgs* f, unsigned int* done,
                                                                                  hipStream_t stream) {
    hipLaunchKernelGGL(ctpwa::k_finalize, dim3(3 * f->nacc), dim3(ctpwa::ASM_THREADS), ctpwa::assemble_lds_bytes(*f), stream,
                       *f, done);
    return hipGetLastError();
}

extern "C" __attribute__((visibility("hidden"))) hipError_t ctpwa_launch_amp_rows(const ctpwa::AuxArgs* a, long long e0,
                                                                                  long long n, double* X, long long ldx,
                                                                                  int withK, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(ctpwa::k_amp_rows, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, stream, *a, e0, n, X, ldx,
                       withK);
    return hipGetLastError();
}

extern "C" __attribute__((visibility("hidden"))) hipError_t ctpwa_launch_partial_int(const ctpwa::AuxArgs* a, const int* wptr,
                                                                                     const int* wblk, int nwave,
                                                                                     long long e0, long long n, double* out,
                                                                                     hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(ctpwa::k_partial_int, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, stream, *a, wptr, wblk,
                       nwave, e0, n, out);
    return hipGetLastError();
}

extern "C" __attribute__((visibility("hidden"))) hipError_t ctpwa_launch_partial_amp(const ctpwa::AuxArgs* a, const int* wptr,
                                                                                     const int* wblk, int nwave,
                                                                                     long long e0, long long n, double* out,
                                                                                     hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(ctpwa::k_partial_amp, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, stream, *a, wptr, wblk,
                       nwave, e0, n, (double2*)out);
    return hipGetLastError();
}

extern "C" __attribute__((visibility("hidden"))) hipError_t ctpwa_launch_aux(const ctpwa::AuxArgs* a, long long e0, long long n, double* out, int intensity,
                                        hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    int bs = 128;
    long long nb = (n + bs - 1) / bs;
    if (intensity) hipLaunchKernelGGL(ctpwa::k_intensity, dim3((unsigned)nb), dim3(bs), 0, stream, *a, e0, n, out);
    else hipLaunchKernelGGL(ctpwa::k_amplitudes, dim3((unsigned)nb), dim3(bs), 0, stream, *a, e0, n, out);
    return hipGetLastError();
}
