// Plain-old-data tables shared by the host driver and the fixed (non-generated)
// kernels, plus small device math helpers.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace ctpwa {

constexpr int MAXNODE = 8;      // decay steps per chain
constexpr int MAXPART = 16 + 1 + MAXNODE;
constexpr int MAXKIN = 32;
constexpr int MAXT = 64;        // entries of a contracted spin tensor
constexpr int MAXPOL = 64;
constexpr int NLOC = 8;         // free parameters per block
constexpr int REC = 1 + NLOC + NLOC * NLOC;

struct DNode {            // one decay step of a topology
    int dj, dj1, dj2, d1ref, d2ref, has_bf, lmin, kin;
    double bf_d, mm, md1, md2;
};
struct DTopo {
    int nn, node0;        // nodes [node0, node0 + nn)
    int order[MAXNODE];   // nodes >= 1 in an order where daughters come first
};
struct DKin { uint32_t mset, d1set, d2set; int owner_topo, owner_node, pad; };
struct DInv { int kin, kind, L, pad; double d, ma, mb; };
struct DU {               // one SL tensor
    int topo, mask, S[MAXNODE], L[MAXNODE], cg[MAXNODE];
};
struct DBlock { int topo, nloc, nres, res0, nsl, site, nden, den0; int vars[NLOC]; };
struct DRes { int node, model, np, nch; double p[6]; int slot[6]; double ch[8]; };
struct DDen { int block, res, rec, ncoef; };
struct DAmp { int block, slc, freeidx, rec, u, pad; double ratio; };
struct DTerm { int inv, coef; double cval; };   // inv < 0: 1; coef < 0: constant cval
struct DThTerm { int block, rec; };               // theta-gradient accumulator -> coefficient record
// terms [t0, t0+nre) real, then nim imaginary; stat 0: the factor is 1 / (sum re - i sum im),
// stat 1 (table lineshapes): the factor is sum re + i sum im itself
struct DDenTerms { int nre, nim, t0, stat; };

// Blatt-Weisskopf polynomial N_L(z), Bf(L,q,q0,d) = sqrt(N_L(q0 d)/N_L(q d))
// (include/ResModel.cuh:184-211); N is 1 for L = 0 and for L > 6.
__host__ __device__ inline double bf_poly(int L, double z) {
    double z2 = z * z;
    switch (L) {
    case 1: return 1.0 + z2;
    case 2: return 9.0 + z2 * (3.0 + z2);
    case 3: return 225.0 + z2 * (45.0 + z2 * (6.0 + z2));
    case 4: return 11025.0 + z2 * (1575.0 + z2 * (135.0 + z2 * (10.0 + z2)));
    case 5: return 893025.0 + z2 * (99225.0 + z2 * (6300.0 + z2 * (315.0 + z2 * (15.0 + z2))));
    case 6: return 540326025.0 + z2 * (6185025.0 + z2 * (363825.0 + z2 * (17325.0 + z2 * (630.0 + z2 * (21.0 + z2)))));
    default: return 1.0;
    }
}

// breakup momentum from invariant masses, 0 below threshold (src/AmpGen.cu:2733-2737)
__host__ __device__ inline double breakup(double m, double m1, double m2) {
    double q2 = (m * m - (m1 + m2) * (m1 + m2)) * (m * m - (m1 - m2) * (m1 - m2));
    if (q2 <= 0.0) return 0.0;
    return sqrt(q2) / (2.0 * m);
}

// argument blocks of the fixed kernels (prep.hip)
struct PrepArgs {
    const DBlock* blocks;
    const DRes* res;
    const DNode* nodes;      // flattened per topology
    const DTopo* topos;
    const int* slcL;         // [topo slc row][MAXNODE], row offset per topology in slc0
    const int* slc0;         // per topology first row
    const DDen* dens;
    const DAmp* amps;
    int nden, namp, nfree, ntheta;
    long long off_v, off_c, off_rec;
    long long off_th;        // theta values (read by Custom lineshapes in the fused kernel)
    long long off_lth;       // theta of the last NLL / Hessian evaluation (ctpwa_last_theta)
    int remember;            // this launch writes theta to off_lth as well
    // the same tables packed in one blob (16-byte aligned sections) that the prep kernel
    // copies to LDS before use; meta_bytes == 0: read the tables from global memory
    const void* meta;
    int meta_bytes;
    int o_blocks, o_res, o_nodes, o_topos, o_slcL, o_slc0, o_dens, o_amps;
};

// dynamic LDS of the prep kernels: the staged model tables, then the parameter vector
inline size_t prep_lds_bytes(const PrepArgs& a) {
    return (size_t)a.meta_bytes + sizeof(double) * (size_t)(2 * a.nfree + a.ntheta);
}

// one term of a gradient contraction in the final assembly
//   theta entries: idx >= 0  -> scale * dcoef * C[idx]          (propagator coefficient)
//                  idx <  0  -> scale * dK * Re(v_a conj(X_a)), a = -idx - 1
//   coupling entries: idx = amplitude a, scale = r_a, rec -> K_a
struct DGEnt { int idx, rec; double scale; };

struct FinArgs {
    const double* part;      // [nacc][nwg]
    int nwg, nacc;
    int g0[3], g1[3];        // workgroup ranges of phsp, data, bkg
    const double* table;
    long long off_v, off_rec;
    const DAmp* amps;
    const DBlock* blocks;
    const DThTerm* thterms;
    int nth;
    const int* thp;          // [ntheta + 1] CSR row pointers into the
    const DGEnt* the;        //   theta-gradient entries
    const int* cpp;          // [nfree + 1] CSR row pointers into the
    const DGEnt* cpe;        //   coupling-gradient entries
    int ncpe, nthe;          // entry counts of the two lists
    int namp, nfree, ntheta, ngauss;
    const int* g_slot;
    const double* g_mu;
    const double* g_sig;
    double Np, Wd, Wb;
    const double* params;
    double* out;             // [1 + 2 nfree + ntheta]
    double* totals;          // [3][nacc] scratch (kept for inspection)
    // no free resonance parameter: the phase-space totals come from the stored amplitude
    // Gram G[a][b] = sum_e sum_p conj(A'_a) A'_b (amplitudes without couplings or theta
    // coefficients) and the table's c_a (offset off_c) instead of per-event partials
    const double2* gram;     // [namp][namp], or NULL
    long long off_c;
};

// LDS of the assembly kernel: C[nacc] | wcp[ncpe] | wth, vre, vim [nthe] (doubles), then
// icp[ncpe] | ith[nthe] | cpp[nfree + 1] | thp[ntheta + 1] (ints)
inline size_t assemble_lds_bytes(const FinArgs& f) {
    return (size_t)(f.nacc + f.ncpe + 3 * f.nthe) * sizeof(double) +
           (size_t)(f.ncpe + f.nthe + f.nfree + f.ntheta + 2) * sizeof(int);
}

// Per-event inputs of a handle: SL tensors [NU][nps][ld] and invariants [K][ld].
// precision 0: everything FP64.  precision 1 (mixed): SL tensors FP32, invariants FP32
// except the squared masses s, which stay FP64 (m0^2 - s cancels near a pole).
// nps < npol (mirror-image polarisations): polarisation p >= nps of SL tensor u is
// psign[u] conj(polarisation npol - 1 - p).
struct EvStore {
    const double2* sl;       // precision 0
    const float2* slf;       // precision 1
    const double* inv;       // precision 0: all invariants; 1: the FP64 ones
    const float* invf;       // precision 1: the FP32 invariants
    const int* imap;         // precision 1: invariant i -> row in inv (>= 0) or -(row in invf) - 1
    long long ld;
    int mixed;
    int nps;                 // stored polarisations per SL tensor
    const int* psign;        // [NU] sign of the mirrored polarisation (nps < npol)
};
__device__ inline double2 ev_sl_row(const EvStore& s, long long row, long long e) {
    if (s.mixed) {
        float2 f = s.slf[row * s.ld + e];
        return make_double2((double)f.x, (double)f.y);
    }
    return s.sl[row * s.ld + e];
}
// polarisation p of SL tensor u for event e (npol: the model's polarisation count)
__device__ inline double2 ev_sl(const EvStore& s, int u, int p, int npol, long long e) {
    if (p < s.nps) return ev_sl_row(s, (long long)u * s.nps + p, e);
    const double2 m = ev_sl_row(s, (long long)u * s.nps + (npol - 1 - p), e);
    const double g = (double)s.psign[u];
    return make_double2(g * m.x, -g * m.y);
}
__device__ inline double ev_inv(const EvStore& s, int i, long long e) {
    if (s.mixed) {
        int m = s.imap[i];
        return (m >= 0) ? s.inv[(long long)m * s.ld + e] : (double)s.invf[(long long)(-m - 1) * s.ld + e];
    }
    return s.inv[(long long)i * s.ld + e];
}

struct AuxArgs {
    EvStore st;
    const double* table;
    long long off_c, off_rec;
    const DAmp* amps;
    const DBlock* blocks;
    const DDen* dens;
    const DDenTerms* dterms;
    const DTerm* terms;
    int namp, npol;
    // Custom factor values per stored event (rows c * ncf * cld and the next one: real,
    // imaginary), block b's factors bcl[bcp[b] .. bcp[b+1])
    const double* cfb;
    long long cld;
    int ncf;
    const int* bcp;
    const int* bcl;
};

// Hessian assembly from the per-class sums
struct HessAsm {
    // per class c: G[c] (upper-triangle list over 2C), Uu[c] (upper triangle over NU),
    // Z[c] (NZ), g[c] (NU)
    const double* G[3];
    const double* Uu[3];
    const double* Zs[3];
    const double* gs[3];
    int C, NU, NZ;
    int namp, nfree, ntheta, nblock;
    const DAmp* amps;
    const DBlock* blocks;
    const double* table;
    long long off_v;
    double Np, Wd, Wb;
    int ngauss;
    const int* g_slot;
    const double* g_sig;
    // fit parameter q = (Re p_f, Im p_f, theta) -> extended entries map_x[map_ptr[q] ..
    // map_ptr[q + 1]) with factors map_r (v_a = r_a p_f(a))
    const int* map_ptr;
    const int* map_x;
    const double* map_r;
    double* Hext;            // [NU][NU] scratch
    double* out;             // [np][np]
};

}  // namespace ctpwa
