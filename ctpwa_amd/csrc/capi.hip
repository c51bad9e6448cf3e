// C ABI of the native likelihood library: handle construction (plan -> device
// layout -> one-time SL/invariant kernels -> generated fused kernel via hipRTC) and
// the per-evaluation entry points.  See include/ctpwa_amd.h for the contract.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ctpwa_amd.h"
#include "device.hpp"
#include "model.hpp"

namespace ctpwa {
std::string generate_eval_kernel(const Model_& M);
std::string generate_custom2_kernel(const Model_& M);
std::string generate_hess_kernel(const Model_& M, std::vector<int>* perm);
}

extern "C" hipError_t ctpwa_launch_init(const double*, long long, int, const ctpwa::DTopo*, int, const ctpwa::DNode*,
                                         const ctpwa::DKin*, int, const ctpwa::DU*, int, const double*, const int*, int,
                                         const ctpwa::DInv*, int, const double*, double2*, double*, long long, long long,
                                         hipStream_t);
extern "C" hipError_t ctpwa_launch_to_mixed(const double2*, const double*, long long, int, int, const int*, float2*,
                                             double*, float*, long long, long long, hipStream_t);
extern "C" hipError_t ctpwa_launch_polsym(const double2*, long long, long long, long long, int, unsigned long long*,
                                          hipStream_t);
extern "C" hipError_t ctpwa_launch_prep(const double*, double*, const ctpwa::PrepArgs*, int, int, hipStream_t);
extern "C" hipError_t ctpwa_launch_finalize(const ctpwa::FinArgs*, unsigned int*, hipStream_t);
extern "C" hipError_t ctpwa_launch_prep_batch(const double* const*, double* const*, const ctpwa::PrepArgs*, int,
                                              const ctpwa::PrepArgs*, int, const int*, hipStream_t);
extern "C" hipError_t ctpwa_launch_finalize_batch(const ctpwa::FinArgs*, int, int, size_t, const int*, hipStream_t);
extern "C" hipError_t ctpwa_launch_aux(const ctpwa::AuxArgs*, long long, long long, double*, int, hipStream_t);


extern "C" hipError_t ctpwa_launch_zero(double*, long long, hipStream_t);
extern "C" long long ctpwa_hess_rows_waves(long long);
extern "C" hipError_t ctpwa_launch_accum_cols(const double*, long long, long long, int, const int*, double*, hipStream_t);
extern "C" int ctpwa_gram_tables(int, int4*, int*, int2*);
extern "C" int ctpwa_gram_nwg(int);
extern "C" hipError_t ctpwa_launch_gram_panels(const void*, int, int, long long, long long, const double*, long long,
                                                const int4*, int, int, double*, int, int, hipStream_t);
extern "C" hipError_t ctpwa_launch_accum_panels(const double*, int, int, const int2*, int, double*, hipStream_t);
extern "C" hipError_t ctpwa_launch_amp_rows(const ctpwa::AuxArgs*, long long, long long, double*, long long, int, hipStream_t);
extern "C" hipError_t ctpwa_launch_partial_int(const ctpwa::AuxArgs*, const int*, const int*, int, long long, long long,
                                               double*, hipStream_t);
extern "C" hipError_t ctpwa_launch_partial_amp(const ctpwa::AuxArgs*, const int*, const int*, int, long long, long long,
                                               double*, hipStream_t);
extern "C" hipError_t ctpwa_launch_gram_to_complex(const double*, int, double*, hipStream_t);
extern "C" hipError_t ctpwa_launch_hess_assemble(const ctpwa::HessAsm*, hipStream_t);

using namespace ctpwa;

static thread_local std::string g_err;

static int fail(const std::string& msg) {
    g_err = msg;
    return 1;
}

#define HIPCHECK(x)                                                                          \
    do {                                                                                     \
        hipError_t e_ = (x);                                                                 \
        if (e_ != hipSuccess) throw std::runtime_error(std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

// ---------------------------------------------------------------------------
// host angular-momentum algebra for the Clebsch-Gordan x Y_lm table
// ---------------------------------------------------------------------------
// factorial with the reference's semantics (src/Amplitude.cu:5-91): exact for
// integers and half-integers up to 11.5, Stirling with two correction terms for
// integers >= 12
static double host_fact(double n) {
    if (n < 0) return 0.0;
    double ri = std::round(n);
    bool is_int = std::fabs(n - ri) < 1e-10;
    if (is_int && ri <= 11) {
        double f = 1.0;
        for (int i = 2; i <= (int)ri; ++i) f *= i;
        return f;
    }
    if (is_int) {
        int m = (int)ri;
        return std::sqrt(2.0 * M_PI * m) * std::pow(m / M_E, m) * (1.0 + 1.0 / (12.0 * m) + 1.0 / (288.0 * m * m));
    }
    if (std::fabs(2 * n - std::round(2 * n)) < 1e-10 && n <= 11.5) {
        int k = (int)std::round(n - 0.5);
        double df = 1.0;
        for (int i = 1; i <= 2 * k + 1; i += 2) df *= i;
        return df / std::pow(2.0, k + 1) * std::sqrt(M_PI);
    }
    return std::tgamma(n);
}

// <j1 m1 j2 m2 | J M> (Racah), arguments as 2j+1 and 2m
static double host_cg(int dj1, int m1x2, int dj2, int m2x2, int dJ, int Mx2) {
    double j1 = (dj1 - 1) / 2.0, j2 = (dj2 - 1) / 2.0, J = (dJ - 1) / 2.0;
    double m1 = m1x2 / 2.0, m2 = m2x2 / 2.0, M = Mx2 / 2.0;
    if (std::fabs(M - m1 - m2) > 1e-10) return 0.0;
    if (J < std::fabs(j1 - j2) || J > j1 + j2) return 0.0;
    if (std::fabs(m1) > j1 || std::fabs(m2) > j2 || std::fabs(M) > J) return 0.0;
    double pre = std::sqrt((2 * J + 1) * host_fact(j1 + j2 - J) * host_fact(j1 - j2 + J) * host_fact(-j1 + j2 + J) /
                           host_fact(j1 + j2 + J + 1));
    pre *= std::sqrt(host_fact(J + M) * host_fact(J - M) * host_fact(j1 - m1) * host_fact(j1 + m1) *
                     host_fact(j2 - m2) * host_fact(j2 + m2));
    int kmin = (int)std::ceil(std::max(0.0, std::max(j2 - J - m1, j1 - J + m2)));
    int kmax = (int)std::floor(std::min(j1 + j2 - J, std::min(j1 - m1, j2 + m2)));
    double s = 0;
    for (int k = kmin; k <= kmax; ++k) {
        double den = host_fact(k) * host_fact(j1 + j2 - J - k) * host_fact(j1 - m1 - k) * host_fact(j2 + m2 - k) *
                     host_fact(J - j2 + m1 + k) * host_fact(J - j1 - m2 + k);
        if (std::fabs(den) > 1e-15) s += ((k & 1) ? -1.0 : 1.0) / den;
    }
    return pre * s;
}

// table entry for amp[s][s1][s2] = qs^L / sqrt(2j+1) <j1 s1 j2 s2|S s1+s2> <S s1+s2 L M|j s>
//   N_LM phase(M) P_L^|M|(cos) e^{-iM phi}:  everything but qs^L, P and the phase factor
static void node_cg_table(const PNode& n, int S, int L, std::vector<double>& out) {
    for (int si = 0; si < n.dj; ++si)
        for (int s1i = 0; s1i < n.dj1; ++s1i)
            for (int s2i = 0; s2i < n.dj2; ++s2i) {
                int s = n.dj - 1 - 2 * si, s1 = n.dj1 - 1 - 2 * s1i, s2 = n.dj2 - 1 - 2 * s2i;
                int Mx2 = s - s1 - s2;
                double v = 0.0;
                if (Mx2 % 2 == 0) {
                    int M = Mx2 / 2, am = std::abs(M);
                    if (am <= L) {
                        double cg1 = host_cg(n.dj1, s1, n.dj2, s2, S, s1 + s2);
                        double cg2 = host_cg(S, s1 + s2, 2 * L + 1, Mx2, n.dj, s);
                        double norm = (L == 0) ? 0.5 * std::sqrt(1.0 / M_PI)
                                               : std::sqrt((2.0 * L + 1.0) * host_fact(L - am) / (4.0 * M_PI * host_fact(L + am)));
                        double ph = (M < 0 && (am & 1)) ? -1.0 : 1.0;
                        v = cg1 * cg2 * norm * ph / std::sqrt((double)n.dj);
                    }
                }
                out.push_back(v);
            }
}

// ---------------------------------------------------------------------------
// generated-kernel cache (one module per device and source text)
// ---------------------------------------------------------------------------
static std::mutex g_mod_mu;
static std::map<std::pair<int, std::string>, hipModule_t> g_mods;

// hipRTC options of the generated kernels: gfx950, and no packed FP32 instructions
// (the same target-feature switch as the library's own hipcc build)
static const char* RTC_OPTS[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=fast",
                                 "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
static const int RTC_NOPTS = (int)(sizeof(RTC_OPTS) / sizeof(RTC_OPTS[0]));

static hipModule_t compile_eval(int device, const std::string& src) {
    std::lock_guard<std::mutex> lk(g_mod_mu);
    auto key = std::make_pair(device, src);
    auto it = g_mods.find(key);
    hipModule_t mod;
    if (it != g_mods.end()) {
        mod = it->second;
    } else {
        hiprtcProgram prog;
        if (hiprtcCreateProgram(&prog, src.c_str(), "ctpwa_eval.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS)
            throw std::runtime_error("hiprtcCreateProgram failed");
        hiprtcResult r = hiprtcCompileProgram(prog, RTC_NOPTS, RTC_OPTS);
        if (r != HIPRTC_SUCCESS) {
            size_t ls = 0;
            hiprtcGetProgramLogSize(prog, &ls);
            std::string log(ls + 1, '\0');
            hiprtcGetProgramLog(prog, &log[0]);
            hiprtcDestroyProgram(&prog);
            throw std::runtime_error("hipRTC compile of the fused kernel failed: " + log);
        }
        size_t cs = 0;
        hiprtcGetCodeSize(prog, &cs);
        std::vector<char> code(cs);
        hiprtcGetCode(prog, code.data());
        hiprtcDestroyProgram(&prog);
        HIPCHECK(hipModuleLoadData(&mod, code.data()));
        g_mods[key] = mod;
    }
    return mod;
}

// ---------------------------------------------------------------------------
// handle
// ---------------------------------------------------------------------------
// upload staging, one per device for the process lifetime: two pinned host buffers and
// two device buffers of STAGE_BYTES (momenta chunks), the mixed form's FP64 batch (grown
// on demand), a stream of default priority and an event per buffer.  Handle creations on
// one device take turns on it.
static const size_t STAGE_BYTES = (size_t)32 << 20;
struct Staging {
    std::mutex mu;
    char* host[2] = {nullptr, nullptr};
    double* dev[2] = {nullptr, nullptr};
    double2* sl64 = nullptr;
    double* inv64 = nullptr;
    size_t n_sl64 = 0, n_inv64 = 0;
    hipStream_t s = nullptr;
    hipEvent_t copied[2] = {nullptr, nullptr};    // the DMA out of host[k] has finished
    void reserve_mixed(size_t nsl, size_t ninv) {
        if (nsl > n_sl64) {
            if (sl64) HIPCHECK(hipFree(sl64));
            sl64 = nullptr;
            HIPCHECK(hipMalloc(&sl64, nsl * sizeof(double2)));
            n_sl64 = nsl;
        }
        if (ninv > n_inv64) {
            if (inv64) HIPCHECK(hipFree(inv64));
            inv64 = nullptr;
            HIPCHECK(hipMalloc(&inv64, ninv * sizeof(double)));
            n_inv64 = ninv;
        }
    }
};
static std::mutex g_stage_mu;
static std::map<int, Staging*> g_stage;
static Staging* staging(int device) {          // called with `device` current
    std::lock_guard<std::mutex> lk(g_stage_mu);
    auto it = g_stage.find(device);
    if (it != g_stage.end()) return it->second;
    Staging* S = new Staging();
    for (int k = 0; k < 2; ++k) {
        HIPCHECK(hipHostMalloc((void**)&S->host[k], STAGE_BYTES, hipHostMallocDefault));
        HIPCHECK(hipMalloc((void**)&S->dev[k], STAGE_BYTES));
        HIPCHECK(hipEventCreateWithFlags(&S->copied[k], hipEventDisableTiming));
    }
    HIPCHECK(hipStreamCreateWithFlags(&S->s, hipStreamNonBlocking));
    for (int k = 0; k < 2; ++k) HIPCHECK(hipEventRecord(S->copied[k], S->s));
    g_stage[device] = S;
    return S;
}

// ---------------------------------------------------------------------------
// matrix-core Gram of ncol columns: panel pairs, tile-pair list, row-share workgroups
// and the partials buffer [nwg][ntp][256]
struct GramTabs {
    int ncol = 0, npp = 0, ntp = 0, nwg = 0;
    int4* pp = nullptr;
    int2* tiles = nullptr;
    double* part = nullptr;
};
static long long gram_part_doubles(int ncol) {
    int npp = 0;
    const int ntp = ctpwa_gram_tables(ncol, nullptr, &npp, nullptr);
    return (long long)ctpwa_gram_nwg(npp) * ntp * 256;
}

// serial numbers of handles: a handle created after another was destroyed may reuse its
// address, never its serial
static std::atomic<unsigned long long> g_handle_serial{0};
// handles alive in the process (a handle's member counts itself in and out)
static std::atomic<int> g_live_handles{0};
struct LiveCount {
    LiveCount() { ++g_live_handles; }
    ~LiveCount() { --g_live_handles; }
    LiveCount(const LiveCount&) = delete;
    LiveCount& operator=(const LiveCount&) = delete;
};

// one set of Hessian buffers: a chunk of Jacobian rows (FP64, or FP32 in the mixed
// form), u rows, weights, the rows kernel's per-wave column sums, and the matrix-core
// Gram partials of the Jacobian (2C columns) and of u (NU columns)
struct HessWS {
    long long chunk = 0, cs_ld = 0;
    void* J = nullptr;
    double *U = nullptr, *W = nullptr, *cs = nullptr;
    GramTabs gJ, gU;
};

struct ctpwa_handle {
    LiveCount live;
    int device = 0;
    const unsigned long long serial = ++g_handle_serial;
    Model_ M;
    long long n[3] = {0, 0, 0}, off[3] = {0, 0, 0}, ntot = 0;
    double Wd = 0, Wb = 0;
    int pw[3] = {0, 0, 0};
    double wc[3] = {1.0, 1.0, 1.0};
    std::vector<void*> allocs;
    double2* sl = nullptr;
    double* inv = nullptr;
    double* wts = nullptr;
    double* table = nullptr;
    double* part = nullptr;
    int nwg = 0, g0[3] = {0, 0, 0}, g1[3] = {0, 0, 0};
    int* wg_cls = nullptr;
    long long* wg_rng = nullptr;
    int nacc = 0;
    PrepArgs prep{};
    FinArgs fin{};
    AuxArgs aux{};
    hipModule_t mod = nullptr;
    hipFunction_t fn_eval = nullptr;       // one likelihood
    hipFunction_t fn_group = nullptr;      // many likelihoods of this model in one launch
    hipFunction_t fn_group_p = nullptr;    // the same, persistent form (graph replays)
    hipFunction_t fn_aux = nullptr;        // amplitudes / intensities (models with Custom lineshapes)
    hipFunction_t fn_custom = nullptr;     // Custom factors and derivatives per event (models with Custom lineshapes)
    double* cfb = nullptr;                 // their per-event rows, [ncustom * ncf][ntot]
    int* wave_ptr = nullptr;               // partial wave w = blocks wave_blk[wave_ptr[w] .. wave_ptr[w+1])
    int* wave_blk = nullptr;
    std::string src;
    // Hessian workspace (allocated on the first ctpwa_hessian call): the generated rows
    // kernel, one chunk of Jacobian rows (FP64, or FP32 in the mixed form), u rows,
    // weights, per-wave column sums, and the matrix-core Gram tables and partials of the
    // Jacobian (2C columns) and of u (NU columns)
    // Buffer set 0 serves the phase space on the caller's stream; set 1 serves data and
    // background on a stream of the handle, so the two run concurrently.
    bool hess_ready = false;
    int hC = 0, hNU = 0, hNZ = 0, hnG = 0, hnUu = 0;
    hipModule_t hmod = nullptr;
    hipFunction_t fn_hrows = nullptr;
    HessWS hws[2];
    hipStream_t hs2 = nullptr;                  // default priority, non-blocking
    hipEvent_t hev[2] = {nullptr, nullptr};     // start of the classes' work; end of set 1's work
    double *htot = nullptr, *hHext = nullptr;
    int *hmap_ptr = nullptr, *hmap_x = nullptr;   // fit parameter -> extended entries (CSR)
    int* hperm = nullptr;                         // rows kernel's summed columns -> [Z | u] order
    double* hmap_r = nullptr;
    // models with Custom lineshapes: second-order factor rows for the Hessian, from a
    // module compiled on the first Hessian call
    hipModule_t mod2 = nullptr;
    hipFunction_t fn_custom2 = nullptr;
    double* cfb2 = nullptr;
    int *hbcp = nullptr, *hbcl = nullptr;   // per-block Custom factor lists (CSR)
    // amplitude-Gram workspace (allocated on the first ctpwa_amplitude_gram call)
    bool agram_ready = false;
    long long gchunk = 0;
    double *gX = nullptr, *gW = nullptr, *gtot = nullptr;
    GramTabs gg;
    // no free resonance parameter: the phase-space amplitude Gram without couplings or
    // theta coefficients, built once at creation; the evaluations then scan data and
    // background only (fin.gram points here)
    double2* d_gram = nullptr;
    std::vector<int> used_inv;
    // storage precision (0 FP64, 1 mixed) and the FP32 stores of the mixed form
    int mixed = 0;
    float2* slf = nullptr;
    float* invf = nullptr;
    int* d_imap = nullptr;
    const void* slp = nullptr;      // the SL store the fused kernel reads
    int* d_psign = nullptr;         // [NU] signs of the mirrored polarisation (folded store)
    unsigned int* fin_done = nullptr;   // finishing-workgroup counter of the final assembly (zero between calls)
    EvStore store() const { return EvStore{sl, slf, inv, invf, d_imap, ntot, mixed, M.npstore(), d_psign}; }

    template <class T>
    T* alloc(size_t count) {
        void* p = nullptr;
        if (count == 0) count = 1;
        // 256 bytes of slack after every buffer: whole 16-byte chunks never end past it
        HIPCHECK(hipMalloc(&p, count * sizeof(T) + 256));
        allocs.push_back(p);
        return (T*)p;
    }
    template <class T>
    T* upload(const std::vector<T>& v) {
        T* p = alloc<T>(v.size());
        if (!v.empty()) HIPCHECK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
        return p;
    }
    ~ctpwa_handle() {
        int prev = 0;
        (void)hipGetDevice(&prev);
        (void)hipSetDevice(device);
        if (hs2) (void)hipStreamSynchronize(hs2);
        if (hs2) (void)hipStreamDestroy(hs2);
        for (hipEvent_t ev : hev)
            if (ev) (void)hipEventDestroy(ev);
        // the modules belong to the process-wide compile cache (shared by handles of one
        // model), not to the handle
        for (void* p : allocs) (void)hipFree(p);
        (void)hipSetDevice(prev);
    }
};

// code-generation options for the fused kernel (tuning; defaults are the measured best)
// code-generation options; the environment overrides them for measurements
static void apply_codegen_env(Model_& M) {
    // next-event loads in registers: measured 4 % faster on the mixed form of a model of
    // about ten amplitudes over four SL tensors (216 of 256 VGPRs without); a larger model
    // has no registers to spare (a 20-amplitude four-body model spills with it)
    M.prefetch_sl = (M.mixed && M.P.namp <= 12 && (int)M.us.size() * M.P.npol <= 8) ? 1 : 0;
    if (const char* e = std::getenv("CTPWA_EVAL_WAVES")) M.min_waves = std::atoi(e);
    if (const char* e = std::getenv("CTPWA_EVAL_LDS")) M.lds_pipe = std::atoi(e) != 0;
    if (const char* e = std::getenv("CTPWA_EVAL_INVTREE")) M.inv_tree = std::atoi(e) != 0;
    if (const char* e = std::getenv("CTPWA_EVAL_PREFETCH")) {
        int d = std::atoi(e);
        if (d < 0 || d > 2) throw std::runtime_error("CTPWA_EVAL_PREFETCH must be 0, 1 or 2");
        M.prefetch_sl = d;
    }
    if (const char* e = std::getenv("CTPWA_EVAL_WGS")) {
        int w = std::atoi(e);
        if (w != 64 && w != 128 && w != 256) throw std::runtime_error("CTPWA_EVAL_WGS must be 64, 128 or 256");
        M.wgs = w;
    }
}

static void build_tables(ctpwa_handle* h, std::vector<DTopo>& topos, std::vector<DNode>& nodes,
                         std::vector<DKin>& kins, std::vector<DU>& us, std::vector<double>& cgtab,
                         std::vector<DInv>& invs) {
    const Model_& M = h->M;
    const Plan& P = M.P;
    if ((int)M.kin.size() > MAXKIN) throw std::runtime_error("too many distinct kinematic nodes");
    if (P.npol > MAXPOL) throw std::runtime_error("too many polarisation states");
    if (P.nfinal + 1 + MAXNODE > MAXPART) throw std::runtime_error("too many final-state particles");
    kins.resize(M.kin.size());
    for (size_t i = 0; i < M.kin.size(); ++i)
        kins[i] = DKin{M.kin[i].mset, M.kin[i].d1set, M.kin[i].d2set, -1, -1, 0};
    for (size_t t = 0; t < P.topos.size(); ++t) {
        const PTopo& T = P.topos[t];
        DTopo dt{};
        dt.nn = (int)T.nodes.size();
        dt.node0 = (int)nodes.size();
        // daughters-first order of the intermediates
        std::vector<int> order;
        std::function<void(int)> visit = [&](int ref) {
            if (ref >= 0) return;
            int k = -ref - 1;
            visit(T.nodes[k].d1ref);
            visit(T.nodes[k].d2ref);
            if (std::find(order.begin(), order.end(), k) == order.end()) order.push_back(k);
        };
        visit(T.nodes[0].d1ref);
        visit(T.nodes[0].d2ref);
        if ((int)order.size() != dt.nn - 1) throw std::runtime_error("topology has unreachable nodes");
        for (size_t i = 0; i < order.size(); ++i) dt.order[i] = order[i];
        topos.push_back(dt);
        for (size_t k = 0; k < T.nodes.size(); ++k) {
            const PNode& n = T.nodes[k];
            int kin = M.topo_kin[t][k];
            nodes.push_back(DNode{n.dj, n.dj1, n.dj2, n.d1ref, n.d2ref, n.has_bf, n.lmin, kin, n.bf_d, n.mm, n.md1, n.md2});
            if (kins[kin].owner_topo < 0) { kins[kin].owner_topo = (int)t; kins[kin].owner_node = (int)k; }
            if (n.dj1 > 5 || n.dj2 > 5) throw std::runtime_error("daughter spin above 2 is not built");
            // an intermediate daughter's q0 mass must come from a configured mass or a
            // resonance mass parameter (otherwise q0 would depend on the event)
        }
    }
    for (const USpec& U : M.us) {
        DU du{};
        du.topo = U.topo;
        du.mask = (int)U.mask;
        const PTopo& T = P.topos[U.topo];
        for (size_t k = 0; k < T.nodes.size(); ++k) {
            du.S[k] = T.slc[U.slc][k].first;
            du.L[k] = T.slc[U.slc][k].second;
            if (du.L[k] > 7) throw std::runtime_error("orbital angular momentum above 7 is not built");
            du.cg[k] = (int)cgtab.size();
            node_cg_table(T.nodes[k], du.S[k], du.L[k], cgtab);
        }
        us.push_back(du);
    }
    for (const InvSpec& s : M.inv) invs.push_back(DInv{s.kin, s.kind, s.L, 0, s.d, s.ma, s.mb});
}

static void check_q0_masses(const Model_& M) {
    const Plan& P = M.P;
    for (const PBlock& B : P.blocks) {
        if (B.nsl == 0) continue;
        const PTopo& T = P.topos[B.topo];
        for (const PNode& n : T.nodes) {
            for (int side = 0; side < 2; ++side) {
                double cfg = side ? n.md2 : n.md1;
                int ref = side ? n.d2ref : n.d1ref;
                if (cfg > 0) continue;
                if (ref >= 0 && cfg == 0.0) continue;    // massless final-state particle: q0 mass 0
                bool ok = false;
                // an occupying resonance gives its mass parameter, or 1 when it has none
                // (src/AmpGen.cu:1504-1527)
                if (ref < 0)
                    for (const PRes& r : B.res) if (r.node == -ref - 1) ok = true;
                if (!ok) throw std::runtime_error("a daughter without a configured mass and without a resonance mass "
                                                  "parameter makes q0 event-dependent; not built");
            }
        }
    }
}

// workgroups of the fused kernel: about 1024 (four 128-thread workgroups per CU), at least
// 256 events each, shared out among the classes by event count, each a whole number of
// 256-event steps of one class.  scan_phsp == false (phase-space term from the stored
// Gram): the phase space gets none.
static void setup_partition(ctpwa_handle* h, bool scan_phsp) {
    const long long nev = (scan_phsp ? h->n[0] : 0) + h->n[1] + h->n[2];
    long long target = std::min<long long>(1024, std::max<long long>(1, (nev + 255) / 256));
    // relative cost of a data or background event: the workgroups of each class get events
    // in inverse proportion, so that the classes finish together.  A phase-space event costs
    // about 20 FP64 instructions per amplitude (sums and gradient accumulators), 10 per
    // propagator and 8 per SL tensor and polarisation; a data event of a constant-weight
    // class adds the two frexp of the product of intensities, the reciprocal of I and the
    // clamps, about 40 (a class with per-event weights takes a logarithm per event, about
    // 100 more, as the first form of the kernel did for every data event: there the fastest
    // split of the 3-pion model was at 1.3-1.4).  Environment CTPWA_DATA_COST overrides it.
    const Model_& M = h->M;
    const double base = 20.0 * M.P.namp + 10.0 * (double)M.dens.size() + 8.0 * (double)M.us.size() * M.P.npol + 40.0;
    const bool any_pw = (h->n[1] > 0 && h->pw[1]) || (h->n[2] > 0 && h->pw[2]);
    double wcost = 1.0 + (any_pw ? 140.0 : 40.0) / base;
    if (const char* e = std::getenv("CTPWA_DATA_COST")) wcost = std::atof(e);
    if (!(wcost >= 0.25 && wcost <= 4.0)) throw std::runtime_error("CTPWA_DATA_COST must lie in [0.25, 4]");
    const double weff = (scan_phsp ? (double)h->n[0] : 0.0) + wcost * (double)(h->n[1] + h->n[2]);
    std::vector<int> cls;
    std::vector<long long> rng;
    for (int c = 0; c < 3; ++c) {
        h->g0[c] = (int)cls.size();
        long long n = (c == 0 && !scan_phsp) ? 0 : h->n[c];
        if (n > 0) {
            const double wn = (c == 0 ? 1.0 : wcost) * (double)n;
            long long want = std::max<long long>(1, (long long)std::llround((double)target * wn / weff));
            long long chunk = (n + want - 1) / want;
            const long long gw = h->M.wgs;          // whole iterations for every thread
            chunk = ((chunk + gw - 1) / gw) * gw;
            long long nw = (n + chunk - 1) / chunk;
            for (long long g = 0; g < nw; ++g) {
                cls.push_back(c);
                rng.push_back(h->off[c] + g * chunk);
                rng.push_back(h->off[c] + std::min(n, (g + 1) * chunk));
            }
        }
        h->g1[c] = (int)cls.size();
    }
    h->nwg = (int)cls.size();
    h->wg_cls = h->upload(cls);
    h->wg_rng = h->upload(rng);
}

static bool constant_weights(const double* w, long long n, double* c) {
    if (!w || n == 0) { *c = 1.0; return true; }
    for (long long i = 1; i < n; ++i) if (w[i] != w[0]) return false;
    *c = w[0];
    return true;
}

// ---------------------------------------------------------------------------
// ctpwa_nll_grad_noalloc: the per-workgroup partial sums and the class totals of an
// evaluation live in `scratch`, a file-scope device array of NOALLOC_DOUBLES doubles
// (2 MiB) that every GPU the library runs on holds until the process ends, instead of
// the handle's buffers.  Host code never passes the array's name to a kernel (it is a
// host-side stand-in there): each device's own address comes from hipGetSymbolAddress,
// taken once per device with the device current (at ctpwa_create) and kept in g_scratch.
// Every evaluation overwrites what it reads (each workgroup writes its partial row, the
// reduction writes the totals), so nothing needs clearing between calls; calls that
// overlap on one GPU share the array and race.
static const size_t NOALLOC_DOUBLES = (size_t)1 << 18;
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];
static std::mutex g_scratch_mu;
static std::map<int, double*> g_scratch;

// the device's copy of `scratch` (call with the device current)
static double* scratch_setup(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch.find(device);
    if (it != g_scratch.end()) return it->second;
    double* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        throw std::runtime_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[device] = ptr;
    return ptr;
}
static double* scratch_lookup(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch.find(device);
    return it == g_scratch.end() ? nullptr : it->second;
}

// After the SL tensors are built: when both polarisations of every SL tensor are mirror
// images over all events (polarisation 1 = sign_u conj(polarisation 0) to 1e-12 of the
// tensor's largest component; spin-0 final states of a J/psi with M = +1, -1), keep only
// polarisation 0 ([NU][1][ld]) and the signs: the kernels then read half the SL bytes.
static void fold_polarisations(ctpwa_handle* h, const unsigned long long* d_sym, int nu, hipStream_t s) {
    std::vector<unsigned long long> raw((size_t)4 * nu);
    HIPCHECK(hipMemcpy(raw.data(), d_sym, raw.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    std::vector<int> sign((size_t)nu, 1);
    for (int u = 0; u < nu; ++u) {
        double dp, dm, m0;
        std::memcpy(&dp, &raw[4 * u + 0], 8);
        std::memcpy(&dm, &raw[4 * u + 1], 8);
        std::memcpy(&m0, &raw[4 * u + 2], 8);
        if (!(std::min(dp, dm) <= 1e-12 * m0) && !(dp == 0.0 || dm == 0.0)) return;   // not mirror images
        sign[u] = (dp <= dm) ? 1 : -1;
    }
    const size_t es = h->mixed ? sizeof(float2) : sizeof(double2);
    void* old = h->mixed ? (void*)h->slf : (void*)h->sl;
    void* neu = h->alloc<unsigned char>((size_t)nu * h->ntot * es);
    for (int u = 0; u < nu; ++u)
        HIPCHECK(hipMemcpyAsync((char*)neu + (size_t)u * h->ntot * es, (const char*)old + (size_t)2 * u * h->ntot * es,
                                (size_t)h->ntot * es, hipMemcpyDeviceToDevice, s));
    HIPCHECK(hipStreamSynchronize(s));
    HIPCHECK(hipFree(old));
    h->allocs.erase(std::remove(h->allocs.begin(), h->allocs.end(), old), h->allocs.end());
    if (h->mixed) h->slf = (float2*)neu; else h->sl = (double2*)neu;
    h->slp = neu;
    h->M.nps = 1;
    h->M.psign = sign;
    h->d_psign = h->upload(sign);
}

static bool fixed_gram_applies(const ctpwa_handle* h);
static void build_fixed_gram(ctpwa_handle* h);

extern "C" int ctpwa_create(const char* plan_text, const double* phsp, long long n_phsp, const double* w_phsp,
                            const double* data, long long n_data, const double* w_data,
                            const double* bkg, long long n_bkg, const double* w_bkg, int precision, int device,
                            ctpwa_handle** out) {
    if (!out) return fail("ctpwa_create: out is NULL");
    *out = nullptr;
    ctpwa_handle* h = nullptr;
    int prev = 0;
    (void)hipGetDevice(&prev);
    try {
        if (!plan_text) throw std::runtime_error("plan_text is NULL");
        if (n_phsp < 0 || n_data < 0 || n_bkg < 0) throw std::runtime_error("negative event count");
        if (n_phsp == 0) throw std::runtime_error("phase-space sample is empty");
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) throw std::runtime_error("no HIP device available");
        if (device < 0 || device >= ndev) throw std::runtime_error("device index out of range");
        HIPCHECK(hipSetDevice(device));
        scratch_setup(device);     // this device's address of the no-alloc entry point's scratch
        h = new ctpwa_handle();
        h->device = device;
        if (precision != 0 && precision != 1) throw std::runtime_error("precision must be 0 (FP64) or 1 (mixed)");
        h->M = derive(parse_plan(plan_text));
        h->M.mixed = precision;
        h->mixed = precision;
        apply_codegen_env(h->M);
        check_q0_masses(h->M);
        const Model_& M = h->M;
        const Plan& P = M.P;
        const double* moms[3] = {phsp, data, bkg};
        const double* ws[3] = {w_phsp, w_data, w_bkg};
        long long ns[3] = {n_phsp, n_data, n_bkg};
        for (int c = 0; c < 3; ++c) {
            h->n[c] = ns[c];
            // class offsets and the leading dimension are multiples of 64 events, so
            // every row of a tile starts 16-byte aligned (the pipelined kernel copies
            // 16-byte chunks)
            h->off[c] = (h->ntot + 63) / 64 * 64;
            h->ntot = h->off[c] + ns[c];
            if (ns[c] > 0 && !moms[c]) throw std::runtime_error("event array is NULL");
        }
        h->ntot = (h->ntot + 63) / 64 * 64;
        // weights: constant per class unless they vary
        std::vector<double> wall;
        bool any_var = false;
        for (int c = 1; c < 3; ++c) {
            double cst;
            if (!constant_weights(ws[c], ns[c], &cst)) { any_var = true; h->pw[c] = 1; }
            h->wc[c] = cst;
        }
        if (ws[0] && n_phsp > 0) {
            double cst;
            if (!constant_weights(ws[0], n_phsp, &cst) || cst != 1.0)
                throw std::runtime_error("phase-space weights are not built (the reference's phase-space mean is unweighted)");
        }
        h->Wd = 0;
        for (long long i = 0; i < n_data; ++i) h->Wd += w_data ? w_data[i] : 1.0;
        h->Wb = 0;
        for (long long i = 0; i < n_bkg; ++i) h->Wb += w_bkg ? w_bkg[i] : 1.0;
        if (any_var) {
            wall.assign(h->ntot, 1.0);
            for (int c = 1; c < 3; ++c)
                if (ws[c]) std::copy(ws[c], ws[c] + ns[c], wall.begin() + h->off[c]);
            h->wts = h->upload(wall);
        }

        std::vector<DTopo> topos;
        std::vector<DNode> nodes;
        std::vector<DKin> kins;
        std::vector<DU> us;
        std::vector<double> cgtab;
        std::vector<DInv> invs;
        build_tables(h, topos, nodes, kins, us, cgtab, invs);
        int nu = (int)us.size(), ninv = (int)invs.size();
        std::vector<int> imap = M.imap();
        int nind = 0, ninf = 0;
        for (int m : imap) (m >= 0 ? nind : ninf)++;
        if (h->mixed) {
            h->slf = h->alloc<float2>((size_t)std::max(1, nu) * P.npol * h->ntot);
            h->inv = h->alloc<double>((size_t)std::max(1, nind) * h->ntot);
            h->invf = h->alloc<float>((size_t)std::max(1, ninf) * h->ntot);
            h->d_imap = h->upload(imap);
            h->slp = h->slf;
        } else {
            h->sl = h->alloc<double2>((size_t)std::max(1, nu) * P.npol * h->ntot);
            h->inv = h->alloc<double>((size_t)std::max(1, ninv) * h->ntot);
            h->slp = h->sl;
        }
        {
            DTopo* d_topos = h->upload(topos);
            DNode* d_nodes = h->upload(nodes);
            DKin* d_kins = h->upload(kins);
            DU* d_us = h->upload(us);
            double* d_cg = h->upload(cgtab);
            int* d_pol = h->upload(P.polmap);
            DInv* d_invs = h->upload(invs);
            const double* d_tabs = h->upload(M.tabs);
            // events class by class, in chunks through the device's staging pool: the
            // host copy of a chunk into pinned memory overlaps the DMA and the init
            // kernel of the previous one; only the pool's stream is waited for
            Staging* S = staging(device);
            std::lock_guard<std::mutex> lk(S->mu);
            const size_t ev_bytes = (size_t)P.nfinal * 4 * sizeof(double);
            const long long chunk = std::max<long long>(1, (long long)(STAGE_BYTES / ev_bytes));
            if (h->mixed) S->reserve_mixed((size_t)std::max(1, nu) * P.npol * chunk, (size_t)std::max(1, ninv) * chunk);
            // two polarisations: check over every event whether they are mirror images
            unsigned long long* d_sym = nullptr;
            if (P.npol == 2 && nu > 0) {
                d_sym = h->alloc<unsigned long long>((size_t)4 * nu);
                HIPCHECK(hipMemsetAsync(d_sym, 0, (size_t)4 * nu * sizeof(unsigned long long), S->s));
            }
            int b = 0;
            for (int c = 0; c < 3; ++c) {
                for (long long s0 = 0; s0 < ns[c]; s0 += chunk) {
                    long long cnt = std::min(chunk, ns[c] - s0);
                    HIPCHECK(hipEventSynchronize(S->copied[b]));      // host[b] free again
                    std::memcpy(S->host[b], moms[c] + (size_t)s0 * P.nfinal * 4, (size_t)cnt * ev_bytes);
                    HIPCHECK(hipMemcpyAsync(S->dev[b], S->host[b], (size_t)cnt * ev_bytes, hipMemcpyHostToDevice, S->s));
                    HIPCHECK(hipEventRecord(S->copied[b], S->s));
                    if (h->mixed) {
                        HIPCHECK(ctpwa_launch_init(S->dev[b], cnt, P.nfinal, d_topos, (int)topos.size(), d_nodes, d_kins,
                                                   (int)kins.size(), d_us, nu, d_cg, d_pol, P.npol, d_invs, ninv, d_tabs,
                                                   S->sl64, S->inv64, cnt, 0, S->s));
                        if (d_sym) HIPCHECK(ctpwa_launch_polsym(S->sl64, cnt, 0, cnt, nu, d_sym, S->s));
                        HIPCHECK(ctpwa_launch_to_mixed(S->sl64, S->inv64, cnt, nu * P.npol, ninv, h->d_imap, h->slf,
                                                       h->inv, h->invf, h->ntot, h->off[c] + s0, S->s));
                    } else {
                        HIPCHECK(ctpwa_launch_init(S->dev[b], cnt, P.nfinal, d_topos, (int)topos.size(), d_nodes, d_kins,
                                                   (int)kins.size(), d_us, nu, d_cg, d_pol, P.npol, d_invs, ninv, d_tabs,
                                                   h->sl, h->inv, h->ntot, h->off[c] + s0, S->s));
                        if (d_sym) HIPCHECK(ctpwa_launch_polsym(h->sl, h->ntot, h->off[c] + s0, cnt, nu, d_sym, S->s));
                    }
                    b ^= 1;
                }
            }
            HIPCHECK(hipStreamSynchronize(S->s));
            if (d_sym) fold_polarisations(h, d_sym, nu, S->s);
        }

        // model tables for prep / finalize / aux
        std::vector<DBlock> blocks;
        std::vector<DRes> res;
        std::vector<int> slcL, slc0;
        std::vector<DDen> dens;
        std::vector<DAmp> amps;
        std::vector<DDenTerms> dterms;
        std::vector<DTerm> terms;
        for (size_t t = 0; t < P.topos.size(); ++t) {
            slc0.push_back((int)(slcL.size() / MAXNODE));
            for (const auto& row : P.topos[t].slc) {
                for (int k = 0; k < MAXNODE; ++k) slcL.push_back(k < (int)row.size() ? row[k].second : 0);
            }
        }
        for (size_t b = 0; b < P.blocks.size(); ++b) {
            const PBlock& B = P.blocks[b];
            DBlock db{};
            db.topo = B.topo;
            db.nloc = (int)M.bvars[b].size();
            db.nres = (int)B.res.size();
            db.res0 = (int)res.size();
            db.nsl = B.nsl;
            db.site = B.site;
            db.nden = (int)M.block_dens[b].size();
            db.den0 = db.nden ? M.block_dens[b][0] : 0;
            for (int i = 0; i < NLOC; ++i) db.vars[i] = i < db.nloc ? M.bvars[b][i] : -1;
            for (size_t k = 1; k < M.block_dens[b].size(); ++k)
                if (M.block_dens[b][k] != db.den0 + (int)k) throw std::runtime_error("denominators of a block are not contiguous");
            blocks.push_back(db);
            for (const PRes& r : B.res) {
                DRes dr{};
                dr.node = r.node;
                dr.model = r.model;
                dr.np = (int)r.params.size();
                dr.nch = (int)r.ch.size() / 2;
                if (dr.np > 6 || dr.nch > 4) throw std::runtime_error("too many lineshape parameters or channels");
                for (int i = 0; i < dr.np; ++i) { dr.p[i] = r.params[i]; dr.slot[i] = r.slots[i]; }
                for (int i = 0; i < 2 * dr.nch; ++i) dr.ch[i] = r.ch[i];
                res.push_back(dr);
            }
        }
        for (const DenSpec& D : M.dens) {
            dens.push_back(DDen{D.block, D.res, D.coef_base, D.ncoef});
            DDenTerms dt{(int)D.re.size(), (int)D.im.size(), (int)terms.size(), D.stat ? 1 : 0};
            for (const Term& t : D.re) terms.push_back(DTerm{t.inv, t.coef, t.cval});
            for (const Term& t : D.im) terms.push_back(DTerm{t.inv, t.coef, t.cval});
            dterms.push_back(dt);
        }
        for (int a = 0; a < P.namp; ++a) {
            const AmpSpec& A = M.amps[a];
            amps.push_back(DAmp{A.block, A.slc, P.amap[a], A.coef_base, A.u, 0, P.aratio[a]});
        }
        DBlock* d_blocks = h->upload(blocks);
        DRes* d_res = h->upload(res);
        DNode* d_nodes2 = h->upload(nodes);
        DTopo* d_topos2 = h->upload(topos);
        int* d_slcL = h->upload(slcL);
        int* d_slc0 = h->upload(slc0);
        DDen* d_dens = h->upload(dens);
        DAmp* d_amps = h->upload(amps);
        DDenTerms* d_dterms = h->upload(dterms);
        DTerm* d_terms = h->upload(terms);
        int* d_gslot = h->upload(P.g_slot);
        double* d_gmu = h->upload(P.g_mu);
        double* d_gsig = h->upload(P.g_sig);
        h->table = h->alloc<double>(M.table_size());
        HIPCHECK(hipMemset(h->table, 0, M.table_size() * sizeof(double)));

        h->nacc = M.nacc();
        const bool use_gram = fixed_gram_applies(h);
        setup_partition(h, !use_gram);
        h->part = h->alloc<double>((size_t)h->nacc * std::max(1, h->nwg));
        h->fin_done = h->alloc<unsigned int>(1);
        HIPCHECK(hipMemset(h->fin_done, 0, sizeof(unsigned int)));

        h->prep = PrepArgs{d_blocks, d_res, d_nodes2, d_topos2, d_slcL, d_slc0, d_dens, d_amps,
                           (int)dens.size(), P.namp, P.nfree, P.ntheta,
                           (long long)M.off_v(), (long long)M.off_c(), (long long)M.off_rec(),
                           (long long)M.off_th(), (long long)M.off_lth(), 0, nullptr, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        {
            std::vector<char> blob;
            auto put = [&](const void* p, size_t n) -> int {
                size_t o = (blob.size() + 15) / 16 * 16;
                blob.resize(o + (n + 15) / 16 * 16, 0);
                if (n) std::memcpy(blob.data() + o, p, n);
                return (int)o;
            };
            h->prep.o_blocks = put(blocks.data(), blocks.size() * sizeof(DBlock));
            h->prep.o_res = put(res.data(), res.size() * sizeof(DRes));
            h->prep.o_nodes = put(nodes.data(), nodes.size() * sizeof(DNode));
            h->prep.o_topos = put(topos.data(), topos.size() * sizeof(DTopo));
            h->prep.o_slcL = put(slcL.data(), slcL.size() * sizeof(int));
            h->prep.o_slc0 = put(slc0.data(), slc0.size() * sizeof(int));
            h->prep.o_dens = put(dens.data(), dens.size() * sizeof(DDen));
            h->prep.o_amps = put(amps.data(), amps.size() * sizeof(DAmp));
            if (blob.size() <= 48 * 1024) {
                h->prep.meta = h->upload(blob);
                h->prep.meta_bytes = (int)blob.size();
            }
        }
        h->fin = FinArgs{};
        h->fin.part = h->part;
        h->fin.nwg = h->nwg;
        h->fin.nacc = h->nacc;
        for (int c = 0; c < 3; ++c) { h->fin.g0[c] = h->g0[c]; h->fin.g1[c] = h->g1[c]; }
        h->fin.table = h->table;
        h->fin.off_v = (long long)M.off_v();
        h->fin.off_rec = (long long)M.off_rec();
        h->fin.amps = d_amps;
        h->fin.blocks = d_blocks;
        {
            std::vector<DThTerm> th;
            for (const ThTerm& t : M.thterms) th.push_back(DThTerm{M.dens[t.den].block, M.dens[t.den].coef_base + t.coef});
            h->fin.thterms = h->upload(th);
            h->fin.nth = (int)th.size();
            // CSR lists of the gradient contractions (see DGEnt)
            auto local = [&](int b, int j) {
                for (int i = 0; i < blocks[b].nloc; ++i) if (blocks[b].vars[i] == j) return i;
                return -1;
            };
            std::vector<int> thp{0}, cpp{0};
            std::vector<DGEnt> the, cpe;
            for (int j = 0; j < P.ntheta; ++j) {
                for (size_t t = 0; t < th.size(); ++t) {
                    int li = local(th[t].block, j);
                    if (li >= 0) the.push_back(DGEnt{1 + 2 * P.namp + (int)t, th[t].rec * REC + 1 + li, -2.0});
                }
                for (int a = 0; a < P.namp; ++a) {
                    int li = local(amps[a].block, j);
                    if (li >= 0) the.push_back(DGEnt{-a - 1, amps[a].rec * REC + 1 + li, 2.0});
                }
                // Custom factors: accumulated Re(z dPsi/dtheta_j) directly (no coefficient record)
                for (size_t t = 0; t < M.cterms.size(); ++t) {
                    const CTerm& T = M.cterms[t];
                    if (M.bvars[M.customs[T.custom].block][T.local] == j)
                        the.push_back(DGEnt{1 + 2 * P.namp + (int)th.size() + (int)t, -1, 2.0});
                }
                thp.push_back((int)the.size());
            }
            for (int fi = 0; fi < P.nfree; ++fi) {
                for (int a = 0; a < P.namp; ++a)
                    if (amps[a].freeidx == fi) cpe.push_back(DGEnt{a, amps[a].rec * REC, amps[a].ratio});
                cpp.push_back((int)cpe.size());
            }
            if (the.empty()) the.push_back(DGEnt{0, 0, 0.0});
            if (cpe.empty()) cpe.push_back(DGEnt{0, 0, 0.0});
            h->fin.thp = h->upload(thp);
            h->fin.the = h->upload(the);
            h->fin.cpp = h->upload(cpp);
            h->fin.cpe = h->upload(cpe);
            h->fin.nthe = (int)the.size();
            h->fin.ncpe = (int)cpe.size();
            h->fin.nacc = h->nacc;
            h->fin.nfree = P.nfree;
            h->fin.ntheta = P.ntheta;
            if (assemble_lds_bytes(h->fin) > 60 * 1024)
                throw std::runtime_error("too many gradient accumulators for the assembly kernel");
        }
        h->fin.namp = P.namp;
        h->fin.nfree = P.nfree;
        h->fin.ntheta = P.ntheta;
        h->fin.ngauss = (int)P.g_slot.size();
        h->fin.g_slot = d_gslot;
        h->fin.g_mu = d_gmu;
        h->fin.g_sig = d_gsig;
        h->fin.Np = (double)n_phsp;
        h->fin.Wd = h->Wd;
        h->fin.Wb = h->Wb;
        h->fin.totals = h->alloc<double>((size_t)3 * h->nacc);
        h->aux = AuxArgs{h->store(), h->table, (long long)M.off_c(), (long long)M.off_rec(),
                         d_amps, d_blocks, d_dens, d_dterms, d_terms, P.namp, P.npol};
        {
            std::vector<int> bcp(1, 0), bcl;
            for (size_t b = 0; b < P.blocks.size(); ++b) {
                for (int c : M.block_customs[b]) bcl.push_back(c);
                bcp.push_back((int)bcl.size());
            }
            h->hbcp = h->upload(bcp);
            h->hbcl = h->upload(bcl);
            h->aux.bcp = h->hbcp;
            h->aux.bcl = h->hbcl;
            h->aux.cld = h->ntot;
            h->aux.ncf = M.customs.empty() ? 0 : M.ncf();
        }
        {
            std::vector<int> wptr(P.nwave + 1, 0), wblk;
            for (int w = 0; w < P.nwave; ++w) {
                for (size_t b = 0; b < P.blocks.size(); ++b)
                    if (P.bwave[b] == w) wblk.push_back((int)b);
                wptr[w + 1] = (int)wblk.size();
            }
            h->wave_ptr = h->upload(wptr);
            h->wave_blk = h->upload(wblk);
        }

        h->src = generate_eval_kernel(M);
        h->mod = compile_eval(device, h->src);
        HIPCHECK(hipModuleGetFunction(&h->fn_eval, h->mod, "ctpwa_eval"));
        HIPCHECK(hipModuleGetFunction(&h->fn_group, h->mod, "ctpwa_eval_group"));
        HIPCHECK(hipModuleGetFunction(&h->fn_group_p, h->mod, "ctpwa_eval_group_persist"));
        if (!M.customs.empty()) {
            HIPCHECK(hipModuleGetFunction(&h->fn_aux, h->mod, "ctpwa_aux"));
            HIPCHECK(hipModuleGetFunction(&h->fn_custom, h->mod, "ctpwa_custom"));
            h->cfb = h->alloc<double>((size_t)M.customs.size() * M.ncf() * h->ntot);
            h->aux.cfb = h->cfb;
        }
        h->used_inv.assign(M.inv.size(), 0);
        for (const DenSpec& D : M.dens) {
            for (const Term& t : D.re) if (t.inv >= 0) h->used_inv[t.inv] = 1;
            for (const Term& t : D.im) if (t.inv >= 0) h->used_inv[t.inv] = 1;
        }
        if (use_gram) build_fixed_gram(h);
        HIPCHECK(hipStreamSynchronize(nullptr));
        (void)hipSetDevice(prev);
        *out = h;
        return 0;
    } catch (const std::exception& e) {
        delete h;
        (void)hipSetDevice(prev);
        return fail(std::string("ctpwa_create: ") + e.what());
    }
}

extern "C" void ctpwa_destroy(ctpwa_handle* h) { delete h; }

extern "C" int ctpwa_dims(const ctpwa_handle* h, int* n_amp, int* n_free, int* n_theta, int* n_pol) {
    if (!h) return fail("ctpwa_dims: NULL handle");
    if (n_amp) *n_amp = h->M.P.namp;
    if (n_free) *n_free = h->M.P.nfree;
    if (n_theta) *n_theta = h->M.P.ntheta;
    if (n_pol) *n_pol = h->M.P.npol;
    return 0;
}

extern "C" int ctpwa_param_info(const ctpwa_handle* h, double* init, double* lower, double* upper) {
    if (!h) return fail("ctpwa_param_info: NULL handle");
    const Plan& P = h->M.P;
    for (int j = 0; j < P.ntheta; ++j) {
        if (init) init[j] = P.th_init[j];
        if (lower) lower[j] = P.th_lo[j];
        if (upper) upper[j] = P.th_hi[j];
    }
    return 0;
}

// Custom factors of every stored event at the parameter point the table holds
static hipError_t launch_custom(ctpwa_handle* h, hipStream_t s) {
    if (!h->fn_custom) return hipSuccess;
    void* args[] = {&h->table, &h->inv, &h->invf, &h->ntot, &h->cfb};
    unsigned grid = (unsigned)std::min<long long>(4096, (h->ntot + 255) / 256);
    return hipModuleLaunchKernel(h->fn_custom, grid, 1, 1, 256, 1, 1, 0, s, args, nullptr);
}

static hipError_t launch_eval(ctpwa_handle* h, hipStream_t s, double* part = nullptr) {
    if (h->nwg <= 0) return hipSuccess;       // stored phase-space Gram and no data events
    double* pp = part ? part : h->part;
    void* args[] = {&h->table, &h->slp, &h->inv, &h->invf, &h->wts, &pp, &h->wg_cls, &h->wg_rng, &h->ntot,
                    &h->nwg, &h->wc[1], &h->wc[2], &h->pw[1], &h->pw[2], &h->cfb};
    return hipModuleLaunchKernel(h->fn_eval, h->nwg, 1, 1, h->M.wgs, 1, 1, 0, s, args, nullptr);
}


struct DeviceGuard {
    int prev = 0;
    explicit DeviceGuard(int dev) { (void)hipGetDevice(&prev); (void)hipSetDevice(dev); }
    ~DeviceGuard() { (void)hipSetDevice(prev); }
};

extern "C" int ctpwa_nll_grad(ctpwa_handle* h, const double* params_dev, double* out_dev, void* stream) {
    if (!h || !params_dev || !out_dev) return fail("ctpwa_nll_grad: NULL argument");
    DeviceGuard g(h->device);
    hipStream_t s = (hipStream_t)stream;
    hipError_t e = ctpwa_launch_prep(params_dev, h->table, &h->prep, h->M.maxloc, 2, s);
    if (e == hipSuccess) e = launch_custom(h, s);
    if (e == hipSuccess) e = launch_eval(h, s);
    if (e == hipSuccess) {
        h->fin.params = params_dev;
        h->fin.out = out_dev;
        e = ctpwa_launch_finalize(&h->fin, h->fin_done, s);
    }
    if (e != hipSuccess) return fail(std::string("ctpwa_nll_grad: ") + hipGetErrorString(e));
    return 0;
}

extern "C" int ctpwa_nll_grad_noalloc(ctpwa_handle* h, const double* params_dev, double* out_dev, void* stream) {
    if (!h || !params_dev || !out_dev) return fail("ctpwa_nll_grad_noalloc: NULL argument");
    const size_t need = (size_t)h->nacc * (size_t)std::max(1, h->nwg) + 3 * (size_t)h->nacc;
    if (need > NOALLOC_DOUBLES)
        return fail("ctpwa_nll_grad_noalloc: the evaluation needs " + std::to_string(need) +
                    " doubles of scratch, more than the " + std::to_string(NOALLOC_DOUBLES) + " it holds");
    DeviceGuard g(h->device);
    double* scr = scratch_lookup(h->device);
    if (!scr) return fail("ctpwa_nll_grad_noalloc: no scratch address for this device");
    hipStream_t s = (hipStream_t)stream;
    hipError_t e = ctpwa_launch_prep(params_dev, h->table, &h->prep, h->M.maxloc, 2, s);
    if (e == hipSuccess) e = launch_custom(h, s);
    if (e == hipSuccess) e = launch_eval(h, s, scr);
    if (e == hipSuccess) {
        FinArgs f = h->fin;
        f.part = scr;
        f.totals = scr + (size_t)h->nacc * (size_t)std::max(1, h->nwg);
        f.params = params_dev;
        f.out = out_dev;
        e = ctpwa_launch_finalize(&f, h->fin_done, s);
    }
    if (e != hipSuccess) return fail(std::string("ctpwa_nll_grad_noalloc: ") + hipGetErrorString(e));
    return 0;
}


// Hessian workspace: Jacobian columns C, u-vector length NU, second-part scalars NZ per event
struct HessDims { int C, NU, NZ; };
static HessDims hess_dims(const Model_& M) {
    const Plan& P = M.P;
    HessDims d{P.namp + P.ntheta, 2 * P.namp + P.ntheta, 1};
    for (size_t b = 0; b < P.blocks.size(); ++b) {
        if (P.blocks[b].nsl == 0) continue;
        int nl = (int)M.bvars[b].size();
        d.NZ += P.blocks[b].nsl * 2 * nl + nl * (nl + 1) / 2;
    }
    return d;
}
// events per chunk of Jacobian rows: the rows of one chunk take at most this many bytes
// (environment CTPWA_HESS_JBYTES overrides it for measurements)
static long long hess_jbytes() {
    if (const char* e = std::getenv("CTPWA_HESS_JBYTES")) {
        long long v = std::atoll(e);
        if (v >= (1LL << 20)) return v;
    }
    return 2LL << 30;
}
static long long hess_chunk(const Model_& M, long long maxn) {
    const HessDims d = hess_dims(M);
    const long long row = (long long)M.P.npol * 2 * d.C * 8;
    long long c = std::max<long long>(256, hess_jbytes() / std::max<long long>(1, row) / 256 * 256);
    // a multiple of 4 events: the Gram's 16-byte slab loads stay aligned
    return (std::max<long long>(1, std::min(maxn, c)) + 3) / 4 * 4;
}
// The data and background classes run on a second stream of the handle, concurrently
// with the phase space, when both sides are large enough for the overlap to outweigh the
// two cross-stream waits, and only for a handle that is alone in the process (`others`:
// the other live handles, counted when the buffers are set up and again at every call):
// with several (a fit service) other handles' work fills the GPU, and a second stream
// per handle, its events and its second buffer set cost more than they gain (measured:
// Newton fits/hour of the service down by a quarter).  Environment CTPWA_HESS_STREAMS =
// 1 or 2 forces one or two streams.
static int hess_streams_env() {
    if (const char* e = std::getenv("CTPWA_HESS_STREAMS")) {
        const int k = std::atoi(e);
        if (k != 1 && k != 2) throw std::runtime_error("CTPWA_HESS_STREAMS must be 1 or 2");
        return k;
    }
    return 0;
}
static bool hess_two_streams(const long long n3[3], int others) {
    const int k = hess_streams_env();
    if (k) return k == 2 && n3[0] > 0 && n3[1] + n3[2] > 0;
    return others == 0 && n3[0] >= (1LL << 20) && n3[1] + n3[2] >= (1LL << 17);
}
// events per chunk of each buffer set: with two streams set 0 takes the phase space and
// set 1 data and background; with one stream set 0 takes every class
static void hess_chunks(const Model_& M, const long long n3[3], int others, long long ch[2]) {
    const long long n12 = std::max(n3[1], n3[2]);
    if (!hess_two_streams(n3, others)) {
        ch[0] = hess_chunk(M, std::max(n3[0], n12));
        ch[1] = 0;
        return;
    }
    ch[0] = hess_chunk(M, n3[0]);
    ch[1] = hess_chunk(M, n12);
}
static long long hess_workspace_bytes(const Model_& M, const long long n3[3], long long ntot, int others) {
    HessDims d = hess_dims(M);
    long long ch[2];
    hess_chunks(M, n3, others, ch);
    long long nG = (long long)(2 * d.C) * (2 * d.C + 1) / 2, nUu = (long long)d.NU * (d.NU + 1) / 2;
    long long w = 3 * (nG + nUu + d.NZ + d.NU) + (long long)d.NU * d.NU;
    for (int k = 0; k < 2; ++k) {
        if (ch[k] == 0) continue;
        w += ch[k] * (M.P.npol * 2LL * d.C + d.NU + 2) + (long long)(d.NZ + d.NU) * ctpwa_hess_rows_waves(ch[k]);
        w += gram_part_doubles(2 * d.C) + gram_part_doubles(d.NU);
    }
    w += (long long)M.customs.size() * M.ncf2() * ntot;   // second-order Custom factor rows
    return w * 8 + 64 * 1024;    // + the index tables and allocation slack
}

static void gram_setup(ctpwa_handle* h, GramTabs& g, int ncol) {
    g.ncol = ncol;
    int npp = 0;
    g.ntp = ctpwa_gram_tables(ncol, nullptr, &npp, nullptr);
    std::vector<int4> pp((size_t)npp);
    std::vector<int2> tiles((size_t)g.ntp);
    ctpwa_gram_tables(ncol, pp.data(), &g.npp, tiles.data());
    g.nwg = ctpwa_gram_nwg(g.npp);
    g.pp = h->upload(pp);
    g.tiles = h->upload(tiles);
    g.part = h->alloc<double>((size_t)g.nwg * g.ntp * 256);
}

static void hess_setup(ctpwa_handle* h) {
    const Model_& M = h->M;
    const Plan& P = M.P;
    const HessDims hd = hess_dims(M);
    h->hC = hd.C;
    h->hNU = hd.NU;
    h->hNZ = hd.NZ;
    h->hnG = 2 * hd.C * (2 * hd.C + 1) / 2;
    h->hnUu = hd.NU * (hd.NU + 1) / 2;
    if (P.ntheta > 64) throw std::runtime_error("Hessian: more than 64 free resonance parameters is not built");
    long long ch[2];
    hess_chunks(M, h->n, g_live_handles.load() - 1, ch);
    const size_t es = M.mixed ? sizeof(float) : sizeof(double);
    for (int k = 0; k < 2; ++k) {
        HessWS& ws = h->hws[k];
        ws.chunk = ch[k];
        if (ch[k] == 0) continue;
        ws.J = h->alloc<unsigned char>((size_t)ch[k] * P.npol * 2 * h->hC * es);
        ws.U = h->alloc<double>((size_t)ch[k] * h->hNU);
        ws.cs_ld = ctpwa_hess_rows_waves(ch[k]);
        ws.cs = h->alloc<double>((size_t)(h->hNZ + h->hNU) * ws.cs_ld);
        ws.W = h->alloc<double>((size_t)ch[k] * 2);
        gram_setup(h, ws.gJ, 2 * h->hC);
        gram_setup(h, ws.gU, h->hNU);
    }
    if (ch[1] > 0) {
        HIPCHECK(hipStreamCreateWithFlags(&h->hs2, hipStreamNonBlocking));
        for (hipEvent_t& ev : h->hev) HIPCHECK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    }
    // totals: per class G, Uu, Z, g
    h->htot = h->alloc<double>((size_t)3 * (h->hnG + h->hnUu + h->hNZ + h->hNU));
    h->hHext = h->alloc<double>((size_t)h->hNU * h->hNU);
    // fit parameter -> extended entries: [Re p_f] -> Re v_a, [Im p_f] -> Im v_a with
    // v_a = r_a p_f(a); theta_j -> theta_j
    {
        const int nf = P.nfree, np = 2 * nf + P.ntheta;
        std::vector<int> ptr(np + 1, 0), xs;
        std::vector<double> rs;
        for (int q = 0; q < np; ++q) {
            if (q >= 2 * nf) {
                xs.push_back(2 * P.namp + (q - 2 * nf));
                rs.push_back(1.0);
            } else {
                const int f = q < nf ? q : q - nf, off = q < nf ? 0 : P.namp;
                for (int a = 0; a < P.namp; ++a)
                    if (P.amap[a] == f) {
                        xs.push_back(off + a);
                        rs.push_back(P.aratio[a]);
                    }
            }
            ptr[q + 1] = (int)xs.size();
        }
        h->hmap_ptr = h->upload(ptr);
        h->hmap_x = h->upload(xs);
        h->hmap_r = h->upload(rs);
    }
    std::vector<int> perm;
    h->hmod = compile_eval(h->device, generate_hess_kernel(M, &perm));
    if ((int)perm.size() != h->hNZ + h->hNU) throw std::runtime_error("Hessian: summed-column count mismatch");
    h->hperm = h->upload(perm);
    HIPCHECK(hipModuleGetFunction(&h->fn_hrows, h->hmod, "ctpwa_hess_rows"));
    if (!M.customs.empty()) {
        h->mod2 = compile_eval(h->device, generate_custom2_kernel(M));
        HIPCHECK(hipModuleGetFunction(&h->fn_custom2, h->mod2, "ctpwa_custom2"));
        h->cfb2 = h->alloc<double>((size_t)M.customs.size() * M.ncf2() * h->ntot);
    }
    h->hess_ready = true;
}

// amplitude Gram of one class: Q[l][k] = sum_e sum_p conj(A_l) A_k (amplitudes without
// couplings), the integrals behind fit fractions and the interference matrix
// events per chunk of amplitude rows: at most 2^18, and rows of at most 256 MiB
static long long agram_chunk(const Plan& P, long long maxn) {
    const long long per_event = (long long)P.npol * 2 * P.namp * 8;
    const long long cap = std::max<long long>(4, (256LL << 20) / per_event);
    return (std::max<long long>(1, std::min<long long>(maxn, std::min<long long>(cap, 1 << 18))) + 3) / 4 * 4;
}
static long long agram_workspace_bytes(const Plan& P, long long maxn) {
    const long long n2 = 2 * P.namp, ch = agram_chunk(P, maxn);
    return ch * P.npol * n2 * 8 + ch * 8 + n2 * (n2 + 1) / 2 * 8 + gram_part_doubles((int)n2) * 8;
}
static void agram_setup(ctpwa_handle* h) {
    const Plan& P = h->M.P;
    const int n2 = 2 * P.namp;
    long long maxn = std::max(h->n[0], std::max(h->n[1], h->n[2]));
    h->gchunk = agram_chunk(P, maxn);
    h->gX = h->alloc<double>((size_t)h->gchunk * P.npol * n2);
    h->gW = h->upload(std::vector<double>((size_t)h->gchunk, 1.0));
    h->gtot = h->alloc<double>((size_t)n2 * (n2 + 1) / 2);
    gram_setup(h, h->gg, n2);
    h->agram_ready = true;
}

// Gram of class cls's amplitude rows at the parameter point the table holds, to out
// [namp][namp][2]: with the theta coefficients K_a (withK = 1) or without them (0)
static void class_gram(ctpwa_handle* h, int cls, int withK, double* out, hipStream_t s) {
    if (!h->agram_ready) agram_setup(h);
    const Plan& P = h->M.P;
    const int n2 = 2 * P.namp;
    const long long ldx = h->gchunk * P.npol;
    HIPCHECK(ctpwa_launch_zero(h->gtot, (long long)n2 * (n2 + 1) / 2, s));
    const GramTabs& G = h->gg;
    int k = 0;
    for (long long s0 = 0; s0 < h->n[cls]; s0 += h->gchunk, ++k) {
        long long cnt = std::min(h->gchunk, h->n[cls] - s0);
        HIPCHECK(ctpwa_launch_amp_rows(&h->aux, h->off[cls] + s0, cnt, h->gX, ldx, withK, s));
        HIPCHECK(ctpwa_launch_gram_panels(h->gX, 0, n2, ldx, cnt * P.npol, h->gW, cnt, G.pp, G.npp, G.ntp, G.part,
                                          G.nwg, k > 0, s));
    }
    if (h->n[cls] > 0) HIPCHECK(ctpwa_launch_accum_panels(G.part, G.nwg, G.ntp, G.tiles, n2, h->gtot, s));
    HIPCHECK(ctpwa_launch_gram_to_complex(h->gtot, P.namp, out, s));
}

// No free resonance parameter: the amplitudes without couplings, A'_a = Psi_b sl_u, do
// not change between evaluations, so the phase-space sum of I = sum_p |sum_a c_a A'_a|^2
// is the quadratic form c^H G c of their Gram G, and the phase-space gradient sums
// sum_e conj(A'_a) S are (G c)_a (the reference's stored phase-space matrix and
// quadratic form, src/Analysis.cu:808-826, src/ComputeNLL.cu:575-597).  G is built once
// here in FP64 on the matrix cores; the final reduction forms the phase-space totals from
// it and the fused kernel scans data and background only.  Environment CTPWA_FIXED_GRAM=0
// keeps the per-event phase-space scan (for comparisons).
static bool fixed_gram_applies(const ctpwa_handle* h) {
    const Plan& P = h->M.P;
    if (P.ntheta != 0 || h->nacc != 1 + 2 * P.namp || h->n[0] <= 0) return false;
    const char* e = std::getenv("CTPWA_FIXED_GRAM");
    return !(e && std::string(e) == "0");
}
static void build_fixed_gram(ctpwa_handle* h) {
    const Plan& P = h->M.P;
    hipStream_t s = nullptr;
    // the table's coefficient records depend on theta only (none free): any couplings do
    double* p0 = h->upload(std::vector<double>((size_t)std::max(1, 2 * P.nfree), 0.0));
    double2* G = h->alloc<double2>((size_t)P.namp * P.namp);
    HIPCHECK(ctpwa_launch_prep(p0, h->table, &h->prep, h->M.maxloc, 0, s));
    HIPCHECK(launch_custom(h, s));
    class_gram(h, 0, 0, (double*)G, s);
    h->d_gram = G;
    h->fin.gram = G;
    h->fin.off_c = (long long)h->M.off_c();
}

extern "C" int ctpwa_amplitude_gram(ctpwa_handle* h, const double* params_dev, int cls, double* out_dev, void* stream) {
    if (!h || !params_dev || !out_dev) return fail("ctpwa_amplitude_gram: NULL argument");
    if (cls < 0 || cls > 2) return fail("ctpwa_amplitude_gram: class must be 0 (phsp), 1 (data) or 2 (bkg)");
    DeviceGuard g(h->device);
    hipStream_t s = (hipStream_t)stream;
    try {
        HIPCHECK(ctpwa_launch_prep(params_dev, h->table, &h->prep, h->M.maxloc, 0, s));
        HIPCHECK(launch_custom(h, s));
        class_gram(h, cls, 1, out_dev, s);
        return 0;
    } catch (const std::exception& e) {
        return fail(std::string("ctpwa_amplitude_gram: ") + e.what());
    }
}

extern "C" int ctpwa_partial_intensities(ctpwa_handle* h, const double* params_dev, int cls, double* out_dev, void* stream) {
    if (!h || !params_dev || !out_dev) return fail("ctpwa_partial_intensities: NULL argument");
    if (cls < 0 || cls > 2) return fail("ctpwa_partial_intensities: class must be 0 (phsp), 1 (data) or 2 (bkg)");
    DeviceGuard g(h->device);
    hipStream_t s = (hipStream_t)stream;
    hipError_t e = ctpwa_launch_prep(params_dev, h->table, &h->prep, h->M.maxloc, 0, s);
    if (e == hipSuccess) e = launch_custom(h, s);
    if (e == hipSuccess)
        e = ctpwa_launch_partial_int(&h->aux, h->wave_ptr, h->wave_blk, h->M.P.nwave, h->off[cls], h->n[cls],
                                     out_dev, s);
    if (e != hipSuccess) return fail(std::string("ctpwa_partial_intensities: ") + hipGetErrorString(e));
    return 0;
}

extern "C" int ctpwa_partial_amplitudes(ctpwa_handle* h, const double* params_dev, int cls, double* out_dev, void* stream) {
    if (!h || !params_dev || !out_dev) return fail("ctpwa_partial_amplitudes: NULL argument");
    if (cls < 0 || cls > 2) return fail("ctpwa_partial_amplitudes: class must be 0 (phsp), 1 (data) or 2 (bkg)");
    DeviceGuard g(h->device);
    hipStream_t s = (hipStream_t)stream;
    hipError_t e = ctpwa_launch_prep(params_dev, h->table, &h->prep, h->M.maxloc, 0, s);
    if (e == hipSuccess) e = launch_custom(h, s);
    if (e == hipSuccess)
        e = ctpwa_launch_partial_amp(&h->aux, h->wave_ptr, h->wave_blk, h->M.P.nwave, h->off[cls], h->n[cls],
                                     out_dev, s);
    if (e != hipSuccess) return fail(std::string("ctpwa_partial_amplitudes: ") + hipGetErrorString(e));
    return 0;
}

extern "C" int ctpwa_last_theta(ctpwa_handle* h, double* theta_host, void* stream) {
    if (!h || (!theta_host && h->M.P.ntheta > 0)) return fail("ctpwa_last_theta: NULL argument");
    if (h->M.P.ntheta == 0) return 0;
    DeviceGuard g(h->device);
    hipStream_t s = (hipStream_t)stream;
    hipError_t e = hipMemcpyAsync(theta_host, h->table + h->M.off_lth(), (size_t)h->M.P.ntheta * sizeof(double),
                                  hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return fail(std::string("ctpwa_last_theta: ") + hipGetErrorString(e));
    return 0;
}

extern "C" int ctpwa_hessian(ctpwa_handle* h, const double* params_dev, double* hess_dev, void* stream) {
    if (!h || !params_dev || !hess_dev) return fail("ctpwa_hessian: NULL argument");
    DeviceGuard g(h->device);
    hipStream_t s = (hipStream_t)stream;
    try {
        if (!h->hess_ready) hess_setup(h);
        const Model_& M = h->M;
        const Plan& P = M.P;
        HIPCHECK(ctpwa_launch_prep(params_dev, h->table, &h->prep, M.maxloc, 3, s));
        if (h->fn_custom2) {
            void* args[] = {&h->table, &h->inv, &h->invf, &h->ntot, &h->cfb2};
            unsigned grid = (unsigned)std::min<long long>(16384, (h->ntot + 63) / 64);
            HIPCHECK(hipModuleLaunchKernel(h->fn_custom2, grid, 1, 1, 64, 1, 1, 0, s, args, nullptr));
        }
        size_t per_class = (size_t)h->hnG + h->hnUu + h->hNZ + h->hNU;
        HIPCHECK(ctpwa_launch_zero(h->htot, (long long)(3 * per_class), s));
        HessAsm as{};
        const int f32 = M.mixed;
        // data and background run on the handle's second stream with buffer set 1, after
        // the coefficient table and the zeroed totals on the caller's stream
        const int forced = hess_streams_env();
        const bool two = h->hs2 != nullptr && (h->n[1] > 0 || h->n[2] > 0) &&
                         (forced == 2 || (forced == 0 && g_live_handles.load() == 1));
        if (two) {
            HIPCHECK(hipEventRecord(h->hev[0], s));
            HIPCHECK(hipStreamWaitEvent(h->hs2, h->hev[0], 0));
        }
        for (int c = 0; c < 3; ++c) {
            double* base = h->htot + c * per_class;
            as.G[c] = base;
            as.Uu[c] = base + h->hnG;
            as.Zs[c] = base + h->hnG + h->hnUu;
            as.gs[c] = base + h->hnG + h->hnUu + h->hNZ;
            HessWS& ws = h->hws[c == 0 || !two ? 0 : 1];
            hipStream_t sc = (c == 0 || !two) ? s : h->hs2;
            const long long ldJ = ws.chunk * P.npol;
            const double* wts = (c > 0 && h->pw[c]) ? h->wts : nullptr;
            double wconst = h->wc[c];
            int cls = c;
            int k = 0;
            for (long long s0 = 0; s0 < h->n[c]; s0 += ws.chunk, ++k) {
                long long cnt = std::min(ws.chunk, h->n[c] - s0);
                long long e0 = h->off[c] + s0;
                // per-event rows of the chunk (generated kernel)
                void* args[] = {&h->table, (void*)&h->slp, &h->inv, &h->invf, &h->ntot, (void*)&wts, &wconst, &cls, &e0,
                                &cnt, &ws.J, (void*)&ldJ, &ws.chunk, &ws.U, &ws.W, &ws.cs, &ws.cs_ld, &h->cfb2,
                                &h->ntot};
                const unsigned grid = (unsigned)((cnt + 255) / 256);
                HIPCHECK(hipModuleLaunchKernel(h->fn_hrows, grid, 1, 1, 256, 1, 1, 0, sc, args, nullptr));
                // weighted Grams on the matrix cores, summed over the class's chunks
                HIPCHECK(ctpwa_launch_gram_panels(ws.J, f32, 2 * h->hC, ldJ, cnt * P.npol, ws.W, cnt, ws.gJ.pp,
                                                  ws.gJ.npp, ws.gJ.ntp, ws.gJ.part, ws.gJ.nwg, k > 0, sc));
                if (c > 0)
                    HIPCHECK(ctpwa_launch_gram_panels(ws.U, 0, h->hNU, ws.chunk, cnt, ws.W + ws.chunk, cnt, ws.gU.pp,
                                                      ws.gU.npp, ws.gU.ntp, ws.gU.part, ws.gU.nwg, k > 0, sc));
                // Zs[c] and gs[c] are adjacent: one pass adds both from the rows kernel's sums
                HIPCHECK(ctpwa_launch_accum_cols(ws.cs, ws.cs_ld, cnt, h->hNZ + h->hNU, h->hperm, (double*)as.Zs[c], sc));
            }
            if (h->n[c] > 0) {
                HIPCHECK(ctpwa_launch_accum_panels(ws.gJ.part, ws.gJ.nwg, ws.gJ.ntp, ws.gJ.tiles, 2 * h->hC,
                                                   (double*)as.G[c], sc));
                if (c > 0)
                    HIPCHECK(ctpwa_launch_accum_panels(ws.gU.part, ws.gU.nwg, ws.gU.ntp, ws.gU.tiles, h->hNU,
                                                       (double*)as.Uu[c], sc));
            }
        }
        if (two) {
            HIPCHECK(hipEventRecord(h->hev[1], h->hs2));
            HIPCHECK(hipStreamWaitEvent(s, h->hev[1], 0));
        }
        as.C = h->hC; as.NU = h->hNU; as.NZ = h->hNZ;
        as.namp = P.namp; as.nfree = P.nfree; as.ntheta = P.ntheta; as.nblock = (int)P.blocks.size();
        as.amps = h->aux.amps; as.blocks = h->aux.blocks; as.table = h->table; as.off_v = (long long)M.off_v();
        as.Np = (double)h->n[0]; as.Wd = h->Wd; as.Wb = h->Wb;
        as.ngauss = (int)P.g_slot.size(); as.g_slot = h->fin.g_slot; as.g_sig = h->fin.g_sig;
        as.map_ptr = h->hmap_ptr; as.map_x = h->hmap_x; as.map_r = h->hmap_r;
        as.Hext = h->hHext; as.out = hess_dev;
        HIPCHECK(ctpwa_launch_hess_assemble(&as, s));
        return 0;
    } catch (const std::exception& e) {
        return fail(std::string("ctpwa_hessian: ") + e.what());
    }
}

static int aux_call(ctpwa_handle* h, const double* params_dev, int cls, double* out_dev, void* stream, int intensity) {
    if (!h || !params_dev || !out_dev) return fail("NULL argument");
    if (cls < 0 || cls > 2) return fail("event class must be 0, 1 or 2");
    DeviceGuard g(h->device);
    hipStream_t s = (hipStream_t)stream;
    hipError_t e = ctpwa_launch_prep(params_dev, h->table, &h->prep, h->M.maxloc, 0, s);
    if (e == hipSuccess) e = launch_custom(h, s);
    if (e == hipSuccess && h->fn_aux) {
        // generated kernel of the model (Custom lineshapes)
        long long e0 = h->off[cls], n = h->n[cls];
        if (n > 0) {
            int flag = intensity;
            void* args[] = {&h->table, &h->slp, &h->inv, &h->invf, &h->ntot, &e0, &n, &out_dev, &flag, &h->cfb};
            e = hipModuleLaunchKernel(h->fn_aux, (unsigned)((n + 255) / 256), 1, 1, 256, 1, 1, 0, s, args, nullptr);
        }
    } else if (e == hipSuccess) {
        e = ctpwa_launch_aux(&h->aux, h->off[cls], h->n[cls], out_dev, intensity, s);
    }
    if (e != hipSuccess) return fail(hipGetErrorString(e));
    return 0;
}

extern "C" int ctpwa_intensity(ctpwa_handle* h, const double* params_dev, int cls, double* out_dev, void* stream) {
    return aux_call(h, params_dev, cls, out_dev, stream, 1);
}

extern "C" int ctpwa_amplitudes(ctpwa_handle* h, const double* params_dev, int cls, double* out_dev, void* stream) {
    return aux_call(h, params_dev, cls, out_dev, stream, 0);
}

extern "C" int ctpwa_time_eval_kernel(ctpwa_handle* h, const double* params_dev, int iters, void* stream, double* avg_ms) {
    if (!h || !params_dev || !avg_ms || iters <= 0 || iters > 10000) return fail("ctpwa_time_eval_kernel: bad argument");
    DeviceGuard g(h->device);
    hipStream_t s = (hipStream_t)stream;
    // one event pair around every launch: the average is over the kernels' own durations,
    // not over the gaps between back-to-back launches
    std::vector<hipEvent_t> ev(2 * (size_t)iters, nullptr);
    hipError_t e = hipSuccess;
    for (size_t i = 0; i < ev.size() && e == hipSuccess; ++i) e = hipEventCreate(&ev[i]);
    if (e == hipSuccess) e = ctpwa_launch_prep(params_dev, h->table, &h->prep, h->M.maxloc, 0, s);
    if (e == hipSuccess) e = launch_custom(h, s);
    for (int i = 0; i < iters && e == hipSuccess; ++i) {
        e = hipEventRecord(ev[2 * i], s);
        if (e == hipSuccess) e = launch_eval(h, s);
        if (e == hipSuccess) e = hipEventRecord(ev[2 * i + 1], s);
    }
    if (e == hipSuccess) e = hipEventSynchronize(ev.back());
    double total = 0.0;
    for (int i = 0; i < iters && e == hipSuccess; ++i) {
        float ms = 0.f;
        e = hipEventElapsedTime(&ms, ev[2 * i], ev[2 * i + 1]);
        total += ms;
    }
    for (hipEvent_t x : ev)
        if (x) (void)hipEventDestroy(x);
    if (e != hipSuccess) return fail(std::string("ctpwa_time_eval_kernel: ") + hipGetErrorString(e));
    *avg_ms = total / iters;
    return 0;
}

// ---------------------------------------------------------------------------
// many likelihoods of one model in one set of launches (continuous batching)
// ---------------------------------------------------------------------------
// mirrors `struct EvalJob` of the generated source (codegen.cpp)
struct EvalJob {
    const double* tab; const void* sl; const double* inv; const float* invf; const double* wts;
    double* part; const int* wg_cls; const long long* wg_rng; const double* cf; long long ld; double w1, w2;
    int nwg, pw1, pw2, wg_begin;
};

struct ctpwa_batch {
    int device = 0;
    int cap = 0;
    size_t bytes = 0;
    char* host = nullptr;          // pinned staging of the descriptors
    // device copy: [header: job count, task count | task counters][EvalJob cap][FinArgs cap]
    // [params ptr cap][table ptr cap][PrepArgs cap]
    char* dev = nullptr;
    hipEvent_t copied = nullptr;   // the staging buffer may be rewritten once this completes
    bool pending = false;
    size_t o_hdr = 0, o_ctr = 0, o_jobs = 0, o_fins = 0, o_params = 0, o_tables = 0, o_preps = 0;
    int max_nacc = 0;              // of the uploaded list
    size_t max_lds = 0;
    // Graph replays (opt-in, environment CTPWA_BATCH_GRAPH=1): the descriptors on the device
    // belong to `key`, a list of (handle serial, params, out); a call with the same list
    // reuses them without an upload (serials, not handle addresses: a new handle may take a
    // destroyed one's address) and replays a captured graph of the launches, one per
    // power-of-two row bucket per model module.  Every argument a captured launch holds is
    // in the descriptors or is a constant of the model (each job's prep arguments included:
    // they point into that job's own handle).
    bool use_graphs = false;
    std::vector<unsigned long long> key;
    int rows = 0;
    hipStream_t cs = nullptr;      // capture stream
    hipModule_t gmod = nullptr;
    std::map<int, hipGraphExec_t> execs;
    int persist = 0;               // workgroups of the persistent grouped kernel
};

static size_t align256(size_t x) { return (x + 255) / 256 * 256; }

extern "C" int ctpwa_batch_create(int device, int max_jobs, ctpwa_batch** out) {
    if (!out) return fail("ctpwa_batch_create: out is NULL");
    *out = nullptr;
    if (max_jobs <= 0 || max_jobs > 65535) return fail("ctpwa_batch_create: max_jobs must be in [1, 65535]");
    ctpwa_batch* b = new ctpwa_batch();
    DeviceGuard g(device);
    try {
        b->device = device;
        b->cap = max_jobs;
        b->o_hdr = 0;
        b->o_ctr = 256;
        b->o_jobs = 512;
        b->o_fins = b->o_jobs + align256(sizeof(EvalJob) * max_jobs);
        b->o_params = b->o_fins + align256(sizeof(FinArgs) * max_jobs);
        b->o_tables = b->o_params + align256(sizeof(void*) * max_jobs);
        b->o_preps = b->o_tables + align256(sizeof(void*) * max_jobs);
        b->bytes = b->o_preps + align256(sizeof(PrepArgs) * max_jobs);
        const char* e = std::getenv("CTPWA_BATCH_GRAPH");
        b->use_graphs = e && std::string(e) == "1";
        HIPCHECK(hipSetDevice(device));
        HIPCHECK(hipHostMalloc((void**)&b->host, b->bytes, hipHostMallocDefault));
        std::memset(b->host, 0, b->bytes);
        HIPCHECK(hipMalloc((void**)&b->dev, b->bytes));
        // the task counters start at zero; the persistent grouped kernel leaves them at zero
        HIPCHECK(hipMemset(b->dev, 0, 512));
        HIPCHECK(hipEventCreateWithFlags(&b->copied, hipEventDisableTiming));
        if (b->use_graphs) HIPCHECK(hipStreamCreateWithFlags(&b->cs, hipStreamNonBlocking));
        *out = b;
        return 0;
    } catch (const std::exception& e) {
        if (b->host) (void)hipHostFree(b->host);
        if (b->dev) (void)hipFree(b->dev);
        if (b->copied) (void)hipEventDestroy(b->copied);
        delete b;
        return fail(std::string("ctpwa_batch_create: ") + e.what());
    }
}

static void batch_drop_graphs(ctpwa_batch* b) {
    for (auto& kv : b->execs) (void)hipGraphExecDestroy(kv.second);
    b->execs.clear();
}

extern "C" void ctpwa_batch_destroy(ctpwa_batch* b) {
    if (!b) return;
    DeviceGuard g(b->device);
    if (!b->execs.empty()) (void)hipDeviceSynchronize();   // no replay still reads the buffers
    else if (b->pending) (void)hipEventSynchronize(b->copied);
    batch_drop_graphs(b);
    if (b->cs) (void)hipStreamDestroy(b->cs);
    (void)hipEventDestroy(b->copied);
    (void)hipHostFree(b->host);
    (void)hipFree(b->dev);
    delete b;
}

// graph replays: the launches over the uploaded descriptors with the job and task counts
// read on the device, `rows` (>= the job count) workgroup rows of prep and reduction, the
// persistent grouped kernel, each job's own prep arguments
static void batch_launches_graph(ctpwa_batch* b, const ctpwa_handle* h0, int rows, hipStream_t s) {
    const int* d_hdr = (const int*)(b->dev + b->o_hdr);
    unsigned int* d_ctr = (unsigned int*)(b->dev + b->o_ctr);
    HIPCHECK(ctpwa_launch_prep_batch((const double* const*)(b->dev + b->o_params), (double* const*)(b->dev + b->o_tables),
                                     &h0->prep, h0->M.maxloc, (const PrepArgs*)(b->dev + b->o_preps), rows, d_hdr, s));
    const EvalJob* d_jobs = (const EvalJob*)(b->dev + b->o_jobs);
    void* args[] = {&d_jobs, &d_hdr, &d_ctr};
    HIPCHECK(hipModuleLaunchKernel(h0->fn_group_p, b->persist, 1, 1, h0->M.wgs, 1, 1, 0, s, args, nullptr));
    HIPCHECK(ctpwa_launch_finalize_batch((const FinArgs*)(b->dev + b->o_fins), rows, b->max_nacc, b->max_lds, d_hdr, s));
}

extern "C" int ctpwa_batch_nll_grad(ctpwa_batch* b, ctpwa_handle* const* hs, const double* const* params_dev,
                                    double* const* out_dev, int n, void* stream) {
    if (!b || !hs || !params_dev || !out_dev) return fail("ctpwa_batch_nll_grad: NULL argument");
    if (n <= 0) return 0;
    if (n > b->cap) return fail("ctpwa_batch_nll_grad: more likelihoods than the batch was created for");
    DeviceGuard g(b->device);
    hipStream_t s = (hipStream_t)stream;
    try {
        const ctpwa_handle* h0 = hs[0];
        for (int j = 0; j < n; ++j) {
            const ctpwa_handle* h = hs[j];
            if (!h || !params_dev[j] || !out_dev[j]) throw std::runtime_error("NULL handle, parameter or output pointer");
            if (h->device != b->device) throw std::runtime_error("a handle lives on another device than the batch");
            if (h->mod != h0->mod) throw std::runtime_error("the handles of one batch must share one model and precision");
        }
        // graph replays only for models without Custom lineshapes (those launch one factor
        // kernel per job)
        const bool graph = b->use_graphs && h0->M.customs.empty();
        std::vector<unsigned long long> key;
        if (graph) {
            key.resize((size_t)3 * n);
            for (int j = 0; j < n; ++j) {
                key[3 * j] = hs[j]->serial;
                key[3 * j + 1] = (unsigned long long)(uintptr_t)params_dev[j];
                key[3 * j + 2] = (unsigned long long)(uintptr_t)out_dev[j];
            }
        }
        int wg = 0;
        if (!graph || key != b->key) {
            {
                // a handle holds one coefficient table and one set of partial sums
                std::vector<const ctpwa_handle*> seen(hs, hs + n);
                std::sort(seen.begin(), seen.end());
                if (std::adjacent_find(seen.begin(), seen.end()) != seen.end())
                    throw std::runtime_error("a handle appears twice in one batch (each job needs a handle of its own)");
            }
            b->key.clear();
            // the staging buffer is reused: wait for the previous upload from it
            if (b->pending) HIPCHECK(hipEventSynchronize(b->copied));
            int* hdr = (int*)(b->host + b->o_hdr);
            EvalJob* jobs = (EvalJob*)(b->host + b->o_jobs);
            FinArgs* fins = (FinArgs*)(b->host + b->o_fins);
            const double** pp = (const double**)(b->host + b->o_params);
            double** tp = (double**)(b->host + b->o_tables);
            PrepArgs* pa = (PrepArgs*)(b->host + b->o_preps);
            int max_nacc = 0;
            size_t max_lds = 0;
            for (int j = 0; j < n; ++j) {
                ctpwa_handle* h = hs[j];
                max_nacc = std::max(max_nacc, h->nacc);
                max_lds = std::max(max_lds, assemble_lds_bytes(h->fin));
                jobs[j] = EvalJob{h->table, h->slp, h->inv, h->invf, h->wts, h->part, h->wg_cls, h->wg_rng, h->cfb,
                                  h->ntot, h->wc[1], h->wc[2], h->nwg, h->pw[1], h->pw[2], wg};
                wg += h->nwg;
                fins[j] = h->fin;
                fins[j].params = params_dev[j];
                fins[j].out = out_dev[j];
                pp[j] = params_dev[j];
                tp[j] = h->table;
                pa[j] = h->prep;
            }
            hdr[0] = n;
            hdr[1] = wg;
            // header and descriptors (the task counters stay as the kernel left them)
            HIPCHECK(hipMemcpyAsync(b->dev + b->o_hdr, b->host + b->o_hdr, 2 * sizeof(int), hipMemcpyHostToDevice, s));
            HIPCHECK(hipMemcpyAsync(b->dev + b->o_jobs, b->host + b->o_jobs, b->bytes - b->o_jobs, hipMemcpyHostToDevice, s));
            HIPCHECK(hipEventRecord(b->copied, s));
            b->pending = true;
            b->max_nacc = max_nacc;
            b->max_lds = max_lds;
            if (graph) {
                b->key.swap(key);
                int rows = 1;
                while (rows < n) rows <<= 1;
                b->rows = std::min(rows, b->cap);
            }
        }
        if (!graph) {
            // one batched prep launch, each job's Custom factors, one grouped launch whose
            // workgroups are shared out among the jobs, one batched reduction and assembly
            HIPCHECK(ctpwa_launch_prep_batch((const double* const*)(b->dev + b->o_params),
                                             (double* const*)(b->dev + b->o_tables), &h0->prep, h0->M.maxloc, nullptr, n,
                                             nullptr, s));
            for (int j = 0; j < n; ++j) HIPCHECK(launch_custom(hs[j], s));
            const EvalJob* d_jobs = (const EvalJob*)(b->dev + b->o_jobs);
            void* args[] = {&d_jobs, &n};
            if (wg > 0) HIPCHECK(hipModuleLaunchKernel(h0->fn_group, wg, 1, 1, h0->M.wgs, 1, 1, 0, s, args, nullptr));
            HIPCHECK(ctpwa_launch_finalize_batch((const FinArgs*)(b->dev + b->o_fins), n, b->max_nacc, b->max_lds, nullptr,
                                                 s));
            return 0;
        }
        if (b->gmod != h0->mod) {
            // a new model: its grouped kernel's persistent grid, and no graph of another model
            batch_drop_graphs(b);
            int per_cu = 0, ncu = 0;
            HIPCHECK(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, h0->fn_group_p, h0->M.wgs, 0));
            HIPCHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, b->device));
            b->persist = std::max(1, per_cu) * std::max(1, ncu);
            b->gmod = h0->mod;
        }
        auto it = b->execs.find(b->rows);
        if (it == b->execs.end()) {
            hipGraph_t graph_ = nullptr;
            hipGraphExec_t exec = nullptr;
            HIPCHECK(hipStreamBeginCapture(b->cs, hipStreamCaptureModeThreadLocal));
            try {
                batch_launches_graph(b, h0, b->rows, b->cs);
            } catch (...) {
                (void)hipStreamEndCapture(b->cs, &graph_);
                if (graph_) (void)hipGraphDestroy(graph_);
                throw;
            }
            HIPCHECK(hipStreamEndCapture(b->cs, &graph_));
            hipError_t e = hipGraphInstantiate(&exec, graph_, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph_);
            HIPCHECK(e);
            it = b->execs.emplace(b->rows, exec).first;
        }
        HIPCHECK(hipGraphLaunch(it->second, s));
        return 0;
    } catch (const std::exception& e) {
        b->key.clear();                    // upload again on the next call
        return fail(std::string("ctpwa_batch_nll_grad: ") + e.what());
    }
}

// per-event bytes of the SL tensors and the invariants a model stores at a precision
static long long event_bytes(const Model_& M, const std::vector<int>* used) {
    std::vector<int> imap = M.imap();
    long long b = (long long)M.us.size() * M.npstore() * (M.mixed ? 8 : 16);
    for (size_t i = 0; i < M.inv.size(); ++i)
        if (!used || (*used)[i]) b += (imap[i] >= 0) ? 8 : 4;
    b += (long long)M.customs.size() * M.ncf() * 8;      // Custom factor rows
    return b;
}

extern "C" long long ctpwa_footprint_bytes(const char* plan_text, long long n_phsp, long long n_data, long long n_bkg,
                                           int precision) {
    try {
        if (precision != 0 && precision != 1) throw std::runtime_error("precision must be 0 (FP64) or 1 (mixed)");
        Model_ M = derive(parse_plan(plan_text));
        M.mixed = precision;
        long long nt = n_phsp + n_data + n_bkg;
        long long b = nt * event_bytes(M, nullptr);
        b += (long long)M.table_size() * 8 + (long long)(1 + 2 * M.P.namp + M.P.ntheta) * 1024 * 8;
        b += nt * (long long)M.customs.size() * M.ncf() * 8;     // Custom factor rows
        // no free resonance parameter: the stored phase-space Gram and its build workspace
        if (M.P.ntheta == 0)
            b += (long long)M.P.namp * M.P.namp * 16 +
                 agram_workspace_bytes(M.P, std::max(n_phsp, std::max(n_data, n_bkg)));
        // (the upload staging is one pool per device for the process, not per handle)
        return b;
    } catch (const std::exception& e) {
        g_err = std::string("ctpwa_footprint_bytes: ") + e.what();
        return -1;
    }
}

extern "C" long long ctpwa_hessian_workspace_bytes(const char* plan_text, long long n_phsp, long long n_data,
                                                   long long n_bkg) {
    try {
        Model_ M = derive(parse_plan(plan_text));
        const long long n3[3] = {n_phsp, n_data, n_bkg};
        // the handle these buffers are for does not exist yet: every live handle is another
        return hess_workspace_bytes(M, n3, n_phsp + n_data + n_bkg, g_live_handles.load());
    } catch (const std::exception& e) {
        g_err = std::string("ctpwa_hessian_workspace_bytes: ") + e.what();
        return -1;
    }
}

extern "C" int ctpwa_layout(const ctpwa_handle* h, long long* n_events3, long long* bytes_per_eval, int* n_sl, int* n_inv,
                            int* n_wg, int* n_pol_stored) {
    if (!h) return fail("ctpwa_layout: NULL handle");
    if (n_events3) for (int c = 0; c < 3; ++c) n_events3[c] = h->n[c];
    int used = 0;
    for (int u : h->used_inv) used += u;
    if (bytes_per_eval) {
        // with the stored phase-space Gram an evaluation reads data and background only
        long long b = ((h->d_gram ? 0 : h->n[0]) + h->n[1] + h->n[2]) * event_bytes(h->M, &h->used_inv);
        for (int c = 1; c < 3; ++c) if (h->pw[c]) b += h->n[c] * 8;
        *bytes_per_eval = b;
    }
    if (n_sl) *n_sl = (int)h->M.us.size();
    if (n_inv) *n_inv = used;
    if (n_wg) *n_wg = h->nwg;
    if (n_pol_stored) *n_pol_stored = h->M.npstore();
    return 0;
}

extern "C" const char* ctpwa_kernel_source(const ctpwa_handle* h) { return h ? h->src.c_str() : ""; }

// source of the fused kernel for a plan, without a device (buf may be NULL to query the size)
extern "C" long long ctpwa_generate_source(const char* plan_text, int precision, char* buf, long long cap) {
    try {
        if (precision != 0 && precision != 1) throw std::runtime_error("precision must be 0 (FP64) or 1 (mixed)");
        Model_ M = derive(parse_plan(plan_text));
        M.mixed = precision;
        apply_codegen_env(M);
        std::string src = generate_eval_kernel(M);
        if (buf && cap > 0) {
            long long n = std::min<long long>(cap - 1, (long long)src.size());
            std::memcpy(buf, src.data(), (size_t)n);
            buf[n] = '\0';
        }
        return (long long)src.size() + 1;
    } catch (const std::exception& e) {
        g_err = std::string("ctpwa_generate_source: ") + e.what();
        return -1;
    }
}

// source of the generated Hessian rows kernel for a plan, without a device
extern "C" long long ctpwa_generate_hessian_source(const char* plan_text, int precision, char* buf, long long cap) {
    try {
        if (precision != 0 && precision != 1) throw std::runtime_error("precision must be 0 (FP64) or 1 (mixed)");
        Model_ M = derive(parse_plan(plan_text));
        M.mixed = precision;
        std::string src = generate_hess_kernel(M, nullptr);
        // models with Custom lineshapes: the second-order factor module follows
        if (!M.customs.empty()) src += "\n// ---- module: ctpwa_custom2 ----\n" + generate_custom2_kernel(M);
        if (buf && cap > 0) {
            long long n = std::min<long long>(cap - 1, (long long)src.size());
            std::memcpy(buf, src.data(), (size_t)n);
            buf[n] = '\0';
        }
        return (long long)src.size() + 1;
    } catch (const std::exception& e) {
        g_err = std::string("ctpwa_generate_hessian_source: ") + e.what();
        return -1;
    }
}

// compile the fused kernel for a plan with hipRTC (no device needed); returns the
// code-object size in bytes, or -1 with the compiler log in ctpwa_last_error()
extern "C" long long ctpwa_compile_check(const char* plan_text, int precision) {
    try {
        if (precision != 0 && precision != 1) throw std::runtime_error("precision must be 0 (FP64) or 1 (mixed)");
        Model_ M = derive(parse_plan(plan_text));
        M.mixed = precision;
        apply_codegen_env(M);
        auto compile = [](const std::string& src) -> long long {
            hiprtcProgram prog;
            if (hiprtcCreateProgram(&prog, src.c_str(), "ctpwa_eval.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS)
                throw std::runtime_error("hiprtcCreateProgram failed");
            hiprtcResult r = hiprtcCompileProgram(prog, RTC_NOPTS, RTC_OPTS);
            size_t ls = 0;
            hiprtcGetProgramLogSize(prog, &ls);
            std::string log(ls + 1, '\0');
            hiprtcGetProgramLog(prog, &log[0]);
            size_t cs = 0;
            if (r == HIPRTC_SUCCESS) hiprtcGetCodeSize(prog, &cs);
            hiprtcDestroyProgram(&prog);
            if (r != HIPRTC_SUCCESS) throw std::runtime_error(log);
            return (long long)cs;
        };
        long long cs = compile(generate_eval_kernel(M));
        cs += compile(generate_hess_kernel(M, nullptr));
        // models with Custom lineshapes: also the Hessian's second-order factor module
        if (!M.customs.empty()) cs += compile(generate_custom2_kernel(M));
        return cs;
    } catch (const std::exception& e) {
        g_err = std::string("ctpwa_compile_check: ") + e.what();
        return -1;
    }
}
extern "C" const char* ctpwa_last_error(void) { return g_err.c_str(); }
extern "C" const char* ctpwa_version(void) { return "ctpwa_amd 0.1.0 (gfx950)"; }
