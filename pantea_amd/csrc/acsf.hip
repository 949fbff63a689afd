// (b) ACSF descriptors, their adjoint and their tangent on gfx950.
//
// One wavefront (64 lanes) per centre.  The centre's neighbour row is staged
// in LDS (relative vectors R_ij = r_i - r_j under the minimum image with r_ij
// or 1/r_ij, and per mode the cutoff values), compacted to r_ij <= rc_max and
// r_ij > 0 (atoms/neighbor.py:102-107).  Rows arrive sorted by type, so each
// neighbour element is a contiguous LDS segment.
//
// Radial groups (descriptors/acsf/radial.py:38-61, acsf.py:231-246):
//   G2_s = sum_j exp(-eta_s (r_ij - rs_s)^2) fc(r_ij), lanes over neighbours.
// Angular groups (angular.py:50-107, acsf.py:249-330): every unordered pair
// {j, k} of the group's segments is visited once (for equal neighbour
// elements this equals the reference's ordered sum halved).  With
//   S = r_ij^2 + r_ik^2 + r_jk^2,  F = fc(r_ij) fc(r_ik) fc(r_jk)
// (G9: S and F without the r_jk terms), each function of the group is a cell
// (eta e, angular term a) of an outer product:
//   term(e, a) = exp(-eta_e S) F * pre_a (1 + lam_a cos)^zeta_a .
// r_jk = |MIC(R_ij - R_ik)| as in the reference (the MIC is skipped when the
// box exceeds 4 rc, where it cannot act), and pairs with r_jk = 0 are
// excluded (acsf.py:325).  Pairs are first screened (r_jk < rc for G3) and
// the live ones compacted into a per-wave LDS queue, so the heavy per-pair
// work runs on full wavefronts.
//
// The adjoint (MODE_ADJ) reverses the same loops with w = dE/dG.  Per pair
// only two eta sums are needed, SP = sum_e P_e and SEP = sum_e eta_e P_e with
// P_e = exp(-eta_e S) sum_a w_ea ang_a, from which dT/dr_ij, dT/dr_ik,
// dT/dr_jk follow; dT/dcos uses sum_e exp(-eta_e S) sum_a w_ea dang_a.  These
// become gradients g_ij, g_ik w.r.t. R_ij and R_ik; -g goes into the LDS slot
// of each neighbour (LDS atomics) and g into the centre.  The neighbour slots
// become the per-edge buffer dE_i/dr_j; the centre part is -F_self
// (potentials/nnp/force.py:16-43).  The tangent (MODE_TAN) pushes a
// displacement field delta forward: T_is = sum_m dG_is/dr_m . delta_m.
#include <algorithm>
#include <cmath>

#include "common.h"

namespace pt {
namespace {

enum { MODE_FWD = 0, MODE_ADJ = 1, MODE_TAN = 2 };

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// ---- precision-specific math -------------------------------------------------
// ex2k(x): exponential of x scaled by kexp (FP32: v_exp_f32 = 2^x with the
// log2(e) factor folded into the caller's constants; FP64: libm exp).
template <typename R> struct M;
// FP64 elementary functions for the pair loops (the kernels' accuracy goal is
// ~1e-15 relative, far inside the 1e-10 parity bound):
//  * exp: Tang's table-driven method (ACM TOMS 15(2), 1989): x = (32 m + j)
//    ln2/32 + r with |r| <= ln2/64 (two-constant Cody-Waite reduction),
//    e^x = 2^m * 2^(j/32) * e^r, e^r by its degree-6 Taylor polynomial
//    (remainder < 4e-18); the 32 values 2^(j/32) live in LDS;
//  * 1/x and 1/sqrt(x): the hardware estimates refined by two Newton steps.
__constant__ double kExp2Tab[32] = {
    0x1.0000000000000p+0, 0x1.059b0d3158574p+0, 0x1.0b5586cf9890fp+0, 0x1.11301d0125b51p+0,
    0x1.172b83c7d517bp+0, 0x1.1d4873168b9aap+0, 0x1.2387a6e756238p+0, 0x1.29e9df51fdee1p+0,
    0x1.306fe0a31b715p+0, 0x1.371a7373aa9cbp+0, 0x1.3dea64c123422p+0, 0x1.44e086061892dp+0,
    0x1.4bfdad5362a27p+0, 0x1.5342b569d4f82p+0, 0x1.5ab07dd485429p+0, 0x1.6247eb03a5585p+0,
    0x1.6a09e667f3bcdp+0, 0x1.71f75e8ec5f74p+0, 0x1.7a11473eb0187p+0, 0x1.82589994cce13p+0,
    0x1.8ace5422aa0dbp+0, 0x1.93737b0cdc5e5p+0, 0x1.9c49182a3f090p+0, 0x1.a5503b23e255dp+0,
    0x1.ae89f995ad3adp+0, 0x1.b7f76f2fb5e47p+0, 0x1.c199bdd85529cp+0, 0x1.cb720dcef9069p+0,
    0x1.d5818dcfba487p+0, 0x1.dfc97337b9b5fp+0, 0x1.ea4afa2a490dap+0, 0x1.f50765b6e4540p+0,
};
__shared__ double g_exp2_tab[32];

// every thread of the block calls this before any early exit
__device__ __forceinline__ void exp2_tab_load() {
    if (threadIdx.x < 32) g_exp2_tab[threadIdx.x] = kExp2Tab[threadIdx.x];
    __syncthreads();
}

__device__ __forceinline__ double exp_tab(double x) {
    x = fmax(x, -708.0);
    const double nf = __builtin_rint(x * 0x1.71547652b82fep+5);          // 32 / ln2
    double r = fma(nf, -0x1.62e42fee00000p-6, x);                       // ln2/32, high part
    r = fma(nf, -0x1.a39ef35793c76p-38, r);                             // ln2/32, low part
    const int n = (int)nf;
    double p = fma(r, 1.0 / 720.0, 1.0 / 120.0);
    p = fma(p, r, 1.0 / 24.0);
    p = fma(p, r, 1.0 / 6.0);
    p = fma(p, r, 0.5);
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    return __builtin_amdgcn_ldexp(g_exp2_tab[n & 31] * p, n >> 5);
}

__device__ __forceinline__ double rcp_nr(double x) {
    double y = __builtin_amdgcn_rcp(x);
    y = fma(y, fma(-x, y, 1.0), y);
    y = fma(y, fma(-x, y, 1.0), y);
    return y;
}

__device__ __forceinline__ double rsq_nr(double x) {
    double y = __builtin_amdgcn_rsq(x);
    const double h = 0.5 * x;
    y = y * fma(-h, y * y, 1.5);
    y = y * fma(-h, y * y, 1.5);
    return y;
}

template <> struct M<double> {
    static constexpr double kexp = 1.0;
    static __device__ __forceinline__ double ex2k(double x) { return exp_tab(x); }
    static __device__ __forceinline__ double ex(double x) { return exp_tab(x); }
    static __device__ __forceinline__ double th(double x) {
        // tanh(u) = (1 - e^{-2u}) / (1 + e^{-2u}); used for u in [-1, 1]
        const double e = exp_tab(-2.0 * x);
        return (1.0 - e) * rcp_nr(1.0 + e);
    }
    static __device__ __forceinline__ double co(double x) { return cos(x); }
    static __device__ __forceinline__ double si(double x) { return sin(x); }
    static __device__ __forceinline__ double sq(double x) { return sqrt(x); }
    static __device__ __forceinline__ double pw(double b, double z) { return pow(b, z); }
    static __device__ __forceinline__ double rcp(double x) { return rcp_nr(x); }
    // r and 1/r from r^2 > 0
    static __device__ __forceinline__ void root(double r2, double& r, double& ir) {
        ir = rsq_nr(r2);
        r = r2 * ir;
    }
};
template <> struct M<float> {
    static constexpr float kexp = 1.4426950408889634f;  // log2(e)
    static __device__ __forceinline__ float ex2k(float x) { return __builtin_amdgcn_exp2f(x); }
    static __device__ __forceinline__ float ex(float x) { return __builtin_amdgcn_exp2f(x * kexp); }
    static __device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
    static __device__ __forceinline__ float th(float x) {
        // tanh(u) = (1 - e^{-2u}) / (1 + e^{-2u}); used for u in [-1, 1]
        const float e = __builtin_amdgcn_exp2f(x * (-2.0f * kexp));
        return (1.0f - e) * rcp(1.0f + e);
    }
    static __device__ __forceinline__ float co(float x) { return __cosf(x); }
    static __device__ __forceinline__ float si(float x) { return __sinf(x); }
    static __device__ __forceinline__ float sq(float x) { return __builtin_amdgcn_sqrtf(x); }
    static __device__ __forceinline__ float pw(float b, float z) { return powf(b, z); }
    static __device__ __forceinline__ void root(float r2, float& r, float& ir) {
        ir = __builtin_amdgcn_rsqf(r2);
        r = r2 * ir;
    }
};

// Cutoff value and derivative (descriptors/acsf/cutoff.py:66-110).
template <typename R>
__device__ __forceinline__ void cutoff(int type, R rc, R inv_rc, R tanh_pre, R r, R& f, R& df) {
    if (!(r < rc)) { f = R(0); df = R(0); return; }
    switch (type) {
    case PT_CUT_HARD: f = R(1); df = R(0); return;
    case PT_CUT_TANHU:
    case PT_CUT_TANH: {
        const R t = M<R>::th(R(1) - r * inv_rc);
        const R pre = (type == PT_CUT_TANH) ? tanh_pre : R(1);
        const R t2 = t * t;
        f = pre * t2 * t;
        df = -pre * R(3) * t2 * (R(1) - t2) * inv_rc;
        return;
    }
    case PT_CUT_COS: {
        const R x = R(M_PI) * r * inv_rc;
        f = R(0.5) * (M<R>::co(x) + R(1));
        df = R(-0.5) * M<R>::si(x) * R(M_PI) * inv_rc;
        return;
    }
    case PT_CUT_EXP: {
        const R x = r * inv_rc, q = R(1) - x * x;
        const R iq = M<R>::rcp(q);
        f = M<R>::ex(R(1) - iq);
        df = f * (R(-2) * x * iq * iq) * inv_rc;
        return;
    }
    case PT_CUT_POLY1:
        f = (R(2) * r - R(3)) * r * r + R(1);
        df = R(6) * r * r - R(6) * r;
        return;
    case PT_CUT_POLY2:
        f = ((R(15) - R(6) * r) * r - R(10)) * r * r * r + R(1);
        df = ((R(-30) * r + R(60)) * r - R(30)) * r * r;
        return;
    }
    f = R(0); df = R(0);
}

// Cutoff specialised per group: CUTK_TANH (tanh / tanhu, cpre = prefactor),
// CUTK_COS, or the generic switch.  Branch-free for the first two.
enum { CUTK_TANH = 0, CUTK_COS = 1, CUTK_GEN = 2 };

template <typename R, int CUT>
__device__ __forceinline__ void cutoff_k(int type, R rc, R irc, R cpre, R tanh_pre, R r, R& f, R& df) {
    if (CUT == CUTK_TANH) {
        const R t = M<R>::th(R(1) - r * irc);
        const R t2 = t * t;
        const bool in = r < rc;
        f = in ? cpre * t2 * t : R(0);
        df = in ? R(-3) * cpre * t2 * (R(1) - t2) * irc : R(0);
    } else if (CUT == CUTK_COS) {
        const R x = R(M_PI) * r * irc;
        const bool in = r < rc;
        f = in ? R(0.5) * (M<R>::co(x) + R(1)) : R(0);
        df = in ? R(-0.5 * M_PI) * irc * M<R>::si(x) : R(0);
    } else {
        cutoff<R>(type, rc, irc, tanh_pre, r, f, df);
    }
}

// pre (1 + lam c)^zeta and its derivative w.r.t. c (angular.py:60-65).
template <typename R>
__device__ __forceinline__ void angular_term(R c, R lam, R zeta, int izeta, R pre, R& v, R& dv) {
    const R b = R(1) + lam * c;
    if (izeta >= 0) {
        // pm1 = b^(izeta - 1) by square-and-multiply over the bits of izeta - 1 (<= 15)
        const int z1 = izeta > 0 ? izeta - 1 : 0;
        R pm1 = R(1), q = b;
        if (z1 & 1) pm1 *= q;
        q *= q;
        if (z1 & 2) pm1 *= q;
        q *= q;
        if (z1 & 4) pm1 *= q;
        q *= q;
        if (z1 & 8) pm1 *= q;
        v = pre * (izeta > 0 ? pm1 * b : R(1));
        dv = (izeta == 0) ? R(0) : pre * R(izeta) * lam * pm1;
    } else {
        v = pre * M<R>::pw(b, zeta);
        dv = pre * zeta * lam * M<R>::pw(b, zeta - R(1));
    }
}

// Angular-term evaluation specialised per group: ANGK_Z1 (every zeta = 1),
// ANGK_Z1N2 (every zeta = 1, at most two terms), ANGK_P4 (every zeta in
// {1, 2, 4}), ANGK_Z14 (exactly the four terms zeta in {1, 4} x lambda in
// {+1, -1}, evaluated in that canonical order from the shared powers of
// 1 +- cos; the group's columns are mapped to it) or the generic
// square-and-multiply / pow.  The specialised forms leave the factor
// pre = 2^(1 - zeta) out of the pair loop: it scales the group's sums
// (forward) or dE/dG (adjoint) once per centre.
enum { ANGK_Z1 = 0, ANGK_P4 = 1, ANGK_GEN = 2, ANGK_Z1N2 = 3, ANGK_Z14 = 4 };

template <int ANG>
__host__ __device__ constexpr int ang_terms() { return ANG == ANGK_Z1N2 ? 2 : PT_AG_ANG; }

template <typename R> struct alignas(4 * sizeof(R)) V4 { R x, y, z, w; };
template <typename R> struct alignas(2 * sizeof(R)) V2 { R x, y; };

// Live-pair queue of one wavefront: (a | b << 16) indices of pairs that
// pass the r_jk screen; a batch of 64 is taken as soon as it is full, so it
// never holds more than 63 + 64 entries.
template <int MODE>
__host__ __device__ constexpr int queue_len() { return 128; }

template <typename R>
struct Stage {
    V4<R>* P;      // (x, y, z, 1/r) of R_ij = r_i - r_j
    V4<R>* Q;      // TAN: (r, r^2, fc, dfc/dr) -- cutoff of the current group
    V2<R>* Q2;     // FWD: (r, fc)
    V2<R>* QF;     // ADJ: (fc, dfc / r) of the current group
    R* D;          // [3][cap] TAN: delta_i - delta_j
    double* DA;    // [cap][3] ADJ: per-edge accumulators (FP64: fast LDS atomics; x, y, z
                   // adjacent so the three adds of an edge share one address)
    unsigned long long* keep;   // ADJ: the staging ballots, one per 64 row entries
    unsigned* queue;
    R* W;          // ADJ, FP64: the group's dE/dG weights (read by broadcast)
};

// The adjoint keeps a compact record (x, y, z, 1/r) + per group (fc, dfc / r)
// and no per-neighbour row slot (r^2 is recomputed per pair from x, y, z,
// edge positions come back from the staging ballots): the LDS per wave then
// allows three waves per SIMD in FP64 and four and a half in FP32.

template <typename R, int MODE>
__host__ __device__ constexpr size_t stage_bytes_per_nbr() {
    return MODE == MODE_ADJ ? sizeof(double) * 3 + sizeof(R) * 6
                            : sizeof(R) * (MODE == MODE_FWD ? 6 : 11);
}

// cap is a multiple of 4 and >= the row length; the per-wave area is a
// multiple of 32 bytes
template <typename R, int MODE>
__host__ __device__ constexpr size_t stage_bytes(int cap) {
    return ((size_t)cap * stage_bytes_per_nbr<R, MODE>() + queue_len<MODE>() * sizeof(unsigned) +
            (MODE == MODE_ADJ ? PT_AG_ETA * PT_AG_ANG * sizeof(R) +
                                    (size_t)((cap + kWave - 1) / kWave) * sizeof(unsigned long long)
                              : 0) + 31) / 32 * 32;
}

template <typename R, int MODE>
__device__ __forceinline__ Stage<R> carve(char* base, int cap) {
    Stage<R> s;
    char* p = base;
    if (MODE == MODE_ADJ) { s.DA = reinterpret_cast<double*>(p); p += 3 * cap * sizeof(double); } else { s.DA = nullptr; }
    s.P = reinterpret_cast<V4<R>*>(p); p += cap * sizeof(V4<R>);
    s.Q = nullptr; s.Q2 = nullptr; s.QF = nullptr;
    if ((MODE == MODE_ADJ)) {
        s.QF = reinterpret_cast<V2<R>*>(p); p += cap * sizeof(V2<R>);
    } else if (MODE == MODE_FWD) {
        s.Q2 = reinterpret_cast<V2<R>*>(p); p += cap * sizeof(V2<R>);
    } else {
        s.Q = reinterpret_cast<V4<R>*>(p); p += cap * sizeof(V4<R>);
    }
    if (MODE == MODE_TAN) { s.D = reinterpret_cast<R*>(p); p += 3 * cap * sizeof(R); } else { s.D = nullptr; }
    s.queue = reinterpret_cast<unsigned*>(p);
    p += queue_len<MODE>() * sizeof(unsigned);
    s.W = nullptr;
    s.keep = nullptr;
    if (MODE == MODE_ADJ) {
        s.W = reinterpret_cast<R*>(p);
        p += PT_AG_ETA * PT_AG_ANG * sizeof(R);
        s.keep = reinterpret_cast<unsigned long long*>(p);
    }
    return s;
}

__device__ __forceinline__ void lds_add(double* p, double v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

template <typename R>
__device__ __forceinline__ R mic_r(R d, R L, R half) {
    d = (d > half) ? d - L : d;
    d = (d < -half) ? d + L : d;
    return d;
}

template <typename R>
__device__ __forceinline__ void mic3(const BoxArg& b, R& x, R& y, R& z) {
    x = mic_r(x, (R)b.L[0], (R)b.half[0]);
    y = mic_r(y, (R)b.L[1], (R)b.half[1]);
    z = mic_r(z, (R)b.L[2], (R)b.half[2]);
}

// Staged neighbours with type < t for t = 0 .. kSegTypes (rows are type-sorted,
// so each element is one contiguous segment).  Types are 1-based.
constexpr int kSegTypes = 9;

// the counts live in lanes 0 .. kSegTypes of one VGPR (no private array)
__device__ __forceinline__ int seg_below(int segv, int t) {
    return __builtin_amdgcn_readlane(segv, t < 0 ? 0 : (t > kSegTypes ? kSegTypes : t));
}

struct KArgs {
    const double* pos4;
    BoxArg box;
    const long long* row_off;
    const int* nbr;
    int n_centers, center_begin;
    const int* centers;
    const pt_sf_table* table;
    void* out;            // FWD: G, TAN: T (n_centers x ld)
    int ld;
    const void* w;        // ADJ: dE/dG (n_centers x ldw)
    int ldw;
    double* f_self;       // ADJ (n_centers x 3)
    void* edge_grad;      // ADJ (n_pairs x 3)
    const int* rev;       // ADJ: when set, edge e's gradient goes to slot rev[e] (the reverse edge)
    const double* delta;  // TAN (n x 3, internal order)
    int center_only;
    int cap;
};

// Uniform description of one angular group inside the kernel.
template <typename R>
struct AngGroup {
    int bj, nj, bk, nk;
    bool same, g3, mic;
    int n_eta, n_ang, cut;
    R rc, irc, cpre, tanh_pre;
    const BoxArg* box;
    R keta[PT_AG_ETA];     // -eta * kexp: exp(-eta S) = ex2k(keta S)
    R lam[PT_AG_ANG], zeta[PT_AG_ANG], pre[PT_AG_ANG], zl[PT_AG_ANG];   // zl = zeta * lambda
    int izeta[PT_AG_ANG];
};

template <typename R, int ANG, bool DER>
__device__ __forceinline__ void ang_eval(const AngGroup<R>& X, R cs, R (&v)[PT_AG_ANG], R (&dv)[PT_AG_ANG]) {
    // terms q >= n_ang carry pre = lam = zl = 0 and izeta = 1; in the
    // specialised forms v = (1 + lam c)^zeta and dv = zeta lam (1 + lam c)^(zeta - 1)
    // without pre, in the generic one both include pre
    constexpr int NA = ang_terms<ANG>();
    if constexpr (ANG == ANGK_Z14) {
        // (1 + c), (1 - c), (1 + c)^4, (1 - c)^4 and their derivatives in c
        const R bp = R(1) + cs, bm = R(1) - cs;
        const R bp2 = bp * bp, bm2 = bm * bm;
        v[0] = bp; v[1] = bm; v[2] = bp2 * bp2; v[3] = bm2 * bm2;
        dv[0] = DER ? R(1) : R(0);
        dv[1] = DER ? R(-1) : R(0);
        dv[2] = DER ? R(4) * bp * bp2 : R(0);
        dv[3] = DER ? R(-4) * bm * bm2 : R(0);
        return;
    }
#pragma unroll
    for (int q = 0; q < NA; ++q) {
        if (ANG == ANGK_Z1 || ANG == ANGK_Z1N2) {
            v[q] = R(1) + X.lam[q] * cs;
            dv[q] = DER ? X.zl[q] : R(0);
        } else if (ANG == ANGK_P4) {
            const R b = R(1) + X.lam[q] * cs;
            const R b2 = b * b;
            const int z = X.izeta[q];
            v[q] = z == 1 ? b : (z == 2 ? b2 : b2 * b2);
            dv[q] = DER ? X.zl[q] * (z == 1 ? R(1) : (z == 2 ? b : b * b2)) : R(0);
        } else {
            angular_term<R>(cs, X.lam[q], X.zeta[q], X.izeta[q], X.pre[q], v[q], dv[q]);
        }
    }
}

// All unordered pairs {j, k} of one angular group: screening on r_jk, a
// per-wave queue of live pairs, and the pair body in full-wave batches.
template <typename R, int MODE, int CUT, int ANG>
__device__ __forceinline__ void angular_group(const Stage<R>& S, const AngGroup<R>& X,
                                              const R (&Win)[PT_AG_ETA][PT_AG_ANG],
                                              R (&acc)[PT_AG_ETA][PT_AG_ANG], int cap) {
    const int lane = wave_lane();
    constexpr int NA = ang_terms<ANG>();
    constexpr bool kPreOut = ANG != ANGK_GEN;   // pre applied outside the pair loop
    // FP64 adjoint: the weights live in LDS (broadcast reads), which frees
    // 32 VGPRs in the pair loop
    constexpr bool kWLds = MODE == MODE_ADJ && sizeof(R) == 8;
    R W[PT_AG_ETA][PT_AG_ANG];
#pragma unroll
    for (int e = 0; e < PT_AG_ETA; ++e)
#pragma unroll
        for (int q = 0; q < PT_AG_ANG; ++q) W[e][q] = kPreOut ? Win[e][q] * X.pre[q] : Win[e][q];
    // Adjoint of the zeta {1, 4} x lambda +-1 class: the weighted angular sum
    // of one eta, U(c) = w0 (1 + c) + w1 (1 - c) + w2 (1 + c)^4 + w3 (1 - c)^4,
    // is a quartic in c.  With sp = w2 + w3, sm = w2 - w3, W[e] becomes
    // (u0, u1, sp, sm4) = (w0 + w1 + sp, w0 - w1 + 4 sm, sp, 4 sm) and
    //   U = u0 + u1 c + sp c^2 (6 + c^2) + sm4 c^3,
    //   U' = u1 + sp (12 c + 4 c^3) + 3 sm4 c^2,
    // so the pair body needs no angular terms, and the powers of c are
    // shared by the eta values.
    constexpr bool kQuartic = MODE == MODE_ADJ && ANG == ANGK_Z14;
    if (kQuartic) {
#pragma unroll
        for (int e = 0; e < PT_AG_ETA; ++e) {
            const R w0 = W[e][0], w1 = W[e][1], sp = W[e][2] + W[e][3], sm = W[e][2] - W[e][3];
            W[e][0] = (w0 + w1) + sp;
            W[e][1] = (w0 - w1) + R(4) * sm;
            W[e][2] = sp;
            W[e][3] = R(4) * sm;
        }
    }
    if (kWLds) {
        if (lane == 0) {
#pragma unroll
            for (int e = 0; e < PT_AG_ETA; ++e)
#pragma unroll
                for (int q = 0; q < PT_AG_ANG; ++q) S.W[e * PT_AG_ANG + q] = W[e][q];
        }
    }
    // cutoff values of the group's neighbours
    for (int pass = 0; pass < (X.same ? 1 : 2); ++pass) {
        const int sb = pass == 0 ? X.bj : X.bk, sn = pass == 0 ? X.nj : X.nk;
        for (int s = sb + lane; s < sb + sn; s += kWave) {
            R f, df;
            if constexpr ((MODE == MODE_ADJ)) {
                const R ir = S.P[s].w;
                cutoff_k<R, CUT>(X.cut, X.rc, X.irc, X.cpre, X.tanh_pre, M<R>::rcp(ir), f, df);
                V2<R> fv;
                fv.x = f; fv.y = df * ir;   // dfc / r
                S.QF[s] = fv;
            } else if constexpr (MODE == MODE_FWD) {
                cutoff_k<R, CUT>(X.cut, X.rc, X.irc, X.cpre, X.tanh_pre, S.Q2[s].x, f, df);
                S.Q2[s].y = f;
            } else {
                cutoff_k<R, CUT>(X.cut, X.rc, X.irc, X.cpre, X.tanh_pre, S.Q[s].x, f, df);
                S.Q[s].z = f;
                S.Q[s].w = df;
            }
        }
    }
    wave_sync();

    // one live pair (ia, ib)
    auto body = [&](const int ia, const int ib) {
        const V4<R> pa = S.P[ia], pb = S.P[ib];
        V4<R> qa, qb;   // (r, r^2, fc, dfc); ADJ: (unused, r^2, fc, dfc / r)
        R ira, irb;     // 1 / r
        if constexpr ((MODE == MODE_ADJ)) {
            const V2<R> ua = S.QF[ia], ub = S.QF[ib];
            ira = pa.w;
            irb = pb.w;
            qa.y = pa.x * pa.x + pa.y * pa.y + pa.z * pa.z;
            qb.y = pb.x * pb.x + pb.y * pb.y + pb.z * pb.z;
            qa.x = R(0); qa.z = ua.x; qa.w = ua.y;
            qb.x = R(0); qb.z = ub.x; qb.w = ub.y;
        } else if constexpr (MODE == MODE_FWD) {
            const V2<R> ua = S.Q2[ia], ub = S.Q2[ib];
            qa.x = ua.x; qa.y = ua.x * ua.x; qa.z = ua.y; qa.w = R(0);
            qb.x = ub.x; qb.y = ub.x * ub.x; qb.z = ub.y; qb.w = R(0);
        } else {
            qa = S.Q[ia];
            qb = S.Q[ib];
        }
        if constexpr ((MODE != MODE_ADJ)) {
            ira = pa.w;
            irb = pb.w;
        }
        R jx = pa.x - pb.x, jy = pa.y - pb.y, jz = pa.z - pb.z;
        if (X.mic) mic3(*X.box, jx, jy, jz);
        const R rjk2 = jx * jx + jy * jy + jz * jz;   // > 0 (screened)
        R rjk, irjk;
        M<R>::root(rjk2, rjk, irjk);
        R fjk = R(1), dfjk = R(0);
        if (X.g3) cutoff_k<R, CUT>(X.cut, X.rc, X.irc, X.cpre, X.tanh_pre, rjk, fjk, dfjk);
        const R Ssum = X.g3 ? qa.y + qb.y + rjk2 : qa.y + qb.y;
        const R F = qa.z * qb.z * fjk;
        const R iab = ira * irb;
        const R cs = (pa.x * pb.x + pa.y * pb.y + pa.z * pb.z) * iab;
        R ang[PT_AG_ANG], dang[PT_AG_ANG];
        if constexpr (!kQuartic) ang_eval<R, ANG, MODE != MODE_FWD>(X, cs, ang, dang);
        if (MODE == MODE_FWD) {
#pragma unroll
            for (int e = 0; e < PT_AG_ETA; ++e) {
                if (e < X.n_eta) {
                    const R rad = M<R>::ex2k(X.keta[e] * Ssum) * F;
#pragma unroll
                    for (int q = 0; q < NA; ++q) acc[e][q] += rad * ang[q];
                }
            }
            return;
        }
        // derivative building blocks
        const R ra = qa.x, rb = qb.x;
        const R ira2 = ira * ira, irb2 = irb * irb;
        if (!X.g3) irjk = R(0);
        // d F / d r_ij etc. (without the exp factor)
        const R Fa = qa.w * qb.z * fjk, Fb = qa.z * qb.w * fjk, Fj = qa.z * qb.z * dfjk;
        if (MODE == MODE_ADJ) {
            R SP = R(0), SKP = R(0), SV = R(0);
            R c3 = R(0), t4 = R(0), dt4 = R(0), c2x3 = R(0);
            if constexpr (kQuartic) {
                const R c2 = cs * cs;
                c3 = c2 * cs;
                t4 = c2 * (R(6) + c2);
                dt4 = R(12) * cs + R(4) * c3;
                c2x3 = R(3) * c2;
            }
#pragma unroll
            for (int e = 0; e < PT_AG_ETA; ++e) {
                if (e < X.n_eta) {
                    const R ex = M<R>::ex2k(X.keta[e] * Ssum);
                    R U = R(0), V = R(0);
                    if constexpr (kQuartic) {
                        const R u0 = kWLds ? S.W[e * PT_AG_ANG + 0] : W[e][0];
                        const R u1 = kWLds ? S.W[e * PT_AG_ANG + 1] : W[e][1];
                        const R sp = kWLds ? S.W[e * PT_AG_ANG + 2] : W[e][2];
                        const R sm4 = kWLds ? S.W[e * PT_AG_ANG + 3] : W[e][3];
                        U = u0 + u1 * cs + sp * t4 + sm4 * c3;
                        V = u1 + sp * dt4 + sm4 * c2x3;
                    } else {
#pragma unroll
                        for (int q = 0; q < NA; ++q) {
                            const R wq = kWLds ? S.W[e * PT_AG_ANG + q] : W[e][q];
                            U += wq * ang[q];
                            V += wq * dang[q];
                        }
                    }
                    const R P = ex * U;
                    SP += P;
                    SKP += X.keta[e] * P;
                    SV += ex * V;
                }
            }
            // dT/dr = d/dr [exp(-eta S) F]: F' sum_e P_e - 2 r F sum_e eta_e P_e,
            // sum_e eta_e P_e = -SKP / kexp.  With Fa, Fb built from dfc / r,
            // d T / d r_ij / r_ij = Fa SP + F2 SKP (and the same for r_ik).
            const R F2S = F * R(2.0 / M<R>::kexp) * SKP;
            const R d_rjk = X.g3 ? (Fj * SP + rjk * F2S) : R(0);
            const R d_c = F * SV;
            // d cos / d R_ij = R_ik/(r_ij r_ik) - cos R_ij / r_ij^2 (and a <-> b), folded:
            // g_a = (d_rij / r_ij - d_c cos / r_ij^2) R_ij + (d_c / (r_ij r_ik)) R_ik + s_j R_jk
            const R sj = d_rjk * irjk;
            const R dcc = d_c * cs;
            const R al = Fa * SP + F2S - dcc * ira2;
            const R bl = Fb * SP + F2S - dcc * irb2;
            const R be = d_c * iab;
            const R gax = al * pa.x + be * pb.x + sj * jx;
            const R gay = al * pa.y + be * pb.y + sj * jy;
            const R gaz = al * pa.z + be * pb.z + sj * jz;
            const R gbx = bl * pb.x + be * pa.x - sj * jx;
            const R gby = bl * pb.y + be * pa.y - sj * jy;
            const R gbz = bl * pb.z + be * pa.z - sj * jz;
            lds_add(&S.DA[3 * ia], -(double)gax); lds_add(&S.DA[3 * ia + 1], -(double)gay);
            lds_add(&S.DA[3 * ia + 2], -(double)gaz);
            lds_add(&S.DA[3 * ib], -(double)gbx); lds_add(&S.DA[3 * ib + 1], -(double)gby);
            lds_add(&S.DA[3 * ib + 2], -(double)gbz);
        } else {  // MODE_TAN
            const R cax = pb.x * iab - cs * pa.x * ira2, cay = pb.y * iab - cs * pa.y * ira2,
                    caz = pb.z * iab - cs * pa.z * ira2;  // d cos / d R_ij
            const R cbx = pa.x * iab - cs * pb.x * irb2, cby = pa.y * iab - cs * pb.y * irb2,
                    cbz = pa.z * iab - cs * pb.z * irb2;  // d cos / d R_ik
            const R ax = S.D[ia], ay = S.D[cap + ia], az = S.D[2 * cap + ia];
            const R bx = S.D[ib], by = S.D[cap + ib], bz = S.D[2 * cap + ib];
            const R tij = (pa.x * ax + pa.y * ay + pa.z * az) * ira;
            const R tik = (pb.x * bx + pb.y * by + pb.z * bz) * irb;
            const R tjk = X.g3 ? (jx * (ax - bx) + jy * (ay - by) + jz * (az - bz)) * irjk : R(0);
            const R cdot = cax * ax + cay * ay + caz * az + cbx * bx + cby * by + cbz * bz;
            // d/dt [exp(-eta S) F] = exp(-eta S) (X0 - 2 eta X1)
            const R X0 = Fa * tij + Fb * tik + Fj * tjk;
            const R X1 = F * (ra * tij + rb * tik + rjk * tjk);
#pragma unroll
            for (int e = 0; e < PT_AG_ETA; ++e) {
                if (e < X.n_eta) {
                    const R ex = M<R>::ex2k(X.keta[e] * Ssum);
                    const R rdot = ex * (X0 + R(2.0 / M<R>::kexp) * X.keta[e] * X1);
                    const R radc = ex * F * cdot;
#pragma unroll
                    for (int q = 0; q < NA; ++q) acc[e][q] += rdot * ang[q] + radc * dang[q];
                }
            }
        }
    };

    // Pair enumeration, ordered so that the 64 lanes of one step touch
    // distinct neighbour slots (few LDS atomic conflicts in the adjoint):
    //  - cross segments: lanes run over the shorter segment q, the other
    //    index is the diagonal k = (q + d) mod n_long, d = 0 .. n_long-1;
    //  - one segment: k = (q + 1 + d) mod n for d < n/2 -- each unordered
    //    pair once (see below for the last diagonal of an even n).
    // Pair p = d * nrow + q; each lane advances p by 64 per step.
    const bool jrow = X.nj <= X.nk;
    const int nrow = X.same ? X.nj : (jrow ? X.nj : X.nk);
    const int ncol = X.same ? X.nj : (jrow ? X.nk : X.nj);
    const int brow = X.same ? X.bj : (jrow ? X.bj : X.bk);
    const int bcol = X.same ? X.bj : (jrow ? X.bk : X.bj);
    // one segment: h = n/2 diagonals d (k = q + 1 + d mod n); for even n the
    // last diagonal pairs q with q + n/2 and holds each pair twice, so only
    // q < n/2 is taken there
    const int h = X.same ? X.nj / 2 : ncol;
    const int npairs = nrow * h;
    if (npairs <= 0) return;
    const int half_d = (X.same && X.nj % 2 == 0) ? h - 1 : -1;   // the doubled diagonal
    const int half_q = X.nj / 2;
    const int d64 = uniform(kWave / nrow), r64 = uniform(kWave - d64 * nrow);
    int d = lane / nrow, q = lane - (lane / nrow) * nrow;
    // screen r_jk^2 < rc^2 (loosened by 1e-4 relative); the exact r_jk < rc
    // test is the cutoff's own (it is 0 beyond rc)
    const R rc2 = X.g3 ? X.rc * X.rc * R(1.0001) : R(INFINITY);
    const int sk = X.same ? 1 : 0;
    int qn = 0, base = 0;
    // one batch of up to 64 queued pairs through the body; the rest moves up
    auto drain = [&]() {
        wave_sync();
        const int take = qn < kWave ? qn : kWave;
        const unsigned e = lane < take ? S.queue[lane] : 0u;
        const int rem = qn - take;
        const unsigned mv0 = lane < rem ? S.queue[kWave + lane] : 0u;
        const unsigned mv1 = lane + kWave < rem ? S.queue[2 * kWave + lane] : 0u;
        wave_sync();
        if (lane < rem) S.queue[lane] = mv0;
        if (lane + kWave < rem) S.queue[kWave + lane] = mv1;
        qn = rem;
        if (lane < take) body((int)(e & 0xffffu), (int)(e >> 16));
    };
    // kDiag diagonals per lane and step: lane (D, q) screens d = kDiag D + u,
    // u = 0 .. kDiag - 1, which share the row slot and the index arithmetic;
    // the queue is drained whenever it holds a full batch, so it never holds
    // more than 63 + 64 entries.  (Measured on gfx950: 4 beats 2 and 8 in
    // FP32, and 2 in the FP64 adjoint by 2.5 %.)
    constexpr int kDiag = 4;
    const int hh = (h + kDiag - 1) / kDiag;
    const int npairs2 = nrow * hh;
    while (true) {
        if (base < npairs2) {
            const int ia = brow + q;
            const int d0 = kDiag * d;
            // the lane's diagonals d0 + u exist for u < nu; the doubled diagonal
            // of an even segment (the last one, half_d = h - 1) only for q < n/2
            int nu = min(max(h - d0, 0), kDiag);
            if (q >= half_q && half_d >= d0 && half_d < d0 + kDiag) nu = half_d - d0;
            if (base + lane >= npairs2) nu = 0;
            // column k = (q + sk + d0 + u) mod ncol: q + sk + d0 + u < 2 ncol for
            // every existing diagonal, so k0 + u wraps at most once; a lane with
            // no diagonal reads from the segment start (loads stay in the stage)
            int k0 = q + sk + d0;
            k0 = k0 >= ncol ? k0 - ncol : k0;
            k0 = nu > 0 ? k0 : 0;
            const V4<R>* pseg = S.P + bcol;
            const unsigned eb = (unsigned)ia | ((unsigned)bcol << 16);
#pragma unroll
            for (int u = 0; u < kDiag; ++u) {
                // (k0 + u) mod ncol as an unsigned minimum: k0 + u - ncol is
                // negative (huge) unless the column wraps; at most 2 past the
                // segment start for a lane beyond nu
                const unsigned kk = min((unsigned)(k0 + u), (unsigned)(k0 + u - ncol));
                const V4<R> pa = S.P[ia], pb = pseg[kk];
                R jx = pa.x - pb.x, jy = pa.y - pb.y, jz = pa.z - pb.z;
                if (X.mic) mic3(*X.box, jx, jy, jz);
                const R r2 = jx * jx + jy * jy + jz * jz;
                const bool live = (u < nu) & (r2 > R(0)) & (r2 < rc2);
                // (a ballot per compare: no lane mask rebuilt from the bool)
                const unsigned long long b = __builtin_amdgcn_ballot_w64(u < nu) &
                                             __builtin_amdgcn_ballot_w64(r2 > R(0)) &
                                             __builtin_amdgcn_ballot_w64(r2 < rc2);
                if (live) S.queue[qn + lanes_below(b)] = eb + (kk << 16);
                qn += __popcll(b);
                if (qn >= kWave) drain();
            }
            base += kWave;
            q += r64;
            d += d64;
            if (q >= nrow) { q -= nrow; ++d; }
        }
        const bool last = base >= npairs2;
        while (qn >= kWave || (last && qn > 0)) drain();
        if (last && qn == 0) break;
    }
    if (kPreOut && MODE != MODE_ADJ) {
#pragma unroll
        for (int e = 0; e < PT_AG_ETA; ++e)
#pragma unroll
            for (int q = 0; q < PT_AG_ANG; ++q) acc[e][q] *= X.pre[q];
    }
    wave_sync();
}

// 1 when the group's terms are exactly zeta {1, 4} x lambda {+1, -1} (any
// order); qmap[k] then gives the term of canonical slot k = 2 [zeta = 4] +
// [lambda = -1].
__device__ __forceinline__ int z14_order(const pt_angular_group* G, int (&qmap)[PT_AG_ANG]) {
    if (uniform(G->n_ang) != 4) return 0;
    int seen = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int iz = uniform(G->izeta[q]);
        const double lam = G->lam[q];
        if (!((iz == 1 || iz == 4) && (lam == 1.0 || lam == -1.0))) return 0;
        const int k = (iz == 4 ? 2 : 0) + (lam < 0.0 ? 1 : 0);
        seen |= 1 << k;
        qmap[k] = q;
    }
    return seen == 15 ? 1 : 0;
}

template <typename R, int MODE, int KIND>
__device__ __forceinline__ void acsf_centre(const KArgs& a, char* smem) {
    if (sizeof(R) == 8) exp2_tab_load();
    const int lane = wave_lane();
    const int wv = uniform(threadIdx.x >> 6);
    const int wpb = blockDim.x >> 6;
    const int c = uniform(xcd_block() * wpb + wv);
    if (c >= a.n_centers) return;
    const int cap = a.cap;
    Stage<R> S = carve<R, MODE>(smem + (size_t)wv * stage_bytes<R, MODE>(cap), cap);
    const pt_sf_table* __restrict__ tab = a.table;
    const int i = uniform(a.centers ? a.centers[c] : a.center_begin + c);
    const double4 pi = reinterpret_cast<const double4*>(a.pos4)[i];
    const long long off = a.row_off[i];
    const int len = (int)(a.row_off[i + 1] - off);
    if (len > cap) {
        // the row does not fit the staging slice sized by the caller's max_row:
        // stage nothing and report the centre as NaN (outputs, f_self, the row's edges)
        const R nan = __builtin_nan("");
        if (MODE != MODE_ADJ) {
            R* out = (R*)a.out;
            for (int q = lane; q < a.ld; q += kWave) out[(size_t)c * a.ld + q] = nan;
        } else {
            if (a.edge_grad) {
                R* eg = (R*)a.edge_grad;
                for (int t = lane; t < len; t += kWave) {
                    const long long e = a.rev ? (long long)a.rev[off + t] : off + t;
                    if (e >= 0) { eg[3 * e] = nan; eg[3 * e + 1] = nan; eg[3 * e + 2] = nan; }
                }
            }
            if (a.f_self && lane < 3) a.f_self[3 * c + lane] = (double)nan;
        }
        return;
    }
    const double rcmax = tab->rc_max;
    const R tanh_pre = (R)tab->tanh_pre;
    // |R_ij - R_ik| <= 2 rc_max: the image shift can only act when L/2 < 2 rc_max
    const double lim = 4.0 * rcmax * (1.0 + 1e-9);
    const bool mic_jk = a.box.periodic && (!(a.box.L[0] > lim) || !(a.box.L[1] > lim) || !(a.box.L[2] > lim));

    double dix = 0, diy = 0, diz = 0;
    if (MODE == MODE_TAN) {
        dix = a.delta[3 * i]; diy = a.delta[3 * i + 1]; diz = a.delta[3 * i + 2];
    }

    // segment counts are needed up to one past the largest type the table uses
    int tmax = 0;
    for (int g = 0; g < uniform(tab->n_rg); ++g) tmax = max(tmax, uniform(tab->rg[g].elem));
    for (int g = 0; g < uniform(tab->n_ag); ++g)
        tmax = max(tmax, max(uniform(tab->ag[g].ej), uniform(tab->ag[g].ek)));
    const int nseg = min(tmax + 1, kSegTypes);

    // ---- stage the row ----------------------------------------------------
    int m = 0;
    int seg = 0;   // lane q: staged neighbours with type < q
    for (int t0 = 0; t0 < len; t0 += kWave) {
        const int t = t0 + lane;
        bool keep = false;
        int j = 0;
        double dx = 0, dy = 0, dz = 0, r = 0, tj = 0;
        if (t < len) {
            j = a.nbr[off + t];
            const double4 pj = reinterpret_cast<const double4*>(a.pos4)[j];
            dx = pi.x - pj.x; dy = pi.y - pj.y; dz = pi.z - pj.z;
            r = exact_dist(dx, dy, dz, a.box);
            tj = pj.w;
            keep = (r <= rcmax) && (r > 0.0);
        }
        const unsigned long long bal = __ballot(keep);
        if (MODE == MODE_ADJ && lane == 0) S.keep[t0 / kWave] = bal;
        if (keep) {
            const int s = m + lanes_below(bal);
            V4<R> pv, qv;
            pv.x = (R)dx; pv.y = (R)dy; pv.z = (R)dz;
            pv.w = (R)(1.0 / r);
            qv.x = (R)r; qv.y = (R)r * (R)r; qv.z = R(0); qv.w = R(0);
            S.P[s] = pv;
            if constexpr ((MODE == MODE_ADJ)) {
                // (fc, dfc) are written per group
            } else if constexpr (MODE == MODE_FWD) {
                V2<R> uv;
                uv.x = qv.x; uv.y = R(0);
                S.Q2[s] = uv;
            } else {
                S.Q[s] = qv;
            }
            if (MODE == MODE_ADJ) { S.DA[3 * s] = 0.0; S.DA[3 * s + 1] = 0.0; S.DA[3 * s + 2] = 0.0; }
            if (MODE == MODE_TAN) {
                double ex = dix, ey = diy, ez = diz;
                if (!a.center_only) {
                    ex -= a.delta[3 * j]; ey -= a.delta[3 * j + 1]; ez -= a.delta[3 * j + 2];
                }
                S.D[s] = (R)ex; S.D[cap + s] = (R)ey; S.D[2 * cap + s] = (R)ez;
            }
        } else if (MODE == MODE_ADJ && t < len && a.edge_grad) {
            // (a reverse index of -1 would mean an asymmetric list: no slot to write)
            const long long e = a.rev ? (long long)a.rev[off + t] : off + t;
            if (e >= 0) {
                R* eg = (R*)a.edge_grad + 3 * e;
                eg[0] = R(0); eg[1] = R(0); eg[2] = R(0);
            }
        }
        m += __popcll(bal);
        const int tk = (int)tj;
        for (int q = 1; q <= nseg; ++q) {
            const int cq = __popcll(__ballot(keep && tk < q));
            seg += lane == q ? cq : 0;
        }
    }
    m = uniform(m);
    wave_sync();

    R* out = (R*)a.out;
    const R* w = (const R*)a.w;

    // ---- radial groups -----------------------------------------------------
    const int n_rg = uniform(tab->n_rg);
    for (int g = 0; g < n_rg; ++g) {
        const pt_radial_group* G = &tab->rg[g];
        const int el = uniform(G->elem);
        const int b = seg_below(seg, el), e = seg_below(seg, el + 1);
        const int nfun = uniform(G->n);
        const int cut = uniform(G->cut);
        const R rc = (R)G->rc, irc = (R)(1.0 / G->rc);
        R wgt[PT_RG_MAX], acc[PT_RG_MAX], keta[PT_RG_MAX], eta[PT_RG_MAX], rs[PT_RG_MAX];
#pragma unroll
        for (int q = 0; q < PT_RG_MAX; ++q) {
            acc[q] = R(0);
            eta[q] = (q < nfun) ? (R)G->eta[q] : R(0);
            keta[q] = -eta[q] * M<R>::kexp;
            rs[q] = (q < nfun) ? (R)G->rs[q] : R(0);
            wgt[q] = (MODE == MODE_ADJ && q < nfun) ? w[(size_t)c * a.ldw + G->col[q]] : R(0);
        }
        for (int s = b + lane; s < e; s += kWave) {
            R r;
            if constexpr ((MODE == MODE_ADJ)) r = M<R>::rcp(S.P[s].w);
            else if constexpr (MODE == MODE_FWD) r = S.Q2[s].x;
            else r = S.Q[s].x;
            R f, df;
            cutoff<R>(cut, rc, irc, tanh_pre, r, f, df);
            R rdot = R(0);
            V4<R> pv;
            if (MODE != MODE_FWD) pv = S.P[s];
            if (MODE == MODE_TAN)
                rdot = (pv.x * S.D[s] + pv.y * S.D[cap + s] + pv.z * S.D[2 * cap + s]) * pv.w;
            R dsum = R(0);
#pragma unroll
            for (int q = 0; q < PT_RG_MAX; ++q) {
                if (q < nfun) {
                    const R d = r - rs[q];
                    const R ex = M<R>::ex2k(keta[q] * d * d);
                    if (MODE == MODE_FWD) acc[q] += ex * f;
                    if (MODE == MODE_ADJ) dsum += wgt[q] * ex * (df - R(2) * eta[q] * d * f);
                    if (MODE == MODE_TAN) acc[q] += ex * (df - R(2) * eta[q] * d * f) * rdot;
                }
            }
            if (MODE == MODE_ADJ) {
                const R sc = dsum * pv.w;
                const R gx = sc * pv.x, gy = sc * pv.y, gz = sc * pv.z;
                S.DA[3 * s] -= (double)gx; S.DA[3 * s + 1] -= (double)gy; S.DA[3 * s + 2] -= (double)gz;
            }
        }
        if (MODE != MODE_ADJ) {
#pragma unroll
            for (int q = 0; q < PT_RG_MAX; ++q)
                if (q < nfun) {
                    const R v = wave_sum(acc[q]);
                    if (lane == 0) out[(size_t)c * a.ld + G->col[q]] = v;
                }
        }
        wave_sync();
    }

    // ---- angular groups ----------------------------------------------------
    const int n_ag = uniform(tab->n_ag);
    bool bad = false;   // a group outside this kernel's table class
    for (int g = 0; g < n_ag; ++g) {
        const pt_angular_group* G = &tab->ag[g];
        AngGroup<R> X;
        const int ej = uniform(G->ej), ek = uniform(G->ek);
        X.bj = seg_below(seg, ej);
        X.nj = seg_below(seg, ej + 1) - X.bj;
        X.bk = seg_below(seg, ek);
        X.nk = seg_below(seg, ek + 1) - X.bk;
        X.same = ej == ek;
        X.g3 = uniform(G->kind) == 3;
        X.mic = mic_jk;
        X.n_eta = uniform(G->n_eta);
        X.n_ang = uniform(G->n_ang);
        X.cut = uniform(G->cut);
        X.rc = (R)G->rc;
        X.irc = (R)(1.0 / G->rc);
        X.cpre = X.cut == PT_CUT_TANH ? tanh_pre : R(1);
        X.tanh_pre = tanh_pre;
        X.box = &a.box;
        // qmap[k]: the table term evaluated in slot k (the identity, or the
        // canonical order of a ZETA {1, 4} x LAMBDA {+1, -1} group)
        int qmap[PT_AG_ANG];
        const int z14 = (KIND == PT_ACSF_TANH_Z14) ? z14_order(G, qmap) : 0;
        if (!z14) {
#pragma unroll
            for (int q = 0; q < PT_AG_ANG; ++q) qmap[q] = q;
        }
        R W[PT_AG_ETA][PT_AG_ANG], acc[PT_AG_ETA][PT_AG_ANG];
#pragma unroll
        for (int e = 0; e < PT_AG_ETA; ++e) {
            X.keta[e] = (e < X.n_eta) ? (R)(-G->eta[e] * M<R>::kexp) : R(0);
#pragma unroll
            for (int q = 0; q < PT_AG_ANG; ++q) {
                acc[e][q] = R(0);
                const int col = G->col[e * PT_AG_ANG + qmap[q]];
                W[e][q] = (MODE == MODE_ADJ && e < X.n_eta && q < X.n_ang && col >= 0)
                              ? w[(size_t)c * a.ldw + col]
                              : R(0);
            }
        }
        bool z1 = true, p4 = true;
#pragma unroll
        for (int q = 0; q < PT_AG_ANG; ++q) {
            const bool on = q < X.n_ang;
            const int tq = qmap[q];
            X.lam[q] = on ? (R)G->lam[tq] : R(0);
            X.zeta[q] = on ? (R)G->zeta[tq] : R(1);
            X.pre[q] = on ? (R)G->pre[tq] : R(0);
            X.izeta[q] = on ? uniform(G->izeta[tq]) : 1;
            X.zl[q] = on ? (R)(G->zeta[tq] * G->lam[tq]) : R(0);
            z1 = z1 && X.izeta[q] == 1;
            p4 = p4 && (X.izeta[q] == 1 || X.izeta[q] == 2 || X.izeta[q] == 4);
        }
        const int ck = (X.cut == PT_CUT_TANH || X.cut == PT_CUT_TANHU) ? CUTK_TANH
                       : (X.cut == PT_CUT_COS ? CUTK_COS : CUTK_GEN);
        const int ak = z14 ? ANGK_Z14
                           : (z1 ? (X.n_ang <= 2 ? ANGK_Z1N2 : ANGK_Z1) : (p4 ? ANGK_P4 : ANGK_GEN));
        wave_sync();
        if constexpr (KIND == PT_ACSF_TANH_Z14) {
            if (ck == CUTK_TANH && ak == ANGK_Z14)
                angular_group<R, MODE, CUTK_TANH, ANGK_Z14>(S, X, W, acc, cap);
            else
                bad = true;
        } else if constexpr (KIND == PT_ACSF_TANH_P4) {
            if (ck == CUTK_TANH && ak != ANGK_GEN)
                angular_group<R, MODE, CUTK_TANH, ANGK_P4>(S, X, W, acc, cap);
            else
                bad = true;
        } else if constexpr (KIND == PT_ACSF_TANH_Z1N2) {
            if (ck == CUTK_TANH && ak == ANGK_Z1N2)
                angular_group<R, MODE, CUTK_TANH, ANGK_Z1N2>(S, X, W, acc, cap);
            else
                bad = true;
        } else if (MODE == MODE_TAN || (sizeof(R) == 8 && ck != CUTK_TANH)) {
            angular_group<R, MODE, CUTK_GEN, ANGK_GEN>(S, X, W, acc, cap);
        } else if (sizeof(R) == 8) {
            if (ak == ANGK_GEN) angular_group<R, MODE, CUTK_TANH, ANGK_GEN>(S, X, W, acc, cap);
            else angular_group<R, MODE, CUTK_TANH, ANGK_P4>(S, X, W, acc, cap);
        } else if (ck == CUTK_TANH) {
            if (ak == ANGK_Z1N2) angular_group<R, MODE, CUTK_TANH, ANGK_Z1N2>(S, X, W, acc, cap);
            else if (ak == ANGK_Z1) angular_group<R, MODE, CUTK_TANH, ANGK_Z1>(S, X, W, acc, cap);
            else if (ak == ANGK_P4) angular_group<R, MODE, CUTK_TANH, ANGK_P4>(S, X, W, acc, cap);
            else angular_group<R, MODE, CUTK_TANH, ANGK_GEN>(S, X, W, acc, cap);
        } else if (ck == CUTK_COS) {
            if (ak == ANGK_Z1N2) angular_group<R, MODE, CUTK_COS, ANGK_Z1N2>(S, X, W, acc, cap);
            else if (ak == ANGK_Z1) angular_group<R, MODE, CUTK_COS, ANGK_Z1>(S, X, W, acc, cap);
            else if (ak == ANGK_P4) angular_group<R, MODE, CUTK_COS, ANGK_P4>(S, X, W, acc, cap);
            else angular_group<R, MODE, CUTK_COS, ANGK_GEN>(S, X, W, acc, cap);
        } else {
            angular_group<R, MODE, CUTK_GEN, ANGK_GEN>(S, X, W, acc, cap);
        }

        if (MODE != MODE_ADJ) {
            // reduce only the group's cells that own a column (uniform branches
            // keep all lanes active)
#pragma unroll
            for (int e = 0; e < PT_AG_ETA; ++e)
#pragma unroll
                for (int q = 0; q < PT_AG_ANG; ++q) {
                    const int col = (e < X.n_eta && q < X.n_ang) ? uniform(G->col[e * PT_AG_ANG + qmap[q]]) : -1;
                    if (col >= 0) {
                        const R v = wave_sum(acc[e][q]);
                        if (lane == 0) out[(size_t)c * a.ld + col] = v;
                    }
                }
        }
        wave_sync();
    }

    const R nan = __builtin_nan("");
    if (MODE != MODE_ADJ && bad) {
        for (int q = lane; q < a.ld; q += kWave) out[(size_t)c * a.ld + q] = nan;
    }
    if (MODE == MODE_ADJ) {
        if (a.edge_grad) {
            R* eg = (R*)a.edge_grad;
            // the staging ballots give each kept row entry its staged slot
            int ms = 0;
            for (int t0 = 0; t0 < len; t0 += kWave) {
                const unsigned long long bal = S.keep[t0 / kWave];
                if ((bal >> lane) & 1ull) {
                    const int s = ms + lanes_below(bal);
                    const long long e = a.rev ? (long long)a.rev[off + t0 + lane] : off + t0 + lane;
                    if (e >= 0) {
                        eg[3 * e + 0] = bad ? nan : (R)S.DA[3 * s];
                        eg[3 * e + 1] = bad ? nan : (R)S.DA[3 * s + 1];
                        eg[3 * e + 2] = bad ? nan : (R)S.DA[3 * s + 2];
                    }
                }
                ms += __popcll(bal);
            }
        }
        if (a.f_self) {
            // G_i depends on r_i only through R_ij = r_i - r_j, so dE_i/dr_i = -sum_j dE_i/dr_j
            // and F_self = -dE_i/dr_i = sum_j dE_i/dr_j, the per-edge accumulators
            double fx = 0.0, fy = 0.0, fz = 0.0;
            for (int s = lane; s < m; s += kWave) {
                fx += S.DA[3 * s];
                fy += S.DA[3 * s + 1];
                fz += S.DA[3 * s + 2];
            }
            const double sx = wave_sum(fx), sy = wave_sum(fy), sz = wave_sum(fz);
            if (lane == 0) {
                a.f_self[3 * c + 0] = bad ? (double)nan : sx;
                a.f_self[3 * c + 1] = bad ? (double)nan : sy;
                a.f_self[3 * c + 2] = bad ? (double)nan : sz;
            }
        }
    }
}

// Entry kernels over the same body: the others ask the compiler for four
// (<= 128 VGPRs) or eight (<= 64 VGPRs) resident waves per SIMD; the FP64
// forwards of the specialised table classes and the FP32 forwards of the
// zeta {1, 4} and zeta = 1 classes fit those without spills.
template <typename R, int MODE, int KIND>
__global__ __launch_bounds__(256) void k_acsf(KArgs a) {
    extern __shared__ __align__(32) char smem[];
    acsf_centre<R, MODE, KIND>(a, smem);
}

template <typename R, int MODE, int KIND>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_acsf_w4(KArgs a) {
    extern __shared__ __align__(32) char smem[];
    acsf_centre<R, MODE, KIND>(a, smem);
}

template <typename R, int MODE, int KIND>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void k_acsf_w8(KArgs a) {
    extern __shared__ __align__(32) char smem[];
    acsf_centre<R, MODE, KIND>(a, smem);
}

template <typename R, int MODE, int KIND>
int launch(KArgs a, hipStream_t stream) {
    if (a.n_centers <= 0) return PT_OK;
    a.cap = (std::max(a.cap, 1) + 3) / 4 * 4;
    if (a.cap > 65535) {
        set_error("acsf: neighbour rows longer than 65535 entries are not supported");
        return PT_ERR_UNSUPPORTED;
    }
    const size_t per_wave = stage_bytes<R, MODE>(a.cap);
    const size_t lds_budget = lds_limit();
    if (per_wave > lds_budget) {
        set_error("acsf: neighbour row of %d entries exceeds the LDS budget", a.cap);
        return PT_ERR_UNSUPPORTED;
    }
    // waves per workgroup: the most resident waves per CU under the LDS limit,
    // ties to the smaller workgroup (a workgroup's LDS is held until its
    // slowest wave ends; one-wave groups ran the FP32 adjoint 2.5 % faster
    // than two-wave groups at the same 18 resident waves)
    const size_t cu = std::max(lds_per_cu(), lds_budget);
    int wpb = 1;
    size_t best = 0;
    for (int w = 1; w <= 4; w *= 2) {
        if ((size_t)w * per_wave > lds_budget) break;
        const size_t waves = (size_t)w * (cu / ((size_t)w * per_wave));
        if (waves > best) { best = waves; wpb = w; }
    }
    const int blocks = (a.n_centers + wpb - 1) / wpb;
    if constexpr (sizeof(R) == 8 && MODE == MODE_FWD && KIND != PT_ACSF_GENERIC)
        hipLaunchKernelGGL((k_acsf_w4<R, MODE, KIND>), dim3(blocks), dim3(64 * wpb), wpb * per_wave, stream, a);
    else if constexpr (sizeof(R) == 4 && MODE == MODE_FWD && (KIND == PT_ACSF_TANH_Z14 || KIND == PT_ACSF_TANH_Z1N2))
        hipLaunchKernelGGL((k_acsf_w8<R, MODE, KIND>), dim3(blocks), dim3(64 * wpb), wpb * per_wave, stream, a);
    else
        hipLaunchKernelGGL((k_acsf<R, MODE, KIND>), dim3(blocks), dim3(64 * wpb), wpb * per_wave, stream, a);
    return launch_check("k_acsf");
}

template <typename R>
int run_kind(int mode, int kind, KArgs& a, hipStream_t st) {
    if (mode == MODE_TAN) {
        switch (kind) {
        case PT_ACSF_GENERIC: return launch<R, MODE_TAN, PT_ACSF_GENERIC>(a, st);
        case PT_ACSF_TANH_P4: return launch<R, MODE_TAN, PT_ACSF_TANH_P4>(a, st);
        case PT_ACSF_TANH_Z1N2: return launch<R, MODE_TAN, PT_ACSF_TANH_Z1N2>(a, st);
        case PT_ACSF_TANH_Z14: return launch<R, MODE_TAN, PT_ACSF_TANH_Z14>(a, st);
        }
        set_error("acsf: unknown kernel kind %d", kind);
        return PT_ERR_ARG;
    }
    switch (kind) {
    case PT_ACSF_GENERIC:
        return mode == MODE_FWD ? launch<R, MODE_FWD, PT_ACSF_GENERIC>(a, st) : launch<R, MODE_ADJ, PT_ACSF_GENERIC>(a, st);
    case PT_ACSF_TANH_P4:
        return mode == MODE_FWD ? launch<R, MODE_FWD, PT_ACSF_TANH_P4>(a, st) : launch<R, MODE_ADJ, PT_ACSF_TANH_P4>(a, st);
    case PT_ACSF_TANH_Z1N2:
        return mode == MODE_FWD ? launch<R, MODE_FWD, PT_ACSF_TANH_Z1N2>(a, st)
                                : launch<R, MODE_ADJ, PT_ACSF_TANH_Z1N2>(a, st);
    case PT_ACSF_TANH_Z14:
        return mode == MODE_FWD ? launch<R, MODE_FWD, PT_ACSF_TANH_Z14>(a, st)
                                : launch<R, MODE_ADJ, PT_ACSF_TANH_Z14>(a, st);
    }
    set_error("acsf: unknown kernel kind %d", kind);
    return PT_ERR_ARG;
}

int run(int precision_kind, int mode, KArgs& a, hipStream_t st) {
    const int precision = precision_kind & ((1 << PT_ACSF_KIND_SHIFT) - 1);
    const int kind = precision_kind >> PT_ACSF_KIND_SHIFT;
    if (precision == PT_F64) return run_kind<double>(mode, kind, a, st);
    if (precision == PT_F32) return run_kind<float>(mode, kind, a, st);
    set_error("acsf: unknown precision %d", precision);
    return PT_ERR_ARG;
}

KArgs common_args(const double* pos4, const double* box, const long long* row_off, const int* nbr,
                  int max_row, int n_centers, int center_begin, const int* centers,
                  const pt_sf_table* table) {
    KArgs a{};
    a.pos4 = pos4;
    a.box.periodic = box != nullptr;
    for (int d = 0; d < 3; ++d) {
        a.box.L[d] = box ? box[d] : 0.0;
        a.box.half[d] = box ? 0.5 * box[d] : 0.0;
    }
    a.row_off = row_off;
    a.nbr = nbr;
    a.n_centers = n_centers;
    a.center_begin = center_begin;
    a.centers = centers;
    a.table = table;
    a.cap = max_row;
    return a;
}

}  // namespace
}  // namespace pt

using namespace pt;

extern "C" int pt_acsf_forward(int precision, const double* pos4, const double* box,
                               const long long* row_off, const int* nbr, int max_row,
                               int n_centers, int center_begin, const int* centers,
                               const pt_sf_table* table, void* G, int ldg, pt_stream_t stream) {
    PT_REQUIRE(pos4 && row_off && table && (G || n_centers == 0), "pt_acsf_forward: null argument");
    PT_REQUIRE(n_centers >= 0 && ldg >= 0, "pt_acsf_forward: bad sizes");
    hipStream_t st = (hipStream_t)stream;
    KArgs a = common_args(pos4, box, row_off, nbr, max_row, n_centers, center_begin, centers, table);
    a.out = G;
    a.ld = ldg;
    return run(precision, MODE_FWD, a, st);
}

extern "C" int pt_acsf_adjoint(int precision, const double* pos4, const double* box,
                               const long long* row_off, const int* nbr, int max_row,
                               int n_centers, int center_begin, const int* centers,
                               const pt_sf_table* table, const void* w, int ldw, double* f_self,
                               void* edge_grad, pt_stream_t stream) {
    PT_REQUIRE(pos4 && row_off && table && (w || n_centers == 0), "pt_acsf_adjoint: null argument");
    hipStream_t st = (hipStream_t)stream;
    KArgs a = common_args(pos4, box, row_off, nbr, max_row, n_centers, center_begin, centers, table);
    a.w = w;
    a.ldw = ldw;
    a.f_self = f_self;
    a.edge_grad = edge_grad;
    return run(precision, MODE_ADJ, a, st);
}

extern "C" int pt_acsf_adjoint_transposed(int precision, const double* pos4, const double* box,
                                          const long long* row_off, const int* nbr, const int* rev, int max_row,
                                          int n_centers, int center_begin, const int* centers,
                                          const pt_sf_table* table, const void* w, int ldw, double* f_self,
                                          void* edge_grad_t, pt_stream_t stream) {
    PT_REQUIRE(pos4 && row_off && rev && table && (w || n_centers == 0),
               "pt_acsf_adjoint_transposed: null argument");
    hipStream_t st = (hipStream_t)stream;
    KArgs a = common_args(pos4, box, row_off, nbr, max_row, n_centers, center_begin, centers, table);
    a.w = w;
    a.ldw = ldw;
    a.f_self = f_self;
    a.edge_grad = edge_grad_t;
    a.rev = rev;
    return run(precision, MODE_ADJ, a, st);
}

extern "C" int pt_acsf_tangent(int precision, const double* pos4, const double* box,
                               const long long* row_off, const int* nbr, int max_row,
                               int n_centers, int center_begin, const int* centers,
                               const pt_sf_table* table, const double* delta, int center_only,
                               void* T, int ldt, pt_stream_t stream) {
    PT_REQUIRE(pos4 && row_off && table && delta && (T || n_centers == 0),
               "pt_acsf_tangent: null argument");
    hipStream_t st = (hipStream_t)stream;
    KArgs a = common_args(pos4, box, row_off, nbr, max_row, n_centers, center_begin, centers, table);
    a.out = T;
    a.ld = ldt;
    a.delta = delta;
    a.center_only = center_only;
    return run(precision, MODE_TAN, a, st);
}

extern "C" int pt_sf_table_kind(const pt_sf_table* t) {
    if (!t) return PT_ACSF_GENERIC;
    if (t->n_ag <= 0) return PT_ACSF_GENERIC;
    bool p4 = true, z1n2 = true, z14 = true;
    for (int g = 0; g < t->n_ag && g < PT_MAX_GROUPS; ++g) {
        const pt_angular_group& G = t->ag[g];
        const bool tanh = G.cut == PT_CUT_TANH || G.cut == PT_CUT_TANHU;
        bool allz1 = true, allp4 = true;
        int seen = 0;
        for (int q = 0; q < G.n_ang; ++q) {
            allz1 = allz1 && G.izeta[q] == 1;
            allp4 = allp4 && (G.izeta[q] == 1 || G.izeta[q] == 2 || G.izeta[q] == 4);
            // the kernel's test (z14_order): zeta 1 or 4, lambda exactly +1 or -1, each pair once
            if ((G.izeta[q] == 1 || G.izeta[q] == 4) && (G.lam[q] == 1.0 || G.lam[q] == -1.0))
                seen |= 1 << ((G.izeta[q] == 4 ? 2 : 0) + (G.lam[q] < 0.0 ? 1 : 0));
            else
                seen |= 1 << 4;
        }
        p4 = p4 && tanh && allp4;
        z1n2 = z1n2 && tanh && allz1 && G.n_ang <= 2;
        z14 = z14 && tanh && G.n_ang == 4 && seen == 15;
    }
    return z1n2 ? PT_ACSF_TANH_Z1N2 : (z14 ? PT_ACSF_TANH_Z14 : (p4 ? PT_ACSF_TANH_P4 : PT_ACSF_GENERIC));
}
