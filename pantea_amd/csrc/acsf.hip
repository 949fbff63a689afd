// (b) ACSF descriptors, their adjoint and their tangent on gfx950.
//
// One wavefront (64 lanes) per centre.  The centre's neighbour row is staged
// in LDS (relative vectors R_ij = r_i - r_j under the minimum image, r_ij,
// 1/r_ij, r_ij^2, type and row slot), compacted to r_ij <= rc_max and
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
template <typename R> struct M;
template <> struct M<double> {
    static __device__ __forceinline__ double ex(double x) { return exp(x); }
    static __device__ __forceinline__ double th(double x) { return tanh(x); }
    static __device__ __forceinline__ double co(double x) { return cos(x); }
    static __device__ __forceinline__ double si(double x) { return sin(x); }
    static __device__ __forceinline__ double sq(double x) { return sqrt(x); }
    static __device__ __forceinline__ double pw(double b, double z) { return pow(b, z); }
    static __device__ __forceinline__ double rcp(double x) { return 1.0 / x; }
};
template <> struct M<float> {
    static __device__ __forceinline__ float ex(float x) { return __expf(x); }
    static __device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
    static __device__ __forceinline__ float th(float x) {
        // tanh(u) = (1 - e^{-2u}) / (1 + e^{-2u}); used for u in [0, 1]
        const float e = __expf(-2.0f * x);
        return (1.0f - e) * rcp(1.0f + e);
    }
    static __device__ __forceinline__ float co(float x) { return __cosf(x); }
    static __device__ __forceinline__ float si(float x) { return __sinf(x); }
    static __device__ __forceinline__ float sq(float x) { return __builtin_sqrtf(x); }
    static __device__ __forceinline__ float pw(float b, float z) { return powf(b, z); }
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

// Live-pair queue of one wavefront: (a | b << 16) indices of pairs that
// pass the r_jk screen.
constexpr int kQueue = 256;

template <typename R>
struct Stage {
    R* rx; R* ry; R* rz; R* r; R* ir; R* r2;
    R* f; R* df;   // cutoff of the current group and its derivative
    R* D;          // [3][cap] TAN: delta_i - delta_j
    double* DA;    // [3][cap] ADJ: per-edge accumulators (FP64: fast LDS atomics)
    int* el; int* slot;
    unsigned* queue;
};

template <typename R, int MODE>
__host__ __device__ constexpr size_t stage_bytes_per_nbr() {
    return sizeof(double) * (MODE == MODE_ADJ ? 3 : 0) + sizeof(R) * (8 + (MODE == MODE_TAN ? 3 : 0)) +
           2 * sizeof(int);
}

template <typename R, int MODE>
__host__ __device__ constexpr size_t stage_bytes(int cap) {
    return (size_t)cap * stage_bytes_per_nbr<R, MODE>() + kQueue * sizeof(unsigned);
}

template <typename R, int MODE>
__device__ __forceinline__ Stage<R> carve(char* base, int cap) {
    Stage<R> s;
    double* dp = reinterpret_cast<double*>(base);
    if (MODE == MODE_ADJ) { s.DA = dp; dp += 3 * cap; } else { s.DA = nullptr; }
    R* p = reinterpret_cast<R*>(dp);
    s.rx = p; p += cap;
    s.ry = p; p += cap;
    s.rz = p; p += cap;
    s.r = p; p += cap;
    s.ir = p; p += cap;
    s.r2 = p; p += cap;
    s.f = p; p += cap;
    s.df = p; p += cap;
    if (MODE == MODE_TAN) { s.D = p; p += 3 * cap; } else { s.D = nullptr; }
    int* q = reinterpret_cast<int*>(p);
    s.el = q; q += cap;
    s.slot = q; q += cap;
    s.queue = reinterpret_cast<unsigned*>(q);
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

// Number of staged neighbours with type < t (rows are type-sorted).
template <typename R>
__device__ __forceinline__ int count_below(const Stage<R>& S, int m, int t) {
    int cnt = 0;
    for (int s0 = 0; s0 < m; s0 += kWave) {
        const int s = s0 + wave_lane();
        cnt += __popcll(__ballot(s < m && S.el[s] < t));
    }
    return uniform(cnt);
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
    const double* delta;  // TAN (n x 3, internal order)
    int center_only;
    int cap;
};

template <typename R, int MODE>
__global__ __launch_bounds__(256) void k_acsf(KArgs a) {
    extern __shared__ __align__(16) char smem[];
    const int lane = wave_lane();
    const int wv = threadIdx.x >> 6;
    const int wpb = blockDim.x >> 6;
    const int c = uniform(blockIdx.x * wpb + wv);
    if (c >= a.n_centers) return;
    const int cap = a.cap;
    Stage<R> S = carve<R, MODE>(smem + (size_t)wv * stage_bytes<R, MODE>(cap), cap);
    const pt_sf_table* __restrict__ tab = a.table;
    const int i = uniform(a.centers ? a.centers[c] : a.center_begin + c);
    const double4 pi = reinterpret_cast<const double4*>(a.pos4)[i];
    const long long off = a.row_off[i];
    const int len = (int)(a.row_off[i + 1] - off);
    const double rcmax = tab->rc_max;
    const R tanh_pre = (R)tab->tanh_pre;
    const R Lx = (R)a.box.L[0], Ly = (R)a.box.L[1], Lz = (R)a.box.L[2];
    const R Hx = (R)a.box.half[0], Hy = (R)a.box.half[1], Hz = (R)a.box.half[2];
    // |R_ij - R_ik| <= 2 rc_max: the image shift can only act when L/2 < 2 rc_max
    const double lim = 4.0 * rcmax * (1.0 + 1e-9);
    const bool mic_jk = a.box.periodic && (!(a.box.L[0] > lim) || !(a.box.L[1] > lim) || !(a.box.L[2] > lim));

    double dix = 0, diy = 0, diz = 0;
    if (MODE == MODE_TAN) {
        dix = a.delta[3 * i]; diy = a.delta[3 * i + 1]; diz = a.delta[3 * i + 2];
    }

    // ---- stage the row ----------------------------------------------------
    int m = 0;
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
        if (keep) {
            const int s = m + lanes_below(bal);
            S.rx[s] = (R)dx; S.ry[s] = (R)dy; S.rz[s] = (R)dz; S.r[s] = (R)r;
            S.ir[s] = (R)(1.0 / r);
            S.r2[s] = (R)r * (R)r;
            S.el[s] = (int)tj;
            S.slot[s] = t;
            if (MODE == MODE_ADJ) { S.DA[s] = 0.0; S.DA[cap + s] = 0.0; S.DA[2 * cap + s] = 0.0; }
            if (MODE == MODE_TAN) {
                double ex = dix, ey = diy, ez = diz;
                if (!a.center_only) {
                    ex -= a.delta[3 * j]; ey -= a.delta[3 * j + 1]; ez -= a.delta[3 * j + 2];
                }
                S.D[s] = (R)ex; S.D[cap + s] = (R)ey; S.D[2 * cap + s] = (R)ez;
            }
        } else if (MODE == MODE_ADJ && t < len && a.edge_grad) {
            R* eg = (R*)a.edge_grad + 3 * (off + t);
            eg[0] = R(0); eg[1] = R(0); eg[2] = R(0);
        }
        m += __popcll(bal);
    }
    m = uniform(m);
    wave_sync();

    R* out = (R*)a.out;
    const R* w = (const R*)a.w;
    R cgx = R(0), cgy = R(0), cgz = R(0);  // ADJ: d E_i / d r_i (per-lane part)

    // ---- radial groups -----------------------------------------------------
    const int n_rg = uniform(tab->n_rg);
    for (int g = 0; g < n_rg; ++g) {
        const pt_radial_group* G = &tab->rg[g];
        const int el = uniform(G->elem);
        const int b = count_below(S, m, el), e = count_below(S, m, el + 1);
        const int nfun = uniform(G->n);
        const int cut = uniform(G->cut);
        const R rc = (R)G->rc, irc = (R)(1.0 / G->rc);
        R wgt[PT_RG_MAX], acc[PT_RG_MAX], eta[PT_RG_MAX], rs[PT_RG_MAX];
#pragma unroll
        for (int q = 0; q < PT_RG_MAX; ++q) {
            acc[q] = R(0);
            eta[q] = (q < nfun) ? (R)G->eta[q] : R(0);
            rs[q] = (q < nfun) ? (R)G->rs[q] : R(0);
            wgt[q] = (MODE == MODE_ADJ && q < nfun) ? w[(size_t)c * a.ldw + G->col[q]] : R(0);
        }
        for (int s = b + lane; s < e; s += kWave) {
            const R r = S.r[s];
            R f, df;
            cutoff<R>(cut, rc, irc, tanh_pre, r, f, df);
            R rdot = R(0);
            if (MODE == MODE_TAN)
                rdot = (S.rx[s] * S.D[s] + S.ry[s] * S.D[cap + s] + S.rz[s] * S.D[2 * cap + s]) * S.ir[s];
            R dsum = R(0);
#pragma unroll
            for (int q = 0; q < PT_RG_MAX; ++q) {
                if (q < nfun) {
                    const R d = r - rs[q];
                    const R ex = M<R>::ex(-eta[q] * d * d);
                    if (MODE == MODE_FWD) acc[q] += ex * f;
                    if (MODE == MODE_ADJ) dsum += wgt[q] * ex * (df - R(2) * eta[q] * d * f);
                    if (MODE == MODE_TAN) acc[q] += ex * (df - R(2) * eta[q] * d * f) * rdot;
                }
            }
            if (MODE == MODE_ADJ) {
                const R sc = dsum * S.ir[s];
                const R gx = sc * S.rx[s], gy = sc * S.ry[s], gz = sc * S.rz[s];
                cgx += gx; cgy += gy; cgz += gz;
                S.DA[s] -= (double)gx; S.DA[cap + s] -= (double)gy; S.DA[2 * cap + s] -= (double)gz;
            }
        }
        if (MODE != MODE_ADJ) {
#pragma unroll
            for (int q = 0; q < PT_RG_MAX; ++q) acc[q] = wave_sum(acc[q]);
            if (lane == 0) {
#pragma unroll
                for (int q = 0; q < PT_RG_MAX; ++q)
                    if (q < nfun) out[(size_t)c * a.ld + G->col[q]] = acc[q];
            }
        }
        wave_sync();
    }

    // ---- angular groups ----------------------------------------------------
    const int n_ag = uniform(tab->n_ag);
    for (int g = 0; g < n_ag; ++g) {
        const pt_angular_group* G = &tab->ag[g];
        const int ej = uniform(G->ej), ek = uniform(G->ek);
        const int bj = count_below(S, m, ej), nj = count_below(S, m, ej + 1) - bj;
        const int bk = count_below(S, m, ek), nk = count_below(S, m, ek + 1) - bk;
        const int cut = uniform(G->cut);
        const bool g3 = uniform(G->kind) == 3;
        const int n_eta = uniform(G->n_eta), n_ang = uniform(G->n_ang);
        const R rc = (R)G->rc, irc = (R)(1.0 / G->rc);

        // cutoff values of the group's neighbours
        for (int pass = 0; pass < (ej == ek ? 1 : 2); ++pass) {
            const int sb = pass == 0 ? bj : bk, sn = pass == 0 ? nj : nk;
            for (int s = sb + lane; s < sb + sn; s += kWave) {
                R f, df;
                cutoff<R>(cut, rc, irc, tanh_pre, S.r[s], f, df);
                S.f[s] = f;
                S.df[s] = df;
            }
        }

        R eta[PT_AG_ETA], W[PT_AG_ETA][PT_AG_ANG], acc[PT_AG_ETA][PT_AG_ANG];
#pragma unroll
        for (int e = 0; e < PT_AG_ETA; ++e) {
            eta[e] = (e < n_eta) ? (R)G->eta[e] : R(0);
#pragma unroll
            for (int q = 0; q < PT_AG_ANG; ++q) {
                acc[e][q] = R(0);
                const int col = G->col[e * PT_AG_ANG + q];
                W[e][q] = (MODE == MODE_ADJ && e < n_eta && q < n_ang && col >= 0)
                              ? w[(size_t)c * a.ldw + col]
                              : R(0);
            }
        }
        R lam[PT_AG_ANG], zeta[PT_AG_ANG], pre[PT_AG_ANG];
        int izeta[PT_AG_ANG];
#pragma unroll
        for (int q = 0; q < PT_AG_ANG; ++q) {
            lam[q] = (R)G->lam[q]; zeta[q] = (R)G->zeta[q]; pre[q] = (R)G->pre[q];
            izeta[q] = uniform(G->izeta[q]);
        }
        wave_sync();

        // one live pair (ia, ib)
        auto pair_body = [&](const int ia, const int ib) {
            const R xa = S.rx[ia], ya = S.ry[ia], za = S.rz[ia], ira = S.ir[ia];
            const R xb = S.rx[ib], yb = S.ry[ib], zb = S.rz[ib], irb = S.ir[ib];
            R jx = xa - xb, jy = ya - yb, jz = za - zb;
            if (mic_jk) { jx = mic_r(jx, Lx, Hx); jy = mic_r(jy, Ly, Hy); jz = mic_r(jz, Lz, Hz); }
            const R rjk2 = jx * jx + jy * jy + jz * jz;
            if (!(rjk2 > R(0))) return;
            const R rjk = M<R>::sq(rjk2);
            R fjk = R(1), dfjk = R(0);
            if (g3) {
                if (!(rjk < rc)) return;
                cutoff<R>(cut, rc, irc, tanh_pre, rjk, fjk, dfjk);
            }
            const R fa = S.f[ia], fb = S.f[ib];
            const R Ssum = g3 ? S.r2[ia] + S.r2[ib] + rjk2 : S.r2[ia] + S.r2[ib];
            const R F = fa * fb * fjk;
            const R iab = ira * irb;
            const R cs = (xa * xb + ya * yb + za * zb) * iab;
            R ang[PT_AG_ANG], dang[PT_AG_ANG];
#pragma unroll
            for (int q = 0; q < PT_AG_ANG; ++q) {
                ang[q] = R(0); dang[q] = R(0);
                if (q < n_ang) angular_term<R>(cs, lam[q], zeta[q], izeta[q], pre[q], ang[q], dang[q]);
            }
            if (MODE == MODE_FWD) {
#pragma unroll
                for (int e = 0; e < PT_AG_ETA; ++e) {
                    if (e < n_eta) {
                        const R rad = M<R>::ex(-eta[e] * Ssum) * F;
#pragma unroll
                        for (int q = 0; q < PT_AG_ANG; ++q) acc[e][q] += rad * ang[q];
                    }
                }
                return;
            }
            // derivative building blocks
            const R dfa = S.df[ia], dfb = S.df[ib];
            const R ra = S.r[ia], rb = S.r[ib];
            const R ira2 = ira * ira, irb2 = irb * irb;
            const R cax = xb * iab - cs * xa * ira2, cay = yb * iab - cs * ya * ira2,
                    caz = zb * iab - cs * za * ira2;  // d cos / d R_ij
            const R cbx = xa * iab - cs * xb * irb2, cby = ya * iab - cs * yb * irb2,
                    cbz = za * iab - cs * zb * irb2;  // d cos / d R_ik
            const R irjk = g3 ? M<R>::rcp(rjk) : R(0);
            // d F / d r_ij etc. (without the exp factor)
            const R Fa = dfa * fb * fjk, Fb = fa * dfb * fjk, Fj = fa * fb * dfjk;
            if (MODE == MODE_ADJ) {
                R SP = R(0), SEP = R(0), SV = R(0);
#pragma unroll
                for (int e = 0; e < PT_AG_ETA; ++e) {
                    if (e < n_eta) {
                        const R ex = M<R>::ex(-eta[e] * Ssum);
                        R U = R(0), V = R(0);
#pragma unroll
                        for (int q = 0; q < PT_AG_ANG; ++q) { U += W[e][q] * ang[q]; V += W[e][q] * dang[q]; }
                        const R P = ex * U;
                        SP += P;
                        SEP += eta[e] * P;
                        SV += ex * V;
                    }
                }
                // dT/dr = d/dr [exp(-eta S) F]: F' sum_e P_e - 2 r F sum_e eta_e P_e
                const R d_rij = Fa * SP - R(2) * ra * F * SEP;
                const R d_rik = Fb * SP - R(2) * rb * F * SEP;
                const R d_rjk = g3 ? (Fj * SP - R(2) * rjk * F * SEP) : R(0);
                const R d_c = F * SV;
                const R sa = d_rij * ira, sb = d_rik * irb, sj = d_rjk * irjk;
                const R gax = sa * xa + d_c * cax + sj * jx;
                const R gay = sa * ya + d_c * cay + sj * jy;
                const R gaz = sa * za + d_c * caz + sj * jz;
                const R gbx = sb * xb + d_c * cbx - sj * jx;
                const R gby = sb * yb + d_c * cby - sj * jy;
                const R gbz = sb * zb + d_c * cbz - sj * jz;
                cgx += gax + gbx; cgy += gay + gby; cgz += gaz + gbz;
                lds_add(&S.DA[ia], -(double)gax); lds_add(&S.DA[cap + ia], -(double)gay);
                lds_add(&S.DA[2 * cap + ia], -(double)gaz);
                lds_add(&S.DA[ib], -(double)gbx); lds_add(&S.DA[cap + ib], -(double)gby);
                lds_add(&S.DA[2 * cap + ib], -(double)gbz);
            } else {  // MODE_TAN
                const R ax = S.D[ia], ay = S.D[cap + ia], az = S.D[2 * cap + ia];
                const R bx = S.D[ib], by = S.D[cap + ib], bz = S.D[2 * cap + ib];
                const R tij = (xa * ax + ya * ay + za * az) * ira;
                const R tik = (xb * bx + yb * by + zb * bz) * irb;
                const R tjk = g3 ? (jx * (ax - bx) + jy * (ay - by) + jz * (az - bz)) * irjk : R(0);
                const R cdot = cax * ax + cay * ay + caz * az + cbx * bx + cby * by + cbz * bz;
                // d/dt [exp(-eta S) F] = exp(-eta S) (X0 - 2 eta X1)
                const R X0 = Fa * tij + Fb * tik + Fj * tjk;
                const R X1 = F * (ra * tij + rb * tik + rjk * tjk);
#pragma unroll
                for (int e = 0; e < PT_AG_ETA; ++e) {
                    if (e < n_eta) {
                        const R ex = M<R>::ex(-eta[e] * Ssum);
                        const R rdot = ex * (X0 - R(2) * eta[e] * X1);
                        const R radc = ex * F * cdot;
#pragma unroll
                        for (int q = 0; q < PT_AG_ANG; ++q) acc[e][q] += rdot * ang[q] + radc * dang[q];
                    }
                }
            }
        };

        // Pair enumeration, ordered so that the 64 lanes of one step touch
        // distinct neighbour slots (no LDS atomic conflicts in the adjoint):
        //  - cross segments: lanes run over the shorter segment q, the other
        //    index is the diagonal k = (q + d) mod n_long, d = 0 .. n_long-1;
        //  - one segment: k = (q + 1 + t) mod n for t < (n-1)/2, plus the
        //    n/2 pairs (q, q + n/2) when n is even -- each unordered pair once.
        const bool same = ej == ek;
        const bool jrow = nj <= nk;
        const int nrow = same ? nj : (jrow ? nj : nk);
        const int ncol = same ? nj : (jrow ? nk : nj);
        const int brow = same ? bj : (jrow ? bj : bk);
        const int bcol = same ? bj : (jrow ? bk : bj);
        const int h = same ? (nj - 1) / 2 : ncol;
        const int p1 = nrow * h;
        const int npairs = same ? p1 + ((nj % 2 == 0) ? nj / 2 : 0) : p1;
        const float inv_n = nrow > 0 ? 1.0f / (float)nrow : 0.0f;
        // screen r_jk^2 < rc^2 (loosened by 1e-4 relative); the exact rjk < rc test is in pair_body
        const R rc2 = rc * rc * R(1.0001);
        int qn = 0;
        for (int base = 0; base < npairs; base += kWave) {
            const int p = base + lane;
            bool live = false;
            int ia = 0, ib = 0;
            if (p < npairs) {
                if (p < p1) {
                    int d = (int)(((float)p + 0.5f) * inv_n);
                    int q = p - d * nrow;
                    if (q < 0) { --d; q += nrow; } else if (q >= nrow) { ++d; q -= nrow; }
                    int k = same ? q + 1 + d : q + d;
                    if (k >= ncol) k -= ncol;
                    ia = brow + q; ib = bcol + k;
                } else {
                    const int q = p - p1;
                    ia = bj + q; ib = bj + q + nj / 2;
                }
                R jx = S.rx[ia] - S.rx[ib], jy = S.ry[ia] - S.ry[ib], jz = S.rz[ia] - S.rz[ib];
                if (mic_jk) { jx = mic_r(jx, Lx, Hx); jy = mic_r(jy, Ly, Hy); jz = mic_r(jz, Lz, Hz); }
                const R r2 = jx * jx + jy * jy + jz * jz;
                live = (r2 > R(0)) && (!g3 || r2 < rc2);
            }
            const unsigned long long bal = __ballot(live);
            if (live) S.queue[qn + lanes_below(bal)] = (unsigned)ia | ((unsigned)ib << 16);
            qn += __popcll(bal);
            if (qn > kQueue - kWave || base + kWave >= npairs) {
                wave_sync();
                for (int q = lane; q < qn; q += kWave) {
                    const unsigned e = S.queue[q];
                    pair_body((int)(e & 0xffffu), (int)(e >> 16));
                }
                qn = 0;
                wave_sync();
            }
        }

        if (MODE != MODE_ADJ) {
#pragma unroll
            for (int e = 0; e < PT_AG_ETA; ++e)
#pragma unroll
                for (int q = 0; q < PT_AG_ANG; ++q) acc[e][q] = wave_sum(acc[e][q]);
            if (lane == 0) {
#pragma unroll
                for (int e = 0; e < PT_AG_ETA; ++e)
#pragma unroll
                    for (int q = 0; q < PT_AG_ANG; ++q) {
                        const int col = G->col[e * PT_AG_ANG + q];
                        if (e < n_eta && q < n_ang && col >= 0) out[(size_t)c * a.ld + col] = acc[e][q];
                    }
            }
        }
        wave_sync();
    }

    if (MODE == MODE_ADJ) {
        if (a.edge_grad) {
            R* eg = (R*)a.edge_grad;
            for (int s = lane; s < m; s += kWave) {
                const long long e = off + S.slot[s];
                eg[3 * e + 0] = (R)S.DA[s];
                eg[3 * e + 1] = (R)S.DA[cap + s];
                eg[3 * e + 2] = (R)S.DA[2 * cap + s];
            }
        }
        if (a.f_self) {
            const double sx = wave_sum((double)cgx), sy = wave_sum((double)cgy), sz = wave_sum((double)cgz);
            if (lane == 0) {
                a.f_self[3 * c + 0] = -sx;
                a.f_self[3 * c + 1] = -sy;
                a.f_self[3 * c + 2] = -sz;
            }
        }
    }
}

template <typename R, int MODE>
int launch(KArgs a, hipStream_t stream) {
    if (a.n_centers <= 0) return PT_OK;
    a.cap = std::max(a.cap, 1);
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
    const int wpb = (int)std::min<size_t>(4, lds_budget / per_wave);
    const int blocks = (a.n_centers + wpb - 1) / wpb;
    hipLaunchKernelGGL((k_acsf<R, MODE>), dim3(blocks), dim3(64 * wpb), wpb * per_wave, stream, a);
    return launch_check("k_acsf");
}

int run(int precision, int mode, KArgs& a, hipStream_t st) {
    if (precision == PT_F64) {
        if (mode == MODE_FWD) return launch<double, MODE_FWD>(a, st);
        if (mode == MODE_ADJ) return launch<double, MODE_ADJ>(a, st);
        return launch<double, MODE_TAN>(a, st);
    }
    if (precision == PT_F32) {
        if (mode == MODE_FWD) return launch<float, MODE_FWD>(a, st);
        if (mode == MODE_ADJ) return launch<float, MODE_ADJ>(a, st);
        return launch<float, MODE_TAN>(a, st);
    }
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
