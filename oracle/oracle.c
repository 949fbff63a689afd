/*
 * oracle.c -- FP64 CPU restatement of pantea's NNP hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker, never the product:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load the library built from it.  The GPU product path (pantea_amd) never
 * links, loads or calls it.
 *
 * Every function restates the reference algorithm as written in
 * pantea 0.11.0; citations are path:line in the reference tree.
 *
 *   mic1 / dist3        atoms/box.py:112-117, atoms/distance.py:63-78
 *   wrap1               atoms/box.py:123-126 (jnp.remainder semantics)
 *   nbr mask            atoms/neighbor.py:102-107  ((r <= rc) & (r > 0))
 *   fc_val              descriptors/acsf/cutoff.py:66-110
 *   G1/G2               descriptors/acsf/radial.py:38-61, acsf.py:231-246
 *   G3/G9               descriptors/acsf/angular.py:50-107, acsf.py:249-330
 *   descriptor layout   descriptors/acsf/acsf.py:163-202 (radial then angular,
 *                       in insertion order)
 *   self-term forces    potentials/nnp/force.py:16-43 (gradient w.r.t. the
 *                       central positions only; neighbours frozen)
 *
 * Derivatives are analytic (hand-derived from the formulas above); they are
 * pinned in tests by central finite differences of the value functions here
 * and by the reference's golden forces (tests/test_nnp.py:45-81).
 *
 * Symmetry-function record (10 doubles per function):
 *   [kind, ej, ek, eta, rs, zeta, lambda, rc, cutoff_type, unused]
 *   kind: 1=G1 2=G2 3=G3 9=G9; ej/ek: 1-based atom types (ek unused for
 *   radial); cutoff_type: 0 hard 1 cos 2 tanhu 3 tanh 4 exp 5 poly1 6 poly2
 *   (the n2p2 codes of potentials/nnp/settings.py:43-51).
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define SF_STRIDE 10

enum { CUT_HARD = 0, CUT_COS = 1, CUT_TANHU = 2, CUT_TANH = 3, CUT_EXP = 4,
       CUT_POLY1 = 5, CUT_POLY2 = 6 };

/* cutoff.py:82 -- ((e + 1/e) / (e - 1/e)) ** 3 evaluated as Python does. */
static double TANH_PRE;
__attribute__((constructor)) static void init_tanh_pre(void) {
    double e = 2.718281828459045;
    TANH_PRE = pow((e + 1.0 / e) / (e - 1.0 / e), 3.0);
}
double or_tanh_prefactor(void) { return TANH_PRE; }

/* box.py:112-117: single-shift minimum image. */
static inline double mic1(double dx, double L) {
    if (dx > 0.5 * L) dx = dx - L;
    if (dx < -0.5 * L) dx = dx + L;
    return dx;
}

/* box.py:123-126: jnp.remainder(x, L) = fmod with sign fix-up. */
static inline double wrap1(double x, double L) {
    double m = fmod(x, L);
    if (m != 0.0 && ((m < 0.0) != (L < 0.0))) m = m + L;
    return m;
}

/* distance.py:63-78: dx = a - b (MIC when periodic), r = ||dx||_2 with the
 * exact-zero guard. */
static inline double dist3(const double *a, const double *b, const double *box,
                           double *dx) {
    dx[0] = a[0] - b[0];
    dx[1] = a[1] - b[1];
    dx[2] = a[2] - b[2];
    if (box) {
        dx[0] = mic1(dx[0], box[0]);
        dx[1] = mic1(dx[1], box[1]);
        dx[2] = mic1(dx[2], box[2]);
    }
    if (dx[0] == 0.0 && dx[1] == 0.0 && dx[2] == 0.0) return 0.0;
    return sqrt((dx[0] * dx[0] + dx[1] * dx[1]) + dx[2] * dx[2]);
}

/* cutoff.py:66-72 and :85-110. */
static double fc_val(int type, double rc, double r) {
    if (!(r < rc)) return 0.0;
    switch (type) {
    case CUT_HARD: return 1.0;
    case CUT_TANHU: { double t = tanh(1.0 - r / rc); return (t * t) * t; }
    case CUT_TANH: { double t = tanh(1.0 - r / rc); return TANH_PRE * ((t * t) * t); }
    case CUT_COS: return 0.5 * (cos((M_PI * r) / rc) + 1.0);
    case CUT_EXP: { double x = r / rc; return exp(1.0 - 1.0 / (1.0 - x * x)); }
    case CUT_POLY1: return (2.0 * r - 3.0) * (r * r) + 1.0;
    case CUT_POLY2: return ((15.0 - 6.0 * r) * r - 10.0) * ((r * r) * r) + 1.0;
    }
    return 0.0;
}

/* d fc / d r for the formulas above (zero outside r < rc). */
static double fc_der(int type, double rc, double r) {
    if (!(r < rc)) return 0.0;
    switch (type) {
    case CUT_HARD: return 0.0;
    case CUT_TANHU: case CUT_TANH: {
        double t = tanh(1.0 - r / rc);
        double d = 3.0 * t * t * (1.0 - t * t) * (-1.0 / rc);
        return type == CUT_TANH ? TANH_PRE * d : d;
    }
    case CUT_COS: return -0.5 * sin((M_PI * r) / rc) * (M_PI / rc);
    case CUT_EXP: {
        double x = r / rc, q = 1.0 - x * x;
        double f = exp(1.0 - 1.0 / q);
        return f * (-2.0 * x / (q * q)) / rc;
    }
    case CUT_POLY1: return 6.0 * r * r - 6.0 * r;
    case CUT_POLY2: return ((-30.0 * r + 60.0) * r - 30.0) * r * r;
    }
    return 0.0;
}

double or_cutoff(int type, double rc, double r) { return fc_val(type, rc, r); }
double or_cutoff_der(int type, double rc, double r) { return fc_der(type, rc, r); }
double or_mic(double dx, double L) { return mic1(dx, L); }
double or_wrap(double x, double L) { return wrap1(x, L); }

/* ------------------------------------------------------------------------
 * Dense neighbour masks (neighbor.py:102-107 over distance.py:63-78):
 * row i = { j : r_ij <= rc and r_ij > 0 }, ascending j.  Returns the number
 * of directed pairs, or -(needed) if cap is too small.
 * ---------------------------------------------------------------------- */
long or_neighbors(int n, const double *pos, const double *box, double rc,
                  int *counts, int *idx, long cap) {
    long total = 0;
    for (int i = 0; i < n; ++i) {
        int c = 0;
        for (int j = 0; j < n; ++j) {
            double dx[3];
            double r = dist3(pos + 3 * i, pos + 3 * j, box, dx);
            if (r <= rc && r > 0.0) {
                if (total + c < cap) idx[total + c] = j;
                ++c;
            }
        }
        counts[i] = c;
        total += c;
    }
    return total <= cap ? total : -total;
}

/* All pairwise MIC distances from one point (distance.py:63-78). */
void or_distances_from(int n, const double *pos, const double *box,
                       const double *p, double *r_out) {
    for (int j = 0; j < n; ++j) {
        double dx[3];
        r_out[j] = dist3(p, pos + 3 * j, box, dx);
    }
}

void or_wrap_positions(int n, double *pos, const double *box) {
    for (int i = 0; i < 3 * n; ++i) pos[i] = wrap1(pos[i], box[i % 3]);
}

/* ------------------------------------------------------------------------
 * Per-centre neighbour scratch (the masked entries of the dense row).
 * ---------------------------------------------------------------------- */
typedef struct {
    int j;        /* atom index */
    int type;     /* 1-based atom type */
    double R[3];  /* R_ij = r_i - r_j (MIC) */
    double r;
} nb_t;

static int gather_row(int n, const double *pos, const int *types, const double *box,
                      const double *ci, int self_idx, double rc, nb_t *row) {
    int m = 0;
    for (int j = 0; j < n; ++j) {
        if (j == self_idx) continue;
        double dx[3];
        double r = dist3(ci, pos + 3 * j, box, dx);
        if (r <= rc && r > 0.0) {
            row[m].j = j;
            row[m].type = types[j];
            row[m].R[0] = dx[0]; row[m].R[1] = dx[1]; row[m].R[2] = dx[2];
            row[m].r = r;
            ++m;
        }
    }
    return m;
}

static double max_rc(int nsf, const double *sf) {
    double m = 0.0;
    for (int s = 0; s < nsf; ++s) if (sf[SF_STRIDE * s + 7] > m) m = sf[SF_STRIDE * s + 7];
    return m;
}

/* Angular pair geometry: acsf.py:307-320 (cos with the zero guard) and the
 * r_jk of distance.py applied to (R_ij - R_ik). */
static inline double pair_geom(const nb_t *a, const nb_t *b, const double *box,
                               double *cost, double *Rjk) {
    double operand = a->r * b->r;
    double dot = (a->R[0] * b->R[0] + a->R[1] * b->R[1]) + a->R[2] * b->R[2];
    *cost = (operand == 0.0) ? 0.0 : dot / operand;
    return dist3(a->R, b->R, box, Rjk);
}

/* ------------------------------------------------------------------------
 * ACSF descriptor of one centre at position ci (acsf.py:163-202).
 * Loops follow the reference: per symmetry function, mask over all
 * neighbours (radial) or all ordered pairs (j, k) (angular), halving when
 * the two neighbour elements coincide (acsf.py:285-290).
 * ---------------------------------------------------------------------- */
static void acsf_row(int m, const nb_t *row, const double *box, int nsf, const double *sf,
                     double *out) {
    for (int s = 0; s < nsf; ++s) {
        const double *p = sf + SF_STRIDE * s;
        int kind = (int)p[0], ej = (int)p[1], ek = (int)p[2], ct = (int)p[8];
        double eta = p[3], rs = p[4], zeta = p[5], lam = p[6], rc = p[7];
        double acc = 0.0;
        if (kind == 1 || kind == 2) {
            for (int a = 0; a < m; ++a) {
                if (row[a].type != ej || !(row[a].r <= rc)) continue;
                double r = row[a].r;
                double g = (kind == 1) ? fc_val(ct, rc, r)
                                       : exp(-eta * ((r - rs) * (r - rs))) * fc_val(ct, rc, r);
                acc += g;
            }
        } else {
            double pre = pow(2.0, 1.0 - zeta);
            for (int a = 0; a < m; ++a) {
                if (row[a].type != ej || !(row[a].r <= rc)) continue;
                double v = 0.0;
                for (int b = 0; b < m; ++b) {
                    if (row[b].type != ek || !(row[b].r <= rc)) continue;
                    double cost, Rjk[3];
                    double rjk = pair_geom(&row[a], &row[b], box, &cost, Rjk);
                    if (!(rjk > 0.0)) continue;
                    double rij = row[a].r, rik = row[b].r, t;
                    if (kind == 3)
                        t = pre * pow(1.0 + lam * cost, zeta) *
                            exp(-eta * (rij * rij + rik * rik + rjk * rjk)) *
                            fc_val(ct, rc, rij) * fc_val(ct, rc, rik) * fc_val(ct, rc, rjk);
                    else
                        t = pre * pow(1.0 + lam * cost, zeta) *
                            exp(-eta * (rij * rij + rik * rik)) *
                            fc_val(ct, rc, rij) * fc_val(ct, rc, rik);
                    v += t;
                }
                acc += v;
            }
            if (ej == ek) acc = acc * 0.5;
        }
        out[s] = acc;
    }
}

static void acsf_one(int n, const double *pos, const int *types, const double *box,
                     int nsf, const double *sf, const double *ci, int self_idx,
                     nb_t *row, double *out) {
    int m = gather_row(n, pos, types, box, ci, self_idx, max_rc(nsf, sf), row);
    acsf_row(m, row, box, nsf, sf, out);
}

/* Descriptor rows for centres at explicit positions.  self_idx[c] is the
 * index of the centre in pos (excluded from its own neighbours) or -1. */
int or_acsf(int n, const double *pos, const int *types, const double *box,
            int nsf, const double *sf, int nc, const double *cpos,
            const int *self_idx, double *out) {
    int err = 0;
#pragma omp parallel
    {
        nb_t *row = (nb_t *)malloc(sizeof(nb_t) * (size_t)(n > 0 ? n : 1));
        if (!row) {
#pragma omp atomic write
            err = -1;
        } else {
#pragma omp for schedule(dynamic, 4)
            for (int c = 0; c < nc; ++c)
                acsf_one(n, pos, types, box, nsf, sf, cpos + 3 * c,
                         self_idx ? self_idx[c] : -1, row, out + (size_t)nsf * c);
            free(row);
        }
    }
    return err;
}

/* ------------------------------------------------------------------------
 * Vector-Jacobian product of one centre's descriptor.
 * Given w[s] = dE_i/dG_is, accumulates
 *   self[3]       += sum_s w_s dG_is/dr_i            (centre derivative)
 *   nbr_grad[m]   += sum_s w_s dG_is/dr_m  for m != i (through row[].j)
 * The per-neighbour part is returned in row-local storage d[a][3].
 * ---------------------------------------------------------------------- */
static void acsf_vjp_one(int m, const nb_t *row, const double *box, int nsf,
                         const double *sf, const double *w, double *self, double *d) {
    for (int a = 0; a < 3 * m; ++a) d[a] = 0.0;
    self[0] = self[1] = self[2] = 0.0;
    for (int s = 0; s < nsf; ++s) {
        if (w[s] == 0.0) continue;
        const double *p = sf + SF_STRIDE * s;
        int kind = (int)p[0], ej = (int)p[1], ek = (int)p[2], ct = (int)p[8];
        double eta = p[3], rs = p[4], zeta = p[5], lam = p[6], rc = p[7];
        if (kind == 1 || kind == 2) {
            for (int a = 0; a < m; ++a) {
                if (row[a].type != ej || !(row[a].r <= rc)) continue;
                double r = row[a].r;
                double f = fc_val(ct, rc, r), fp = fc_der(ct, rc, r), dg;
                if (kind == 1) dg = fp;
                else {
                    double ex = exp(-eta * ((r - rs) * (r - rs)));
                    dg = ex * (fp - 2.0 * eta * (r - rs) * f);
                }
                double sc = w[s] * dg / r;
                for (int q = 0; q < 3; ++q) {
                    self[q] += sc * row[a].R[q];
                    d[3 * a + q] -= sc * row[a].R[q];
                }
            }
        } else {
            double pre = pow(2.0, 1.0 - zeta);
            double half = (ej == ek) ? 0.5 : 1.0;
            for (int a = 0; a < m; ++a) {
                if (row[a].type != ej || !(row[a].r <= rc)) continue;
                for (int b = 0; b < m; ++b) {
                    if (row[b].type != ek || !(row[b].r <= rc)) continue;
                    double cost, Rjk[3];
                    double rjk = pair_geom(&row[a], &row[b], box, &cost, Rjk);
                    if (!(rjk > 0.0)) continue;
                    double rij = row[a].r, rik = row[b].r;
                    double fij = fc_val(ct, rc, rij), fik = fc_val(ct, rc, rik);
                    double dfij = fc_der(ct, rc, rij), dfik = fc_der(ct, rc, rik);
                    double fjk = 1.0, dfjk = 0.0, X;
                    if (kind == 3) {
                        fjk = fc_val(ct, rc, rjk); dfjk = fc_der(ct, rc, rjk);
                        X = exp(-eta * (rij * rij + rik * rik + rjk * rjk));
                    } else {
                        X = exp(-eta * (rij * rij + rik * rik));
                    }
                    double base = 1.0 + lam * cost;
                    double P = pow(base, zeta);
                    double dP = (zeta == 0.0) ? 0.0 : zeta * lam * pow(base, zeta - 1.0);
                    double W = w[s] * half * pre;
                    double dT_dc = W * dP * X * fij * fik * fjk;
                    double dT_drij = W * P * X * fik * fjk * (dfij - 2.0 * eta * rij * fij);
                    double dT_drik = W * P * X * fij * fjk * (dfik - 2.0 * eta * rik * fik);
                    double dT_drjk = (kind == 3) ? W * P * X * fij * fik * (dfjk - 2.0 * eta * rjk * fjk) : 0.0;
                    double inv = 1.0 / (rij * rik);
                    for (int q = 0; q < 3; ++q) {
                        double dc_dRij = row[b].R[q] * inv - cost * row[a].R[q] / (rij * rij);
                        double dc_dRik = row[a].R[q] * inv - cost * row[b].R[q] / (rik * rik);
                        double gij = dT_drij * row[a].R[q] / rij + dT_dc * dc_dRij +
                                     dT_drjk * Rjk[q] / rjk;
                        double gik = dT_drik * row[b].R[q] / rik + dT_dc * dc_dRik -
                                     dT_drjk * Rjk[q] / rjk;
                        self[q] += gij + gik;
                        d[3 * a + q] -= gij;
                        d[3 * b + q] -= gik;
                    }
                }
            }
        }
    }
}

/* Self-Jacobian dG_i/dr_i for centres at explicit positions (acsf.py:215-228
 * semantics: jacfwd w.r.t. the central position only).  out: nc x nsf x 3. */
int or_acsf_selfjac(int n, const double *pos, const int *types, const double *box,
                    int nsf, const double *sf, int nc, const double *cpos,
                    const int *self_idx, double *out) {
    int err = 0;
#pragma omp parallel
    {
        nb_t *row = (nb_t *)malloc(sizeof(nb_t) * (size_t)(n > 0 ? n : 1));
        double *d = (double *)malloc(sizeof(double) * 3 * (size_t)(n > 0 ? n : 1));
        double *w = (double *)malloc(sizeof(double) * (size_t)(nsf > 0 ? nsf : 1));
        if (!row || !d || !w) {
#pragma omp atomic write
            err = -1;
        } else {
#pragma omp for schedule(dynamic, 4)
            for (int c = 0; c < nc; ++c) {
                int m = gather_row(n, pos, types, box, cpos + 3 * c,
                                   self_idx ? self_idx[c] : -1, max_rc(nsf, sf), row);
                for (int s = 0; s < nsf; ++s) {
                    for (int q = 0; q < nsf; ++q) w[q] = (q == s) ? 1.0 : 0.0;
                    double self[3];
                    acsf_vjp_one(m, row, box, nsf, sf, w, self, d);
                    for (int q = 0; q < 3; ++q) out[((size_t)c * nsf + s) * 3 + q] = self[q];
                }
            }
        }
        free(row); free(d); free(w);
    }
    return err;
}

/* ------------------------------------------------------------------------
 * Forces for the NNP: given per-atom weights w (n x nsf_max, row i holds
 * dE_i/dG_i for the descriptor table of atom i's element), computes
 *   f_self[i] = -sum_s w_is dG_is/dr_i            (force.py:16-43 semantics)
 *   f_full[m] = -sum_i sum_s w_is dG_is/dr_m      (-grad E)
 * Tables: tab_of_type[t-1] indexes (sf_off, nsf) for 1-based type t.
 * Only centres with want_center[i] != 0 (or all if NULL) contribute; this is
 * used for halo-restricted checks on large structures.
 * ---------------------------------------------------------------------- */
int or_nnp_forces(int n, const double *pos, const int *types, const double *box,
                  int ntypes, const int *sf_off, const int *sf_cnt, const double *sf_all,
                  int w_stride, const double *w, const unsigned char *want_center,
                  double *f_self, double *f_full) {
    int nthreads = 1;
#ifdef _OPENMP
    nthreads = omp_get_max_threads();
#endif
    double *part = (double *)calloc((size_t)nthreads * 3 * (size_t)n, sizeof(double));
    if (!part) return -1;
    int err = 0;
    double rcmax = 0.0;
    for (int t = 0; t < ntypes; ++t) {
        double r = max_rc(sf_cnt[t], sf_all + SF_STRIDE * sf_off[t]);
        if (r > rcmax) rcmax = r;
    }
#pragma omp parallel
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        double *mine = part + (size_t)tid * 3 * n;
        nb_t *row = (nb_t *)malloc(sizeof(nb_t) * (size_t)(n > 0 ? n : 1));
        double *d = (double *)malloc(sizeof(double) * 3 * (size_t)(n > 0 ? n : 1));
        if (!row || !d) {
#pragma omp atomic write
            err = -1;
        } else {
#pragma omp for schedule(dynamic, 8)
            for (int i = 0; i < n; ++i) {
                if (want_center && !want_center[i]) {
                    if (f_self) f_self[3 * i] = f_self[3 * i + 1] = f_self[3 * i + 2] = 0.0;
                    continue;
                }
                int t = types[i] - 1;
                const double *sf = sf_all + SF_STRIDE * sf_off[t];
                int m = gather_row(n, pos, types, box, pos + 3 * i, i, rcmax, row);
                double self[3];
                acsf_vjp_one(m, row, box, sf_cnt[t], sf, w + (size_t)w_stride * i, self, d);
                if (f_self) for (int q = 0; q < 3; ++q) f_self[3 * i + q] = -self[q];
                for (int q = 0; q < 3; ++q) mine[3 * i + q] -= self[q];
                for (int a = 0; a < m; ++a)
                    for (int q = 0; q < 3; ++q) mine[3 * row[a].j + q] -= d[3 * a + q];
            }
        }
        free(row); free(d);
    }
    if (f_full) {
        for (size_t k = 0; k < 3 * (size_t)n; ++k) {
            double s = 0.0;
            for (int t = 0; t < nthreads; ++t) s += part[(size_t)t * 3 * n + k];
            f_full[k] = s;
        }
    }
    free(part);
    return err;
}

/* ------------------------------------------------------------------------
 * The reference ALGORITHM, for timing the CPU baseline (not a checker: the
 * functions above are).  Per centre i (acsf.py:163-202): MIC differences and
 * distances to all N atoms; per symmetry function a masked sum over all N
 * atoms (radial, acsf.py:231-246) or over all N x N ordered pairs (angular:
 * lax.scan over j with a vector over k, acsf.py:249-330).  Like the
 * vectorised JAX code, every (j, k) term is evaluated and the mask applied
 * to the sum, so the angular work is O(N^2) per function and centre.
 * With w != NULL the same loops also accumulate the gradient of
 * sum_s w_s G_is with respect to the centre (force.py:16-43: jax.grad w.r.t.
 * the central positions), written to grad (nc x 3).
 * ---------------------------------------------------------------------- */
static void dense_centre(int n, const double *pos, const int *types, const double *box,
                         int nsf, const double *sf, int i, double *R, double *r,
                         double *G, const double *w, double *grad) {
    double gself[3] = {0.0, 0.0, 0.0};
    for (int j = 0; j < n; ++j) r[j] = dist3(pos + 3 * i, pos + 3 * j, box, R + 3 * j);
    for (int s = 0; s < nsf; ++s) {
        const double *p = sf + SF_STRIDE * s;
        int kind = (int)p[0], ej = (int)p[1], ek = (int)p[2], ct = (int)p[8];
        double eta = p[3], rs = p[4], zeta = p[5], lam = p[6], rc = p[7];
        double acc = 0.0;
        double ws = w ? w[s] : 0.0;
        if (kind == 1 || kind == 2) {
            for (int j = 0; j < n; ++j) {
                double rj = r[j];
                double f = fc_val(ct, rc, rj);
                double ex = (kind == 1) ? 1.0 : exp(-eta * ((rj - rs) * (rj - rs)));
                int mask = (rj <= rc) && (rj > 0.0) && types[j] == ej;
                acc += mask ? ex * f : 0.0;
                if (w) {
                    double fp = fc_der(ct, rc, rj);
                    double dg = (kind == 1) ? fp : ex * (fp - 2.0 * eta * (rj - rs) * f);
                    double sc = mask ? ws * dg / rj : 0.0;
                    for (int q = 0; q < 3; ++q) gself[q] += sc * R[3 * j + q];
                }
            }
        } else {
            double pre = pow(2.0, 1.0 - zeta);
            for (int j = 0; j < n; ++j) {
                double rj = r[j];
                int mj = (rj <= rc) && (rj > 0.0) && types[j] == ej;
                double fij = fc_val(ct, rc, rj), dfij = w ? fc_der(ct, rc, rj) : 0.0;
                double v = 0.0;
                for (int k = 0; k < n; ++k) {
                    double rk = r[k];
                    int mk = (rk <= rc) && (rk > 0.0) && types[k] == ek;
                    double operand = rj * rk;
                    double dot = (R[3 * j] * R[3 * k] + R[3 * j + 1] * R[3 * k + 1]) + R[3 * j + 2] * R[3 * k + 2];
                    double cost = mk ? (operand == 0.0 ? 0.0 : dot / operand) : 1.0;
                    double Rjk[3];
                    double rjk = mk ? dist3(R + 3 * j, R + 3 * k, box, Rjk) : 0.0;
                    if (!mk) Rjk[0] = Rjk[1] = Rjk[2] = 0.0;
                    double fik = fc_val(ct, rc, rk);
                    double fjk = (kind == 3) ? fc_val(ct, rc, rjk) : 1.0;
                    double S = (kind == 3) ? rj * rj + rk * rk + rjk * rjk : rj * rj + rk * rk;
                    double X = exp(-eta * S);
                    double base = 1.0 + lam * cost;
                    double P = pow(base, zeta);
                    double t = pre * P * X * fij * fik * fjk;
                    int m = mk && (rjk > 0.0);
                    v += m ? t : 0.0;
                    if (w && mj && m) {
                        double dfik = fc_der(ct, rc, rk);
                        double dP = (zeta == 0.0) ? 0.0 : zeta * lam * pow(base, zeta - 1.0);
                        double W = ws * pre * ((ej == ek) ? 0.5 : 1.0);
                        double dT_dc = W * dP * X * fij * fik * fjk;
                        double dT_drij = W * P * X * fik * fjk * (dfij - 2.0 * eta * rj * fij);
                        double dT_drik = W * P * X * fij * fjk * (dfik - 2.0 * eta * rk * fik);
                        double inv = 1.0 / (rj * rk);
                        for (int q = 0; q < 3; ++q) {
                            double dc_dRij = R[3 * k + q] * inv - cost * R[3 * j + q] / (rj * rj);
                            double dc_dRik = R[3 * j + q] * inv - cost * R[3 * k + q] / (rk * rk);
                            /* R_jk = R_ij - R_ik does not move with r_i */
                            gself[q] += dT_drij * R[3 * j + q] / rj + dT_drik * R[3 * k + q] / rk +
                                        dT_dc * (dc_dRij + dc_dRik);
                        }
                    }
                }
                acc += mj ? v : 0.0;
            }
            if (ej == ek) acc = acc * 0.5;
        }
        if (G) G[s] = acc;
    }
    if (grad) for (int q = 0; q < 3; ++q) grad[q] = gself[q];
}

int or_reference_algorithm(int n, const double *pos, const int *types, const double *box,
                           int nsf, const double *sf, int nc, const int *centres,
                           double *G, const double *w, double *grad) {
    int err = 0;
#pragma omp parallel
    {
        double *R = (double *)malloc(sizeof(double) * 3 * (size_t)(n > 0 ? n : 1));
        double *r = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
        if (!R || !r) {
#pragma omp atomic write
            err = -1;
        } else {
#pragma omp for schedule(dynamic, 1)
            for (int c = 0; c < nc; ++c)
                dense_centre(n, pos, types, box, nsf, sf, centres[c], R, r,
                             G ? G + (size_t)nsf * c : NULL, w ? w + (size_t)nsf * c : NULL,
                             grad ? grad + 3 * (size_t)c : NULL);
        }
        free(R); free(r);
    }
    return err;
}

/* ------------------------------------------------------------------------
 * Cell-list neighbour rows (same rows as or_neighbors: { j : r_ij <= rc,
 * r_ij > 0 } under the single-shift minimum image, ascending j) in O(N) for
 * periodic boxes with at least 3 cells of width >= rc per dimension; other
 * boxes fall back to the dense scan.  For timing the cell-list CPU baseline.
 * ---------------------------------------------------------------------- */
static int cmp_int(const void *a, const void *b) {
    int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}

long or_cell_neighbors(int n, const double *pos, const double *box, double rc,
                       long long *row_off, int *idx, long cap) {
    int nc[3];
    int ok = box != NULL;
    for (int d = 0; d < 3 && ok; ++d) {
        nc[d] = (int)floor(box[d] / rc);
        if (nc[d] < 3) ok = 0;
    }
    if (!ok) {
        int *counts = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
        if (!counts) return 0;
        long t = or_neighbors(n, pos, box, rc, counts, idx, cap);
        row_off[0] = 0;
        for (int i = 0; i < n; ++i) row_off[i + 1] = row_off[i] + counts[i];
        free(counts);
        return t;
    }
    long ncell = (long)nc[0] * nc[1] * nc[2];
    int *head = (int *)malloc(sizeof(int) * (size_t)(ncell + 1));
    int *order = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    int *cell = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    int *cnt = (int *)calloc((size_t)(n > 0 ? n : 1), sizeof(int));
    if (!head || !order || !cell || !cnt) { free(head); free(order); free(cell); free(cnt); return 0; }
    for (long c = 0; c <= ncell; ++c) head[c] = 0;
    for (int i = 0; i < n; ++i) {
        int ci[3];
        for (int d = 0; d < 3; ++d) {
            double x = wrap1(pos[3 * i + d], box[d]);
            int k = (int)(x / box[d] * nc[d]);
            ci[d] = k < 0 ? 0 : (k >= nc[d] ? nc[d] - 1 : k);
        }
        cell[i] = (ci[0] * nc[1] + ci[1]) * nc[2] + ci[2];
        head[cell[i] + 1]++;
    }
    for (long c = 0; c < ncell; ++c) head[c + 1] += head[c];
    {
        int *fill = (int *)malloc(sizeof(int) * (size_t)ncell);
        if (!fill) { free(head); free(order); free(cell); free(cnt); return 0; }
        for (long c = 0; c < ncell; ++c) fill[c] = head[c];
        for (int i = 0; i < n; ++i) order[fill[cell[i]]++] = i;
        free(fill);
    }
    /* count, then fill each row (sorted ascending) */
    for (int pass = 0; pass < 2; ++pass) {
        if (pass == 1) {
            row_off[0] = 0;
            for (int i = 0; i < n; ++i) row_off[i + 1] = row_off[i] + cnt[i];
            if (row_off[n] > cap) break;
        }
#pragma omp parallel for schedule(dynamic, 64)
        for (int i = 0; i < n; ++i) {
            int c0 = cell[i];
            int cz = c0 % nc[2], cy = (c0 / nc[2]) % nc[1], cx = c0 / (nc[1] * nc[2]);
            int m = 0;
            int *out = pass ? idx + row_off[i] : NULL;
            for (int dx = -1; dx <= 1; ++dx)
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dz = -1; dz <= 1; ++dz) {
                        int c = (((cx + dx + nc[0]) % nc[0]) * nc[1] + (cy + dy + nc[1]) % nc[1]) * nc[2] +
                                (cz + dz + nc[2]) % nc[2];
                        for (int q = head[c]; q < head[c + 1]; ++q) {
                            int j = order[q];
                            double d3[3];
                            double r = dist3(pos + 3 * i, pos + 3 * j, box, d3);
                            if (r <= rc && r > 0.0) {
                                if (pass) out[m] = j;
                                ++m;
                            }
                        }
                    }
            if (pass) qsort(out, (size_t)m, sizeof(int), cmp_int);
            else cnt[i] = m;
        }
    }
    long total = 0;
    for (int i = 0; i < n; ++i) total += cnt[i];
    free(head); free(order); free(cell); free(cnt);
    return total <= cap ? total : -total;
}

/* Row of centre i from CSR neighbour rows (the masked entries of the dense
 * row, in the same order). */
static int csr_row(const double *pos, const int *types, const double *box, int i,
                   const long long *row_off, const int *idx, nb_t *row) {
    int m = 0;
    for (long long e = row_off[i]; e < row_off[i + 1]; ++e) {
        int j = idx[e];
        double dx[3];
        double r = dist3(pos + 3 * i, pos + 3 * j, box, dx);
        row[m].j = j;
        row[m].type = types[j];
        row[m].R[0] = dx[0]; row[m].R[1] = dx[1]; row[m].R[2] = dx[2];
        row[m].r = r;
        ++m;
    }
    return m;
}

/* Descriptors, then (with w) the per-centre VJP, of the centres `centres`
 * from CSR rows: the cell-list CPU baseline (same arithmetic as or_acsf /
 * or_nnp_forces).  G: nc x nsf; self_grad: nc x 3 (sum_s w_s dG_is/dr_i). */
int or_celllist_centres(int n, const double *pos, const int *types, const double *box,
                        const long long *row_off, const int *idx, int nsf, const double *sf,
                        int nc, const int *centres, double *G, const double *w, double *self_grad) {
    int err = 0;
    int maxrow = 0;
    for (int c = 0; c < nc; ++c) {
        int len = (int)(row_off[centres[c] + 1] - row_off[centres[c]]);
        if (len > maxrow) maxrow = len;
    }
#pragma omp parallel
    {
        nb_t *row = (nb_t *)malloc(sizeof(nb_t) * (size_t)(maxrow > 0 ? maxrow : 1));
        double *d = (double *)malloc(sizeof(double) * 3 * (size_t)(maxrow > 0 ? maxrow : 1));
        if (!row || !d) {
#pragma omp atomic write
            err = -1;
        } else {
#pragma omp for schedule(dynamic, 4)
            for (int c = 0; c < nc; ++c) {
                int i = centres[c];
                int m = csr_row(pos, types, box, i, row_off, idx, row);
                if (G) acsf_row(m, row, box, nsf, sf, G + (size_t)nsf * c);
                if (w && self_grad) {
                    double self[3];
                    acsf_vjp_one(m, row, box, nsf, sf, w + (size_t)nsf * c, self, d);
                    for (int q = 0; q < 3; ++q) self_grad[3 * (size_t)c + q] = self[q];
                }
            }
        }
        free(row); free(d);
    }
    return err;
}
