/*
 * CPU oracle for the Bargmann->Fock hot path.  TEST INFRASTRUCTURE ONLY.
 *
 * This file restates, statement by statement, the reference's numba kernels
 *   mrmustard/math/lattice/strategies/vanilla/core.py:22        SQRT table
 *   mrmustard/math/lattice/strategies/vanilla/core.py:64-124    vanilla_numba
 *   mrmustard/math/lattice/strategies/vanilla/core.py:164-213   stable_numba
 *   mrmustard/math/lattice/strategies/vanilla/gradients.py:41-82   vanilla_vjp_numba
 *   mrmustard/math/lattice/strategies/vanilla/batch.py:54-61       vanilla_batch_numba
 *   mrmustard/math/lattice/strategies/vanilla/gradients.py:104-116 vanilla_batch_vjp_numba
 * in plain C, so that the GPU kernels can be checked against it on identical
 * inputs.  Only tests, the smoke check and the CPU-baseline leg of the
 * benchmark load it; the product path never does.
 *
 * Rounding fidelity.  numba compiles the reference without fastmath, so there
 * is no FMA contraction (build with -ffp-contract=off) and:
 *   - complex*float promotes the float to complex(x, 0.0) and multiplies as
 *     (ac-bd, ad+bc): see c_mul / c_scale below;
 *   - complex/float and complex/int promote the divisor and use CPython's
 *     _Py_c_quot algorithm: see c_div;
 *   - negative flat indices wrap around (numpy-style) when G is read: see
 *     rd().  The wrapped reads are always multiplied by SQRT[0] == 0.
 *
 * The VJP also has a compensated-sum variant (oracle_vjp_kahan) that gives the
 * exactly-rounded-looking reference value for very large lattices, where the
 * reference's own left-to-right sum carries ~sqrt(N)*eps relative error.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define SQRT_LEN 100000

typedef struct {
    double re, im;
} cplx;

static double SQRT[SQRT_LEN];
static int sqrt_ready = 0;

/* core.py:22  SQRT = np.sqrt(np.arange(100000)) ; C sqrt is correctly rounded. */
static void init_sqrt(void) {
    if (sqrt_ready) return;
    for (int n = 0; n < SQRT_LEN; ++n) SQRT[n] = sqrt((double)n);
    sqrt_ready = 1;
}

/* numpy-style index: SQRT[-1] is SQRT[99999]. */
static inline double sq(int64_t n) { return SQRT[n < 0 ? n + SQRT_LEN : n]; }

/* numba complex multiply: (a+bi)(c+di) = (ac-bd) + (ad+bc)i, no FMA. */
static inline cplx c_mul(cplx x, cplx y) {
    cplx z;
    z.re = x.re * y.re - x.im * y.im;
    z.im = x.re * y.im + x.im * y.re;
    return z;
}

/* complex * float: the float is promoted to complex(s, 0.0) first. */
static inline cplx c_scale(cplx x, double s) {
    cplx y = {s, 0.0};
    return c_mul(x, y);
}

/* float * complex (float on the left): promoted the same way. */
static inline cplx s_mul(double s, cplx x) {
    cplx y = {s, 0.0};
    return c_mul(y, x);
}

static inline cplx c_add(cplx x, cplx y) {
    cplx z = {x.re + y.re, x.im + y.im};
    return z;
}

/* CPython _Py_c_quot as used by numba for complex division. */
static inline cplx c_div(cplx a, cplx b) {
    cplx z;
    double abs_br = fabs(b.re), abs_bi = fabs(b.im);
    if (abs_br >= abs_bi) {
        if (b.re == 0.0 && b.im == 0.0) {
            z.re = NAN;
            z.im = NAN;
            return z;
        }
        double ratio = b.im / b.re;
        double denom = b.re + b.im * ratio;
        z.re = (a.re + a.im * ratio) / denom;
        z.im = (a.im - a.re * ratio) / denom;
    } else {
        double ratio = b.re / b.im;
        double denom = b.re * ratio + b.im;
        z.re = (a.re * ratio + a.im) / denom;
        z.im = (a.im * ratio - a.re) / denom;
    }
    return z;
}

static inline cplx c_div_real(cplx a, double s) {
    cplx b = {s, 0.0};
    return c_div(a, b);
}

/* G[idx] with numpy wrap-around for negative idx (only ever multiplied by 0). */
static inline cplx rd(const cplx *G, int64_t idx, int64_t N) {
    if (idx < 0) {
        idx %= N;
        if (idx < 0) idx += N;
    }
    return G[idx];
}

static void strides_of(const int64_t *shape, int D, int64_t *strides) {
    strides[D - 1] = 1;
    for (int i = D - 1; i > 0; --i) strides[i - 1] = strides[i] * shape[i];
}

static int64_t prod_of(const int64_t *shape, int D) {
    int64_t n = 1;
    for (int i = 0; i < D; ++i) n *= shape[i];
    return n;
}

/* np.ndindex advance (C order, last axis fastest). */
static inline void next_index(int64_t *idx, const int64_t *shape, int D) {
    for (int a = D - 1; a >= 0; --a) {
        if (++idx[a] < shape[a]) return;
        idx[a] = 0;
    }
}

#define MAXD 32

/* core.py:64-124.  G must hold prod(shape) entries, zero-initialised by the
 * caller unless it stands for a caller-provided `out` buffer.  Returns 0, or
 * -1 on invalid input. */
int oracle_vanilla(const int64_t *shape, int D, const double *A_, const double *b_,
                   const double *c_, double *G_) {
    if (D < 1 || D > MAXD) return -1;
    init_sqrt();
    const cplx *A = (const cplx *)A_;
    const cplx *b = (const cplx *)b_;
    cplx *G = (cplx *)G_;
    int64_t strides[MAXD], index[MAXD];
    strides_of(shape, D, strides);
    int64_t N = prod_of(shape, D);
    for (int a = 0; a < D; ++a) index[a] = 0;

    G[0] = *(const cplx *)c_; /* core.py:79 */

    /* core.py:85-104: flat indices below the largest stride, with pivot search */
    for (int64_t flat = 1; flat < strides[0] && flat < N; ++flat) {
        next_index(index, shape, D);
        int i = 0;
        int64_t pivot = 0;
        for (int s = 0; s < D; ++s) {
            pivot = flat - strides[s];
            if (pivot >= 0) break;
            i += 1;
        }
        cplx v = c_mul(b[i], G[pivot]);
        v = c_add(v, c_mul(c_scale(A[i * D + i], sq(index[i] - 1)), rd(G, pivot - strides[i], N)));
        for (int j = i + 1; j < D; ++j)
            v = c_add(v, c_mul(c_scale(A[i * D + j], sq(index[j])), rd(G, pivot - strides[j], N)));
        G[flat] = c_div_real(v, sq(index[i]));
    }
    /* core.py:108-122: the rest, pivot always on axis 0 */
    for (int64_t flat = strides[0]; flat < N; ++flat) {
        next_index(index, shape, D);
        int64_t pivot = flat - strides[0];
        cplx v = c_mul(b[0], G[pivot]);
        v = c_add(v, c_mul(c_scale(A[0], sq(index[0] - 1)), rd(G, pivot - strides[0], N)));
        for (int j = 1; j < D; ++j)
            v = c_add(v, c_mul(c_scale(A[j], sq(index[j])), rd(G, pivot - strides[j], N)));
        G[flat] = c_div_real(v, sq(index[0]));
    }
    return 0;
}

/* core.py:164-213 */
int oracle_stable(const int64_t *shape, int D, const double *A_, const double *b_,
                  const double *c_, double *G_) {
    if (D < 1 || D > MAXD) return -1;
    init_sqrt();
    const cplx *A = (const cplx *)A_;
    const cplx *b = (const cplx *)b_;
    cplx *G = (cplx *)G_;
    int64_t strides[MAXD], nd[MAXD];
    strides_of(shape, D, strides);
    int64_t N = prod_of(shape, D);
    for (int a = 0; a < D; ++a) nd[a] = 0;

    G[0] = *(const cplx *)c_;
    for (int64_t flat = 1; flat < N; ++flat) {
        next_index(nd, shape, D);
        int64_t num_pivots = 0;
        cplx vals = {0.0, 0.0}; /* `vals = 0` unified to complex128 */
        for (int i = 0; i < D; ++i) {
            if (nd[i] == 0) continue;
            num_pivots += 1;
            int64_t pivot = flat - strides[i];
            cplx val = c_mul(b[i], G[pivot]);
            for (int j = 0; j < i; ++j)
                val = c_add(val, c_mul(c_scale(A[i * D + j], sq(nd[j])), rd(G, pivot - strides[j], N)));
            val = c_add(val, c_mul(c_scale(A[i * D + i], sq(nd[i] - 1)), rd(G, pivot - strides[i], N)));
            for (int j = i + 1; j < D; ++j)
                val = c_add(val, c_mul(c_scale(A[i * D + j], sq(nd[j])), rd(G, pivot - strides[j], N)));
            vals = c_add(vals, c_div_real(val, sq(nd[i])));
        }
        G[flat] = c_div_real(vals, (double)num_pivots);
    }
    return 0;
}

/* Two-sum based compensated accumulator (used only by the _kahan variant). */
typedef struct {
    double s, c;
} ksum;

static inline void kadd(ksum *k, double x) {
    double t = k->s + x;
    double bp = t - k->s;
    double err = (k->s - (t - bp)) + (x - bp);
    k->s = t;
    k->c += err;
}

/* gradients.py:41-82.  Outputs dA (D*D complex), db (D), dc (1).
 * kahan == 0: the reference's left-to-right sums, term for term.
 * kahan == 1: the same terms summed with compensation (accurate value). */
static int vjp_impl(const int64_t *shape, int D, const double *G_, const double *c_,
                    const double *g_, double *dA_out, double *db_out, double *dc_out,
                    int kahan) {
    if (D < 1 || D > MAXD) return -1;
    init_sqrt();
    const cplx *G = (const cplx *)G_;
    const cplx *g = (const cplx *)g_;
    int64_t strides[MAXD], idx[MAXD];
    strides_of(shape, D, strides);
    int64_t N = prod_of(shape, D);
    for (int a = 0; a < D; ++a) idx[a] = 0;

    cplx dA[MAXD * MAXD], db[MAXD], dLdA[MAXD * MAXD], dLdb[MAXD];
    ksum kA[MAXD * MAXD][2], kb[MAXD][2];
    memset(dA, 0, sizeof(dA));
    memset(db, 0, sizeof(db));
    memset(dLdA, 0, sizeof(dLdA));
    memset(dLdb, 0, sizeof(dLdb));
    memset(kA, 0, sizeof(kA));
    memset(kb, 0, sizeof(kb));

    for (int64_t flat = 1; flat < N; ++flat) {
        next_index(idx, shape, D);
        for (int i = 0; i < D; ++i) {
            int64_t pivot = flat - strides[i];
            db[i] = s_mul(sq(idx[i]), rd(G, pivot, N));
            if (idx[i] > 1) {
                double w = 0.5 * sq(idx[i]) * sq(idx[i] - 1);
                dA[i * D + i] = s_mul(w, rd(G, pivot - strides[i], N));
            } else {
                dA[i * D + i].re = 0.0;
                dA[i * D + i].im = 0.0;
            }
            for (int j = i + 1; j < D; ++j) {
                double w = sq(idx[i]) * sq(idx[j]);
                dA[i * D + j] = s_mul(w, rd(G, pivot - strides[j], N));
            }
        }
        cplx gk = g[flat];
        for (int e = 0; e < D * D; ++e) {
            cplx t = c_mul(dA[e], gk);
            if (kahan) {
                kadd(&kA[e][0], t.re);
                kadd(&kA[e][1], t.im);
            } else {
                dLdA[e] = c_add(dLdA[e], t);
            }
        }
        for (int i = 0; i < D; ++i) {
            cplx t = c_mul(db[i], gk);
            if (kahan) {
                kadd(&kb[i][0], t.re);
                kadd(&kb[i][1], t.im);
            } else {
                dLdb[i] = c_add(dLdb[i], t);
            }
        }
    }
    if (kahan) {
        for (int e = 0; e < D * D; ++e) {
            dLdA[e].re = kA[e][0].s + kA[e][0].c;
            dLdA[e].im = kA[e][1].s + kA[e][1].c;
        }
        for (int i = 0; i < D; ++i) {
            dLdb[i].re = kb[i][0].s + kb[i][0].c;
            dLdb[i].im = kb[i][1].s + kb[i][1].c;
        }
    }

    /* gradients.py:80  dLdc = np.sum(G * dLdG) / c */
    cplx s = {0.0, 0.0};
    ksum ks[2];
    memset(ks, 0, sizeof(ks));
    for (int64_t f = 0; f < N; ++f) {
        cplx t = c_mul(G[f], g[f]);
        if (kahan) {
            kadd(&ks[0], t.re);
            kadd(&ks[1], t.im);
        } else {
            s = c_add(s, t);
        }
    }
    if (kahan) {
        s.re = ks[0].s + ks[0].c;
        s.im = ks[1].s + ks[1].c;
    }
    cplx dc = c_div(s, *(const cplx *)c_);

    /* gradients.py:82  (dLdA + dLdA.T) / 2 */
    cplx *dA_o = (cplx *)dA_out;
    for (int i = 0; i < D; ++i)
        for (int j = 0; j < D; ++j) dA_o[i * D + j] = c_div_real(c_add(dLdA[i * D + j], dLdA[j * D + i]), 2.0);
    memcpy(db_out, dLdb, sizeof(cplx) * (size_t)D);
    memcpy(dc_out, &dc, sizeof(cplx));
    return 0;
}

int oracle_vjp(const int64_t *shape, int D, const double *G, const double *c, const double *g,
               double *dA, double *db, double *dc) {
    return vjp_impl(shape, D, G, c, g, dA, db, dc, 0);
}

int oracle_vjp_kahan(const int64_t *shape, int D, const double *G, const double *c,
                     const double *g, double *dA, double *db, double *dc) {
    return vjp_impl(shape, D, G, c, g, dA, db, dc, 1);
}

/* The compensated sums of oracle_vjp_kahan split over OpenMP threads: the
 * flat range is cut into fixed chunks, each chunk sums its terms with
 * compensation, and the chunk sums are added with compensation in chunk
 * order.  Deterministic for a given chunk count; same accuracy class as the
 * serial compensated sum.  For lattices too large for a serial check. */
int oracle_vjp_kahan_par(const int64_t *shape, int D, const double *G_, const double *c_,
                         const double *g_, double *dA_out, double *db_out, double *dc_out) {
    if (D < 1 || D > MAXD) return -1;
    init_sqrt();
    const cplx *G = (const cplx *)G_;
    const cplx *g = (const cplx *)g_;
    int64_t strides[MAXD];
    strides_of(shape, D, strides);
    const int64_t N = prod_of(shape, D);
    const int NCH = 256;
    const int NT = D * D + D + 1; /* dA (full D x D), db, sum G g */
    double *part = (double *)calloc((size_t)NCH * NT * 2, sizeof(double));
    if (!part) return -1;
#pragma omp parallel for schedule(dynamic, 1)
    for (int ch = 0; ch < NCH; ++ch) {
        const int64_t f0 = N * ch / NCH, f1 = N * (ch + 1) / NCH;
        ksum acc[MAXD * MAXD + MAXD + 1][2];
        memset(acc, 0, sizeof(ksum) * (size_t)NT * 2);
        int64_t idx[MAXD];
        int64_t rem = f0;
        for (int a = D - 1; a >= 0; --a) {
            idx[a] = rem % shape[a];
            rem /= shape[a];
        }
        for (int64_t flat = f0; flat < f1; ++flat) {
            const cplx gk = g[flat];
            const cplx tg = c_mul(G[flat], gk);
            kadd(&acc[NT - 1][0], tg.re);
            kadd(&acc[NT - 1][1], tg.im);
            if (flat > 0) {
                for (int i = 0; i < D; ++i) {
                    const int64_t pivot = flat - strides[i];
                    const cplx tb = c_mul(s_mul(sq(idx[i]), rd(G, pivot, N)), gk);
                    kadd(&acc[D * D + i][0], tb.re);
                    kadd(&acc[D * D + i][1], tb.im);
                    if (idx[i] > 1) {
                        const double w = 0.5 * sq(idx[i]) * sq(idx[i] - 1);
                        const cplx t = c_mul(s_mul(w, rd(G, pivot - strides[i], N)), gk);
                        kadd(&acc[i * D + i][0], t.re);
                        kadd(&acc[i * D + i][1], t.im);
                    }
                    for (int j = i + 1; j < D; ++j) {
                        const double w = sq(idx[i]) * sq(idx[j]);
                        const cplx t = c_mul(s_mul(w, rd(G, pivot - strides[j], N)), gk);
                        kadd(&acc[i * D + j][0], t.re);
                        kadd(&acc[i * D + j][1], t.im);
                    }
                }
            }
            next_index(idx, shape, D);
        }
        for (int e = 0; e < NT; ++e) {
            part[((size_t)ch * NT + e) * 2 + 0] = acc[e][0].s + acc[e][0].c;
            part[((size_t)ch * NT + e) * 2 + 1] = acc[e][1].s + acc[e][1].c;
        }
    }
    cplx tot[MAXD * MAXD + MAXD + 1];
    for (int e = 0; e < NT; ++e) {
        ksum k[2];
        memset(k, 0, sizeof(k));
        for (int ch = 0; ch < NCH; ++ch) {
            kadd(&k[0], part[((size_t)ch * NT + e) * 2 + 0]);
            kadd(&k[1], part[((size_t)ch * NT + e) * 2 + 1]);
        }
        tot[e].re = k[0].s + k[0].c;
        tot[e].im = k[1].s + k[1].c;
    }
    free(part);
    cplx *dA_o = (cplx *)dA_out;
    for (int i = 0; i < D; ++i)
        for (int j = 0; j < D; ++j) dA_o[i * D + j] = c_div_real(c_add(tot[i * D + j], tot[j * D + i]), 2.0);
    memcpy(db_out, &tot[D * D], sizeof(cplx) * (size_t)D);
    cplx dc = c_div(tot[NT - 1], *(const cplx *)c_);
    memcpy(dc_out, &dc, sizeof(cplx));
    return 0;
}

/* batch.py:54-61: prange over the batch (OpenMP here).  Every triple shares
 * one shape; G is (B, *shape), zero-initialised by the caller. */
int oracle_vanilla_batch(const int64_t *shape, int D, int64_t B, const double *A,
                         const double *b, const double *c, int stable, double *G) {
    int64_t N = prod_of(shape, D);
    int err = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(| : err)
    for (int64_t k = 0; k < B; ++k) {
        const double *Ak = A + 2 * k * D * D;
        const double *bk = b + 2 * k * D;
        const double *ck = c + 2 * k;
        double *Gk = G + 2 * k * N;
        err |= stable ? oracle_stable(shape, D, Ak, bk, ck, Gk) : oracle_vanilla(shape, D, Ak, bk, ck, Gk);
    }
    return err ? -1 : 0;
}

/* gradients.py:104-116 */
int oracle_vjp_batch(const int64_t *shape, int D, int64_t B, const double *G, const double *c,
                     const double *g, double *dA, double *db, double *dc) {
    int64_t N = prod_of(shape, D);
    int err = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(| : err)
    for (int64_t k = 0; k < B; ++k)
        err |= oracle_vjp(shape, D, G + 2 * k * N, c + 2 * k, g + 2 * k * N, dA + 2 * k * D * D,
                          db + 2 * k * D, dc + 2 * k);
    return err ? -1 : 0;
}

/* Uneven batch (one shape per triple), forward then VJP per triple, as a
 * Python loop of unbatched reference calls would do; triples run on OpenMP
 * threads.  offsets[k] is triple k's start in G/g (complex elements);
 * a_off/b_off likewise into A and b.  g == NULL means forward only. */
int oracle_uneven(int64_t n, const int32_t *dims, const int64_t *shapes, const int64_t *a_off,
                  const int64_t *b_off, const int64_t *g_off, const double *A, const double *b,
                  const double *c, double *G, const double *g, double *dA, double *db,
                  double *dc) {
    int err = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(| : err)
    for (int64_t k = 0; k < n; ++k) {
        int D = dims[k];
        const int64_t *shp = shapes + (size_t)k * 8;
        int e = oracle_vanilla(shp, D, A + 2 * a_off[k], b + 2 * b_off[k], c + 2 * k, G + 2 * g_off[k]);
        if (!e && g)
            e = oracle_vjp(shp, D, G + 2 * g_off[k], c + 2 * k, g + 2 * g_off[k], dA + 2 * a_off[k],
                           db + 2 * b_off[k], dc + 2 * k);
        err |= e;
    }
    return err ? -1 : 0;
}

int oracle_max_threads(void) {
#ifdef _OPENMP
    extern int omp_get_max_threads(void);
    return omp_get_max_threads();
#else
    return 1;
#endif
}
