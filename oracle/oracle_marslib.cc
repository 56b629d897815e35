/*
 *  ORACLE -- test infrastructure only, never part of the product.
 *
 *  CPU restatement of the reference MARS forward pass (aleon1138/mars), used
 *  as the FP64 correctness referee for the HIP path; it is itself checked
 *  against the reference compiled from its own sources (oracle/_ref) and the
 *  golden fixtures that build wrote.  Only tests/, __graft_entry__.smoke()
 *  and bench.py's cpu_baseline leg load it.
 *
 *  Every arithmetic step keeps the reference's rounding points (which
 *  products are rounded to f32, which sums run in f64) and, where the
 *  reference uses 4-lane AVX accumulators, the same lane assignment and
 *  horizontal-sum shape (a0+a1)+(a2+a3); products the reference compiler
 *  contracts into FMA are explicit std::fma (this file is compiled
 *  -ffp-contract=off).  Results are not bitwise identical to the reference
 *  build: g++ vectorises some of the reference's sum-of-squares loops as
 *  unfused fold-left sums and others with FMA, so its last bits depend on
 *  per-loop compiler choices.  The restatement reproduces the reference's
 *  term sequences exactly and its per-step mse within 1e-10 relative.
 *
 *  Module `oracle_marslib` mirrors the reference binding marslib.cc:152-179
 *  (MarsAlgo(X,y,w,max_terms), eval, nbasis, yvar, append) and adds
 *  introspection hooks used by the tests (knot indices, basis state, the
 *  kernel-level functions).
 */
#include <pybind11/pybind11.h>
#include <pybind11/numpy.h>
#include <omp.h>
#include <xmmintrin.h>

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

namespace py = pybind11;

namespace oracle {

/* kernels.h:17 -- residual energy below 1/9 of projected energy re-runs GS. */
constexpr double kGateRatioSq = 9.0;

static void check(bool ok, const char *msg)
{
    if (!ok) throw std::runtime_error(msg);
}

/* (a0+a1)+(a2+a3): the reference's horizontal sum (kernels.cc:17-20). */
static inline double hsum(const double a[4])
{
    return (a[0] + a[1]) + (a[2] + a[3]);
}

/*
 *  sum_k (double)a[k] * b[k] with two interleaved 4-lane accumulators over
 *  8-wide blocks, one extra 4-wide block into the first, a lane-wise merge,
 *  horizontal sum, then a scalar FMA tail.  Follows kernels.cc:56-78.
 */
template <typename TA>
static double dot_two_acc(const TA *a, const double *b, size_t m)
{
    double A[4] = {0, 0, 0, 0}, C[4] = {0, 0, 0, 0};
    size_t k = 0;
    for (; k + 8 <= m; k += 8) {
        for (int q = 0; q < 4; ++q) A[q] = std::fma((double)a[k + q], b[k + q], A[q]);
        for (int q = 0; q < 4; ++q) C[q] = std::fma((double)a[k + 4 + q], b[k + 4 + q], C[q]);
    }
    if (k + 4 <= m) {
        for (int q = 0; q < 4; ++q) A[q] = std::fma((double)a[k + q], b[k + q], A[q]);
        k += 4;
    }
    double S[4];
    for (int q = 0; q < 4; ++q) S[q] = A[q] + C[q];
    double acc = hsum(S);
    for (; k < m; ++k) acc = std::fma((double)a[k], b[k], acc);
    return acc;
}

/* Four columns sharing one widened row, one 4-lane accumulator each
 * (kernels.cc:86-118). */
static void dot_four(const float *row, const double *t0, const double *t1,
                     const double *t2, const double *t3, size_t m, double out[4])
{
    const double *t[4] = {t0, t1, t2, t3};
    double acc[4][4] = {{0}};
    size_t k = 0;
    for (; k + 4 <= m; k += 4)
        for (int c = 0; c < 4; ++c)
            for (int q = 0; q < 4; ++q)
                acc[c][q] = std::fma((double)row[k + q], t[c][k + q], acc[c][q]);
    double s[4];
    for (int c = 0; c < 4; ++c) s[c] = hsum(acc[c]);
    for (; k < m; ++k) {
        const double r = (double)row[k];
        for (int c = 0; c < 4; ++c) s[c] = std::fma(r, t[c][k], s[c]);
    }
    for (int c = 0; c < 4; ++c) out[c] = s[c];
}

/* f32 x f32 dot accumulated in f64: two 4-lane accumulators over 8-wide
 * blocks, no 4-wide step, scalar tail (kernels.cc:215-237). */
double dot_widen(const float *a, const float *b, size_t n)
{
    double A[4] = {0, 0, 0, 0}, C[4] = {0, 0, 0, 0};
    size_t i = 0;
    for (; i + 8 <= n; i += 8) {
        for (int q = 0; q < 4; ++q) A[q] = std::fma((double)a[i + q], (double)b[i + q], A[q]);
        for (int q = 0; q < 4; ++q) C[q] = std::fma((double)a[i + 4 + q], (double)b[i + 4 + q], C[q]);
    }
    double S[4];
    for (int q = 0; q < 4; ++q) S[q] = A[q] + C[q];
    double s = hsum(S);
    for (; i < n; ++i) s = std::fma((double)a[i], (double)b[i], s);
    return s;
}

/* t[k] += bo[k]*c (elementwise; kernels.cc:38-52). */
static inline void axpy_row(double *t, const float *bo, double c, size_t m)
{
    for (size_t k = 0; k < m; ++k) t[k] = std::fma((double)bo[k], c, t[k]);
}

/* Subtract Bo*tc from one column, round to f32, return sum of the rounded
 * squares (kernels.cc:154-169). */
static double subtract_round_norm(size_t n, size_t m, const float *Bo, size_t ldBo,
                                  const double *tc, float *bx)
{
    double s = 0.0;
    for (size_t i = 0; i < n; ++i) {
        const double v = (double)bx[i] - dot_two_acc(Bo + i * ldBo, tc, m);
        const double vr = (double)(float)v;
        bx[i] = (float)v;
        s = std::fma(vr, vr, s);
    }
    return s;
}

/* One modified Gram-Schmidt sweep of v against Bo's m columns
 * (kernels.cc:194-211). Returns the projected energy. */
static double mgs_sweep(size_t n, size_t m, double *v, const float *Bo, size_t ldBo)
{
    double e = 0.0;
    for (size_t j = 0; j < m; ++j) {
        double c = 0.0;
        for (size_t i = 0; i < n; ++i) c = std::fma((double)Bo[i * ldBo + j], v[i], c);
        for (size_t i = 0; i < n; ++i) v[i] = std::fma(-c, (double)Bo[i * ldBo + j], v[i]);
        e = std::fma(c, c, e);
    }
    return e;
}

/* kernels.cc:239-281 */
double orthonormalize_col(size_t n, size_t m, double *v, const float *Bo, size_t ldBo,
                          double tol, long *fired)
{
    const double proj = mgs_sweep(n, m, v, Bo, ldBo);
    double vv = 0.0;
    for (size_t i = 0; i < n; ++i) vv = std::fma(v[i], v[i], vv);
    if (vv > tol && vv * kGateRatioSq < proj) {
        if (fired) ++*fired;
        mgs_sweep(n, m, v, Bo, ldBo);
        vv = 0.0;
        for (size_t i = 0; i < n; ++i) vv = std::fma(v[i], v[i], vv);
    }
    const double w = std::sqrt(vv);
    if (w * w > tol)
        for (size_t i = 0; i < n; ++i) v[i] /= w;
    return w;
}

/* kernels.cc:283-417 -- batched CGS with one DGKS retry per column. */
void orthonormalize(size_t n, size_t m, size_t p, const float *B, size_t ldB,
                    const float *x, const int *mask, const float *Bo, size_t ldBo,
                    float *Bx, size_t ldBx, double *T, size_t ldT, double *s,
                    double tol, long *fired)
{
    for (size_t j = 0; j < p; ++j) {
        const float *b = B + (size_t)mask[j] * ldB;
        float *bx = Bx + j * ldBx;
        for (size_t i = 0; i < n; ++i) bx[i] = b[i] * x[i];
    }
    for (size_t j = 0; j < p; ++j) std::fill_n(T + j * ldT, m, 0.0);
    for (size_t i = 0; i < n; ++i) {
        const float *row = Bo + i * ldBo;
        for (size_t j = 0; j < p; ++j) axpy_row(T + j * ldT, row, (double)Bx[i + j * ldBx], m);
    }
    std::fill_n(s, p, 0.0);
    for (size_t i = 0; i < n; ++i) {
        const float *row = Bo + i * ldBo;
        size_t j = 0;
        for (; j + 4 <= p; j += 4) {
            double pr[4];
            dot_four(row, T + j * ldT, T + (j + 1) * ldT, T + (j + 2) * ldT, T + (j + 3) * ldT, m, pr);
            for (int c = 0; c < 4; ++c) {
                float *bx = Bx + (j + c) * ldBx;
                const double v = (double)bx[i] - pr[c];
                const double vr = (double)(float)v;
                bx[i] = (float)v;
                s[j + c] = std::fma(vr, vr, s[j + c]);
            }
        }
        for (; j < p; ++j) {
            float *bx = Bx + j * ldBx;
            const double v = (double)bx[i] - dot_two_acc(row, T + j * ldT, m);
            const double vr = (double)(float)v;
            bx[i] = (float)v;
            s[j] = std::fma(vr, vr, s[j]);
        }
    }
    for (size_t j = 0; j < p; ++j) {
        double *tc = T + j * ldT;
        float *bx = Bx + j * ldBx;
        const double tn = dot_two_acc(tc, tc, m);
        if (s[j] > tol && s[j] * kGateRatioSq < tn) {
            if (fired) ++*fired;
            std::fill_n(tc, m, 0.0);
            for (size_t i = 0; i < n; ++i) axpy_row(tc, Bo + i * ldBo, (double)bx[i], m);
            s[j] = subtract_round_norm(n, m, Bo, ldBo, tc, bx);
        }
    }
    for (size_t j = 0; j < p; ++j) {
        float *bx = Bx + j * ldBx;
        const float scale = (s[j] > tol) ? (float)(1.0 / std::sqrt(s[j] + tol)) : 0.0f;
        for (size_t i = 0; i < n; ++i) bx[i] *= scale;
    }
}

/* Stable descending order of v (marsalgo.cc:33-39): equal keys (including
 * -0.0 == +0.0) keep their row order. */
void argsort_desc(int32_t *idx, const float *v, size_t n)
{
    std::iota(idx, idx + n, 0);
    std::stable_sort(idx, idx + n, [v](int32_t a, int32_t b) { return v[a] > v[b]; });
}

struct Cov {
    double ff, fy;
};

/*
 *  One sorted-row step of the hinge sweep accumulators (marsalgo.cc:91-151):
 *  f += k0*g, g += k1*q over lanes [0,m) (q = Bo row) and lane m (q = xm),
 *  optionally returning ff = |f|^2 and fy = f.[y; ym].  Lanes [0, m - m%4)
 *  use four lane accumulators for ff/fy, the rest continue in scalar order.
 */
template <bool kSse>
static Cov sweep_step(double *f, double *g, const float *q, const double *y, double xm,
                      double ym, double k0, float k1, size_t m)
{
    Cov o = {0, 0};
    const double K1 = (double)k1;
    const size_t m4 = m - m % 4;
    double S0[4] = {0, 0, 0, 0}, S1[4] = {0, 0, 0, 0};
    for (size_t i = 0; i < m4; i += 4) {
        for (int l = 0; l < 4; ++l) {
            const double fv = std::fma(k0, g[i + l], f[i + l]);
            const double gv = std::fma(K1, (double)q[i + l], g[i + l]);
            if (kSse) {
                S0[l] = std::fma(fv, fv, S0[l]);
                S1[l] = std::fma(fv, y[i + l], S1[l]);
            }
            f[i + l] = fv;
            g[i + l] = gv;
        }
    }
    if (kSse) {
        o.ff = hsum(S0);
        o.fy = hsum(S1);
    }
    for (size_t i = m4; i < m; ++i) {
        f[i] = std::fma(k0, g[i], f[i]);
        g[i] = std::fma(K1, (double)q[i], g[i]);
        if (kSse) {
            o.ff = std::fma(f[i], f[i], o.ff);
            o.fy = std::fma(f[i], y[i], o.fy);
        }
    }
    f[m] = std::fma(k0, g[m], f[m]);
    g[m] = std::fma(K1, xm, g[m]);
    if (kSse) {
        o.ff = std::fma(f[m], f[m], o.ff);
        o.fy = std::fma(f[m], ym, o.fy);
    }
    return o;
}

Cov covariates(double *f, double *g, const float *q, const double *y, double xm, double ym,
               double k0, float k1, size_t m)
{
    return sweep_step<true>(f, g, q, y, xm, ym, k0, k1, m);
}

/*
 *  Fit state and the forward-pass operations (marsalgo.cc:185-875). Bo is
 *  kept row-major with a fixed row stride of max_terms (the reference grows
 *  the stride geometrically; the stride does not enter any arithmetic).
 */
class Algo {
public:
    Algo(const float *X, const float *y, const float *w, size_t n, size_t px, size_t max_terms,
         size_t ldx)
        : n_(n), px_(px), ldx_(ldx), X_(X), max_terms_(max_terms),
          tol_((n * 0.02) * DBL_EPSILON), B_(n * max_terms, 0.0f), Bo_(n * max_terms, 0.0f)
    {
        check(!std::isfinite(NAN), "NAN check is disabled, recompile without --fast-math");
        check(max_terms >= 1, "max_terms must be >= 1");
        /* marsalgo.cc:278-307: sqrt(w) weighting, non-finite rows dropped, unit norm. */
        std::vector<double> yd(n), sw(n);
        for (size_t i = 0; i < n; ++i) {
            double yi = (double)y[i];
            double si = std::sqrt((double)w[i]);
            if (!std::isfinite(yi)) yi = si = 0.0;
            sw[i] = si;
            yd[i] = yi * si;
        }
        double yy = 0.0, ww = 0.0;
        for (size_t i = 0; i < n; ++i) {
            yy = std::fma(yd[i], yd[i], yy);
            ww = std::fma(sw[i], sw[i], ww);
        }
        const double yn = std::sqrt(yy), wn = std::sqrt(ww);
        check(yn > 0.0 && wn > 0.0, "target Y is all zero or NANs");
        for (size_t i = 0; i < n; ++i) {
            yd[i] /= yn;
            sw[i] /= wn;
        }
        /* marsalgo.cc:314-337: f32 target, intercept column, its projection. */
        y_.resize(n);
        for (size_t i = 0; i < n; ++i) y_[i] = (float)yd[i];
        for (size_t i = 0; i < n; ++i) B_[i] = Bo_[i * max_terms_] = (float)sw[i];
        double yb = 0.0;
        for (size_t i = 0; i < n; ++i) yb = std::fma((double)Bo_[i * max_terms_], (double)y_[i], yb);
        ybo_.assign(1, yb);
        /* marsalgo.cc:342-352: variance of the normalized target. */
        double ms = 0.0;
        for (size_t i = 0; i < n; ++i) ms += yd[i];
        const double mean = ms / n;
        double vs = 0.0;
        for (size_t i = 0; i < n; ++i) {
            const double dv = yd[i] - mean;
            vs = std::fma(dv, dv, vs);
        }
        yvar_ = vs / n;
        /* marsalgo.cc:359-369: per-column 1/rms scale. */
        s_.resize(px);
        for (size_t j = 0; j < px; ++j) {
            const float *xj = X + j * ldx;
            double sq = 0.0;
            for (size_t i = 0; i < n; ++i) sq = std::fma((double)xj[i], (double)xj[i], sq);
            const double rms = std::sqrt(sq / n);
            check(std::isfinite(rms), "not all columns in X are finite");
            s_[j] = rms > 0.0 ? (float)(1.0 / rms) : 1.0f;
        }
    }

    size_t n() const { return n_; }
    size_t px() const { return px_; }
    size_t nbasis() const { return m_; }
    size_t max_terms() const { return max_terms_; }
    double yvar() const { return yvar_; }
    double tol() const { return tol_; }

    double dsse() const
    {
        double s = 0.0;
        for (double v : ybo_) s = std::fma(v, v, s);
        return s;
    }

    /*
     *  Score one X column against the masked parents (marsalgo.cc:402-662).
     *  Outputs are (m) rows; `knot` (optional) receives the winning sorted
     *  position per parent (-1 if none). `fired` counts DGKS retries.
     */
    void eval(double *lin, double *hin, double *cut, int32_t *knot, int xcol, const bool *mask,
              int min_span, int end_span, bool linear_only, long *fired,
              double *cond = nullptr) const
    {
        check(xcol >= 0 && (size_t)xcol < px_, "invalid X column index");
        check(min_span >= 1, "min_span must be >= 1");
        const size_t m = m_, n = n_, L = max_terms_;
        std::fill_n(lin, m, 0.0);
        std::fill_n(hin, m, 0.0);
        std::fill_n(cut, m, NAN);
        if (knot) std::fill_n(knot, m, -1);
        if (cond) std::fill_n(cond, m, 0.0);

        std::vector<int> par;
        for (size_t j = 0; j < m; ++j)
            if (mask[j]) par.push_back((int)j);
        const size_t p = par.size();
        if (p == 0) return;

        const unsigned csr = _mm_getcsr();
        _mm_setcsr(csr | 0x8040); /* FTZ|DAZ as marsalgo.cc:428-431 */

        std::vector<float> x(n);
        const float *xc = X_ + (size_t)xcol * ldx_;
        const float sx = s_[xcol];
        for (size_t i = 0; i < n; ++i) x[i] = xc[i] * sx;

        std::vector<float> Bx(n * p);
        std::vector<double> T(m * p), ss(p), ybx(p);
        orthonormalize(n, m, p, B_.data(), n, x.data(), par.data(), Bo_.data(), L, Bx.data(), n,
                       T.data(), m, ss.data(), tol_, fired);
        for (size_t j = 0; j < p; ++j) {
            ybx[j] = dot_widen(Bx.data() + j * n, y_.data(), n);
            lin[par[j]] = ybx[j] * ybx[j];
        }

        if (!linear_only) {
            std::vector<int32_t> k(n);
            argsort_desc(k.data(), xc, n);
            std::vector<float> d(n, 0.0f);
            for (size_t i = 1; i < n; ++i) d[i] = x[k[i - 1]] - x[k[i]];
            const size_t head = (size_t)end_span;
            const size_t tail = n - (size_t)end_span;
            std::vector<double> f(m + 1), g(m + 1);
            std::vector<int> best_i(p, -1);
            std::vector<double> best(p, 0.0), best_cond(p, 0.0);
            for (size_t j = 0; j < p; ++j) {
                std::fill(f.begin(), f.end(), 0.0);
                std::fill(g.begin(), g.end(), 0.0);
                const float *b = B_.data() + (size_t)par[j] * n;
                const float *bx = Bx.data() + j * n;
                double bk = b[k[0]], bxk = bx[k[0]], yk = y_[k[0]];
                sweep_step<false>(f.data(), g.data(), Bo_.data() + (size_t)k[0] * L, ybo_.data(),
                                  bxk, ybx[j], 0, (float)bk, m);
                double k0 = 0, k1 = 0, w = 0, bd = 0;
                double vb = bk * yk, b2 = bk * bk;
                size_t grid = head + 1;
                for (size_t i = 1; i < tail; ++i) {
                    bk = b[k[i]];
                    bxk = bx[k[i]];
                    yk = y_[k[i]];
                    const double di = d[i];
                    const bool on = (i == grid);
                    if (on) grid += min_span;
                    const float *q = Bo_.data() + (size_t)k[i] * L;
                    const Cov o = on ? sweep_step<true>(f.data(), g.data(), q, ybo_.data(), bxk,
                                                        ybx[j], di, (float)bk, m)
                                     : sweep_step<false>(f.data(), g.data(), q, ybo_.data(), bxk,
                                                         ybx[j], di, (float)bk, m);
                    k0 = std::fma(di * di, b2, k0);
                    k1 = std::fma(di * 2, bd, k1);
                    w = std::fma(di, vb, w);
                    bd = std::fma(di, b2, bd);
                    b2 = std::fma(bk, bk, b2);
                    vb = std::fma(yk, bk, vb);
                    if (on) {
                        const double uw = o.fy - w;
                        const double den = (k0 + k1) - o.ff;
                        const double sse = den > tol_ ? (uw * uw) / (den + tol_) : 0;
                        if (sse > best[j]) {
                            best[j] = sse;
                            best_i[j] = (int)i;
                            best_cond[j] = (k0 + k1) / den;
                        }
                    }
                }
            }
            for (size_t j = 0; j < p; ++j) {
                if (best_i[j] >= 0) {
                    hin[par[j]] = lin[par[j]] + best[j];
                    cut[par[j]] = X_[(size_t)xcol * ldx_ + k[best_i[j]]];
                    if (knot) knot[par[j]] = best_i[j];
                    if (cond) cond[par[j]] = best_cond[j];
                }
            }
        }
        _mm_setcsr(csr);
    }

    /* marsalgo.cc:760-875 */
    double append(char type, int xcol, int bcol, float h, long *fired)
    {
        if (m_ >= max_terms_) throw std::runtime_error("basis matrix is full");
        if (bcol < 0 || (size_t)bcol >= m_) {
            char msg[80];
            snprintf(msg, sizeof(msg), "invalid basis column number: %d", bcol);
            throw std::runtime_error(msg);
        }
        check(xcol >= 0 && (size_t)xcol < px_, "invalid X column index");
        const size_t n = n_, L = max_terms_;
        const float s = s_[xcol];
        const float *bp = B_.data() + (size_t)bcol * n;
        const float *xp = X_ + (size_t)xcol * ldx_;
        float *Bm = B_.data() + m_ * n;
        switch (type) {
        case 'l':
            for (size_t i = 0; i < n; ++i) Bm[i] = s * bp[i] * xp[i];
            break;
        case '+':
            for (size_t i = 0; i < n; ++i) Bm[i] = s * bp[i] * std::max(xp[i] - h, 0.0f);
            break;
        case '-':
            for (size_t i = 0; i < n; ++i) Bm[i] = s * bp[i] * std::max(h - xp[i], 0.0f);
            break;
        default:
            throw std::runtime_error("invalid basis type");
        }
        std::vector<double> v(n);
        for (size_t i = 0; i < n; ++i) v[i] = (double)Bm[i];
        double r2 = 0.0;
        for (size_t i = 0; i < n; ++i) r2 = std::fma(v[i], v[i], r2);
        const float rn = (float)std::sqrt(r2);
        for (size_t i = 0; i < n; ++i) Bm[i] /= rn;
        const double w = orthonormalize_col(n, m_, v.data(), Bo_.data(), L, tol_, fired);
        if (w * w > tol_) {
            for (size_t i = 0; i < n; ++i) Bo_[i * L + m_] = (float)v[i];
            double yb = 0.0;
            for (size_t i = 0; i < n; ++i) yb = std::fma((double)Bo_[i * L + m_], (double)y_[i], yb);
            double sq = yb * yb;
            for (double e : ybo_) sq = std::fma(e, e, sq);
            const double mse = (1. - sq) / n;
            if (mse >= -tol_) {
                m_ += 1;
                ybo_.push_back(yb);
                return std::max(mse, 0.0);
            }
        }
        return -1.;
    }

    /* Test hook: replace the basis by m columns given as B (n, m) and Bo (n, m), both
     * C-order, and recompute y . Bo[:, k] as append does, so that eval scores candidates
     * against a basis built elsewhere (the device's). */
    void load_basis(const float *B, const float *Bo, size_t m)
    {
        check(m >= 1 && m <= max_terms_, "basis size out of range");
        const size_t n = n_, L = max_terms_;
        ybo_.clear();
        for (size_t j = 0; j < m; ++j) {
            double yb = 0.0;
            for (size_t i = 0; i < n; ++i) {
                B_[j * n + i] = B[i * m + j];
                Bo_[i * L + j] = Bo[i * m + j];
                yb = std::fma((double)Bo_[i * L + j], (double)y_[i], yb);
            }
            ybo_.push_back(yb);
        }
        m_ = m;
    }

    const std::vector<float> &B() const { return B_; }
    const std::vector<float> &Bo() const { return Bo_; }
    const std::vector<float> &y() const { return y_; }
    const std::vector<double> &ybo() const { return ybo_; }
    const std::vector<float> &s() const { return s_; }

private:
    size_t n_, px_, ldx_;
    const float *X_;
    size_t max_terms_;
    double tol_;
    size_t m_ = 1;
    double yvar_ = 0;
    std::vector<float> B_, Bo_, y_, s_;
    std::vector<double> ybo_;
};

} // namespace oracle

/* ------------------------------------------------------------------------- */
/*  Python binding (mirrors marslib.cc:12-179)                               */
/* ------------------------------------------------------------------------- */

using oracle::Algo;

struct PyAlgo {
    py::array_t<float> X;  /* keeps the borrowed X alive */
    std::unique_ptr<Algo> a;
    long fired = 0;
};

static PyAlgo *new_algo(py::array_t<float> X, py::array_t<float> y, py::array_t<float> w,
                        int max_terms)
{
    py::buffer_info xi = X.request(), yi = y.request(), wi = w.request();
    oracle::check(xi.ndim == 2, "expected 2D array for X");
    oracle::check(yi.ndim == 1, "expected 1D array for y");
    oracle::check(wi.ndim == 1, "expected 1D array for w");
    oracle::check(xi.strides[0] == sizeof(float), "X must be column-major");
    oracle::check(yi.strides[0] == sizeof(float), "y must be column-major");
    oracle::check(wi.strides[0] == sizeof(float), "w must be column-major");
    oracle::check(xi.shape[0] == yi.shape[0], "invalid dimension for y");
    oracle::check(xi.shape[0] == wi.shape[0], "invalid dimension for w");
    auto *p = new PyAlgo;
    p->X = X;
    p->a.reset(new Algo((const float *)xi.ptr, (const float *)yi.ptr, (const float *)wi.ptr,
                        xi.shape[0], xi.shape[1], max_terms, xi.strides[1] / sizeof(float)));
    return p;
}

static py::tuple eval_impl(PyAlgo &pa, py::array_t<bool> mask, int min_span, int endspan,
                           bool linear, int threads, bool cuda, bool with_knots)
{
    Algo &a = *pa.a;
    py::buffer_info mi = mask.request();
    oracle::check(mi.ndim == 2, "expected 2D array for mask");
    oracle::check(mi.strides[1] == sizeof(bool), "mask must be row-major");
    oracle::check(mi.shape[1] == (ssize_t)a.nbasis(), "invalid dimension for mask");
    oracle::check(min_span >= 1, "min_span must be >= 1");
    oracle::check(!cuda, "the oracle is CPU-only");
    const ssize_t rows = mi.shape[0], cols = mi.shape[1];
    const ssize_t rs = mi.strides[0] / sizeof(bool);
    py::array_t<double> d1({rows, cols}), d2({rows, cols}), hc({rows, cols});
    py::array_t<int32_t> kn({rows, cols});
    py::array_t<double> cd({rows, cols});
    const int nprocs = omp_get_num_procs();
    threads = threads <= 0 ? std::min(64, nprocs) : std::min(threads, nprocs);
    threads = std::max(1, (int)std::min<ssize_t>(threads, rows));
    const bool *mb = (const bool *)mi.ptr;
    double *p1 = d1.mutable_data(), *p2 = d2.mutable_data(), *pc = hc.mutable_data();
    int32_t *pk = kn.mutable_data();
    double *pcd = cd.mutable_data();
    std::atomic<bool> ok{true};
    std::atomic<long> fired{0};
    std::string err;
    {
        py::gil_scoped_release nogil;
#pragma omp parallel for num_threads(threads) schedule(static)
        for (ssize_t i = 0; i < rows; ++i) {
            if (!ok.load()) continue;
            try {
                long f = 0;
                a.eval(p1 + i * cols, p2 + i * cols, pc + i * cols, pk + i * cols, (int)i,
                       mb + i * rs, min_span, endspan, linear, &f, pcd + i * cols);
                fired += f;
            } catch (const std::exception &e) {
#pragma omp critical
                if (err.empty()) err = e.what();
                ok = false;
            }
        }
    }
    if (!ok) throw std::runtime_error(err);
    pa.fired += fired.load();
    if (with_knots) return py::make_tuple(a.dsse(), d1, d2, hc, kn, cd);
    return py::make_tuple(a.dsse(), d1, d2, hc);
}

template <typename T>
static py::array_t<T> vec_copy(const std::vector<T> &v)
{
    py::array_t<T> out(v.size());
    std::copy(v.begin(), v.end(), out.mutable_data());
    return out;
}

PYBIND11_MODULE(oracle_marslib, m)
{
    m.doc() = "ORACLE (test infrastructure): CPU restatement of the reference MARS forward pass";
    m.attr("__version__") = "0.12";

    py::class_<PyAlgo>(m, "MarsAlgo")
        .def(py::init(&new_algo), py::arg("X").noconvert(), py::arg("y").noconvert(),
             py::arg("w").noconvert(), py::arg("max_terms"))
        .def(
            "eval",
            [](PyAlgo &pa, py::array_t<bool> mask, int min_span, int endspan, bool linear_only,
               int threads, bool cuda) {
                return eval_impl(pa, mask, min_span, endspan, linear_only, threads, cuda, false);
            },
            py::arg("mask").noconvert(), py::arg("min_span"), py::arg("endspan"),
            py::arg("linear_only"), py::arg("threads"), py::arg("cuda") = false)
        .def(
            "eval_knots",
            [](PyAlgo &pa, py::array_t<bool> mask, int min_span, int endspan, bool linear_only,
               int threads) {
                return eval_impl(pa, mask, min_span, endspan, linear_only, threads, false, true);
            },
            py::arg("mask").noconvert(), py::arg("min_span"), py::arg("endspan"),
            py::arg("linear_only"), py::arg("threads") = -1)
        .def("nbasis", [](PyAlgo &pa) { return (int)pa.a->nbasis(); })
        .def("yvar", [](PyAlgo &pa) { return pa.a->yvar(); })
        .def("dsse", [](PyAlgo &pa) { return pa.a->dsse(); })
        .def("tol", [](PyAlgo &pa) { return pa.a->tol(); })
        .def("dgks_fired", [](PyAlgo &pa) { return pa.fired; })
        .def("append",
             [](PyAlgo &pa, char type, int xcol, int bcol, float h) {
                 return pa.a->append(type, xcol, bcol, h, &pa.fired);
             })
        .def("y", [](PyAlgo &pa) { return vec_copy(pa.a->y()); })
        .def("ybo", [](PyAlgo &pa) { return vec_copy(pa.a->ybo()); })
        .def("scales", [](PyAlgo &pa) { return vec_copy(pa.a->s()); })
        .def("load_basis",
             [](PyAlgo &pa, py::array_t<float, py::array::c_style | py::array::forcecast> B,
                py::array_t<float, py::array::c_style | py::array::forcecast> Bo) {
                 oracle::check(B.ndim() == 2 && Bo.ndim() == 2 && B.shape(0) == Bo.shape(0) &&
                                   B.shape(1) == Bo.shape(1) && (size_t)B.shape(0) == pa.a->n(),
                               "load_basis: B and Bo must both be (n, m)");
                 pa.a->load_basis(B.data(), Bo.data(), B.shape(1));
             })
        .def("basis",
             [](PyAlgo &pa) {
                 /* (B, Bo) restricted to the live columns, both (n, m) C-order. */
                 const Algo &a = *pa.a;
                 const size_t n = a.n(), m = a.nbasis(), L = a.max_terms();
                 py::array_t<float> B({(ssize_t)n, (ssize_t)m}), Bo({(ssize_t)n, (ssize_t)m});
                 float *pb = B.mutable_data(), *po = Bo.mutable_data();
                 for (size_t i = 0; i < n; ++i)
                     for (size_t j = 0; j < m; ++j) {
                         pb[i * m + j] = a.B()[j * n + i];
                         po[i * m + j] = a.Bo()[i * L + j];
                     }
                 return py::make_tuple(B, Bo);
             });

    /* Kernel-level hooks for the known-answer tests. */
    m.def("dot_widen", [](py::array_t<float, py::array::c_style> a, py::array_t<float, py::array::c_style> b) {
        oracle::check(a.size() == b.size(), "length mismatch");
        return oracle::dot_widen(a.data(), b.data(), a.size());
    });
    m.def("argsort", [](py::array_t<float, py::array::c_style> v) {
        py::array_t<int32_t> k(v.size());
        oracle::argsort_desc(k.mutable_data(), v.data(), v.size());
        return k;
    });
    m.def(
        "orthonormalize",
        [](py::array_t<float, py::array::f_style> B, py::array_t<float, py::array::c_style> x,
           py::array_t<int32_t, py::array::c_style> mask, py::array_t<float, py::array::c_style> Bo,
           double tol) {
            /* B (n, k) F-order; Bo (n, m) C-order; returns (Bx (n,p) F-order, fired). */
            const size_t n = B.shape(0), p = mask.size(), m = Bo.shape(1);
            oracle::check((size_t)x.size() == n && (size_t)Bo.shape(0) == n, "shape mismatch");
            for (size_t j = 0; j < p; ++j)
                oracle::check(mask.data()[j] >= 0 && mask.data()[j] < B.shape(1), "mask out of range");
            py::array_t<float, py::array::f_style> Bx({(ssize_t)n, (ssize_t)p});
            std::vector<double> T(std::max<size_t>(1, m * p)), s(std::max<size_t>(1, p));
            long fired = 0;
            oracle::orthonormalize(n, m, p, B.data(), n, x.data(), mask.data(), Bo.data(), m,
                                   Bx.mutable_data(), n, T.data(), m, s.data(), tol, &fired);
            return py::make_tuple(Bx, fired);
        });
    m.def("orthonormalize_col",
          [](py::array_t<double, py::array::c_style> v, py::array_t<float, py::array::c_style> Bo,
             double tol) {
              const size_t n = v.size(), m = Bo.shape(1);
              py::array_t<double> out(n);
              std::copy(v.data(), v.data() + n, out.mutable_data());
              long fired = 0;
              const double w = oracle::orthonormalize_col(n, m, out.mutable_data(), Bo.data(), m,
                                                          tol, &fired);
              return py::make_tuple(out, w, fired);
          });
    m.def("covariates", [](py::array_t<double, py::array::c_style> f, py::array_t<double, py::array::c_style> g,
                           py::array_t<float, py::array::c_style> q, py::array_t<double, py::array::c_style> y,
                           double xm, double ym, double k0, float k1) {
        const size_t m = q.size();
        oracle::check((size_t)f.size() == m + 1 && (size_t)g.size() == m + 1 && (size_t)y.size() == m,
                      "shape mismatch");
        oracle::Cov o = oracle::covariates(f.mutable_data(), g.mutable_data(), q.data(), y.data(),
                                           xm, ym, k0, k1, m);
        return py::make_tuple(o.ff, o.fy);
    });
}
