// Backward stepwise selection of the basis terms (Friedman 1991, section 3.4)
// and the final coefficient solve, on the host, behind the C ABI
// (mars_hip_prune_gram).  The O(n M^2) Gram products that feed it run on the
// device (mars_hip_gram); what is left here is O(M^3) work on M x M matrices,
// M a few hundred at most.
//
// Selection rule (the one the reference's mars.prune applies, mars.py:493-698):
//   * columns are standardised to a unit Gram diagonal; zero-norm columns are
//     never active; `ridge` loads the standardised diagonal;
//   * the state of an active set is H = G^-1 (explicit, G = XX_act + ridge I),
//     beta = G^-1 XY_act and SSE = max(YY - XY_act . beta, 0), from a Cholesky
//     factor; a failed factorisation is retried once with the diagonal loaded
//     by 1e-8 (trace / k + 1);
//   * a step removes the active column j with the smallest beta_j^2 / H_jj (the
//     SSE increase of dropping it) among those with H_jj above
//     1e-12 max(max_j H_jj, 1); the intercept (global column 0) is never
//     removed; SSE grows by that amount, and H, beta are downdated by the
//     rank-1 Schur complement of column j, then H is symmetrised and its
//     diagonal clamped at 0;
//   * every max(64, M / 4) removals the state is rebuilt from a fresh factor;
//   * the subset kept is the one with the lowest GCV = SSE / (1 - dof / n)^2,
//     dof = k + penalty (k - 1); only a strictly lower GCV replaces it;
//   * the coefficients of the kept subset are the minimum-norm least-squares
//     solution of G beta = XY (ridge-loaded G), back in the original scale.
#include "../../include/mars_hip.h"

#include <algorithm>
#include <cmath>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

namespace mh {

/* The calling thread's last-error text (defined with the rest of the C ABI). */
void set_last_error(const char *msg);

namespace {

/* Square matrix, row-major, k x k. */
struct Mat {
    int k = 0;
    std::vector<double> a;
    explicit Mat(int n = 0) : k(n), a((size_t)n * n, 0.0) {}
    double &operator()(int i, int j) { return a[(size_t)i * k + j]; }
    double operator()(int i, int j) const { return a[(size_t)i * k + j]; }
};

/* Lower Cholesky factor of S in place (upper triangle left as is); false when a pivot is not
 * positive (S not numerically positive definite). */
bool cholesky(Mat &S)
{
    const int k = S.k;
    for (int j = 0; j < k; ++j) {
        double d = S(j, j);
        for (int p = 0; p < j; ++p) d -= S(j, p) * S(j, p);
        if (!(d > 0.0)) return false;
        const double l = std::sqrt(d);
        S(j, j) = l;
        for (int i = j + 1; i < k; ++i) {
            double s = S(i, j);
            for (int p = 0; p < j; ++p) s -= S(i, p) * S(j, p);
            S(i, j) = s / l;
        }
    }
    return true;
}

/* x <- (L L^T)^-1 x. */
void chol_solve(const Mat &L, double *x)
{
    const int k = L.k;
    for (int i = 0; i < k; ++i) {
        double s = x[i];
        for (int p = 0; p < i; ++p) s -= L(i, p) * x[p];
        x[i] = s / L(i, i);
    }
    for (int i = k - 1; i >= 0; --i) {
        double s = x[i];
        for (int p = i + 1; p < k; ++p) s -= L(p, i) * x[p];
        x[i] = s / L(i, i);
    }
}

/* State of one active set: its global columns, H = G^-1, beta, SSE. */
struct Active {
    std::vector<int> col;
    Mat H;
    std::vector<double> beta;
    double sse = 0.0;
};

/* Rebuild H, beta and SSE of `st.col` from a fresh Cholesky factor. */
void refactor(const double *XX, int M, const double *XY, double YY, double ridge, Active &st)
{
    const int k = (int)st.col.size();
    st.H = Mat(k);
    st.beta.assign(k, 0.0);
    if (k == 0) {
        st.sse = YY;
        return;
    }
    Mat G(k);
    double tr = 0.0;
    for (int i = 0; i < k; ++i)
        for (int j = 0; j < k; ++j) G(i, j) = XX[(size_t)st.col[i] * M + st.col[j]];
    if (ridge > 0)
        for (int i = 0; i < k; ++i) G(i, i) += ridge;
    for (int i = 0; i < k; ++i) tr += G(i, i);
    Mat L = G;
    if (!cholesky(L)) {
        // not numerically positive definite (collinear columns): load the diagonal once so the
        // explicit inverse exists and the rank-1 downdates stay valid
        L = G;
        const double jit = 1e-8 * (tr / k + 1.0);
        for (int i = 0; i < k; ++i) L(i, i) += jit;
        if (!cholesky(L)) throw std::runtime_error("prune: Gram matrix is not positive semi-definite");
    }
    std::vector<double> b(k);
    for (int i = 0; i < k; ++i) b[i] = XY[st.col[i]];
    st.beta = b;
    chol_solve(L, st.beta.data());
    std::vector<double> e(k);
    for (int c = 0; c < k; ++c) {
        std::fill(e.begin(), e.end(), 0.0);
        e[c] = 1.0;
        chol_solve(L, e.data());
        for (int r = 0; r < k; ++r) st.H(r, c) = e[r];
    }
    double fit = 0.0;
    for (int i = 0; i < k; ++i) fit += b[i] * st.beta[i];
    st.sse = std::max(YY - fit, 0.0);
}

/* Remove active position j: SSE rise beta_j^2 / H_jj was added by the caller; downdate H and
 * beta by the Schur complement of j, symmetrise, clamp the diagonal, and drop row/column j. */
void remove_column(Active &st, int j)
{
    const int k = st.H.k;
    std::vector<double> hj(k), v(k);
    for (int i = 0; i < k; ++i) hj[i] = st.H(i, j);
    const double pivot = hj[j];
    for (int i = 0; i < k; ++i) v[i] = hj[i] / pivot;
    const double bj = st.beta[j];
    for (int i = 0; i < k; ++i) st.beta[i] -= bj * v[i];
    Mat N(k - 1);
    int ri = 0;
    for (int r = 0; r < k; ++r) {
        if (r == j) continue;
        int ci = 0;
        for (int c = 0; c < k; ++c) {
            if (c == j) continue;
            const double x = st.H(r, c) - v[r] * hj[c];
            const double y = st.H(c, r) - v[c] * hj[r];
            double s = 0.5 * (x + y);
            if (r == c && s < 0.0) s = 0.0;
            N(ri, ci) = s;
            ++ci;
        }
        ++ri;
    }
    st.H = std::move(N);
    st.beta.erase(st.beta.begin() + j);
    st.col.erase(st.col.begin() + j);
}

/* Minimum-norm least-squares solution of the symmetric system S x = b (S positive
 * semi-definite up to rounding) through a cyclic Jacobi eigendecomposition: eigenvalues with
 * |lambda| <= eps * k * max|lambda| are treated as zero, the cut-off of an SVD-based solve. */
std::vector<double> sym_lstsq(Mat S, const std::vector<double> &b)
{
    const int k = S.k;
    Mat V(k);
    for (int i = 0; i < k; ++i) V(i, i) = 1.0;
    for (int sweep = 0; sweep < 64; ++sweep) {
        double off = 0.0, tot = 0.0;
        for (int i = 0; i < k; ++i)
            for (int j = 0; j < k; ++j) {
                tot += S(i, j) * S(i, j);
                if (i != j) off += S(i, j) * S(i, j);
            }
        if (off <= 1e-30 * tot || off == 0.0) break;
        for (int p = 0; p < k - 1; ++p)
            for (int q = p + 1; q < k; ++q) {
                const double apq = S(p, q);
                if (apq == 0.0) continue;
                const double theta = (S(q, q) - S(p, p)) / (2.0 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int r = 0; r < k; ++r) {  // S <- S J (columns p, q)
                    const double sp = S(r, p), sq = S(r, q);
                    S(r, p) = c * sp - s * sq;
                    S(r, q) = s * sp + c * sq;
                }
                for (int r = 0; r < k; ++r) {  // S <- J^T S (rows p, q)
                    const double sp = S(p, r), sq = S(q, r);
                    S(p, r) = c * sp - s * sq;
                    S(q, r) = s * sp + c * sq;
                }
                for (int r = 0; r < k; ++r) {
                    const double vp = V(r, p), vq = V(r, q);
                    V(r, p) = c * vp - s * vq;
                    V(r, q) = s * vp + c * vq;
                }
            }
    }
    double lmax = 0.0;
    for (int i = 0; i < k; ++i) lmax = std::max(lmax, std::fabs(S(i, i)));
    const double cut = std::numeric_limits<double>::epsilon() * k * lmax;
    std::vector<double> x(k, 0.0);
    for (int e = 0; e < k; ++e) {
        const double lam = S(e, e);
        if (!(std::fabs(lam) > cut)) continue;
        double proj = 0.0;
        for (int r = 0; r < k; ++r) proj += V(r, e) * b[r];
        proj /= lam;
        for (int r = 0; r < k; ++r) x[r] += V(r, e) * proj;
    }
    return x;
}

double gcv(double sse, int k, double penalty, double n_true)
{
    const double denom = 1.0 - (k + penalty * (k - 1)) / n_true;
    return denom > 0 ? sse / (denom * denom) : std::numeric_limits<double>::infinity();
}

/* Elimination over the standardised Gram (XX M x M row-major, XY, YY); `keep` in: candidate
 * columns, out: the lowest-GCV subset. */
void eliminate(const double *XX, int M, const double *XY, double YY, double penalty, double n_true, double ridge,
               std::vector<char> &keep)
{
    Active st;
    for (int i = 0; i < M; ++i)
        if (keep[i]) st.col.push_back(i);
    if (st.col.empty()) return;
    refactor(XX, M, XY, YY, ridge, st);
    std::vector<char> cur = keep;
    double best = gcv(st.sse, (int)st.col.size(), penalty, n_true);
    const int period = std::max(64, M / 4);
    int since = 0;
    const double inf = std::numeric_limits<double>::infinity();
    while (!st.col.empty()) {
        const int k = (int)st.col.size();
        double hmax = -inf;
        for (int i = 0; i < k; ++i) hmax = std::max(hmax, st.H(i, i));
        const double floor = 1e-12 * std::max(hmax, 1.0);
        int j = -1;
        double rise = inf;
        for (int i = 0; i < k; ++i) {
            if (i == 0 && st.col[0] == 0) continue;  // the intercept stays
            const double h = st.H(i, i);
            if (!(h > floor)) continue;
            const double r = st.beta[i] * st.beta[i] / h;
            if (r < rise) {  // first of equal minima
                rise = r;
                j = i;
            }
        }
        if (j < 0 || !(rise < inf)) break;
        st.sse += std::max(rise, 0.0);
        cur[st.col[j]] = 0;
        remove_column(st, j);
        if (++since >= period && !st.col.empty()) {
            refactor(XX, M, XY, YY, ridge, st);
            since = 0;
        }
        const double g = gcv(st.sse, (int)st.col.size(), penalty, n_true);
        if (g < best) {
            best = g;
            keep = cur;
        }
    }
}

} // namespace
} // namespace mh

extern "C" int mars_hip_prune_gram(const double *XX, size_t ldXX, const double *XY, double YY, size_t M,
                                   const unsigned char *mask, double penalty, double n_true, double ridge,
                                   double *beta, unsigned char *chosen)
{
    using namespace mh;
    try {
        if (!XX || !XY || !beta || (M > 0 && ldXX < M)) throw std::invalid_argument("prune: bad argument");
        if (!(n_true > 0)) throw std::invalid_argument("prune: n_true must be positive");
        const int m = (int)M;
        // standardise: unit diagonal on the columns with nonzero norm
        std::vector<double> scale(M, 1.0);
        std::vector<char> keep(M, 0);
        for (int i = 0; i < m; ++i) {
            const double d = std::sqrt(XX[(size_t)i * ldXX + i]);
            if (d > 0) {
                scale[i] = d;
                keep[i] = mask ? (mask[i] != 0) : 1;
            }
        }
        std::vector<double> S((size_t)M * M), b(M);
        for (int i = 0; i < m; ++i) {
            for (int j = 0; j < m; ++j) S[(size_t)i * M + j] = XX[(size_t)i * ldXX + j] / scale[j] / scale[i];
            b[i] = XY[i] / scale[i];
        }
        eliminate(S.data(), m, b.data(), YY, penalty, n_true, ridge, keep);
        std::vector<int> sel;
        for (int i = 0; i < m; ++i)
            if (keep[i]) sel.push_back(i);
        std::fill(beta, beta + M, 0.0);
        if (!sel.empty()) {
            const int k = (int)sel.size();
            Mat G(k);
            std::vector<double> rhs(k);
            for (int r = 0; r < k; ++r) {
                for (int c = 0; c < k; ++c) G(r, c) = S[(size_t)sel[r] * M + sel[c]];
                if (ridge > 0) G(r, r) += ridge;
                rhs[r] = b[sel[r]];
            }
            const std::vector<double> x = sym_lstsq(G, rhs);
            for (int r = 0; r < k; ++r) beta[sel[r]] = x[r] / scale[sel[r]];
        }
        if (chosen)
            for (int i = 0; i < m; ++i) chosen[i] = keep[i] ? 1 : 0;
        set_last_error("");
        return MARS_HIP_OK;
    } catch (const std::exception &e) {
        set_last_error(e.what());
        return MARS_HIP_EINVAL;
    }
}
