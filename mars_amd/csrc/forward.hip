// The forward pass of one resident fit, driven in C++: mars.py:285-406 for the
// default configuration of mars.fit (Fast-MARS ranking off: max_basis and
// max_inputs None; no per-candidate filter callback).
//
// Per step: the candidate mask from the degree / self-interaction filter
// (mars.py:237-239, 319-323), built from each basis function's variable set;
// one device eval with its argmax (fit_eval: scoring, hinge sweep, first-maximum
// winners; the (p_x, m) score arrays stay on the device); the GCV choice between
// the linear term and the hinge pair (mars.py:342-357); the appends on the device
// (mars.py:359-384); the stopping rules (mars.py:393-406).  The model fields r2
// and r2_cv are float32, and numpy 2 compares them with Python floats in
// float32 (a Python float is a weak scalar), so the rules compare in float.
//
// With Fast-MARS ranking on, mars.fit keeps its numpy loop: the ranking's
// argsorts order ties by numpy's own sort implementation (its SIMD sorting
// networks on AVX-512 hosts), which a restatement could not promise to match.
#include "internal.h"

#include <chrono>
#include <cmath>

namespace mh {
namespace {

double append_one(mars_hip_ctx *ctx, int type, int xcol, int bcol, float h)
{
    std::vector<AppendArg> args{AppendArg{0, xcol, bcol, 1, h, 0}};
    args[0].type = type;
    std::vector<AppendOut> out;
    fit_append_batch(ctx, args, out);
    return out[0].mse;
}

void fit_forward(mars_hip_ctx *ctx, const mars_hip_forward_opts &o, mars_hip_model_row *rows, int *nrows,
                 mars_hip_step *steps, int steps_cap, int *nsteps)
{
    using clock = std::chrono::steady_clock;
    const auto t0 = clock::now();
    require(ctx->resident && ctx->fits.size() == 1, "mars_hip_fit_forward needs a single-fit resident context");
    FitState &fs = ctx->fits[0];
    require(fs.m == 1, "mars_hip_fit_forward starts from a fresh fit (nbasis 1)");
    require(o.max_terms >= 1 && (size_t)o.max_terms <= ctx->max_terms, "max_terms out of range");
    require(o.max_degree >= 0, "max_degree must be >= 0");
    require(o.n_true > 0, "n_true must be > 0");
    require(o.min_span >= 1, "min_span must be >= 1");
    require(o.end_span >= 0 && (size_t)o.end_span < fs.n, "end_span must be in [0, n)");
    require(rows != nullptr && nrows != nullptr, "null output");
    require(steps == nullptr || steps_cap >= 0, "negative step capacity");

    const size_t px = fs.px;
    const double n = (double)fs.n;  // len(X): the mse of a choice is an SSE over the rows (mars.py:351-352)
    auto dof = [&](double m) { return m + o.penalty * (m - 1); };
    auto gcv = [&](double mse, double m) { return mse / std::pow(1.0 - dof(m) / o.n_true, 2); };
    auto seconds = [&] { return std::chrono::duration<double>(clock::now() - t0).count(); };

    // variables of each basis function, as a degree and a membership row
    std::vector<int> deg(1, 0);
    std::vector<unsigned char> has(px, 0);  // has[b * px + i]: X column i occurs in basis function b
    std::vector<unsigned char> mask;
    int R = 0, S = 0;
    rows[R++] = mars_hip_model_row{'i', 0, 0, 0, std::nan(""), 0.0, 0.0f, 0.0f, 0.0f, 0};

    while ((int)fs.m < o.max_terms) {
        const size_t m = fs.m;
        mask.assign(px * m, 0);
        int n_cand = 0;
        for (size_t b = 0; b < m; ++b) {
            if (deg[b] >= o.max_degree) continue;
            const unsigned char *hb = &has[b * px];
            for (size_t i = 0; i < px; ++i)
                if (o.self_interact || !hb[i]) {
                    mask[i * m + b] = 1;
                    ++n_cand;
                }
        }
        mars_hip_best best;
        fit_eval(ctx, mask.data(), px, m, o.min_span, o.end_span, o.linear_only, nullptr, nullptr, nullptr, &best);
        const double dt = seconds();
        if (best.sse1 <= 0 && best.sse2 <= 0) break;  // all inputs filtered or constant (mars.py:345-346)

        const size_t m0 = m;
        const double mse1 = gcv((1.0 - best.sse0 - best.sse1) / n, (double)m + 1);
        const double mse2 = gcv((1.0 - best.sse0 - best.sse2) / n, (double)m + 2);
        const bool linear = mse1 <= mse2;
        const int xcol = linear ? best.xcol1 : best.xcol2, bcol = linear ? best.bcol1 : best.bcol2;
        const double cut = linear ? std::nan("") : best.cut2;
        mars_hip_step st{};
        st.kind = linear ? 'l' : 'h';
        st.xcol = xcol;
        st.bcol = bcol;
        st.knot = linear ? -1 : best.knot2;
        st.n_cand = n_cand;
        st.cut = cut;
        st.sse0 = best.sse0;
        st.sse1 = best.sse1;
        st.sse2 = best.sse2;
        const int types[2] = {linear ? 'l' : '+', '-'};
        for (int t = 0; t < (linear ? 1 : 2); ++t) {
            if ((int)fs.m >= o.max_terms) break;
            const int code = types[t] == 'l' ? 0 : (types[t] == '+' ? 1 : 2);
            const double mse = append_one(ctx, code, xcol, bcol, (float)cut);
            st.mse[st.n_mse++] = mse;
            if (mse < 0) continue;  // degenerate column: the basis is unchanged (marsalgo.cc:840-846)
            const double mm = (double)(fs.m - 1);  // the row's index (mars.py:366-368)
            deg.push_back(deg[bcol] + 1);
            has.resize(has.size() + px, 0);
            unsigned char *hn = &has[(deg.size() - 1) * px];
            for (size_t i = 0; i < px; ++i) hn[i] = has[(size_t)bcol * px + i];
            hn[xcol] = 1;
            mars_hip_model_row r{};
            r.type = types[t];
            r.basis = bcol;
            r.input = xcol;
            r.order = deg.back();
            r.hinge = cut;
            r.mse = mse;
            r.r2 = (float)(1.0 - mse / fs.yvar);
            r.r2_cv = (float)(1.0 - gcv(mse, mm + 1) / fs.yvar);
            r.time = (float)dt;
            rows[R++] = r;
        }
        if (steps && S < steps_cap) steps[S] = st;
        ++S;

        // stopping rules (mars.py:393-406)
        if (fs.m == m0) break;
        if (rows[R - 1].r2 > (float)(1.0 - o.r2_thresh)) break;
        if (dt > o.max_runtime) break;
        if (R > o.r2_window) {
            float all = -INFINITY, last = -INFINITY;
            for (int r = 0; r < R; ++r) {
                all = std::max(all, rows[r].r2_cv);
                if (r + o.r2_window >= R) last = std::max(last, rows[r].r2_cv);
            }
            if (last < all - (float)o.r2_thresh) break;
        }
    }
    *nrows = R;
    if (nsteps) *nsteps = S;
}

} // namespace
} // namespace mh

using namespace mh;

extern "C" int mars_hip_fit_forward(mars_hip_ctx *ctx, const mars_hip_forward_opts *opts, mars_hip_model_row *rows,
                                    int *nrows, mars_hip_step *steps, int steps_capacity, int *nsteps)
{
    return guarded([&] {
        require(ctx != nullptr && opts != nullptr, "null argument");
        fit_forward(ctx, *opts, rows, nrows, steps, steps_capacity, nsteps);
    });
}
