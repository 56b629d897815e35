/*
 *  marslib -- Python binding of the device-resident MarsAlgo.
 *
 *  Same module surface as the reference binding (marslib.cc:152-179):
 *  MarsAlgo(X, y, w, max_terms) with eval / nbasis / yvar / append, the same
 *  argument names, dtypes, layout checks and error texts.  The work is done by
 *  libmars_hip.so through its C ABI (include/mars_hip.h); this file holds no
 *  arithmetic.  `cuda=True` is the device route, as in the reference; this
 *  library has no CPU scoring path, so `cuda=False` raises.
 */
#include <pybind11/pybind11.h>
#include <pybind11/numpy.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "mars_hip.h"

namespace py = pybind11;

namespace {

void verify(bool ok, const char *msg)
{
    if (!ok) throw std::runtime_error(msg);
}

void check(int rc)
{
    if (rc != MARS_HIP_OK) throw std::runtime_error(mars_hip_last_error());
}

int parse_form(const std::string &f)
{
    if (f == "fp64" || f == "plain") return MARS_HIP_FORM_PLAIN;
    if (f == "mixed") return MARS_HIP_FORM_MIXED;
    if (f == "mixed-bf16") return MARS_HIP_FORM_MIXED_BF16;
    throw std::runtime_error("unknown precision form '" + f + "' (expected 'fp64', 'mixed' or 'mixed-bf16')");
}

py::dict tuning_dict(const mars_hip_tuning &t)
{
    py::dict d;
    d["pass_items"] = t.pass_items;
    d["retry_items"] = t.retry_items;
    d["update_waves"] = t.update_waves;
    d["sweep_waves"] = t.sweep_waves;
    d["sweep_min_rows"] = t.sweep_min_rows;
    d["append_rows"] = t.append_rows;
    d["narrow_avg"] = t.narrow_avg;
    return d;
}

/* Start from `base`, override the keys `d` names (unknown keys are an error). */
mars_hip_tuning tuning_from(const mars_hip_tuning &base, const py::dict &d)
{
    mars_hip_tuning t = base;
    for (auto kv : d) {
        const std::string k = py::cast<std::string>(kv.first);
        if (k == "pass_items") t.pass_items = py::cast<double>(kv.second);
        else if (k == "retry_items") t.retry_items = py::cast<double>(kv.second);
        else if (k == "update_waves") t.update_waves = py::cast<double>(kv.second);
        else if (k == "sweep_waves") t.sweep_waves = py::cast<double>(kv.second);
        else if (k == "sweep_min_rows") t.sweep_min_rows = py::cast<long>(kv.second);
        else if (k == "append_rows") t.append_rows = py::cast<long>(kv.second);
        else if (k == "narrow_avg") t.narrow_avg = py::cast<int>(kv.second);
        else throw std::runtime_error("unknown tuning parameter '" + k + "'");
    }
    return t;
}

class MarsAlgo {
public:
    /* new_algo (marslib.cc:12-37) + MarsAlgo::MarsAlgo (marsalgo.cc:266-370) on the device */
    MarsAlgo(py::array_t<float> X, py::array_t<float> y, py::array_t<float> w, int max_terms,
             const std::string &precision)
    {
        py::buffer_info xi = X.request(), yi = y.request(), wi = w.request();
        verify(xi.ndim == 2, "expected 2D array for X");
        verify(yi.ndim == 1, "expected 1D array for y");
        verify(wi.ndim == 1, "expected 1D array for w");
        verify(xi.strides[0] == sizeof(float), "X must be column-major");
        verify(yi.strides[0] == sizeof(float), "y must be column-major");
        verify(wi.strides[0] == sizeof(float), "w must be column-major");
        verify(xi.shape[0] == yi.shape[0], "invalid dimension for y");
        verify(xi.shape[0] == wi.shape[0], "invalid dimension for w");
        verify(max_terms >= 1, "max_terms must be >= 1");
        n_ = (size_t)xi.shape[0];
        px_ = (size_t)xi.shape[1];
        const size_t ldx = xi.shape[1] > 1 ? (size_t)(xi.strides[1] / sizeof(float)) : n_;
        int rc;
        {
            py::gil_scoped_release nogil;
            rc = mars_hip_fit_create((const float *)xi.ptr, ldx, (const float *)yi.ptr,
                                     (const float *)wi.ptr, n_, px_, (size_t)max_terms,
                                     parse_form(precision), &ctx_);
        }
        check(rc);
    }

    ~MarsAlgo() { close(); }

    py::dict get_tuning() const
    {
        mars_hip_tuning t;
        check(mars_hip_get_tuning(ctx_, &t));
        return tuning_dict(t);
    }

    void set_tuning(py::dict d)
    {
        mars_hip_tuning t;
        check(mars_hip_get_tuning(ctx_, &t));
        t = tuning_from(t, d);
        check(mars_hip_set_tuning(ctx_, &t));
    }

    /* Free the device state now (the object is unusable afterwards). */
    void close()
    {
        if (ctx_) mars_hip_context_destroy(ctx_);
        ctx_ = nullptr;
    }

    int nbasis() const
    {
        int m = 0;
        check(mars_hip_fit_info(ctx_, &m, nullptr, nullptr));
        return m;
    }

    double yvar() const
    {
        double v = 0;
        check(mars_hip_fit_info(ctx_, nullptr, &v, nullptr));
        return v;
    }

    double dsse() const
    {
        double v = 0;
        check(mars_hip_fit_info(ctx_, nullptr, nullptr, &v));
        return v;
    }

    void set_precision(const std::string &f) { check(mars_hip_fit_set_form(ctx_, parse_form(f))); }

    /* binding eval (marslib.cc:41-148); `best` adds the device argmax */
    py::tuple eval(py::array_t<bool> mask, int min_span, int endspan, bool linear, int threads,
                   bool cuda, bool with_best)
    {
        (void)threads;  // one host thread drives the device, as the reference's cuda route
        verify(cuda, "mars_amd scores candidates on the GPU only; pass cuda=True");
        py::buffer_info mi = mask.request();
        verify(mi.ndim == 2, "expected 2D array for mask");
        verify(mi.strides[1] == sizeof(bool), "mask must be row-major");
        verify(mi.shape[1] == nbasis(), "invalid dimension for mask");
        verify(min_span >= 1, "min_span must be >= 1");
        verify(mi.shape[0] == (ssize_t)px_, "invalid dimension for mask");
        const ssize_t rows = mi.shape[0], cols = mi.shape[1];
        py::array_t<double> d1({rows, cols}), d2({rows, cols}), hc({rows, cols});
        mars_hip_best b{};
        int rc;
        {
            py::gil_scoped_release nogil;
            rc = mars_hip_fit_eval(ctx_, (const unsigned char *)mi.ptr, (size_t)rows,
                                   (size_t)(mi.strides[0] / sizeof(bool)), min_span, endspan,
                                   linear ? 1 : 0, d1.mutable_data(), d2.mutable_data(),
                                   hc.mutable_data(), &b);
        }
        check(rc);
        if (!with_best) return py::make_tuple(b.sse0, d1, d2, hc);
        py::dict best;
        best["sse1"] = b.sse1;
        best["j1"] = py::make_tuple(b.xcol1, b.bcol1);
        best["sse2"] = b.sse2;
        best["j2"] = py::make_tuple(b.xcol2, b.bcol2);
        best["knot2"] = b.knot2;
        best["cut2"] = b.cut2;
        return py::make_tuple(b.sse0, d1, d2, hc, best);
    }

    /* The whole forward pass in C++ (mars_hip_fit_forward, mars.py:285-406 with Fast-MARS ranking
     * off).  Returns (records, mse, steps): the model rows as packed records of mars.MODEL_DTYPE
     * (type S1, basis i4, input i4, hinge f8, beta f4, r2 f4, r2_cv f4, order i4, time f4: 37
     * bytes), each row's append mse, and one dict per step (mars.fit's trace entries). */
    py::tuple forward(int max_terms, int max_degree, double penalty, int min_span, int end_span,
                      bool self_interact, bool linear_only, double n_true, int r2_window, double r2_thresh,
                      double max_runtime)
    {
        verify(max_terms >= 1, "max_terms must be >= 1");
        mars_hip_forward_opts o{max_terms,   max_degree, penalty,   min_span,  end_span,   self_interact ? 1 : 0,
                                linear_only ? 1 : 0, n_true, r2_window, r2_thresh, max_runtime};
        std::vector<mars_hip_model_row> rows((size_t)max_terms);
        std::vector<mars_hip_step> steps((size_t)max_terms);
        int nr = 0, ns = 0, rc;
        {
            py::gil_scoped_release nogil;
            rc = mars_hip_fit_forward(ctx_, &o, rows.data(), &nr, steps.data(), (int)steps.size(), &ns);
        }
        check(rc);
        constexpr size_t kRec = 37;
        std::string rec((size_t)nr * kRec, '\0');
        py::array_t<double> mse(nr);
        char *q = &rec[0];
        const float fnan = std::nanf("");
        for (int r = 0; r < nr; ++r, q += kRec) {
            const mars_hip_model_row &x = rows[(size_t)r];
            const char ty = (char)x.type;
            std::memcpy(q, &ty, 1);
            std::memcpy(q + 1, &x.basis, 4);
            std::memcpy(q + 5, &x.input, 4);
            std::memcpy(q + 9, &x.hinge, 8);
            std::memcpy(q + 17, &fnan, 4);
            std::memcpy(q + 21, &x.r2, 4);
            std::memcpy(q + 25, &x.r2_cv, 4);
            std::memcpy(q + 29, &x.order, 4);
            std::memcpy(q + 33, &x.time, 4);
            mse.mutable_data()[r] = x.mse;
        }
        py::list sl;
        for (int k = 0; k < std::min(ns, (int)steps.size()); ++k) {
            const mars_hip_step &t = steps[(size_t)k];
            py::dict d;
            d["epoch"] = k + 1;
            d["xcol"] = t.xcol;
            d["bcol"] = t.bcol;
            d["cut"] = t.cut;
            d["knot"] = t.knot;
            d["kind"] = std::string(1, (char)t.kind);
            d["sse0"] = t.sse0;
            d["sse1"] = t.sse1;
            d["sse2"] = t.sse2;
            d["n_cand"] = t.n_cand;
            py::list ms;
            for (int j = 0; j < t.n_mse; ++j) ms.append(t.mse[j]);
            d["mse"] = ms;
            sl.append(d);
        }
        return py::make_tuple(py::bytes(rec), mse, sl);
    }

    /* MarsAlgo::append (marsalgo.cc:760-875) */
    double append(char type, int xcol, int bcol, float h)
    {
        double mse = -1.0;
        int rc;
        {
            py::gil_scoped_release nogil;
            rc = mars_hip_fit_append(ctx_, type, xcol, bcol, h, &mse);
        }
        check(rc);
        return mse;
    }

    py::tuple basis() const
    {
        size_t m = 0;
        check(mars_hip_fit_info(ctx_, nullptr, nullptr, nullptr));
        int mm = nbasis();
        py::array_t<float> B({(ssize_t)n_, (ssize_t)mm}), Bo({(ssize_t)n_, (ssize_t)mm});
        check(mars_hip_fit_basis(ctx_, B.mutable_data(), Bo.mutable_data(), &m));
        return py::make_tuple(B, Bo);
    }

    py::dict stats() const
    {
        long fired = 0;
        double ms = 0;
        check(mars_hip_stats(ctx_, &fired, &ms));
        long sc = 0, re = 0, ga = 0, wv = 0, rd = 0, gc = 0, gl = 0;
        check(mars_hip_score_stats(ctx_, &sc, &re, &ga));
        check(mars_hip_check_stats(ctx_, &wv, &rd, &gc, &gl));
        py::dict d;
        d["dgks_fired"] = fired;
        d["eval_ms"] = ms;
        d["scored"] = sc;
        d["rescored"] = re;
        d["gated"] = ga;
        d["window_violations"] = wv;
        d["host_redo_steps"] = rd;
        d["graph_captures"] = gc;
        d["graph_launches"] = gl;
        return d;
    }

    void set_timing(bool on) { check(mars_hip_timing(ctx_, on ? 1 : 0)); }

    py::list kernel_stats() const
    {
        mars_hip_kstat ks[32];
        int cnt = 0;
        check(mars_hip_kernel_stats(ctx_, ks, 32, &cnt));
        py::list out;
        for (int i = 0; i < cnt && i < 32; ++i) {
            py::dict d;
            d["name"] = std::string(ks[i].name);
            d["launches"] = ks[i].launches;
            d["ms"] = ks[i].ms;
            d["flops"] = ks[i].flops;
            d["bytes"] = ks[i].bytes;
            out.append(d);
        }
        return out;
    }

private:
    mars_hip_ctx *ctx_ = nullptr;
    size_t n_ = 0, px_ = 0;
};

/* Many independent linear_only fits advanced together (batching form). */
class Batch {
public:
    Batch(py::list Xs, py::list ys, py::object ws, int max_terms, const std::string &precision)
    {
        const size_t F = Xs.size();
        verify(F >= 1 && ys.size() == F, "need one y per X");
        std::vector<const float *> xp(F), yp(F), wp(F, nullptr);
        std::vector<size_t> ld(F), n(F), px(F);
        for (size_t f = 0; f < F; ++f) {
            auto X = py::cast<py::array_t<float>>(Xs[f]);
            auto y = py::cast<py::array_t<float>>(ys[f]);
            py::buffer_info xi = X.request(), yi = y.request();
            verify(xi.ndim == 2, "expected 2D array for X");
            verify(yi.ndim == 1, "expected 1D array for y");
            verify(xi.strides[0] == sizeof(float), "X must be column-major");
            verify(yi.strides[0] == sizeof(float), "y must be column-major");
            verify(xi.shape[0] == yi.shape[0], "invalid dimension for y");
            keep_.append(X);
            keep_.append(y);
            xp[f] = (const float *)xi.ptr;
            yp[f] = (const float *)yi.ptr;
            n[f] = (size_t)xi.shape[0];
            px[f] = (size_t)xi.shape[1];
            ld[f] = xi.shape[1] > 1 ? (size_t)(xi.strides[1] / sizeof(float)) : n[f];
            if (!ws.is_none()) {
                auto w = py::cast<py::array_t<float>>(py::list(ws)[f]);
                py::buffer_info wi = w.request();
                verify(wi.ndim == 1 && wi.strides[0] == sizeof(float) && wi.shape[0] == xi.shape[0],
                       "invalid dimension for w");
                keep_.append(w);
                wp[f] = (const float *)wi.ptr;
            }
        }
        max_terms_ = (size_t)max_terms;
        int rc;
        {
            py::gil_scoped_release nogil;
            rc = mars_hip_batch_create((int)F, xp.data(), ld.data(), yp.data(), wp.data(), n.data(), px.data(),
                                       max_terms_, parse_form(precision), &ctx_);
        }
        check(rc);
        nfits_ = F;
    }

    ~Batch() { close(); }

    py::dict get_tuning() const
    {
        mars_hip_tuning t;
        check(mars_hip_get_tuning(ctx_, &t));
        return tuning_dict(t);
    }

    void set_tuning(py::dict d)
    {
        mars_hip_tuning t;
        check(mars_hip_get_tuning(ctx_, &t));
        t = tuning_from(t, d);
        check(mars_hip_set_tuning(ctx_, &t));
    }

    /* Free the device state now (the object is unusable afterwards). */
    void close()
    {
        if (ctx_) mars_hip_context_destroy(ctx_);
        ctx_ = nullptr;
    }

    /* returns (terms list of per-fit lists of tuples, steps) */
    py::tuple fit(int max_degree, double penalty, bool self_interact, int r2_window, double r2_thresh)
    {
        mars_hip_batch_opts o{max_degree, penalty, self_interact ? 1 : 0, r2_window, r2_thresh};
        std::vector<mars_hip_term> t(nfits_ * max_terms_);
        std::vector<int> nt(nfits_);
        int steps = 0, rc;
        {
            py::gil_scoped_release nogil;
            rc = mars_hip_batch_fit(ctx_, &o, t.data(), nt.data(), &steps);
        }
        check(rc);
        // model rows as packed records of mars.MODEL_DTYPE (type S1, basis i4, input i4, hinge f8,
        // beta f4, r2 f4, r2_cv f4, order i4, time f4: 37 bytes), fit by fit, plus the mse of each
        // row: the caller views them with numpy without a Python loop over rows
        constexpr size_t kRec = 37;
        size_t total = 0;
        for (size_t f = 0; f < nfits_; ++f) total += (size_t)nt[f];
        std::string rec(total * kRec, '\0');
        py::array_t<double> mse(total);
        py::array_t<int> counts(nfits_);
        double *pm = mse.mutable_data();
        char *q = &rec[0];
        const double nan = std::nan("");
        const float zero = 0.0f, fnan = std::nanf("");
        for (size_t f = 0; f < nfits_; ++f) {
            counts.mutable_data()[f] = nt[f];
            for (int r = 0; r < nt[f]; ++r) {
                const mars_hip_term &x = t[f * max_terms_ + r];
                const char ty = (char)x.type;
                std::memcpy(q, &ty, 1);
                std::memcpy(q + 1, &x.basis, 4);
                std::memcpy(q + 5, &x.input, 4);
                std::memcpy(q + 9, &nan, 8);
                std::memcpy(q + 17, &fnan, 4);
                std::memcpy(q + 21, &x.r2, 4);
                std::memcpy(q + 25, &x.r2_cv, 4);
                std::memcpy(q + 29, &x.order, 4);
                std::memcpy(q + 33, &zero, 4);
                q += kRec;
                *pm++ = x.mse;
            }
        }
        return py::make_tuple(py::bytes(rec), counts, mse, steps);
    }

    void set_timing(bool on) { check(mars_hip_timing(ctx_, on ? 1 : 0)); }

    py::list kernel_stats() const
    {
        mars_hip_kstat ks[32];
        int cnt = 0;
        check(mars_hip_kernel_stats(ctx_, ks, 32, &cnt));
        py::list out;
        for (int i = 0; i < cnt && i < 32; ++i) {
            py::dict d;
            d["name"] = std::string(ks[i].name);
            d["launches"] = ks[i].launches;
            d["ms"] = ks[i].ms;
            d["flops"] = ks[i].flops;
            d["bytes"] = ks[i].bytes;
            out.append(d);
        }
        return out;
    }

    py::dict stats() const
    {
        long fired = 0;
        double ms = 0;
        check(mars_hip_stats(ctx_, &fired, &ms));
        long sc = 0, re = 0, ga = 0, wv = 0, rd = 0, gc = 0, gl = 0;
        check(mars_hip_score_stats(ctx_, &sc, &re, &ga));
        check(mars_hip_check_stats(ctx_, &wv, &rd, &gc, &gl));
        py::dict d;
        d["dgks_fired"] = fired;
        d["eval_ms"] = ms;
        d["scored"] = sc;
        d["rescored"] = re;
        d["gated"] = ga;
        d["window_violations"] = wv;
        d["host_redo_steps"] = rd;
        d["graph_captures"] = gc;
        d["graph_launches"] = gl;
        return d;
    }

private:
    mars_hip_ctx *ctx_ = nullptr;
    size_t nfits_ = 0, max_terms_ = 0;
    py::list keep_;
};

/* mars.expand (mars.py:423-450) on the device.  X (n, p_x) f32 or f64; the
 * model as type codes ('i', 'l', '+', '-'), parent, input and knot arrays. */
template <typename T>
py::array_t<T> expand_impl(py::array_t<T, py::array::f_style | py::array::forcecast> X, const std::string &types,
                           py::array_t<int, py::array::c_style | py::array::forcecast> basis,
                           py::array_t<int, py::array::c_style | py::array::forcecast> input,
                           py::array_t<double, py::array::c_style | py::array::forcecast> hinge)
{
    py::buffer_info xi = X.request();
    verify(xi.ndim == 2, "expected 2D array for X");
    const size_t n = (size_t)xi.shape[0], px = (size_t)xi.shape[1], M = types.size();
    verify((size_t)basis.size() == M && (size_t)input.size() == M && (size_t)hinge.size() == M,
           "model arrays differ in length");
    py::array_t<T, py::array::f_style> B(std::vector<py::ssize_t>{(py::ssize_t)n, (py::ssize_t)M});
    const T *xp = (const T *)xi.ptr;
    T *bp = B.mutable_data();
    int rc;
    {
        py::gil_scoped_release nogil;
        if constexpr (std::is_same<T, float>::value)
            rc = mars_hip_expand(xp, n, n, px, M, types.data(), basis.data(), input.data(), hinge.data(), bp, n);
        else
            rc = mars_hip_expand_f64(xp, n, n, px, M, types.data(), basis.data(), input.data(), hinge.data(), bp, n);
    }
    check(rc);
    return B;
}

/* mars.gram (mars.py:456-490) on the device: A^T B (A^T A when B is None), f64 result. */
template <typename T>
py::array_t<double> gram_impl(py::array_t<T, py::array::f_style | py::array::forcecast> A, py::object Bobj)
{
    py::buffer_info ai = A.request();
    verify(ai.ndim == 2, "expected 2D array for A");
    const size_t n = (size_t)ai.shape[0], ma = (size_t)ai.shape[1];
    py::array_t<T, py::array::f_style | py::array::forcecast> B;
    size_t mb = ma;
    const T *bp = nullptr;
    if (!Bobj.is_none()) {
        B = py::array_t<T, py::array::f_style | py::array::forcecast>::ensure(Bobj);
        verify(B.ptr() != nullptr, "B must be a numeric array");
        py::buffer_info bi = B.request();
        verify(bi.ndim == 2, "expected 2D array for B");
        verify((size_t)bi.shape[0] == n, "shapes are not aligned");
        mb = (size_t)bi.shape[1];
        bp = (const T *)bi.ptr;
    }
    py::array_t<double, py::array::f_style> out(std::vector<py::ssize_t>{(py::ssize_t)ma, (py::ssize_t)mb});
    if (ma == 0 || mb == 0) return out;
    const T *ap = (const T *)ai.ptr;
    double *op = out.mutable_data();
    int rc;
    {
        py::gil_scoped_release nogil;
        if constexpr (std::is_same<T, float>::value)
            rc = mars_hip_gram(ap, n, n, ma, bp, n, mb, op, ma);
        else
            rc = mars_hip_gram_f64(ap, n, n, ma, bp, n, mb, op, ma);
    }
    check(rc);
    return out;
}

} // namespace

PYBIND11_MODULE(_marslib, m)
{
    m.doc() = "Multivariate Adaptive Regression Splines -- MI355X forward pass";
    m.attr("__version__") = "0.12";
    py::class_<MarsAlgo>(m, "MarsAlgo")
        .def(py::init<py::array_t<float>, py::array_t<float>, py::array_t<float>, int,
                      const std::string &>(),
             py::arg("X").noconvert(), py::arg("y").noconvert(), py::arg("w").noconvert(),
             py::arg("max_terms"), py::arg("precision") = "fp64", py::keep_alive<1, 2>())
        .def(
            "eval",
            [](MarsAlgo &a, py::array_t<bool> mask, int min_span, int endspan, bool linear_only,
               int threads, bool cuda) {
                return a.eval(mask, min_span, endspan, linear_only, threads, cuda, false);
            },
            py::arg("mask").noconvert(), py::arg("min_span"), py::arg("endspan"),
            py::arg("linear_only"), py::arg("threads"), py::arg("cuda") = false)
        .def(
            "eval_best",
            [](MarsAlgo &a, py::array_t<bool> mask, int min_span, int endspan, bool linear_only) {
                return a.eval(mask, min_span, endspan, linear_only, 1, true, true);
            },
            py::arg("mask").noconvert(), py::arg("min_span"), py::arg("endspan"),
            py::arg("linear_only"))
        .def("nbasis", &MarsAlgo::nbasis)
        .def("yvar", &MarsAlgo::yvar)
        .def("dsse", &MarsAlgo::dsse)
        .def("append", &MarsAlgo::append)
        .def("forward", &MarsAlgo::forward, py::arg("max_terms"), py::arg("max_degree"), py::arg("penalty"),
             py::arg("min_span"), py::arg("end_span"), py::arg("self_interact"), py::arg("linear_only"),
             py::arg("n_true"), py::arg("r2_window"), py::arg("r2_thresh"), py::arg("max_runtime"))
        .def("set_precision", &MarsAlgo::set_precision)
        .def("basis", &MarsAlgo::basis)
        .def("stats", &MarsAlgo::stats)
        .def("set_timing", &MarsAlgo::set_timing)
        .def("kernel_stats", &MarsAlgo::kernel_stats)
        .def("close", &MarsAlgo::close)
        .def("get_tuning", &MarsAlgo::get_tuning)
        .def("set_tuning", &MarsAlgo::set_tuning);
    py::class_<Batch>(m, "Batch")
        .def(py::init<py::list, py::list, py::object, int, const std::string &>(), py::arg("Xs"), py::arg("ys"),
             py::arg("ws") = py::none(), py::arg("max_terms"), py::arg("precision") = "fp64")
        .def("fit", &Batch::fit, py::arg("max_degree") = 3, py::arg("penalty") = 3.0,
             py::arg("self_interact") = false, py::arg("r2_window") = 32, py::arg("r2_thresh") = 3e-5)
        .def("set_timing", &Batch::set_timing)
        .def("kernel_stats", &Batch::kernel_stats)
        .def("stats", &Batch::stats)
        .def("close", &Batch::close)
        .def("get_tuning", &Batch::get_tuning)
        .def("set_tuning", &Batch::set_tuning);
    m.def("default_tuning", [] {
        mars_hip_tuning t;
        check(mars_hip_default_tuning(&t));
        return tuning_dict(t);
    });
    m.def(
        "prune_gram",
        [](py::array_t<double, py::array::c_style | py::array::forcecast> XX,
           py::array_t<double, py::array::c_style | py::array::forcecast> XY, double YY, py::object mask,
           double penalty, double n_true, double ridge) {
            verify(XX.ndim() == 2 && XX.shape(0) == XX.shape(1), "XX must be square");
            const size_t M = (size_t)XX.shape(0);
            verify(XY.ndim() == 1 && (size_t)XY.shape(0) == M, "XY must have one entry per column");
            std::vector<unsigned char> mk;
            if (!mask.is_none()) {
                auto a = py::cast<py::array_t<bool, py::array::c_style | py::array::forcecast>>(mask);
                verify(a.ndim() == 1 && (size_t)a.shape(0) == M, "mask must have one entry per column");
                mk.assign(a.data(), a.data() + M);
            }
            py::array_t<double> beta(M);
            py::array_t<bool> chosen(M);
            std::vector<unsigned char> ch(M);
            check(mars_hip_prune_gram(XX.data(), M, XY.data(), YY, M, mk.empty() ? nullptr : mk.data(), penalty,
                                      n_true, ridge, beta.mutable_data(), ch.data()));
            for (size_t i = 0; i < M; ++i) chosen.mutable_data()[i] = ch[i] != 0;
            return py::make_tuple(beta, chosen);
        },
        py::arg("XX"), py::arg("XY"), py::arg("YY"), py::arg("mask") = py::none(), py::arg("penalty") = 3.0,
        py::arg("n_true") = 1.0, py::arg("ridge") = 0.0);
    m.def(
        "device_bytes",
        [](bool reset_peak) {
            size_t live = 0, peak = 0;
            check(mars_hip_device_bytes(&live, &peak, reset_peak ? 1 : 0));
            return py::make_tuple(live, peak);
        },
        py::arg("reset_peak") = false);
    m.def("expand_f32", &expand_impl<float>, py::arg("X"), py::arg("types"), py::arg("basis"), py::arg("input"),
          py::arg("hinge"));
    m.def("expand_f64", &expand_impl<double>, py::arg("X"), py::arg("types"), py::arg("basis"), py::arg("input"),
          py::arg("hinge"));
    m.def("gram_f32", &gram_impl<float>, py::arg("A"), py::arg("B") = py::none());
    m.def("gram_f64", &gram_impl<double>, py::arg("A"), py::arg("B") = py::none());
}
