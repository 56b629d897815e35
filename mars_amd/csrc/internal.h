// Host-side entry points shared between the libmars_hip translation units.
#pragma once

#include "common.h"

namespace mh {

/* Candidate-column source for the scoring kernels: column xc of X (ld ldx)
 * scaled by s[xc].  Mirror-mode single-column calls point X at one column
 * with scale 1. */
struct XSrc {
    const float *X;
    size_t ldx;
    const float *s;
};

/* Rows per split / groups used by the plain orthonormalize pipeline. */
struct PlainPlan {
    int P = 0;          // candidates in this block
    int mt = 0;         // 16-wide basis tiles covering m
    int ldT = 0;        // leading dimension of T (mt * 16)
    int ks = 0;         // row splits of the T contraction
    size_t rps = 0;     // rows per split (multiple of 16)
    int groups = 0;     // row groups of the projection / finalize passes
    size_t rpg = 0;     // rows per group (multiple of 64)
    size_t ldr = 0;     // row stride of the residual buffer
};

/* Orthonormalize the current candidate block (ctx->cand, P entries) against
 * the live basis with the reference's FP64 arithmetic and f32 rounding points
 * (kernels.cc:283-417), then ybx = Bxn.y (marsalgo.cc:500-505).  Leaves:
 *   ctx->ybx[c], ctx->lin[c] = ybx^2, and, if keep_bxn, the normalized f32
 *   columns in ctx->rbuf (row-major, row stride plan.ldr). */
PlainPlan plain_orthonormalize(mars_hip_ctx *ctx, const XSrc &xs, int P, bool keep_bxn);

/* Mixed-precision scoring of the current candidate block (linear part only). */
void mixed_orthonormalize(mars_hip_ctx *ctx, const XSrc &xs, int P);

/* Hinge sweep over the candidates of the current block whose normalized
 * columns sit in ctx->rbuf (plan from plain_orthonormalize). Fills ctx->hin,
 * ctx->knot, ctx->krow. */
void hinge_sweep(mars_hip_ctx *ctx, const XSrc &xs, const PlainPlan &plan, int min_span,
                 int end_span);

/* Make the stable descending sort order and f32 deltas of X column xc resident. */
void ensure_sorted(mars_hip_ctx *ctx, const XSrc &xs, int xc);

/* Largest candidate block whose residual buffer fits the scratch budget. */
size_t block_capacity(const mars_hip_ctx *ctx);

void upload_cands(mars_hip_ctx *ctx);

} // namespace mh
