// C ABI of the 2D fiber beam-column elements outside any frame
// (forceBeamColumn2d, dispBeamColumn2d): registration, state determination
// of listed elements on the device, commit, and the force-based basic state.
// The kernels are the frame path's element routines (frame2d.hip).
#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/strut_gpu.h"
#include "ctx.h"
#include "frame2d_types.h"

namespace sgpu {

hipError_t launch_elem2d_state(hipStream_t stream, const Frame2dArgs &args, int n, const int32_t *ids,
                               const double *u6, double *out42, int32_t *status);
hipError_t launch_elem2d_commit(hipStream_t stream, const Frame2dArgs &args, int n, const int32_t *ids,
                                const double *u6);

struct Elem2dHost {
    std::vector<ElemDev> elems;
    std::vector<int32_t> inst0;  // section instance of integration point 0
    int64_t q_total = 0;
    bool uploaded = false;
    ElemDev *d_elems = nullptr;
    int32_t *d_inst0 = nullptr;
    double *d_q = nullptr, *d_last = nullptr;
    unsigned long long *d_sec_evals = nullptr;
    // staging of the host-buffer entry points
    int cap = 0;
    int32_t *d_ids = nullptr, *d_status = nullptr;
    double *d_u = nullptr, *d_out = nullptr;
};

void elem2d_host_free(Elem2dHost *h) {
    if (!h) return;
    for (void *p : {(void *)h->d_ids, (void *)h->d_status, (void *)h->d_u, (void *)h->d_out})
        if (p) (void)hipFree(p);
    delete h;
}

namespace {

Elem2dHost &elems_of(sgpu_ctx *ctx) {
    if (!ctx->elems2d) ctx->elems2d = new Elem2dHost();
    return *ctx->elems2d;
}

int upload_elems(sgpu_ctx *ctx, Elem2dHost &h) {
    if (h.uploaded) return SGPU_OK;
    int rc;
    const size_t n = h.elems.size();
    if ((rc = upload(ctx, &h.d_elems, h.elems))) return rc;
    if ((rc = upload(ctx, &h.d_inst0, h.inst0))) return rc;
    if ((rc = upload(ctx, &h.d_q, std::vector<double>((size_t)std::max<int64_t>(h.q_total, 1), 0.0)))) return rc;
    if ((rc = upload(ctx, &h.d_last, std::vector<double>(2 * kMaxNipDev * std::max<size_t>(n, 1), 0.0)))) return rc;
    if ((rc = upload(ctx, &h.d_sec_evals, std::vector<unsigned long long>(1, 0ull)))) return rc;
    h.uploaded = true;
    return SGPU_OK;
}

Frame2dArgs args_of(sgpu_ctx *ctx, const Elem2dHost &h) {
    Frame2dArgs A{};
    A.elems = h.d_elems;
    A.q = h.d_q;
    A.last = h.d_last;
    A.elem_inst = h.d_inst0;
    A.sec_evals = h.d_sec_evals;
    A.defs = ctx->d_defs;
    A.tb = ctx->tables();
    A.mats = ctx->d_mats;
    A.n_mats = (int32_t)ctx->mats.size();
    A.c = ctx->c;
    A.t = ctx->t;
    return A;
}

// ids in range and distinct within one call (two blocks on one element would race)
int check_ids(sgpu_ctx *ctx, const Elem2dHost &h, int n, const int32_t *ids) {
    std::vector<char> seen(h.elems.size(), 0);
    for (int i = 0; i < n; ++i) {
        if (ids[i] < 0 || ids[i] >= (int)h.elems.size())
            return fail(ctx, SGPU_ERR_ARG, "element id %d out of range", ids[i]);
        if (seen[ids[i]]) return fail(ctx, SGPU_ERR_ARG, "element id %d listed twice", ids[i]);
        seen[ids[i]] = 1;
    }
    return SGPU_OK;
}

int stage(sgpu_ctx *ctx, Elem2dHost &h, int n) {
    if (n <= h.cap) return SGPU_OK;
    for (void *p : {(void *)h.d_ids, (void *)h.d_status, (void *)h.d_u, (void *)h.d_out})
        if (p) HIP_OK(ctx, hipFree(p));
    h.d_ids = h.d_status = nullptr;
    h.d_u = h.d_out = nullptr;
    h.cap = 0;
    HIP_OK(ctx, hipMalloc((void **)&h.d_ids, sizeof(int32_t) * n));
    HIP_OK(ctx, hipMalloc((void **)&h.d_status, sizeof(int32_t) * n));
    HIP_OK(ctx, hipMalloc((void **)&h.d_u, sizeof(double) * 6 * n));
    HIP_OK(ctx, hipMalloc((void **)&h.d_out, sizeof(double) * 42 * n));
    h.cap = n;
    return SGPU_OK;
}

}  // namespace
}  // namespace sgpu

int sgpu_elem2d_add(sgpu_ctx *ctx, const sgpu_elem2d_desc *e, int32_t *elem_id) {
    if (!ctx || !e || !elem_id) return SGPU_ERR_ARG;
    if (ctx->finalized) return fail(ctx, SGPU_ERR_STATE, "2D elements are added before sgpu_sections_finalize");
    if (e->type != SGPU_ELEM2D_FORCE && e->type != SGPU_ELEM2D_DISP)
        return fail(ctx, SGPU_ERR_ARG, "unknown 2D element type %d", e->type);
    if (e->def_id < 0 || e->def_id >= (int)ctx->defs.size() || ctx->defs[e->def_id].dim != 2)
        return fail(ctx, SGPU_ERR_ARG, "the 2D element needs a 2D fiber section definition");
    const SecDefHost &sd = ctx->defs[e->def_id];
    if (e->nip < 1 || e->nip > SGPU_MAX_NIP2D)
        return fail(ctx, SGPU_ERR_ARG, "2D element: %d integration points (1..%d on the device)", e->nip,
                    SGPU_MAX_NIP2D);
    if ((int)sd.batch_chunk_begin.size() - 1 > 63)
        return fail(ctx, SGPU_ERR_ARG, "2D element: section definitions above 63 x 64 slots are not supported");
    if (e->type == SGPU_ELEM2D_FORCE && !sd.init_flex_valid)
        return fail(ctx, SGPU_ERR_ARG, "2D element: section initial flexibility is singular");
    const double dx = e->xy2[0] - e->xy1[0], dy = e->xy2[1] - e->xy1[1];
    ElemDev el{};
    el.L = std::hypot(dx, dy);
    if (el.L == 0.0) return fail(ctx, SGPU_ERR_ARG, "2D element: zero length");
    int32_t inst0 = 0;
    int rc = sgpu_section_instances_alloc(ctx, e->def_id, e->nip, &inst0);
    if (rc) return rc;
    Elem2dHost &h = elems_of(ctx);
    // the frame path's element record (frame2d_api.cpp), all six DOFs free
    bool any_nonlinear = false;
    for (size_t s = 0; s < sd.mat.size(); ++s)
        if (sd.mat[s] >= 0 && ctx->mats[sd.mat[s]].type != kElastic) any_nonlinear = true;
    el.all_elastic = any_nonlinear ? 0 : 1;
    el.nip = e->nip;
    el.def = e->def_id;
    el.type = e->type;
    el.inv_L = 1.0 / el.L;
    el.c = e->type == SGPU_ELEM2D_DISP ? dx / el.L : dx * el.inv_L;
    el.s = e->type == SGPU_ELEM2D_DISP ? dy / el.L : dy * el.inv_L;
    el.L2 = std::sqrt(dx * dx + dy * dy);
    el.c2 = dx / el.L2;
    el.s2 = dy / el.L2;
    for (int ip = 0; ip < e->nip; ++ip) {
        el.xi[ip] = e->xi[ip];
        el.w[ip] = e->w[ip];
    }
    el.if00 = sd.init_flex[0];
    el.if01 = sd.init_flex[1];
    el.if11 = sd.init_flex[2];
    el.if_valid = sd.init_flex_valid;
    for (int k = 0; k < 6; ++k) el.fmap[k] = k;
    el.q_off = (int32_t)h.q_total;
    h.q_total += 3 + 2 * e->nip + 3;  // loader.mojo's force_basic_q: q, eps0[nip], kappa[nip], v
    h.uploaded = false;
    *elem_id = (int32_t)h.elems.size();
    h.elems.push_back(el);
    h.inst0.push_back(inst0);
    return SGPU_OK;
}

int sgpu_elem2d_state(sgpu_ctx *ctx, int n, const int32_t *ids, const double *u6, double *k36, double *f6,
                      int32_t *status) {
    if (!ctx) return SGPU_ERR_ARG;
    if (!ctx->finalized) return fail(ctx, SGPU_ERR_STATE, "sgpu_sections_finalize has not run");
    if (n < 0 || (n > 0 && (!ids || !u6 || !k36 || !f6 || !status)))
        return fail(ctx, SGPU_ERR_ARG, "bad element buffers");
    if (n == 0) return SGPU_OK;
    Elem2dHost &h = elems_of(ctx);
    int rc;
    if ((rc = check_ids(ctx, h, n, ids)) || (rc = upload_elems(ctx, h)) || (rc = stage(ctx, h, n))) return rc;
    HIP_OK(ctx, hipMemcpyAsync(h.d_ids, ids, sizeof(int32_t) * n, hipMemcpyHostToDevice, ctx->stream));
    HIP_OK(ctx, hipMemcpyAsync(h.d_u, u6, sizeof(double) * 6 * n, hipMemcpyHostToDevice, ctx->stream));
    HIP_OK(ctx, launch_elem2d_state(ctx->stream, args_of(ctx, h), n, h.d_ids, h.d_u, h.d_out, h.d_status));
    std::vector<double> out((size_t)42 * n);
    HIP_OK(ctx, hipMemcpyAsync(out.data(), h.d_out, sizeof(double) * 42 * n, hipMemcpyDeviceToHost, ctx->stream));
    HIP_OK(ctx, hipMemcpyAsync(status, h.d_status, sizeof(int32_t) * n, hipMemcpyDeviceToHost, ctx->stream));
    HIP_OK(ctx, hipStreamSynchronize(ctx->stream));
    for (int i = 0; i < n; ++i) {
        for (int k = 0; k < 36; ++k) k36[36 * (size_t)i + k] = out[42 * (size_t)i + k];
        for (int k = 0; k < 6; ++k) f6[6 * (size_t)i + k] = out[42 * (size_t)i + 36 + k];
    }
    for (int i = 0; i < n; ++i)
        if (status[i] != 0) {
            status[i] = SGPU_ERR_LOCAL_NONCONV;
            return fail(ctx, SGPU_ERR_LOCAL_NONCONV, "element %d: forceBeamColumn2d tangent recovery failed", ids[i]);
        }
    return SGPU_OK;
}

int sgpu_elem2d_commit(sgpu_ctx *ctx, int n, const int32_t *ids, const double *u6) {
    if (!ctx) return SGPU_ERR_ARG;
    if (!ctx->finalized) return fail(ctx, SGPU_ERR_STATE, "sgpu_sections_finalize has not run");
    if (n < 0 || (n > 0 && !ids)) return fail(ctx, SGPU_ERR_ARG, "bad element list");
    if (n == 0) return SGPU_OK;
    Elem2dHost &h = elems_of(ctx);
    int rc;
    if ((rc = check_ids(ctx, h, n, ids)) || (rc = upload_elems(ctx, h)) || (rc = stage(ctx, h, n))) return rc;
    HIP_OK(ctx, hipMemcpyAsync(h.d_ids, ids, sizeof(int32_t) * n, hipMemcpyHostToDevice, ctx->stream));
    if (u6) HIP_OK(ctx, hipMemcpyAsync(h.d_u, u6, sizeof(double) * 6 * n, hipMemcpyHostToDevice, ctx->stream));
    HIP_OK(ctx, launch_elem2d_commit(ctx->stream, args_of(ctx, h), n, h.d_ids, u6 ? h.d_u : nullptr));
    HIP_OK(ctx, hipStreamSynchronize(ctx->stream));
    return SGPU_OK;
}

int sgpu_elem2d_basic_state(sgpu_ctx *ctx, int32_t elem_id, double *q) {
    if (!ctx || !q) return SGPU_ERR_ARG;
    if (!ctx->finalized) return fail(ctx, SGPU_ERR_STATE, "sgpu_sections_finalize has not run");
    Elem2dHost &h = elems_of(ctx);
    if (elem_id < 0 || elem_id >= (int)h.elems.size()) return fail(ctx, SGPU_ERR_ARG, "element id out of range");
    const ElemDev &e = h.elems[elem_id];
    if (e.type != SGPU_ELEM2D_FORCE) return fail(ctx, SGPU_ERR_ARG, "not a forceBeamColumn2d element");
    int rc = upload_elems(ctx, h);
    if (rc) return rc;
    HIP_OK(ctx, hipStreamSynchronize(ctx->stream));
    HIP_OK(ctx, hipMemcpy(q, h.d_q + e.q_off, sizeof(double) * (3 + 2 * e.nip + 3), hipMemcpyDeviceToHost));
    return SGPU_OK;
}
