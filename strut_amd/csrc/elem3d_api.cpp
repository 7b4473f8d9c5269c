// C ABI of the 3D fiber beam-column elements (forceBeamColumn3d,
// dispBeamColumn3d): element registration, state determination of listed
// elements on the device, commit, and the force-based basic state.
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/strut_gpu.h"
#include "ctx.h"
#include "elem3d_types.h"

namespace sgpu {

void launch_elem3d(hipStream_t stream, const Elem3dArgs &args);

struct Elem3dHost {
    std::vector<Elem3dDev> elems;
    int64_t q_total = 0;
    bool uploaded = false;
    Elem3dDev *d_elems = nullptr;
    double *d_q = nullptr;
    // staging of the host-buffer entry point
    int cap = 0;
    int32_t *d_ids = nullptr, *d_status = nullptr;
    double *d_u = nullptr, *d_k = nullptr, *d_f = nullptr;
};

int elem3d_count(const sgpu_ctx *ctx) { return ctx && ctx->elems3d ? (int)ctx->elems3d->elems.size() : 0; }

void elem3d_host_free(Elem3dHost *h) {
    if (!h) return;
    for (void *p : {(void *)h->d_ids, (void *)h->d_status, (void *)h->d_u, (void *)h->d_k, (void *)h->d_f})
        if (p) (void)hipFree(p);
    delete h;
}

namespace {

Elem3dHost &elems_of(sgpu_ctx *ctx) {
    if (!ctx->elems3d) ctx->elems3d = new Elem3dHost();
    return *ctx->elems3d;
}

// The element axes of elements/utils.mojo:70-139 and the rotation of
// force_beam_column3d.mojo:236-284 (rows: local x, y, z).
bool rotation(const sgpu_elem3d_desc &e, double R[9], double &L) {
    const double dx = e.xyz2[0] - e.xyz1[0], dy = e.xyz2[1] - e.xyz1[1], dz = e.xyz2[2] - e.xyz1[2];
    L = std::sqrt(dx * dx + dy * dy + dz * dz);
    if (L == 0.0) return false;
    const double n = std::sqrt(dx * dx + dy * dy + dz * dz);
    const double lx = dx / n, ly = dy / n, lz = dz / n;
    double vx, vy, vz;
    if (e.has_vecxz) {
        vx = e.vecxz[0];
        vy = e.vecxz[1];
        vz = e.vecxz[2];
    } else {
        vx = 1.0;
        vy = 0.0;
        vz = 0.0;
        if (std::fabs(lx * vx + ly * vy + lz * vz) >= 0.9) {
            vx = 0.0;
            vy = 1.0;
            vz = 0.0;
            if (std::fabs(lx * vx + ly * vy + lz * vz) >= 0.9) {
                vx = 0.0;
                vy = 0.0;
                vz = 1.0;
            }
        }
    }
    double yx = vy * lz - vz * ly, yy = vz * lx - vx * lz, yz = vx * ly - vy * lx;
    const double ynorm = std::sqrt(yx * yx + yy * yy + yz * yz);
    if (ynorm == 0.0) return false;
    yx /= ynorm;
    yy /= ynorm;
    yz /= ynorm;
    R[0] = lx;
    R[1] = ly;
    R[2] = lz;
    R[3] = yx;
    R[4] = yy;
    R[5] = yz;
    R[6] = ly * yz - lz * yy;
    R[7] = lz * yx - lx * yz;
    R[8] = lx * yy - ly * yx;
    return true;
}

int upload_elems(sgpu_ctx *ctx, Elem3dHost &h) {
    if (h.uploaded) return SGPU_OK;
    int rc;
    if ((rc = upload(ctx, &h.d_elems, h.elems))) return rc;
    if ((rc = upload(ctx, &h.d_q, std::vector<double>((size_t)std::max<int64_t>(h.q_total, 1), 0.0)))) return rc;
    h.uploaded = true;
    return SGPU_OK;
}

int check_ids(sgpu_ctx *ctx, const Elem3dHost &h, int n, const int32_t *ids) {
    for (int i = 0; i < n; ++i)
        if (ids[i] < 0 || ids[i] >= (int)h.elems.size())
            return fail(ctx, SGPU_ERR_ARG, "element id %d out of range", ids[i]);
    return SGPU_OK;
}

}  // namespace
}  // namespace sgpu

int sgpu_elem3d_add(sgpu_ctx *ctx, const sgpu_elem3d_desc *e, int32_t *elem_id) {
    if (!ctx || !e || !elem_id) return SGPU_ERR_ARG;
    if (ctx->finalized) return fail(ctx, SGPU_ERR_STATE, "3D elements are added before sgpu_sections_finalize");
    if (e->type != SGPU_ELEM3D_FORCE && e->type != SGPU_ELEM3D_DISP)
        return fail(ctx, SGPU_ERR_ARG, "unknown 3D element type %d", e->type);
    if (e->def_id < 0 || e->def_id >= (int)ctx->defs.size() || ctx->defs[e->def_id].dim != 3)
        return fail(ctx, SGPU_ERR_ARG, "the 3D element needs a 3D fiber section definition");
    if (e->nip < 1 || e->nip > SGPU_MAX_NIP3D)
        return fail(ctx, SGPU_ERR_ARG, "3D element: %d integration points (1..%d on the device)", e->nip,
                    SGPU_MAX_NIP3D);
    if (e->transf != SGPU_TRANSF_LINEAR && e->transf != SGPU_TRANSF_PDELTA)
        return fail(ctx, SGPU_ERR_ARG, "3D element: transformation %d is not Linear or PDelta", e->transf);
    if ((int)ctx->defs[e->def_id].batch_chunk_begin.size() - 1 > 63)
        return fail(ctx, SGPU_ERR_ARG, "3D element: section definitions above 63 x 64 slots are not supported");
    Elem3dDev d{};
    double L;
    if (!rotation(*e, d.R, L)) return fail(ctx, SGPU_ERR_ARG, "3D element: zero length or invalid vecxz");
    int32_t inst0 = 0;
    int rc = sgpu_section_instances_alloc(ctx, e->def_id, e->nip, &inst0);
    if (rc) return rc;
    Elem3dHost &h = elems_of(ctx);
    d.type = e->type;
    d.def = e->def_id;
    d.nip = e->nip;
    d.transf = e->transf;
    d.inst0 = inst0;
    d.L = L;
    d.G = e->G;
    d.J = e->J;
    for (int k = 0; k < e->nip; ++k) {
        d.xi[k] = e->xi[k];
        d.w[k] = e->w[k];
    }
    d.q_off = h.q_total;
    if (e->type == SGPU_ELEM3D_FORCE) h.q_total += 5 + 3 * e->nip;  // loader.mojo:1826-1841
    h.uploaded = false;
    *elem_id = (int32_t)h.elems.size();
    h.elems.push_back(d);
    return SGPU_OK;
}

int sgpu_elem3d_state_device(sgpu_ctx *ctx, int n, const int32_t *d_ids, const double *d_u, double *d_k,
                             double *d_f, int32_t *d_status) {
    if (!ctx) return SGPU_ERR_ARG;
    if (!ctx->finalized) return fail(ctx, SGPU_ERR_STATE, "sgpu_sections_finalize has not run");
    if (n < 0 || (n > 0 && (!d_ids || !d_u || !d_k || !d_f || !d_status)))
        return fail(ctx, SGPU_ERR_ARG, "bad element buffers");
    if (n == 0) return SGPU_OK;
    Elem3dHost &h = elems_of(ctx);
    int rc = upload_elems(ctx, h);
    if (rc) return rc;
    Elem3dArgs A{};
    A.defs = ctx->d_defs;
    A.tb = ctx->tables();
    A.mats = ctx->d_mats;
    A.n_mats = (int)ctx->mats.size();
    A.c = ctx->c;
    A.t = ctx->t;
    A.elems = h.d_elems;
    A.n_elems = (int)h.elems.size();
    A.q_state = h.d_q;
    A.n = n;
    A.ids = d_ids;
    A.u = d_u;
    A.k = d_k;
    A.f = d_f;
    A.status = d_status;
    launch_elem3d(ctx->stream, A);
    HIP_OK(ctx, hipGetLastError());
    return SGPU_OK;
}

int sgpu_elem3d_state(sgpu_ctx *ctx, int n, const int32_t *ids, const double *u12, double *k144, double *f12,
                      int32_t *status) {
    if (!ctx) return SGPU_ERR_ARG;
    if (!ctx->finalized) return fail(ctx, SGPU_ERR_STATE, "sgpu_sections_finalize has not run");
    if (n < 0 || (n > 0 && (!ids || !u12 || !k144 || !f12 || !status)))
        return fail(ctx, SGPU_ERR_ARG, "bad element buffers");
    if (n == 0) return SGPU_OK;
    Elem3dHost &h = elems_of(ctx);
    int rc = check_ids(ctx, h, n, ids);
    if (rc) return rc;
    if (n > h.cap) {
        for (void *p : {(void *)h.d_ids, (void *)h.d_status, (void *)h.d_u, (void *)h.d_k, (void *)h.d_f})
            if (p) HIP_OK(ctx, hipFree(p));
        h.d_ids = h.d_status = nullptr;
        h.d_u = h.d_k = h.d_f = nullptr;
        h.cap = 0;
        HIP_OK(ctx, hipMalloc((void **)&h.d_ids, sizeof(int32_t) * n));
        HIP_OK(ctx, hipMalloc((void **)&h.d_status, sizeof(int32_t) * n));
        HIP_OK(ctx, hipMalloc((void **)&h.d_u, sizeof(double) * 12 * n));
        HIP_OK(ctx, hipMalloc((void **)&h.d_k, sizeof(double) * 144 * n));
        HIP_OK(ctx, hipMalloc((void **)&h.d_f, sizeof(double) * 12 * n));
        h.cap = n;
    }
    HIP_OK(ctx, hipMemcpyAsync(h.d_ids, ids, sizeof(int32_t) * n, hipMemcpyHostToDevice, ctx->stream));
    HIP_OK(ctx, hipMemcpyAsync(h.d_u, u12, sizeof(double) * 12 * n, hipMemcpyHostToDevice, ctx->stream));
    if ((rc = sgpu_elem3d_state_device(ctx, n, h.d_ids, h.d_u, h.d_k, h.d_f, h.d_status))) return rc;
    HIP_OK(ctx, hipMemcpyAsync(k144, h.d_k, sizeof(double) * 144 * n, hipMemcpyDeviceToHost, ctx->stream));
    HIP_OK(ctx, hipMemcpyAsync(f12, h.d_f, sizeof(double) * 12 * n, hipMemcpyDeviceToHost, ctx->stream));
    HIP_OK(ctx, hipMemcpyAsync(status, h.d_status, sizeof(int32_t) * n, hipMemcpyDeviceToHost, ctx->stream));
    HIP_OK(ctx, hipStreamSynchronize(ctx->stream));
    for (int i = 0; i < n; ++i)
        if (status[i] != 0)
            return fail(ctx, SGPU_ERR_LOCAL_NONCONV, "element %d: forceBeamColumn3d element compatibility did not converge",
                        ids[i]);
    return SGPU_OK;
}

int sgpu_elem3d_commit(sgpu_ctx *ctx, int n, const int32_t *ids) {
    if (!ctx) return SGPU_ERR_ARG;
    if (!ctx->finalized) return fail(ctx, SGPU_ERR_STATE, "sgpu_sections_finalize has not run");
    if (n < 0 || (n > 0 && !ids)) return fail(ctx, SGPU_ERR_ARG, "bad element list");
    Elem3dHost &h = elems_of(ctx);
    int rc = check_ids(ctx, h, n, ids);
    if (rc) return rc;
    std::vector<int32_t> defs, insts;
    for (int i = 0; i < n; ++i) {
        const Elem3dDev &e = h.elems[ids[i]];
        for (int ip = 0; ip < e.nip; ++ip) {
            defs.push_back(e.def);
            insts.push_back(e.inst0 + ip);
        }
    }
    return sgpu_sections_commit(ctx, (int)defs.size(), defs.data(), insts.data());
}

int sgpu_elem3d_basic_state(sgpu_ctx *ctx, int32_t elem_id, double *q) {
    if (!ctx || !q) return SGPU_ERR_ARG;
    if (!ctx->finalized) return fail(ctx, SGPU_ERR_STATE, "sgpu_sections_finalize has not run");
    Elem3dHost &h = elems_of(ctx);
    if (elem_id < 0 || elem_id >= (int)h.elems.size()) return fail(ctx, SGPU_ERR_ARG, "element id out of range");
    const Elem3dDev &e = h.elems[elem_id];
    if (e.type != SGPU_ELEM3D_FORCE) return fail(ctx, SGPU_ERR_ARG, "not a forceBeamColumn3d element");
    int rc = upload_elems(ctx, h);
    if (rc) return rc;
    HIP_OK(ctx, hipStreamSynchronize(ctx->stream));
    HIP_OK(ctx, hipMemcpy(q, h.d_q + e.q_off, sizeof(double) * (5 + 3 * e.nip), hipMemcpyDeviceToHost));
    return SGPU_OK;
}
