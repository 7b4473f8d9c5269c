// Host side of libstrut_gpu: the C ABI declared in include/strut_gpu.h.
//
// Builds the slot layout of every fiber-section definition (the reference's
// family grouping and W padding, fiber2d.mojo:3146-3455; the 3D fiber order,
// fiber3d.mojo:157-186), the initial material states (core.mojo:203-263),
// and drives the device kernels in fiber_section.hip on one HIP stream.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/strut_gpu.h"
#include "strut_gpu_types.h"

namespace sgpu {
void launch_aux_eval(hipStream_t stream, const SecDefDev *defs, int n_defs, const SlotTables &tb,
                     const MatDev *mats, const StateArrays &c, const StateArrays &t,
                     const sgpu_section_point *pts, int n, double *fsks, int max_aux);
void launch_section_eval(hipStream_t stream, const SecDefDev *defs, int n_defs, const SlotTables &tb,
                         const MatDev *mats, int n_mats, const StateArrays &c, const StateArrays &t,
                         const sgpu_section_point *pts, sgpu_section_result *res, int n, const double *fsks,
                         int max_aux);
void launch_copy_state(hipStream_t stream, const StateArrays &dst, const StateArrays &src, int64_t n_slots,
                       int64_t n_aux);
void launch_copy_instances(hipStream_t stream, const SecDefDev *defs, const int32_t *def_ids,
                           const int32_t *insts, int n, const StateArrays &dst, const StateArrays &src);
}  // namespace sgpu

using namespace sgpu;

#include "ctx.h"

using namespace sgpu;

namespace {

double initial_tangent(const MatDev &m) {  // core.mojo:270-282
    switch (m.type) {
    case kElastic: return m.p[0];
    case kSteel01: return m.p[1];
    case kConcrete01: return (2.0 * m.p[0]) / m.p[1];
    case kSteel02: return m.p[1];
    case kConcrete02: return (2.0 * m.p[0]) / m.p[1];
    default: return 0.0;
    }
}

struct Cell {
    double y, z, area;
    int mat;
};

// Append one slot to the layout.
void push_slot(SecDefHost &d, double y, double z, double area, int mat, bool real, int fiber) {
    d.y.push_back(y);
    d.z.push_back(z);
    d.area.push_back(area);
    d.mat.push_back(mat);
    d.aux.push_back(-1);
    d.chunk.push_back(-1);
    d.slot_real.push_back(real ? 1 : 0);
    if (fiber >= 0) d.fiber_slot[fiber] = (int32_t)d.y.size() - 1;
}

void pad_dead_to_w(SecDefHost &d) {
    while (d.y.size() % kW) push_slot(d, 0.0, 0.0, 0.0, -1, false, -1);
}

// Build the slot layout, reduction chunks and aux map of one definition.
int build_def(sgpu_ctx *ctx, SecDefHost &d, const std::vector<Cell> &cells) {
    const int n = (int)cells.size();
    d.n_fibers = n;
    d.fiber_slot.assign(n, -1);
    // centroid: running sums in fiber order (fiber2d.mojo:3925-3941, fiber3d.mojo:418-470)
    double area_sum = 0.0, qz_sum = 0.0, qy_sum = 0.0;
    for (const Cell &c : cells) {
        area_sum += c.area;
        qz_sum += c.area * c.y;
        qy_sum += c.area * c.z;
    }
    d.y_bar = qz_sum / area_sum;
    d.z_bar = d.dim == 3 ? qy_sum / area_sum : 0.0;
    std::vector<int> elastic, nonlinear;
    double k11 = 0.0, k12 = 0.0, k22 = 0.0;
    for (int i = 0; i < n; ++i) {
        const MatDev &m = ctx->mats[cells[i].mat];
        const double y_rel = cells[i].y - d.y_bar;
        const double ks = initial_tangent(m) * cells[i].area;  // fiber2d.mojo:3341-3344
        k11 += ks;
        k12 += -ks * y_rel;
        k22 += ks * y_rel * y_rel;
        if (m.type == kElastic) elastic.push_back(i);
        else nonlinear.push_back(i);
    }
    if (d.dim == 2) {  // fiber2d.mojo:3448-3454
        const double det = k11 * k22 - k12 * k12;
        if (std::fabs(det) > 1.0e-40) {
            const double inv_det = 1.0 / det;
            d.init_flex_valid = 1;
            d.init_flex[0] = k22 * inv_det;
            d.init_flex[1] = -k12 * inv_det;
            d.init_flex[2] = k11 * inv_det;
        }
    }
    auto yrel = [&](int i) { return cells[i].y - d.y_bar; };
    auto zrel = [&](int i) { return d.dim == 3 ? cells[i].z - d.z_bar : 0.0; };
    // chunks are recorded as (first slot, width)
    std::vector<std::pair<int, int>> chunks;
    // elastic part: W-chunks then scalar tail (vectorize, fiber2d.mojo:3459-3487)
    const int ne = (int)elastic.size();
    int i = 0;
    for (; i + kW <= ne; i += kW) {
        chunks.push_back({(int)d.y.size(), kW});
        for (int l = 0; l < kW; ++l) {
            const int f = elastic[i + l];
            push_slot(d, yrel(f), zrel(f), cells[f].area, cells[f].mat, true, f);
        }
    }
    for (; i < ne; ++i) {
        chunks.push_back({(int)d.y.size(), 1});
        const int f = elastic[i];
        push_slot(d, yrel(f), zrel(f), cells[f].area, cells[f].mat, true, f);
    }
    d.n_elastic_chunks = (int)chunks.size();
    if (ne > 0) {  // uniform modulus rule, fiber2d.mojo:3491-3499, fiber3d.mojo:245-254
        bool uniform = true;
        const double e0 = ctx->mats[cells[elastic[0]].mat].p[0];
        for (int j = 1; j < ne; ++j)
            if (ctx->mats[cells[elastic[j]].mat].p[0] != e0) {
                uniform = false;
                break;
            }
        d.zero_k12 = uniform ? 1 : 0;
    }
    if (d.dim == 2) {
        pad_dead_to_w(d);
        // Steel02 then Concrete02 families, grouped by definition in order of
        // first appearance, each group padded to W (fiber2d.mojo:3146-3228).
        const int fam_types[2] = {kSteel02, kConcrete02};
        for (int ft : fam_types) {
            std::vector<int> uniq;
            for (int f : nonlinear)
                if (ctx->mats[cells[f].mat].type == ft) {
                    bool seen = false;
                    for (int u : uniq) seen = seen || (u == cells[f].mat);
                    if (!seen) uniq.push_back(cells[f].mat);
                }
            for (int gdef : uniq) {
                std::vector<int> members;
                for (int f : nonlinear)
                    if (cells[f].mat == gdef) members.push_back(f);
                const int pad = ((int)members.size() + kW - 1) / kW * kW;
                const int first = (int)d.y.size();
                for (int k = 0; k < pad; ++k) {
                    if (k < (int)members.size()) {
                        const int f = members[k];
                        push_slot(d, yrel(f), 0.0, cells[f].area, gdef, true, f);
                    } else {
                        push_slot(d, 0.0, 0.0, 0.0, gdef, false, -1);  // zero-area padding lane
                    }
                }
                for (int k = 0; k < pad; k += kW) chunks.push_back({first + k, kW});
            }
        }
        // scalar Steel01 then Concrete01 (fiber2d.mojo:4002-4022)
        const int scal_types[2] = {kSteel01, kConcrete01};
        for (int ft : scal_types)
            for (int f : nonlinear)
                if (ctx->mats[cells[f].mat].type == ft) {
                    chunks.push_back({(int)d.y.size(), 1});
                    push_slot(d, yrel(f), 0.0, cells[f].area, cells[f].mat, true, f);
                }
        for (int f : nonlinear) {
            const int ty = ctx->mats[cells[f].mat].type;
            if (ty != kSteel01 && ty != kConcrete01 && ty != kSteel02 && ty != kConcrete02)
                return fail(ctx, SGPU_ERR_ARG, "unsupported uniaxial material type %d", ty);
        }
    } else {
        // 3D: every nonlinear fiber accumulates on its own, in input order
        // (fiber3d.mojo:804-822)
        for (int f : nonlinear) {
            chunks.push_back({(int)d.y.size(), 1});
            push_slot(d, yrel(f), zrel(f), cells[f].area, cells[f].mat, true, f);
        }
    }
    // aux slots: Steel02 and Concrete01 (padding lanes included)
    for (size_t s = 0; s < d.y.size(); ++s) {
        const int mi = d.mat[s];
        if (mi < 0) continue;
        const int ty = ctx->mats[mi].type;
        if (ty == kSteel02 || ty == kConcrete01) {
            d.aux[s] = d.n_aux++;
            d.aux_slot.push_back((int32_t)s);
        }
    }
    // chunk codes and per-batch chunk ranges
    const int n_slots = (int)d.y.size();
    const int n_batches = (n_slots + kBatchSlots - 1) / kBatchSlots;
    d.batch_chunk_begin.assign(n_batches + 1, (int32_t)chunks.size());
    {
        int ci = 0;
        for (int b = 0; b < n_batches; ++b) {
            while (ci < (int)chunks.size() && chunks[ci].first < b * kBatchSlots) ++ci;
            d.batch_chunk_begin[b] = ci;
        }
    }
    for (int ci = 0; ci < (int)chunks.size(); ++ci) {
        const int s = chunks[ci].first, w = chunks[ci].second;
        const int b = s / kBatchSlots;
        if (w == kW && (s % kW) != 0) return fail(ctx, SGPU_ERR_STATE, "internal: misaligned W chunk");
        if ((s + w - 1) / kBatchSlots != b) return fail(ctx, SGPU_ERR_STATE, "internal: chunk straddles batch");
        const int local = ci - d.batch_chunk_begin[b];
        if (local < 0 || local >= kBatchSlots) return fail(ctx, SGPU_ERR_STATE, "internal: chunk index");
        d.chunk[s] = local * 2 + (w == kW ? 1 : 0);
        d.chunk_slot.push_back(s);
    }
    return SGPU_OK;
}

// UniMaterialState(mat_def), core.mojo:203-263, restricted to the fields the
// slot arrays carry.  Padding lanes and filler stay zero (fiber2d.mojo:868-911).
void init_slot_state(const MatDev &m, bool real, double *eps, double *sig, double *tan, double *h0,
                     double *h1, double *a0, int32_t *kon) {
    *eps = *sig = *tan = *h0 = *h1 = 0.0;
    if (a0) *a0 = 0.0;
    if (kon) *kon = 0;
    if (!real || m.type == kElastic) return;
    const double tg = initial_tangent(m);
    *tan = tg;
    if (m.type == kConcrete01 && a0) *a0 = tg;
    if (m.type == kSteel02) {
        const double epsy = m.p[0] / m.p[1];
        *h1 = epsy;   // epsmax
        *h0 = -epsy;  // epsmin
        const double sigini = m.p[10];
        if (sigini != 0.0) {
            *eps = sigini / m.p[1];
            *sig = sigini;
        }
    }
}

int ensure_scratch(sgpu_ctx *ctx, int n) {
    if (n <= ctx->scratch_cap) return SGPU_OK;
    if (ctx->d_pts) (void)hipFree(ctx->d_pts);
    if (ctx->d_res) (void)hipFree(ctx->d_res);
    ctx->d_pts = nullptr;
    ctx->d_res = nullptr;
    int cap = n < 1024 ? 1024 : n;
    HIP_OK(ctx, hipMalloc((void **)&ctx->d_pts, sizeof(sgpu_section_point) * cap));
    HIP_OK(ctx, hipMalloc((void **)&ctx->d_res, sizeof(sgpu_section_result) * cap));
    ctx->scratch_cap = cap;
    return SGPU_OK;
}

int ensure_list(sgpu_ctx *ctx, int n) {
    if (2 * n <= ctx->list_cap) return SGPU_OK;
    if (ctx->d_list) (void)hipFree(ctx->d_list);
    if (ctx->d_fsks) (void)hipFree(ctx->d_fsks);
    ctx->d_list = nullptr;
    int cap = 2 * n < 1024 ? 1024 : 2 * n;
    HIP_OK(ctx, hipMalloc((void **)&ctx->d_list, sizeof(int32_t) * cap));
    ctx->list_cap = cap;
    return SGPU_OK;
}

int check_points(sgpu_ctx *ctx, int n, const sgpu_section_point *pts) {
    for (int p = 0; p < n; ++p) {
        const int d = pts[p].def_id;
        if (d < 0 || d >= (int)ctx->defs.size())
            return fail(ctx, SGPU_ERR_ARG, "point %d: section definition %d out of range", p, d);
        if (pts[p].instance < 0 || pts[p].instance >= ctx->defs[d].n_inst)
            return fail(ctx, SGPU_ERR_ARG, "point %d: instance %d out of range for definition %d", p,
                        pts[p].instance, d);
    }
    return SGPU_OK;
}

// Scratch of sgpu_section_eval_device_noalloc: the aux pre-pass results
// ((sigma A, E_t A) per aux slot per point) in a file-scope device array, so
// that entry point allocates nothing.  Each GPU the library runs on holds
// its own copy for the life of the process; host code reaches a device's
// copy only through the address hipGetSymbolAddress returns on that device
// (the array's name is a host-side stand-in).
constexpr int64_t kNoAllocScratchDoubles = int64_t(1) << 20;  // 8 MiB
static __device__ __align__(256) double scratch[kNoAllocScratchDoubles];

constexpr int kNoAllocMaxDevices = 64;
std::mutex g_scratch_mutex;
double *g_scratch_ptr[kNoAllocMaxDevices] = {};

// Once per device, with that device current (sgpu_ctx_create).
int noalloc_scratch_init(int device) {
    if (device < 0 || device >= kNoAllocMaxDevices) return SGPU_ERR_ARG;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_ptr[device]) return SGPU_OK;
    void *ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return SGPU_ERR_HIP;
    g_scratch_ptr[device] = static_cast<double *>(ptr);
    return SGPU_OK;
}

double *noalloc_scratch(int device) {
    if (device < 0 || device >= kNoAllocMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    return g_scratch_ptr[device];
}

// Both passes of a section evaluation over device points, with the aux
// results in `fsks` (n * max_aux * 2 doubles).  `timed`: per-kernel events
// when profiling is on.
int launch_section_passes(sgpu_ctx *ctx, int n, const sgpu_section_point *d_points, sgpu_section_result *d_results,
                          double *fsks, bool timed) {
    {
        ProfScope ps(ctx, timed ? SGPU_KERNEL_AUX_EVAL : -1);
        launch_aux_eval(ctx->stream, ctx->d_defs, (int)ctx->defs.size(), ctx->tables(), ctx->d_mats, ctx->c, ctx->t,
                        d_points, n, fsks, ctx->max_aux);
    }
    {
        ProfScope ps(ctx, timed ? SGPU_KERNEL_SECTION_EVAL : -1);
        launch_section_eval(ctx->stream, ctx->d_defs, (int)ctx->defs.size(), ctx->tables(), ctx->d_mats,
                            (int)ctx->mats.size(), ctx->c, ctx->t, d_points, d_results, n, fsks, ctx->max_aux);
    }
    HIP_OK(ctx, hipGetLastError());
    return SGPU_OK;
}

}  // namespace

extern "C" {

int sgpu_abi_version(void) { return SGPU_ABI_VERSION; }

int sgpu_ctx_create(int device, sgpu_ctx **out) {
    if (!out) return SGPU_ERR_ARG;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return SGPU_ERR_NODEV;
    if (device < 0 || device >= count) return SGPU_ERR_ARG;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return SGPU_ERR_NODEV;
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return SGPU_ERR_NODEV;
    sgpu_ctx *ctx = new sgpu_ctx();
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&ctx->stream) != hipSuccess) {
        delete ctx;
        return SGPU_ERR_HIP;
    }
    if (noalloc_scratch_init(device) != SGPU_OK) {
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return SGPU_ERR_HIP;
    }
    *out = ctx;
    return SGPU_OK;
}

int sgpu_ctx_destroy(sgpu_ctx *ctx) {
    if (!ctx) return SGPU_ERR_ARG;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (auto &pe : ctx->prof_pending) {
        (void)hipEventDestroy(pe.second.first);
        (void)hipEventDestroy(pe.second.second);
    }
    for (hipEvent_t e : ctx->prof_free) (void)hipEventDestroy(e);
    for (void *p : ctx->allocations) (void)hipFree(p);
    if (ctx->d_pts) (void)hipFree(ctx->d_pts);
    if (ctx->d_res) (void)hipFree(ctx->d_res);
    if (ctx->d_list) (void)hipFree(ctx->d_list);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    sgpu::frame2d_host_free(ctx->frames);
    sgpu::elem3d_host_free(ctx->elems3d);
    sgpu::elem2d_host_free(ctx->elems2d);
    sgpu::mpc_host_free(ctx->mpc);
    sgpu::frame3d_host_free(ctx->frame3d);
    delete ctx;
    return SGPU_OK;
}

const char *sgpu_last_error(const sgpu_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

void *sgpu_stream(sgpu_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

int sgpu_synchronize(sgpu_ctx *ctx) {
    if (!ctx) return SGPU_ERR_ARG;
    HIP_OK(ctx, hipStreamSynchronize(ctx->stream));
    return SGPU_OK;
}

int sgpu_material_defs_upload(sgpu_ctx *ctx, int n, const int32_t *types, const double *params) {
    if (!ctx) return SGPU_ERR_ARG;
    if (ctx->finalized) return fail(ctx, SGPU_ERR_STATE, "materials must be uploaded before finalize");
    if (n < 0 || (n > 0 && (!types || !params))) return fail(ctx, SGPU_ERR_ARG, "bad material table");
    ctx->mats.assign(n, MatDev{});
    for (int i = 0; i < n; ++i) {
        MatDev &m = ctx->mats[i];
        m.type = types[i];
        if (m.type < kElastic || m.type > kConcrete02)
            return fail(ctx, SGPU_ERR_ARG, "material %d: unsupported type %d", i, m.type);
        for (int k = 0; k < 11; ++k) m.p[k] = params[11 * i + k];
        for (int k = 0; k < 5; ++k) m.pre[k] = 0.0;
        if (m.type == kConcrete02) {  // fiber2d.mojo:3285-3308
            const double fc = m.p[0], epsc0 = m.p[1], fcu = m.p[2], epscu = m.p[3], rat = m.p[4];
            const double ft = m.p[5], ets = m.p[6];
            const double ec0 = (2.0 * fc) / epsc0;
            const double epsr = (fcu - rat * ec0 * epscu) / (ec0 * (1.0 - rat));
            m.pre[0] = ec0;
            m.pre[1] = ft / ec0;
            m.pre[2] = ft * (1.0 / ets + 1.0 / ec0);
            m.pre[3] = epsr;
            m.pre[4] = ec0 * epsr;
        } else if (m.type == kSteel02) {  // steel02.mojo:18-24
            m.pre[0] = m.p[2] * m.p[1];
            m.pre[1] = m.p[0] / m.p[1];
            m.pre[2] = m.p[10] / m.p[1];
        }
    }
    return SGPU_OK;
}

int sgpu_section_def_create(sgpu_ctx *ctx, int dim, int n_fibers, const double *y, const double *z,
                            const double *area, const int32_t *mat_index, int32_t *def_id) {
    if (!ctx) return SGPU_ERR_ARG;
    if (ctx->finalized) return fail(ctx, SGPU_ERR_STATE, "section definitions must precede finalize");
    if (dim != 2 && dim != 3) return fail(ctx, SGPU_ERR_ARG, "dim must be 2 or 3");
    if (n_fibers <= 0 || !y || !area || !mat_index || !def_id || (dim == 3 && !z))
        return fail(ctx, SGPU_ERR_ARG, "section definition needs fibers");
    std::vector<Cell> cells(n_fibers);
    double area_sum = 0.0;
    for (int i = 0; i < n_fibers; ++i) {
        if (mat_index[i] < 0 || mat_index[i] >= (int)ctx->mats.size())
            return fail(ctx, SGPU_ERR_ARG, "fiber %d: material index %d out of range", i, mat_index[i]);
        if (!(area[i] > 0.0)) return fail(ctx, SGPU_ERR_ARG, "fiber %d: area must be > 0", i);
        cells[i] = Cell{y[i], dim == 3 ? z[i] : 0.0, area[i], mat_index[i]};
        area_sum += area[i];
    }
    if (!(area_sum > 0.0)) return fail(ctx, SGPU_ERR_ARG, "section total area must be > 0");
    SecDefHost d;
    d.dim = dim;
    int rc = build_def(ctx, d, cells);
    if (rc) return rc;
    *def_id = (int32_t)ctx->defs.size();
    ctx->defs.push_back(std::move(d));
    return SGPU_OK;
}

int sgpu_section_instances_alloc(sgpu_ctx *ctx, int32_t def_id, int n, int32_t *first_instance) {
    if (!ctx) return SGPU_ERR_ARG;
    if (ctx->finalized) return fail(ctx, SGPU_ERR_STATE, "instances must be allocated before finalize");
    if (def_id < 0 || def_id >= (int)ctx->defs.size() || n < 0 || !first_instance)
        return fail(ctx, SGPU_ERR_ARG, "bad instance allocation");
    *first_instance = ctx->defs[def_id].n_inst;
    ctx->defs[def_id].n_inst += n;
    return SGPU_OK;
}

int sgpu_sections_finalize(sgpu_ctx *ctx) {
    if (!ctx) return SGPU_ERR_ARG;
    if (ctx->finalized) return fail(ctx, SGPU_ERR_STATE, "already finalized");
    HIP_OK(ctx, hipSetDevice(ctx->device));
    std::vector<SecDefDev> ddefs;
    std::vector<double> ty, tz, ta;
    std::vector<int32_t> tm, tx, tc, tb, tcs, tas, tam;
    std::vector<double> tay, taz, taa;
    ctx->max_aux = 0;
    int64_t state = 0, aux = 0, fibers = 0;
    for (SecDefHost &d : ctx->defs) {
        d.slot_base = (int64_t)ty.size();
        d.batch_base = (int64_t)tb.size();
        d.chunk_base = (int64_t)tcs.size();
        tcs.insert(tcs.end(), d.chunk_slot.begin(), d.chunk_slot.end());
        d.aux_slot_base = (int64_t)tas.size();
        tas.insert(tas.end(), d.aux_slot.begin(), d.aux_slot.end());
        for (int32_t sl : d.aux_slot) {
            tam.push_back(d.mat[sl]);
            tay.push_back(d.y[sl]);
            taz.push_back(d.z[sl]);
            taa.push_back(d.area[sl]);
        }
        if (d.n_aux > ctx->max_aux) ctx->max_aux = d.n_aux;
        d.state_base = state;
        d.aux_base = aux;
        ty.insert(ty.end(), d.y.begin(), d.y.end());
        tz.insert(tz.end(), d.z.begin(), d.z.end());
        ta.insert(ta.end(), d.area.begin(), d.area.end());
        tm.insert(tm.end(), d.mat.begin(), d.mat.end());
        tx.insert(tx.end(), d.aux.begin(), d.aux.end());
        tc.insert(tc.end(), d.chunk.begin(), d.chunk.end());
        tb.insert(tb.end(), d.batch_chunk_begin.begin(), d.batch_chunk_begin.end());
        state += (int64_t)d.n_inst * (int64_t)d.y.size();
        aux += (int64_t)d.n_inst * d.n_aux;
        fibers += (int64_t)d.n_inst * d.n_fibers;
        SecDefDev dd{};
        dd.dim = d.dim;
        dd.n_slots = (int32_t)d.y.size();
        dd.n_fibers = d.n_fibers;
        dd.n_aux = d.n_aux;
        dd.n_batches = (int32_t)d.batch_chunk_begin.size() - 1;
        dd.n_elastic_chunks = d.n_elastic_chunks;
        dd.zero_k12 = d.zero_k12;
        dd.n_inst = d.n_inst;
        dd.slot_base = d.slot_base;
        dd.batch_base = d.batch_base;
        dd.chunk_base = d.chunk_base;
        dd.state_base = d.state_base;
        dd.aux_base = d.aux_base;
        dd.aux_slot_base = d.aux_slot_base;
        for (int k = 0; k < 3; ++k) dd.init_flex[k] = d.init_flex[k];
        dd.init_flex_valid = d.init_flex_valid;
        ddefs.push_back(dd);
    }
    ctx->total_slots = state;
    ctx->total_aux = aux;
    ctx->total_fibers = fibers;
    ctx->total_table_slots = (int64_t)ty.size();
    int rc;
    if ((rc = upload(ctx, &ctx->d_mats, ctx->mats))) return rc;
    if ((rc = upload(ctx, &ctx->d_defs, ddefs))) return rc;
    if ((rc = upload(ctx, &ctx->d_tab_y, ty))) return rc;
    if ((rc = upload(ctx, &ctx->d_tab_z, tz))) return rc;
    if ((rc = upload(ctx, &ctx->d_tab_area, ta))) return rc;
    if ((rc = upload(ctx, &ctx->d_tab_mat, tm))) return rc;
    if ((rc = upload(ctx, &ctx->d_tab_aux, tx))) return rc;
    if ((rc = upload(ctx, &ctx->d_tab_auxslot, tas))) return rc;
    if ((rc = upload(ctx, &ctx->d_tab_auxmat, tam))) return rc;
    if ((rc = upload(ctx, &ctx->d_tab_auxy, tay))) return rc;
    if ((rc = upload(ctx, &ctx->d_tab_auxz, taz))) return rc;
    if ((rc = upload(ctx, &ctx->d_tab_auxarea, taa))) return rc;
    if ((rc = upload(ctx, &ctx->d_tab_chunk, tc))) return rc;
    if ((rc = upload(ctx, &ctx->d_tab_bcb, tb))) return rc;
    if ((rc = upload(ctx, &ctx->d_tab_cslot, tcs))) return rc;
    // initial states, built per definition and replicated over its instances
    std::vector<double> e(state), s(state), tg(state), h0(state), h1(state), a0(aux, 0.0);
    std::vector<int32_t> kon(aux, 0);
    for (const SecDefHost &d : ctx->defs) {
        const int ns = (int)d.y.size();
        for (int inst = 0; inst < d.n_inst; ++inst) {
            const int64_t sb = d.state_base + (int64_t)inst * ns;
            const int64_t ab = d.aux_base + (int64_t)inst * d.n_aux;
            for (int k = 0; k < ns; ++k) {
                const int mi = d.mat[k];
                const int ax = d.aux[k];
                MatDev dead{};
                dead.type = kDead;
                init_slot_state(mi >= 0 ? ctx->mats[mi] : dead, d.slot_real[k] != 0, &e[sb + k], &s[sb + k],
                                &tg[sb + k], &h0[sb + k], &h1[sb + k], ax >= 0 ? &a0[ab + ax] : nullptr,
                                ax >= 0 ? &kon[ab + ax] : nullptr);
            }
        }
    }
    StateArrays *arr[2] = {&ctx->c, &ctx->t};
    for (StateArrays *A : arr) {
        if ((rc = upload(ctx, &A->eps, e))) return rc;
        if ((rc = upload(ctx, &A->sig, s))) return rc;
        if ((rc = upload(ctx, &A->tan, tg))) return rc;
        if ((rc = upload(ctx, &A->h0, h0))) return rc;
        if ((rc = upload(ctx, &A->h1, h1))) return rc;
        if ((rc = upload(ctx, &A->a0, a0))) return rc;
        std::vector<double> zero(aux, 0.0);
        if ((rc = upload(ctx, &A->a1, zero))) return rc;
        if ((rc = upload(ctx, &A->a2, zero))) return rc;
        if ((rc = upload(ctx, &A->a3, zero))) return rc;
        if ((rc = upload(ctx, &A->a4, zero))) return rc;
        if ((rc = upload(ctx, &A->kon, kon))) return rc;
    }
    ctx->finalized = true;
    return SGPU_OK;
}

int sgpu_section_def_info(sgpu_ctx *ctx, int32_t def_id, double *out8) {
    if (!ctx) return SGPU_ERR_ARG;
    if (def_id < 0 || def_id >= (int)ctx->defs.size() || !out8) return fail(ctx, SGPU_ERR_ARG, "bad def id");
    const SecDefHost &d = ctx->defs[def_id];
    out8[0] = d.n_fibers;
    out8[1] = (double)d.y.size();
    out8[2] = d.y_bar;
    out8[3] = d.z_bar;
    out8[4] = d.init_flex_valid;
    out8[5] = d.init_flex[0];
    out8[6] = d.init_flex[1];
    out8[7] = d.init_flex[2];
    return SGPU_OK;
}

int sgpu_section_eval_device(sgpu_ctx *ctx, int n, const sgpu_section_point *d_points,
                             sgpu_section_result *d_results) {
    if (!ctx) return SGPU_ERR_ARG;
    if (!ctx->finalized) return fail(ctx, SGPU_ERR_STATE, "sgpu_sections_finalize has not run");
    if (n < 0 || (n > 0 && (!d_points || !d_results))) return fail(ctx, SGPU_ERR_ARG, "bad point buffers");
    if (n == 0) return SGPU_OK;
    const int64_t need = (int64_t)n * ctx->max_aux * 2;
    if (need > ctx->fsks_cap) {
        if (ctx->d_fsks) HIP_OK(ctx, hipFree(ctx->d_fsks));
        ctx->d_fsks = nullptr;
        ctx->fsks_cap = 0;
        HIP_OK(ctx, hipMalloc((void **)&ctx->d_fsks, sizeof(double) * need));
        ctx->fsks_cap = need;
    }
    return launch_section_passes(ctx, n, d_points, d_results, ctx->d_fsks, true);
}

int sgpu_section_eval_device_noalloc(sgpu_ctx *ctx, int n, const sgpu_section_point *d_points,
                                     sgpu_section_result *d_results) {
    if (!ctx) return SGPU_ERR_ARG;
    if (!ctx->finalized) return fail(ctx, SGPU_ERR_STATE, "sgpu_sections_finalize has not run");
    if (n < 0 || (n > 0 && (!d_points || !d_results))) return fail(ctx, SGPU_ERR_ARG, "bad point buffers");
    if (n == 0) return SGPU_OK;
    const int64_t need = (int64_t)n * ctx->max_aux * 2;
    if (need > kNoAllocScratchDoubles)
        return fail(ctx, SGPU_ERR_ARG,
                    "%d points x %d aux slots need %lld scratch doubles; the fixed scratch holds %lld", n,
                    ctx->max_aux, (long long)need, (long long)kNoAllocScratchDoubles);
    double *fsks = noalloc_scratch(ctx->device);
    if (!fsks) return fail(ctx, SGPU_ERR_STATE, "no scratch address for device %d", ctx->device);
    return launch_section_passes(ctx, n, d_points, d_results, fsks, false);
}

int sgpu_section_eval(sgpu_ctx *ctx, int n, const sgpu_section_point *points, sgpu_section_result *results) {
    if (!ctx) return SGPU_ERR_ARG;
    if (!ctx->finalized) return fail(ctx, SGPU_ERR_STATE, "sgpu_sections_finalize has not run");
    if (n < 0 || (n > 0 && (!points || !results))) return fail(ctx, SGPU_ERR_ARG, "bad point buffers");
    if (n == 0) return SGPU_OK;
    int rc = check_points(ctx, n, points);
    if (rc) return rc;
    if ((rc = ensure_scratch(ctx, n))) return rc;
    HIP_OK(ctx, hipMemcpyAsync(ctx->d_pts, points, sizeof(sgpu_section_point) * n, hipMemcpyHostToDevice,
                               ctx->stream));
    if ((rc = sgpu_section_eval_device(ctx, n, ctx->d_pts, ctx->d_res))) return rc;
    HIP_OK(ctx, hipMemcpyAsync(results, ctx->d_res, sizeof(sgpu_section_result) * n, hipMemcpyDeviceToHost,
                               ctx->stream));
    HIP_OK(ctx, hipStreamSynchronize(ctx->stream));
    for (int p = 0; p < n; ++p)
        if (results[p].valid == 0) return SGPU_SINGULAR_SECTION;
    return SGPU_OK;
}

int sgpu_sections_commit_all(sgpu_ctx *ctx) {
    if (!ctx) return SGPU_ERR_ARG;
    if (!ctx->finalized) return fail(ctx, SGPU_ERR_STATE, "sgpu_sections_finalize has not run");
    {
        ProfScope ps(ctx, SGPU_KERNEL_COMMIT);
        launch_copy_state(ctx->stream, ctx->c, ctx->t, ctx->total_slots, ctx->total_aux);
    }
    HIP_OK(ctx, hipGetLastError());
    return SGPU_OK;
}

int sgpu_sections_revert_all(sgpu_ctx *ctx) {
    if (!ctx) return SGPU_ERR_ARG;
    if (!ctx->finalized) return fail(ctx, SGPU_ERR_STATE, "sgpu_sections_finalize has not run");
    {
        ProfScope ps(ctx, SGPU_KERNEL_COMMIT);
        launch_copy_state(ctx->stream, ctx->t, ctx->c, ctx->total_slots, ctx->total_aux);
    }
    HIP_OK(ctx, hipGetLastError());
    return SGPU_OK;
}

static int copy_listed(sgpu_ctx *ctx, int n, const int32_t *def_ids, const int32_t *insts, bool commit) {
    if (!ctx) return SGPU_ERR_ARG;
    if (!ctx->finalized) return fail(ctx, SGPU_ERR_STATE, "sgpu_sections_finalize has not run");
    if (n < 0 || (n > 0 && (!def_ids || !insts))) return fail(ctx, SGPU_ERR_ARG, "bad instance list");
    if (n == 0) return SGPU_OK;
    for (int i = 0; i < n; ++i) {
        if (def_ids[i] < 0 || def_ids[i] >= (int)ctx->defs.size() || insts[i] < 0 ||
            insts[i] >= ctx->defs[def_ids[i]].n_inst)
            return fail(ctx, SGPU_ERR_ARG, "list entry %d out of range", i);
    }
    int rc = ensure_list(ctx, n);
    if (rc) return rc;
    HIP_OK(ctx, hipMemcpyAsync(ctx->d_list, def_ids, sizeof(int32_t) * n, hipMemcpyHostToDevice, ctx->stream));
    HIP_OK(ctx, hipMemcpyAsync(ctx->d_list + n, insts, sizeof(int32_t) * n, hipMemcpyHostToDevice, ctx->stream));
    ProfScope ps(ctx, SGPU_KERNEL_COMMIT);
    if (commit) launch_copy_instances(ctx->stream, ctx->d_defs, ctx->d_list, ctx->d_list + n, n, ctx->c, ctx->t);
    else launch_copy_instances(ctx->stream, ctx->d_defs, ctx->d_list, ctx->d_list + n, n, ctx->t, ctx->c);
    HIP_OK(ctx, hipGetLastError());
    HIP_OK(ctx, hipStreamSynchronize(ctx->stream));
    return SGPU_OK;
}

int sgpu_sections_commit(sgpu_ctx *ctx, int n, const int32_t *def_ids, const int32_t *instances) {
    return copy_listed(ctx, n, def_ids, instances, true);
}

int sgpu_sections_revert(sgpu_ctx *ctx, int n, const int32_t *def_ids, const int32_t *instances) {
    return copy_listed(ctx, n, def_ids, instances, false);
}

int sgpu_fiber_state_read(sgpu_ctx *ctx, int32_t def_id, int32_t instance, int which, double *out) {
    if (!ctx) return SGPU_ERR_ARG;
    if (!ctx->finalized) return fail(ctx, SGPU_ERR_STATE, "sgpu_sections_finalize has not run");
    if (def_id < 0 || def_id >= (int)ctx->defs.size() || !out || (which != 0 && which != 1))
        return fail(ctx, SGPU_ERR_ARG, "bad fiber state request");
    const SecDefHost &d = ctx->defs[def_id];
    if (instance < 0 || instance >= d.n_inst) return fail(ctx, SGPU_ERR_ARG, "instance out of range");
    const StateArrays &A = which == 0 ? ctx->t : ctx->c;
    const size_t ns = d.y.size();
    const int64_t sb = d.state_base + (int64_t)instance * (int64_t)ns;
    std::vector<double> e(ns), s(ns), tg(ns);
    HIP_OK(ctx, hipStreamSynchronize(ctx->stream));
    HIP_OK(ctx, hipMemcpy(e.data(), A.eps + sb, ns * sizeof(double), hipMemcpyDeviceToHost));
    HIP_OK(ctx, hipMemcpy(s.data(), A.sig + sb, ns * sizeof(double), hipMemcpyDeviceToHost));
    HIP_OK(ctx, hipMemcpy(tg.data(), A.tan + sb, ns * sizeof(double), hipMemcpyDeviceToHost));
    for (int f = 0; f < d.n_fibers; ++f) {
        const int slot = d.fiber_slot[f];
        const bool elastic = ctx->mats[d.mat[slot]].type == kElastic;
        out[3 * f + 0] = elastic ? 0.0 : e[slot];
        out[3 * f + 1] = elastic ? 0.0 : s[slot];
        out[3 * f + 2] = elastic ? 0.0 : tg[slot];
    }
    return SGPU_OK;
}

int sgpu_fiber_totals(sgpu_ctx *ctx, int64_t *fibers, int64_t *slots) {
    if (!ctx || !fibers || !slots) return SGPU_ERR_ARG;
    *fibers = ctx->total_fibers;
    *slots = ctx->total_slots;
    return SGPU_OK;
}

}  // extern "C"

int sgpu_profile_enable(sgpu_ctx *ctx, int on) {
    if (!ctx) return SGPU_ERR_ARG;
    HIP_OK(ctx, hipStreamSynchronize(ctx->stream));
    for (auto &pe : ctx->prof_pending) {
        ctx->prof_free.push_back(pe.second.first);
        ctx->prof_free.push_back(pe.second.second);
    }
    ctx->prof_pending.clear();
    for (int k = 0; k < SGPU_KERNEL_COUNT; ++k) {
        ctx->prof_ms[k] = 0.0;
        ctx->prof_n[k] = 0;
    }
    ctx->prof_on = on != 0;
    return SGPU_OK;
}

int sgpu_profile_read(sgpu_ctx *ctx, int kernel, double *total_ms, int64_t *launches) {
    if (!ctx) return SGPU_ERR_ARG;
    if (kernel < 0 || kernel >= SGPU_KERNEL_COUNT) return fail(ctx, SGPU_ERR_ARG, "bad kernel id %d", kernel);
    HIP_OK(ctx, hipStreamSynchronize(ctx->stream));
    for (auto &pe : ctx->prof_pending) {
        float ms = 0.0f;
        HIP_OK(ctx, hipEventElapsedTime(&ms, pe.second.first, pe.second.second));
        ctx->prof_ms[pe.first] += ms;
        ctx->prof_n[pe.first] += 1;
        ctx->prof_free.push_back(pe.second.first);
        ctx->prof_free.push_back(pe.second.second);
    }
    ctx->prof_pending.clear();
    if (total_ms) *total_ms = ctx->prof_ms[kernel];
    if (launches) *launches = ctx->prof_n[kernel];
    return SGPU_OK;
}
