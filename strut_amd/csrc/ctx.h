// Internal state of one libstrut_gpu context (shared by the host
// translation units; not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/strut_gpu.h"
#include "strut_gpu_types.h"

struct sgpu_ctx;
namespace sgpu {
struct FrameHost;
void frame2d_host_free(FrameHost *fh);
struct Elem3dHost;
void elem3d_host_free(Elem3dHost *h);
int elem3d_count(const sgpu_ctx *ctx);  // sgpu_elem3d elements of the context
struct MpcHost;
void mpc_host_free(MpcHost *h);
struct Frame3dHost;
void frame3d_host_free(Frame3dHost *h);
struct Elem2dHost;
void elem2d_host_free(Elem2dHost *h);
}

namespace sgpu {

constexpr int kW = 8;       // SIMD width of the reference's family loops
constexpr int kBatchSlots = 64;

struct SecDefHost {
    int dim = 2;
    int n_fibers = 0;
    double y_bar = 0.0, z_bar = 0.0;
    // slot layout
    std::vector<double> y, z, area;
    std::vector<int32_t> mat, aux, chunk;
    std::vector<int32_t> batch_chunk_begin;  // n_batches + 1
    std::vector<int32_t> chunk_slot;         // first slot of each chunk
    std::vector<int32_t> aux_slot;           // slot of each aux index
    std::vector<int32_t> fiber_slot;         // input fiber -> slot
    std::vector<uint8_t> slot_real;          // 1 for real material fibers
    int n_aux = 0;
    int n_elastic_chunks = 0;
    int zero_k12 = 0;
    int init_flex_valid = 0;
    double init_flex[3] = {0, 0, 0};
    int n_inst = 0;
    int64_t slot_base = 0, batch_base = 0, chunk_base = 0, state_base = 0, aux_base = 0, aux_slot_base = 0;
};

}  // namespace sgpu

using namespace sgpu;

struct sgpu_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    std::vector<MatDev> mats;
    std::vector<SecDefHost> defs;
    bool finalized = false;
    int64_t total_slots = 0, total_aux = 0, total_fibers = 0, total_table_slots = 0;
    // device buffers
    MatDev *d_mats = nullptr;
    SecDefDev *d_defs = nullptr;
    double *d_tab_y = nullptr, *d_tab_z = nullptr, *d_tab_area = nullptr;
    int32_t *d_tab_mat = nullptr, *d_tab_aux = nullptr, *d_tab_chunk = nullptr, *d_tab_bcb = nullptr,
            *d_tab_cslot = nullptr, *d_tab_auxslot = nullptr, *d_tab_auxmat = nullptr;
    double *d_tab_auxy = nullptr, *d_tab_auxz = nullptr, *d_tab_auxarea = nullptr;
    int max_aux = 0;  // most aux slots of any definition
    StateArrays c{}, t{};
    std::vector<void *> allocations;
    // scratch for the host-buffer entry points
    sgpu_section_point *d_pts = nullptr;
    sgpu_section_result *d_res = nullptr;
    int scratch_cap = 0;
    int32_t *d_list = nullptr;
    double *d_fsks = nullptr;  // (sigma A, E_t A) of aux slots, per point: the aux pre-pass output
    int64_t fsks_cap = 0;
    int list_cap = 0;
    // per-kernel timing (sgpu_profile_*): event pairs recorded around each launch
    bool prof_on = false;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> prof_pending;
    std::vector<hipEvent_t> prof_free;
    double prof_ms[SGPU_KERNEL_COUNT] = {};
    int64_t prof_n[SGPU_KERNEL_COUNT] = {};
    // 2D frame transient instances (frame2d_api.cpp)
    sgpu::FrameHost *frames = nullptr;
    // 3D fiber beam-column elements (elem3d_api.cpp)
    sgpu::Elem3dHost *elems3d = nullptr;
    // multi-point constraint tables (mpc.hip)
    sgpu::MpcHost *mpc = nullptr;
    // 3D frame transient (frame3d.hip)
    sgpu::Frame3dHost *frame3d = nullptr;
    // 2D elements outside any frame (elem2d_api.cpp)
    sgpu::Elem2dHost *elems2d = nullptr;
    SlotTables tables() const {
        SlotTables tb;
        tb.y = d_tab_y;
        tb.z = d_tab_z;
        tb.area = d_tab_area;
        tb.mat = d_tab_mat;
        tb.aux = d_tab_aux;
        tb.chunk = d_tab_chunk;
        tb.batch_chunk_begin = d_tab_bcb;
        tb.chunk_slot = d_tab_cslot;
        tb.aux_slot = d_tab_auxslot;
        tb.aux_mat = d_tab_auxmat;
        tb.aux_y = d_tab_auxy;
        tb.aux_z = d_tab_auxz;
        tb.aux_area = d_tab_auxarea;
        return tb;
    }
};

namespace sgpu {

inline int fail(sgpu_ctx *ctx, int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (ctx) ctx->err = buf;
    return code;
}

#define HIP_OK(ctx, expr)                                                                         \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess) return fail(ctx, SGPU_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

template <typename T>
inline int dalloc(sgpu_ctx *ctx, T **p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void **)p, count * sizeof(T));
    if (e != hipSuccess) return fail(ctx, SGPU_ERR_HIP, "hipMalloc(%zu bytes): %s", count * sizeof(T), hipGetErrorString(e));
    ctx->allocations.push_back((void *)*p);
    return SGPU_OK;
}

// Record the start (a) / end (b) event of one timed launch when profiling is on.
inline hipEvent_t prof_event(sgpu_ctx *ctx) {
    hipEvent_t e = nullptr;
    if (!ctx->prof_free.empty()) {
        e = ctx->prof_free.back();
        ctx->prof_free.pop_back();
    } else if (hipEventCreate(&e) != hipSuccess) {
        return nullptr;
    }
    return e;
}

struct ProfScope {
    sgpu_ctx *ctx;
    int kernel;
    hipEvent_t a = nullptr;
    ProfScope(sgpu_ctx *c, int k) : ctx(c), kernel(k) {
        // kernel < 0: never timed (the no-allocation entry points create no events)
        if (kernel >= 0 && ctx->prof_on && (a = prof_event(ctx)) != nullptr) (void)hipEventRecord(a, ctx->stream);
    }
    ~ProfScope() {
        if (!a) return;
        hipEvent_t b = prof_event(ctx);
        if (!b) return;
        (void)hipEventRecord(b, ctx->stream);
        ctx->prof_pending.push_back({kernel, {a, b}});
    }
};

template <typename T>
inline int upload(sgpu_ctx *ctx, T **p, const std::vector<T> &v) {
    int rc = dalloc(ctx, p, v.size());
    if (rc) return rc;
    if (!v.empty()) HIP_OK(ctx, hipMemcpy(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return SGPU_OK;
}

}  // namespace sgpu
