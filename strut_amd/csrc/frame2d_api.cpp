// Host side of the 2D fiber-frame transient: loads a frame (plain DOF
// numbering, band half-width from the element free maps, gather lists),
// adds (frame, ground motion) instances and launches frame2d.hip.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "ctx.h"
#include "frame2d_types.h"

namespace sgpu {

const MlpDev *mlp_device_view(const sgpu_mlp *m);
size_t frame2d_lds_bytes(int n_free, int bw);
hipError_t frame2d_set_lds(size_t lds_bytes);
hipError_t launch_frame2d_round(hipStream_t stream, const Frame2dArgs &args, int n_pairs, int n_inst,
                                size_t lds_bytes);
hipError_t launch_frame2d_persistent(hipStream_t stream, const Frame2dArgs &args, int n_inst, int max_steps,
                                     size_t lds_bytes);
hipError_t frame2d_queue_set_lds(size_t lds_bytes, int *blocks_per_cu);
hipError_t launch_frame2d_queue(hipStream_t stream, const Frame2dArgs &args, const QueueDev &Q, int workers,
                                int max_steps, size_t lds_bytes);

struct FrameDef {
    FrameDev dev{};
    int n_nodes = 0, total_dofs = 0;
    std::vector<ElemDev> elems;
    std::vector<int32_t> elem_def;
    std::vector<double> mass_ff, exc_ff;
    std::vector<int32_t> exc_flag, rec_fidx, kg_off, kg, fg_off, fg;
};

struct InstHost {
    int frame = 0;
    int n_steps = 0;
    std::vector<double> ag;
    std::vector<int32_t> elem_inst;
    std::vector<int32_t> range_def, range_first, range_count;  // section instances per definition
};

struct FrameHost {
    std::vector<FrameDef> frames;
    std::vector<InstHost> insts;
    bool finalized = false;
    Frame2dArgs args{};
    size_t lds = 0;
    int max_steps = 0;
    InstState *d_state = nullptr;
    unsigned long long *d_sec_evals = nullptr;  // per instance
    double *d_hist = nullptr;
    int32_t *d_iters = nullptr;
    std::vector<int64_t> hist_base, iters_base;
    int n_pairs = 0;
    int64_t max_rounds = 0;  // rounds after which every instance has finished
    // rounds captured in a HIP graph and replayed
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    int graph_rounds = 0;
    // schedule: requested (SGPU_SCHEDULE_*), the one in use, and the last
    // measured rates of both (GM-steps/s)
    int schedule = SGPU_SCHEDULE_AUTO;
    int chosen = SGPU_SCHEDULE_ROUNDS;
    double rate_rounds = 0.0, rate_persistent = 0.0;
    int tunings = 0;
    // work-queue schedule
    QueueDev queue{};
    int workers = 0;
    double rate_queue = 0.0;
    // learned predictor (per frame) and training-sample recorder
    MlpDev *d_preds = nullptr;
    bool record_samples = false;
    float *d_samples = nullptr;
    std::vector<int64_t> sample_base;
};

void frame2d_host_free(FrameHost *fh) {
    if (!fh) return;
    if (fh->graph_exec) (void)hipGraphExecDestroy(fh->graph_exec);
    if (fh->graph) (void)hipGraphDestroy(fh->graph);
    delete fh;
}

}  // namespace sgpu

using namespace sgpu;

namespace {

FrameHost &frames_of(sgpu_ctx *ctx) {
    if (!ctx->frames) ctx->frames = new FrameHost();
    return *ctx->frames;
}

// time_series.mojo:71-96 (Path with dt)
double path_eval(const sgpu_path_series &ts, double t) {
    const int n = ts.n;
    if (n == 0) return 0.0;
    if (t < ts.start_time) return 0.0;
    const double incr = (t - ts.start_time) / ts.dt;
    const long incr1 = (long)std::floor(incr);
    const long incr2 = incr1 + 1;
    if (incr2 >= n) {
        if (ts.use_last) return ts.factor * ts.values[n - 1];
        return 0.0;
    }
    const double v1 = ts.values[incr1], v2 = ts.values[incr2];
    return ts.factor * (v1 + (v2 - v1) * (incr - (double)incr1));
}

}  // namespace

extern "C" {

int sgpu_frame2d_create(sgpu_ctx *ctx, const sgpu_frame2d_desc *d, int32_t *frame_id) {
    if (!ctx || !d || !frame_id) return SGPU_ERR_ARG;
    FrameHost &fh = frames_of(ctx);
    if (fh.finalized) return fail(ctx, SGPU_ERR_STATE, "frames must be created before sgpu_frame2d_finalize");
    if (d->n_nodes <= 0 || d->n_elem <= 0 || !d->x || !d->y || !d->constrained || !d->mass || !d->elem_nodes ||
        !d->elem_def || !d->elem_nip || !d->elem_xi || !d->elem_w)
        return fail(ctx, SGPU_ERR_ARG, "incomplete frame description");
    if (!(d->dt > 0.0) || !(d->beta > 0.0) || d->max_iters <= 0)
        return fail(ctx, SGPU_ERR_ARG, "transient needs dt > 0, beta > 0, max_iters > 0");
    FrameDef f;
    f.n_nodes = d->n_nodes;
    f.total_dofs = 3 * d->n_nodes;
    std::vector<int> free_idx(f.total_dofs, -1);
    int n = 0;
    for (int i = 0; i < f.total_dofs; ++i)
        if (!d->constrained[i]) free_idx[i] = n++;
    if (n == 0) return fail(ctx, SGPU_ERR_ARG, "no free dofs");
    int hbw = 0;
    for (int e = 0; e < d->n_elem; ++e) {
        ElemDev el{};
        const int ni = d->elem_nodes[2 * e], nj = d->elem_nodes[2 * e + 1];
        if (ni < 0 || nj < 0 || ni >= d->n_nodes || nj >= d->n_nodes)
            return fail(ctx, SGPU_ERR_ARG, "element %d: node index out of range", e);
        const int def = d->elem_def[e];
        if (def < 0 || def >= (int)ctx->defs.size() || ctx->defs[def].dim != 2)
            return fail(ctx, SGPU_ERR_ARG, "element %d: needs a 2D fiber-section definition", e);
        const SecDefHost &sd = ctx->defs[def];
        const int type = d->elem_type ? d->elem_type[e] : 0;
        if (type != 0 && type != 1) return fail(ctx, SGPU_ERR_ARG, "element %d: unknown element type %d", e, type);
        if (type == 0 && !sd.init_flex_valid)
            return fail(ctx, SGPU_ERR_ARG, "element %d: section initial flexibility is singular", e);
        if ((int)sd.batch_chunk_begin.size() - 1 > 63)
            return fail(ctx, SGPU_ERR_ARG, "element %d: sections of more than 63 x 64 slots are not on the frame path", e);
        bool any_nonlinear = false;
        for (size_t s = 0; s < sd.mat.size(); ++s)
            if (sd.mat[s] >= 0 && ctx->mats[sd.mat[s]].type != kElastic) any_nonlinear = true;
        el.all_elastic = any_nonlinear ? 0 : 1;
        const int nip = d->elem_nip[e];
        if (nip < 1 || nip > kMaxNipDev)
            return fail(ctx, SGPU_ERR_ARG, "element %d: 1..%d integration points on the device path", e, kMaxNipDev);
        el.nip = nip;
        el.def = def;
        el.type = type;
        const double dx = d->x[nj] - d->x[ni], dy = d->y[nj] - d->y[ni];
        el.L = std::hypot(dx, dy);
        if (el.L == 0.0) return fail(ctx, SGPU_ERR_ARG, "element %d: zero length", e);
        el.inv_L = 1.0 / el.L;
        el.c = type == 1 ? dx / el.L : dx * el.inv_L;
        el.s = type == 1 ? dy / el.L : dy * el.inv_L;
        el.L2 = std::sqrt(dx * dx + dy * dy);
        el.c2 = dx / el.L2;
        el.s2 = dy / el.L2;
        for (int ip = 0; ip < nip; ++ip) {
            el.xi[ip] = d->elem_xi[10 * e + ip];
            el.w[ip] = d->elem_w[10 * e + ip];
        }
        el.if00 = sd.init_flex[0];
        el.if01 = sd.init_flex[1];
        el.if11 = sd.init_flex[2];
        el.if_valid = sd.init_flex_valid;
        for (int k = 0; k < 3; ++k) {
            el.fmap[k] = free_idx[3 * ni + k];
            el.fmap[3 + k] = free_idx[3 * nj + k];
        }
        for (int a = 0; a < 6; ++a)  // lsb.mojo:1481-1506
            for (int b = a + 1; b < 6; ++b)
                if (el.fmap[a] >= 0 && el.fmap[b] >= 0) hbw = std::max(hbw, std::abs(el.fmap[a] - el.fmap[b]));
        el.q_off = 0;
        f.elems.push_back(el);
        f.elem_def.push_back(def);
    }
    int q_off = 0;
    for (ElemDev &el : f.elems) {
        el.q_off = q_off;
        q_off += 3 + 2 * el.nip + 3;
    }
    const int bw = 2 * hbw + 1;
    // gather lists in element order
    std::vector<std::vector<int32_t>> kl((size_t)n * bw), fl(n);
    for (int e = 0; e < d->n_elem; ++e) {
        const ElemDev &el = f.elems[e];
        for (int a = 0; a < 6; ++a) {
            if (el.fmap[a] < 0) continue;
            for (int b = 0; b < 6; ++b) {
                if (el.fmap[b] < 0) continue;
                const int i = el.fmap[a], j = el.fmap[b];
                kl[(size_t)i * bw + (j - i + hbw)].push_back(42 * e + 6 * a + b);
            }
            fl[el.fmap[a]].push_back(42 * e + 36 + a);
        }
    }
    f.kg_off.push_back(0);
    for (auto &v : kl) {
        f.kg.insert(f.kg.end(), v.begin(), v.end());
        f.kg_off.push_back((int32_t)f.kg.size());
    }
    f.fg_off.push_back(0);
    for (auto &v : fl) {
        f.fg.insert(f.fg.end(), v.begin(), v.end());
        f.fg_off.push_back((int32_t)f.fg.size());
    }
    for (int i = 0; i < f.total_dofs; ++i) {
        if (free_idx[i] < 0) continue;
        f.mass_ff.push_back(d->mass[i]);
        const bool exc = d->exc_dir > 0 && (i % 3) + 1 == d->exc_dir;
        f.exc_ff.push_back(exc ? d->mass[i] : 0.0);
        f.exc_flag.push_back(exc ? 1 : 0);
    }
    bool has_mass = false;
    for (double m : f.mass_ff) has_mass = has_mass || m != 0.0;
    if (!has_mass) return fail(ctx, SGPU_ERR_ARG, "transient_nonlinear requires masses on free dofs");
    for (int r = 0; r < d->n_record; ++r) {
        const int dof = d->record_dofs[r];
        if (dof < 0 || dof >= f.total_dofs) return fail(ctx, SGPU_ERR_ARG, "record dof %d out of range", dof);
        f.rec_fidx.push_back(free_idx[dof]);
    }
    FrameDev &fd = f.dev;
    fd.n_elem = d->n_elem;
    fd.n_free = n;
    fd.hbw = hbw;
    fd.bw = bw;
    fd.n_record = d->n_record;
    fd.max_iters = d->max_iters;
    fd.exc_dir = d->exc_dir;
    fd.has_damping = (d->alpha_m != 0.0 || d->beta_k != 0.0 || d->beta_k_init != 0.0) ? 1 : 0;
    fd.q_len = q_off;
    fd.dt = d->dt;
    fd.gamma = d->gamma;
    fd.beta = d->beta;
    fd.tol = d->tol;
    // transient_nonlinear.mojo:1052-1055, :559, :510
    fd.a0 = 1.0 / (d->beta * d->dt * d->dt);
    fd.a1 = d->gamma / (d->beta * d->dt);
    fd.a2 = 1.0 / (d->beta * d->dt);
    fd.a3 = 1.0 / (2.0 * d->beta) - 1.0;
    fd.coeff = d->dt * d->dt * (0.5 - d->beta);
    fd.omg = 1.0 - d->gamma;
    fd.alpha_m = d->alpha_m;
    fd.beta_k = d->beta_k;
    fd.beta_k_init = d->beta_k_init;
    const size_t lds = frame2d_lds_bytes(n, bw);
    if (lds > 136 * 1024)  // the persistent kernel adds its static element LDS (about 19 KB)
        return fail(ctx, SGPU_ERR_ARG, "frame band of %d x %d does not fit the LDS of one workgroup", n, bw);
    *frame_id = (int32_t)fh.frames.size();
    fh.frames.push_back(std::move(f));
    return SGPU_OK;
}

int sgpu_frame2d_add_instance(sgpu_ctx *ctx, int32_t frame_id, int32_t n_steps, const sgpu_path_series *accel,
                              int32_t *instance_id) {
    if (!ctx || !accel || !instance_id) return SGPU_ERR_ARG;
    FrameHost &fh = frames_of(ctx);
    if (fh.finalized || ctx->finalized)
        return fail(ctx, SGPU_ERR_STATE, "instances must be added before sgpu_frame2d_finalize");
    if (frame_id < 0 || frame_id >= (int)fh.frames.size() || n_steps <= 0)
        return fail(ctx, SGPU_ERR_ARG, "bad frame id or step count");
    if (accel->n < 0 || (accel->n > 0 && (!accel->values || !(accel->dt > 0.0))))
        return fail(ctx, SGPU_ERR_ARG, "bad acceleration series");
    const FrameDef &f = fh.frames[frame_id];
    InstHost in;
    in.frame = frame_id;
    in.n_steps = n_steps;
    in.ag.resize(n_steps);
    for (int s = 0; s < n_steps; ++s) in.ag[s] = path_eval(*accel, (double)(s + 1) * f.dev.dt);
    for (int e = 0; e < f.dev.n_elem; ++e) {
        int32_t first = 0;
        int rc = sgpu_section_instances_alloc(ctx, f.elem_def[e], f.elems[e].nip, &first);
        if (rc) return rc;
        in.elem_inst.push_back(first);
        bool merged = false;
        for (size_t r = 0; r < in.range_def.size(); ++r)
            if (in.range_def[r] == f.elem_def[e] && in.range_first[r] + in.range_count[r] == first) {
                in.range_count[r] += f.elems[e].nip;
                merged = true;
            }
        if (!merged) {
            if (in.range_def.size() >= 4) return fail(ctx, SGPU_ERR_ARG, "at most 4 section definitions per frame");
            in.range_def.push_back(f.elem_def[e]);
            in.range_first.push_back(first);
            in.range_count.push_back(f.elems[e].nip);
        }
    }
    *instance_id = (int32_t)fh.insts.size();
    fh.insts.push_back(std::move(in));
    return SGPU_OK;
}

int sgpu_frame2d_finalize(sgpu_ctx *ctx) {
    if (!ctx) return SGPU_ERR_ARG;
    FrameHost &fh = frames_of(ctx);
    if (fh.finalized) return fail(ctx, SGPU_ERR_STATE, "frames already finalized");
    if (fh.insts.empty()) return fail(ctx, SGPU_ERR_STATE, "no frame instances");
    int rc;
    if (!ctx->finalized && (rc = sgpu_sections_finalize(ctx))) return rc;
    std::vector<FrameDev> fdev;
    std::vector<ElemDev> elems;
    std::vector<double> mass_ff, exc_ff, ag;
    std::vector<int32_t> exc_flag, rec_fidx, kg_off, kg, fg_off, fg, elem_inst;
    size_t lds = 0;
    for (FrameDef &f : fh.frames) {
        FrameDev d = f.dev;
        d.elem_base = (int64_t)elems.size();
        d.free_base = (int64_t)mass_ff.size();
        d.rec_base = (int64_t)rec_fidx.size();
        // offsets of the gather lists are stored relative to the frame's list start
        d.kg_off_base = (int64_t)kg_off.size();
        d.fg_off_base = (int64_t)fg_off.size();
        const int32_t kg0 = (int32_t)kg.size(), fg0 = (int32_t)fg.size();
        for (int32_t o : f.kg_off) kg_off.push_back(o + kg0);
        for (int32_t o : f.fg_off) fg_off.push_back(o + fg0);
        kg.insert(kg.end(), f.kg.begin(), f.kg.end());
        fg.insert(fg.end(), f.fg.begin(), f.fg.end());
        elems.insert(elems.end(), f.elems.begin(), f.elems.end());
        mass_ff.insert(mass_ff.end(), f.mass_ff.begin(), f.mass_ff.end());
        exc_ff.insert(exc_ff.end(), f.exc_ff.begin(), f.exc_ff.end());
        exc_flag.insert(exc_flag.end(), f.exc_flag.begin(), f.exc_flag.end());
        rec_fidx.insert(rec_fidx.end(), f.rec_fidx.begin(), f.rec_fidx.end());
        fdev.push_back(d);
        lds = std::max(lds, frame2d_lds_bytes(d.n_free, d.bw));
    }
    std::vector<InstDev> idev;
    int64_t q = 0, vec = 0, kinit = 0, scratch = 0, hist = 0, iters = 0, dense = 0, piv = 0, last = 0, samples = 0;
    fh.hist_base.clear();
    fh.sample_base.clear();
    fh.iters_base.clear();
    fh.max_steps = 0;
    for (InstHost &in : fh.insts) {
        const FrameDev &d = fdev[in.frame];
        InstDev I{};
        I.frame = in.frame;
        I.n_steps = in.n_steps;
        I.ag_base = (int64_t)ag.size();
        ag.insert(ag.end(), in.ag.begin(), in.ag.end());
        I.elem_inst_base = (int64_t)elem_inst.size();
        elem_inst.insert(elem_inst.end(), in.elem_inst.begin(), in.elem_inst.end());
        I.q_base = q;
        q += d.q_len;
        I.vec_base = vec;
        vec += 10 * d.n_free;
        I.kinit_base = kinit;
        kinit += (int64_t)d.n_free * d.bw;
        I.scratch_base = scratch;
        scratch += 42 * d.n_elem;
        I.dense_base = dense;
        dense += (int64_t)d.n_free * d.n_free + (int64_t)d.n_free * d.bw;
        I.piv_base = piv;
        piv += d.n_free;
        I.last_base = last;
        last += 2 * kMaxNipDev * d.n_elem;
        I.hist_base = hist;
        hist += (int64_t)in.n_steps * d.n_record;
        I.iters_base = iters;
        iters += in.n_steps;
        I.sample_base = -1;
        if (fh.record_samples) {
            I.sample_base = samples;
            samples += (int64_t)in.n_steps * (6 * d.n_free + 2);
        }
        fh.sample_base.push_back(I.sample_base);
        fh.hist_base.push_back(I.hist_base);
        fh.iters_base.push_back(I.iters_base);
        I.n_ranges = (int32_t)in.range_def.size();
        for (int r = 0; r < I.n_ranges; ++r) {
            const SecDefHost &sd = ctx->defs[in.range_def[r]];
            const int64_t ns = (int64_t)sd.y.size();
            I.slot_begin[r] = sd.state_base + (int64_t)in.range_first[r] * ns;
            I.slot_count[r] = (int64_t)in.range_count[r] * ns;
            I.aux_begin[r] = sd.aux_base + (int64_t)in.range_first[r] * sd.n_aux;
            I.aux_count[r] = (int64_t)in.range_count[r] * sd.n_aux;
        }
        fh.max_steps = std::max(fh.max_steps, in.n_steps);
        fh.max_rounds = std::max(fh.max_rounds, 1 + (int64_t)in.n_steps * d.max_iters);
        idev.push_back(I);
    }
    std::vector<int32_t> pair_inst, pair_elem;
    for (size_t k = 0; k < fh.insts.size(); ++k)
        for (int e = 0; e < fdev[fh.insts[k].frame].n_elem; ++e) {
            pair_inst.push_back((int32_t)k);
            pair_elem.push_back(e);
        }
    fh.n_pairs = (int)pair_inst.size();
    Frame2dArgs &A = fh.args;
    FrameDev *d_frames;
    InstDev *d_insts;
    ElemDev *d_elems;
    double *d_mass, *d_exc, *d_ag;
    int32_t *d_excf, *d_rec, *d_kgo, *d_kg, *d_fgo, *d_fg, *d_ei;
    if ((rc = upload(ctx, &d_frames, fdev))) return rc;
    if ((rc = upload(ctx, &d_insts, idev))) return rc;
    if ((rc = upload(ctx, &d_elems, elems))) return rc;
    if ((rc = upload(ctx, &d_mass, mass_ff))) return rc;
    if ((rc = upload(ctx, &d_exc, exc_ff))) return rc;
    if ((rc = upload(ctx, &d_excf, exc_flag))) return rc;
    if ((rc = upload(ctx, &d_rec, rec_fidx))) return rc;
    if ((rc = upload(ctx, &d_kgo, kg_off))) return rc;
    if ((rc = upload(ctx, &d_kg, kg))) return rc;
    if ((rc = upload(ctx, &d_fgo, fg_off))) return rc;
    if ((rc = upload(ctx, &d_fg, fg))) return rc;
    if ((rc = upload(ctx, &d_ei, elem_inst))) return rc;
    if ((rc = upload(ctx, &d_ag, ag))) return rc;
    int32_t *d_pi, *d_pe, *d_ef;
    if ((rc = upload(ctx, &d_pi, pair_inst))) return rc;
    if ((rc = upload(ctx, &d_pe, pair_elem))) return rc;
    if ((rc = upload(ctx, &d_ef, std::vector<int32_t>(fh.insts.size(), 0)))) return rc;
    std::vector<InstState> st(fh.insts.size(), InstState{});
    if ((rc = upload(ctx, &fh.d_state, st))) return rc;
    if ((rc = upload(ctx, &fh.d_sec_evals, std::vector<unsigned long long>(fh.insts.size(), 0ull)))) return rc;
    std::vector<double> zq(q, 0.0), zv(vec, 0.0);
    double *d_q, *d_vec, *d_kinit, *d_scratch, *d_dense, *d_last;
    int32_t *d_piv;
    if ((rc = upload(ctx, &d_q, zq))) return rc;
    if ((rc = upload(ctx, &d_vec, zv))) return rc;
    if ((rc = dalloc(ctx, &d_kinit, (size_t)kinit))) return rc;
    if ((rc = dalloc(ctx, &d_scratch, (size_t)scratch))) return rc;
    if ((rc = dalloc(ctx, &d_dense, (size_t)dense))) return rc;
    if ((rc = dalloc(ctx, &d_piv, (size_t)piv))) return rc;
    if ((rc = upload(ctx, &d_last, std::vector<double>((size_t)last, 0.0)))) return rc;
    if ((rc = dalloc(ctx, &fh.d_hist, (size_t)hist))) return rc;
    if ((rc = dalloc(ctx, &fh.d_iters, (size_t)iters))) return rc;
    HIP_OK(ctx, hipMemset(fh.d_hist, 0, sizeof(double) * (size_t)(hist > 0 ? hist : 1)));
    HIP_OK(ctx, hipMemset(fh.d_iters, 0, sizeof(int32_t) * (size_t)(iters > 0 ? iters : 1)));
    A.frames = d_frames;
    A.insts = d_insts;
    A.state = fh.d_state;
    A.elems = d_elems;
    A.mass_ff = d_mass;
    A.exc_ff = d_exc;
    A.exc_flag = d_excf;
    A.rec_fidx = d_rec;
    A.kg_off = d_kgo;
    A.kg = d_kg;
    A.fg_off = d_fgo;
    A.fg = d_fg;
    A.elem_inst = d_ei;
    A.pair_inst = d_pi;
    A.pair_elem = d_pe;
    A.eflag = d_ef;
    A.sec_evals = fh.d_sec_evals;
    A.ag = d_ag;
    A.q = d_q;
    A.vec = d_vec;
    A.kinit = d_kinit;
    A.scratch = d_scratch;
    A.dense = d_dense;
    A.piv = d_piv;
    A.last = d_last;
    A.hist = fh.d_hist;
    A.iters = fh.d_iters;
    A.defs = ctx->d_defs;
    A.tb = ctx->tables();
    A.mats = ctx->d_mats;
    A.n_mats = (int32_t)ctx->mats.size();
    A.c = ctx->c;
    A.t = ctx->t;
    if ((rc = upload(ctx, &fh.d_preds, std::vector<MlpDev>(fh.frames.size(), MlpDev{})))) return rc;
    A.preds = fh.d_preds;
    A.samples = nullptr;
    if (fh.record_samples) {
        if ((rc = dalloc(ctx, &fh.d_samples, (size_t)samples))) return rc;
        A.samples = fh.d_samples;
    }
    fh.lds = lds;
    HIP_OK(ctx, frame2d_set_lds(lds));
    {  // work queue: room for every instance's largest round plus one claim per worker
        int per_cu = 0, cus = 0;
        HIP_OK(ctx, frame2d_queue_set_lds(lds, &per_cu));
        HIP_OK(ctx, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
        if (per_cu < 1) return fail(ctx, SGPU_ERR_STATE, "the work-queue kernel does not fit on a compute unit");
        fh.workers = per_cu * cus;
        int64_t need = fh.workers + 1024;
        for (const InstHost &in : fh.insts) need += fdev[in.frame].n_elem + 1;
        int64_t cap = 1;
        while (cap < need) cap <<= 1;
        if (fh.insts.size() >= 32768 || cap > (1ll << 30))
            return fail(ctx, SGPU_ERR_ARG, "too many instances for the work queue");
        QueueDev &Q = fh.queue;
        unsigned long long *ht;
        if ((rc = dalloc(ctx, &ht, 2))) return rc;
        Q.head = ht;
        Q.tail = ht + 1;
        if ((rc = dalloc(ctx, &Q.slots, (size_t)cap))) return rc;
        HIP_OK(ctx, hipMemset(Q.slots, 0xff, sizeof(int32_t) * (size_t)cap));
        if ((rc = dalloc(ctx, &Q.pending, fh.insts.size())) || (rc = dalloc(ctx, &Q.live, 1)) ||
            (rc = dalloc(ctx, &Q.step0, fh.insts.size())) || (rc = dalloc(ctx, &Q.abort, 1)) ||
            (rc = dalloc(ctx, &Q.start, 1)))
            return rc;
        Q.cap_mask = (int32_t)(cap - 1);
        Q.n_inst = (int32_t)fh.insts.size();
    }
    fh.finalized = true;
    return SGPU_OK;
}

namespace {

// Capture `rounds` rounds in a HIP graph once; replays reuse it.
int ensure_graph(sgpu_ctx *ctx, FrameHost &fh, int rounds) {
    if (fh.graph_rounds != rounds) {
        if (fh.graph_exec) HIP_OK(ctx, hipGraphExecDestroy(fh.graph_exec));
        if (fh.graph) HIP_OK(ctx, hipGraphDestroy(fh.graph));
        fh.graph_exec = nullptr;
        fh.graph = nullptr;
        fh.graph_rounds = 0;
        HIP_OK(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        hipError_t le = hipSuccess;
        for (int r = 0; r < rounds && le == hipSuccess; ++r)
            le = launch_frame2d_round(ctx->stream, fh.args, fh.n_pairs, (int)fh.insts.size(), fh.lds);
        hipGraph_t g = nullptr;
        const hipError_t ce = hipStreamEndCapture(ctx->stream, &g);
        if (le != hipSuccess || ce != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return fail(ctx, SGPU_ERR_HIP, "capturing the frame rounds: %s",
                        hipGetErrorString(le != hipSuccess ? le : ce));
        }
        fh.graph = g;
        HIP_OK(ctx, hipGraphInstantiate(&fh.graph_exec, g, nullptr, nullptr, 0));
        fh.graph_rounds = rounds;
    }
    return SGPU_OK;
}

}  // namespace

int sgpu_frame2d_launch(sgpu_ctx *ctx, int32_t rounds) {
    if (!ctx) return SGPU_ERR_ARG;
    FrameHost &fh = frames_of(ctx);
    if (!fh.finalized) return fail(ctx, SGPU_ERR_STATE, "sgpu_frame2d_finalize has not run");
    if (rounds <= 0) return fail(ctx, SGPU_ERR_ARG, "rounds must be > 0");
    int rc;
    if ((rc = ensure_graph(ctx, fh, rounds))) return rc;
    {
        ProfScope ps(ctx, SGPU_KERNEL_FRAME2D);
        HIP_OK(ctx, hipGraphLaunch(fh.graph_exec, ctx->stream));
    }
    return SGPU_OK;
}

int sgpu_frame2d_launch_persistent(sgpu_ctx *ctx, int32_t max_steps) {
    if (!ctx) return SGPU_ERR_ARG;
    FrameHost &fh = frames_of(ctx);
    if (!fh.finalized) return fail(ctx, SGPU_ERR_STATE, "sgpu_frame2d_finalize has not run");
    if (max_steps <= 0) return fail(ctx, SGPU_ERR_ARG, "max_steps must be > 0");
    ProfScope ps(ctx, SGPU_KERNEL_FRAME2D);
    HIP_OK(ctx, launch_frame2d_persistent(ctx->stream, fh.args, (int)fh.insts.size(), max_steps, fh.lds));
    return SGPU_OK;
}

namespace {

// A work-queue launch; budget_s > 0 parks every instance at its first
// system task after that many seconds.
int queue_launch(sgpu_ctx *ctx, FrameHost &fh, int32_t max_steps, double budget_s) {
    QueueDev Q = fh.queue;
    Q.budget = budget_s > 0.0 ? (unsigned long long)(budget_s * 1.0e8) + 1ull : 0ull;  // 100 MHz counter
    {
        ProfScope ps(ctx, SGPU_KERNEL_FRAME2D);
        HIP_OK(ctx, launch_frame2d_queue(ctx->stream, fh.args, Q, fh.workers, max_steps, fh.lds));
    }
    int32_t abort_flag = 0;
    HIP_OK(ctx, hipMemcpyAsync(&abort_flag, fh.queue.abort, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
    HIP_OK(ctx, hipStreamSynchronize(ctx->stream));
    if (abort_flag) return fail(ctx, SGPU_ERR_STATE, "work queue: a worker's wait exceeded its bound");
    return SGPU_OK;
}

}  // namespace

int sgpu_frame2d_launch_queue(sgpu_ctx *ctx, int32_t max_steps) {
    if (!ctx) return SGPU_ERR_ARG;
    FrameHost &fh = frames_of(ctx);
    if (!fh.finalized) return fail(ctx, SGPU_ERR_STATE, "sgpu_frame2d_finalize has not run");
    if (max_steps <= 0) return fail(ctx, SGPU_ERR_ARG, "max_steps must be > 0");
    return queue_launch(ctx, fh, max_steps, 0.0);
}

int sgpu_frame2d_set_schedule(sgpu_ctx *ctx, int schedule) {
    if (!ctx) return SGPU_ERR_ARG;
    if (schedule != SGPU_SCHEDULE_AUTO && schedule != SGPU_SCHEDULE_ROUNDS && schedule != SGPU_SCHEDULE_PERSISTENT &&
        schedule != SGPU_SCHEDULE_QUEUE)
        return fail(ctx, SGPU_ERR_ARG, "unknown schedule %d", schedule);
    FrameHost &fh = frames_of(ctx);
    fh.schedule = schedule;
    if (schedule != SGPU_SCHEDULE_AUTO) fh.chosen = schedule;
    return SGPU_OK;
}

int sgpu_frame2d_schedule_info(sgpu_ctx *ctx, double *out5) {
    if (!ctx || !out5) return SGPU_ERR_ARG;
    FrameHost &fh = frames_of(ctx);
    out5[0] = fh.chosen;
    out5[1] = fh.rate_rounds;
    out5[2] = fh.rate_persistent;
    out5[3] = fh.tunings;
    out5[4] = fh.rate_queue;
    return SGPU_OK;
}

namespace {

// How far ahead the rounds schedule's sampled rate must be of the queue's
// for the autotuner to choose rounds.
constexpr double kRoundsMargin = 1.25;

// Launch one chunk of the given schedule, wait, and report the GM-steps the
// instances advanced and whether any instance is still active.
int run_chunk(sgpu_ctx *ctx, FrameHost &fh, int schedule, int rounds, std::vector<InstState> &st, int64_t &advanced,
              int &active, double &seconds, double queue_budget) {
    int64_t before = 0;
    for (const InstState &s : st) before += s.step;
    const auto t0 = std::chrono::steady_clock::now();
    // a timed chunk is a few steps per instance for every schedule
    const int steps = rounds / 4 > 0 ? rounds / 4 : 1;
    // a queue chunk is either timed (a wall-clock budget, so that a slow
    // instance's tail does not set the measured rate) or runs the instances
    // to their ends: a chunk end waits for the slowest instance while the
    // others' workers idle
    int rc = schedule == SGPU_SCHEDULE_PERSISTENT ? sgpu_frame2d_launch_persistent(ctx, steps)
             : schedule == SGPU_SCHEDULE_QUEUE    ? queue_launch(ctx, fh, 1 << 30, queue_budget)
                                                  : sgpu_frame2d_launch(ctx, rounds);
    if (rc) return rc;
    HIP_OK(ctx, hipMemcpyAsync(st.data(), fh.d_state, sizeof(InstState) * st.size(), hipMemcpyDeviceToHost,
                               ctx->stream));
    HIP_OK(ctx, hipStreamSynchronize(ctx->stream));
    seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    int64_t after = 0;
    active = 0;
    for (size_t k = 0; k < st.size(); ++k) {
        after += st[k].step;
        if (st[k].status == 0 && st[k].step < fh.insts[k].n_steps) ++active;
    }
    advanced = after - before;
    return SGPU_OK;
}

}  // namespace

int sgpu_frame2d_run(sgpu_ctx *ctx, int32_t rounds) {
    if (!ctx) return SGPU_ERR_ARG;
    FrameHost &fh = frames_of(ctx);
    if (!fh.finalized) return fail(ctx, SGPU_ERR_STATE, "sgpu_frame2d_finalize has not run");
    if (rounds <= 0) return fail(ctx, SGPU_ERR_ARG, "rounds must be > 0");
    std::vector<InstState> st(fh.insts.size());
    HIP_OK(ctx, hipMemcpy(st.data(), fh.d_state, sizeof(InstState) * st.size(), hipMemcpyDeviceToHost));
    int active = 0;
    for (size_t k = 0; k < st.size(); ++k)
        if (st[k].status == 0 && st[k].step < fh.insts[k].n_steps) ++active;
    // Autotuner: time one chunk of each schedule on the live ensemble (both
    // advance the same per-instance state with the same arithmetic) and keep
    // the faster; measure again when the active set has halved.
    bool tune = fh.schedule == SGPU_SCHEDULE_AUTO;
    int active_at_tune = active;
    int since_tune = 0;
    double tuned_rate = 0.0;  // the chosen schedule's rate when it was chosen
    int64_t chunks = 0;
    const int64_t max_chunks = 4 + 2 * fh.max_rounds / rounds;
    while (active > 0) {
        if (++chunks > max_chunks)
            return fail(ctx, SGPU_ERR_STATE, "instances still active after %lld chunks", (long long)chunks);
        int64_t adv = 0;
        double sec = 0.0;
        int rc;
        if (tune) {
            // the rounds and the work queue (the persistent schedule, one
            // block per instance with its elements in sequence, is slower than
            // the queue wherever both were measured, and is not sampled)
            double rate[2] = {0.0, 0.0};
            const int order[2] = {SGPU_SCHEDULE_ROUNDS, SGPU_SCHEDULE_QUEUE};
            if ((rc = ensure_graph(ctx, fh, rounds))) return rc;
            double budget = 0.002;  // the queue's sample gets the rounds sample's time
            for (int k = 0; k < 2 && active > 0; ++k) {
                if ((rc = run_chunk(ctx, fh, order[k], rounds, st, adv, active, sec, budget))) return rc;
                rate[k] = sec > 0.0 ? (double)adv / sec : 0.0;
                if (k == 0 && sec > budget) budget = sec;
            }
            fh.rate_rounds = rate[0];
            fh.rate_queue = rate[1];
            // Rounds must be clearly ahead to be chosen: lockstep rounds wait
            // for each round's slowest element task, so their rate falls as
            // records go nonlinear and their steps' costs diverge, while the
            // queue's holds.  A sample taken in the nearly linear start of a
            // batch puts rounds a little ahead (1.16x on 256 records of the
            // 3-story frame) and then loses about 5 % of the run to rounds
            // chunks before the rate-drop check measures again; one record
            // alone still picks rounds (3x ahead).
            fh.chosen = rate[0] > kRoundsMargin * rate[1] ? SGPU_SCHEDULE_ROUNDS : SGPU_SCHEDULE_QUEUE;
            tuned_rate = fh.chosen == SGPU_SCHEDULE_ROUNDS ? rate[0] : rate[1];
            fh.tunings += 1;
            tune = false;
            active_at_tune = active;
            continue;
        }
        if ((rc = run_chunk(ctx, fh, fh.chosen, rounds, st, adv, active, sec, 0.0))) return rc;
        // measure again when the active set has halved, every 64 chunks, and
        // when the chosen schedule's rate has fallen well below the rate it
        // was chosen at: the cost of an instance's step changes as its record
        // goes nonlinear, and lockstep rounds then wait for every round's
        // slowest element task while the queue keeps its workers busy
        const double chunk_rate = sec > 0.0 ? (double)adv / sec : 0.0;
        ++since_tune;
        if (fh.schedule == SGPU_SCHEDULE_AUTO && active > 0 &&
            (2 * active <= active_at_tune || since_tune >= 64 || (since_tune >= 2 && chunk_rate < 0.6 * tuned_rate))) {
            tune = true;
            since_tune = 0;
        }
    }
    return SGPU_OK;
}

int sgpu_frame2d_result(sgpu_ctx *ctx, int32_t instance, double *u_hist, int32_t *iters, int32_t *status,
                        int32_t *steps_done) {
    if (!ctx) return SGPU_ERR_ARG;
    FrameHost &fh = frames_of(ctx);
    if (!fh.finalized) return fail(ctx, SGPU_ERR_STATE, "sgpu_frame2d_finalize has not run");
    if (instance < 0 || instance >= (int)fh.insts.size()) return fail(ctx, SGPU_ERR_ARG, "instance out of range");
    const InstHost &in = fh.insts[instance];
    const FrameDef &f = fh.frames[in.frame];
    HIP_OK(ctx, hipStreamSynchronize(ctx->stream));
    InstState st;
    HIP_OK(ctx, hipMemcpy(&st, fh.d_state + instance, sizeof(st), hipMemcpyDeviceToHost));
    if (status) *status = st.status;
    if (steps_done) *steps_done = st.step;
    if (u_hist && f.dev.n_record > 0)
        HIP_OK(ctx, hipMemcpy(u_hist, fh.d_hist + fh.hist_base[instance],
                              sizeof(double) * (size_t)in.n_steps * f.dev.n_record, hipMemcpyDeviceToHost));
    if (iters)
        HIP_OK(ctx, hipMemcpy(iters, fh.d_iters + fh.iters_base[instance], sizeof(int32_t) * (size_t)in.n_steps,
                              hipMemcpyDeviceToHost));
    return SGPU_OK;
}

int sgpu_frame2d_info(sgpu_ctx *ctx, int32_t frame_id, int32_t *out4) {
    if (!ctx || !out4) return SGPU_ERR_ARG;
    FrameHost &fh = frames_of(ctx);
    if (frame_id < 0 || frame_id >= (int)fh.frames.size()) return fail(ctx, SGPU_ERR_ARG, "frame out of range");
    const FrameDef &f = fh.frames[frame_id];
    out4[0] = f.dev.n_free;
    out4[1] = f.dev.hbw;
    out4[2] = f.dev.n_record;
    out4[3] = f.total_dofs;
    return SGPU_OK;
}

int sgpu_frame2d_record_samples(sgpu_ctx *ctx, int32_t on) {
    if (!ctx) return SGPU_ERR_ARG;
    FrameHost &fh = frames_of(ctx);
    if (fh.finalized) return fail(ctx, SGPU_ERR_STATE, "the sample recorder is chosen before sgpu_frame2d_finalize");
    fh.record_samples = on != 0;
    return SGPU_OK;
}

int sgpu_frame2d_samples(sgpu_ctx *ctx, int32_t frame_id, float *x, float *y, int64_t capacity, int64_t *count) {
    if (!ctx || !count) return SGPU_ERR_ARG;
    FrameHost &fh = frames_of(ctx);
    if (!fh.finalized || !fh.record_samples) return fail(ctx, SGPU_ERR_STATE, "no sample recorder");
    if (frame_id < 0 || frame_id >= (int)fh.frames.size()) return fail(ctx, SGPU_ERR_ARG, "frame out of range");
    const int n = fh.frames[frame_id].dev.n_free, n_in = 5 * n + 2;
    std::vector<InstState> st(fh.insts.size());
    HIP_OK(ctx, hipStreamSynchronize(ctx->stream));
    HIP_OK(ctx, hipMemcpy(st.data(), fh.d_state, sizeof(InstState) * st.size(), hipMemcpyDeviceToHost));
    int64_t total = 0;
    for (size_t k = 0; k < fh.insts.size(); ++k) {
        if (fh.insts[k].frame != frame_id) continue;
        const int64_t done = st[k].step;
        if (x && y) {
            if (total + done > capacity) return fail(ctx, SGPU_ERR_ARG, "sample capacity %lld too small", (long long)capacity);
            const float *src = fh.d_samples + fh.sample_base[k];
            HIP_OK(ctx, hipMemcpyAsync(x + total * n_in, src, sizeof(float) * (size_t)(done * n_in),
                                       hipMemcpyDeviceToDevice, ctx->stream));
            HIP_OK(ctx, hipMemcpyAsync(y + total * n, src + (int64_t)fh.insts[k].n_steps * n_in,
                                       sizeof(float) * (size_t)(done * n), hipMemcpyDeviceToDevice, ctx->stream));
        }
        total += done;
    }
    HIP_OK(ctx, hipStreamSynchronize(ctx->stream));
    *count = total;
    return SGPU_OK;
}

int sgpu_frame2d_set_predictor(sgpu_ctx *ctx, int32_t frame_id, const sgpu_mlp *m) {
    if (!ctx) return SGPU_ERR_ARG;
    FrameHost &fh = frames_of(ctx);
    if (!fh.finalized) return fail(ctx, SGPU_ERR_STATE, "sgpu_frame2d_finalize has not run");
    if (frame_id < 0 || frame_id >= (int)fh.frames.size()) return fail(ctx, SGPU_ERR_ARG, "frame out of range");
    MlpDev snap{};
    if (m) {
        const MlpDev &src = *mlp_device_view(m);
        const int n = fh.frames[frame_id].dev.n_free;
        if (src.n_in != 5 * n + 2 || src.n_out != n)
            return fail(ctx, SGPU_ERR_ARG, "network sizes %d -> %d do not match the frame (%d -> %d)", src.n_in,
                        src.n_out, 5 * n + 2, n);
        // a copy owned by this context: the network may be trained further or destroyed
        snap = src;
        int rc;
        for (int l = 0; l < kMlpLayers; ++l) {
            const size_t nk = (size_t)mlp_n(src, l) * mlp_k(src, l);
            uint16_t *w;
            float *b;
            if ((rc = dalloc(ctx, &w, nk)) || (rc = dalloc(ctx, &b, (size_t)mlp_n(src, l)))) return rc;
            HIP_OK(ctx, hipMemcpy(w, src.w[l], sizeof(uint16_t) * nk, hipMemcpyDeviceToDevice));
            HIP_OK(ctx, hipMemcpy(b, src.b[l], sizeof(float) * (size_t)mlp_n(src, l), hipMemcpyDeviceToDevice));
            snap.w[l] = w;
            snap.wt[l] = nullptr;
            snap.b[l] = b;
        }
        float *nrm;
        const size_t nn = 2 * (size_t)src.n_in + 2 * (size_t)src.n_out;
        if ((rc = dalloc(ctx, &nrm, nn))) return rc;
        HIP_OK(ctx, hipMemcpy(nrm, src.x_mean, sizeof(float) * nn, hipMemcpyDeviceToDevice));
        snap.x_mean = nrm;
        snap.x_scale = nrm + src.n_in;
        snap.y_mean = nrm + 2 * src.n_in;
        snap.y_scale = nrm + 2 * src.n_in + src.n_out;
    }
    HIP_OK(ctx, hipStreamSynchronize(ctx->stream));
    HIP_OK(ctx, hipMemcpy(fh.d_preds + frame_id, &snap, sizeof(MlpDev), hipMemcpyHostToDevice));
    return SGPU_OK;
}

int sgpu_frame2d_inst_stats(sgpu_ctx *ctx, int32_t instance, int32_t *out4) {
    if (!ctx || !out4) return SGPU_ERR_ARG;
    FrameHost &fh = frames_of(ctx);
    if (!fh.finalized) return fail(ctx, SGPU_ERR_STATE, "sgpu_frame2d_finalize has not run");
    if (instance < 0 || instance >= (int)fh.insts.size()) return fail(ctx, SGPU_ERR_ARG, "instance out of range");
    InstState st;
    HIP_OK(ctx, hipStreamSynchronize(ctx->stream));
    HIP_OK(ctx, hipMemcpy(&st, fh.d_state + instance, sizeof(st), hipMemcpyDeviceToHost));
    out4[0] = st.total_iters;
    out4[1] = st.proposals;
    out4[2] = st.fallbacks;
    out4[3] = st.dense_fallbacks;
    return SGPU_OK;
}

int sgpu_frame2d_section_evals(sgpu_ctx *ctx, int32_t first, int32_t count, int64_t *out) {
    if (!ctx || count < 0 || (count > 0 && !out)) return SGPU_ERR_ARG;
    FrameHost &fh = frames_of(ctx);
    if (!fh.finalized) return fail(ctx, SGPU_ERR_STATE, "sgpu_frame2d_finalize has not run");
    if (first < 0 || (int64_t)first + count > (int64_t)fh.insts.size())
        return fail(ctx, SGPU_ERR_ARG, "instances [%d, %d) out of range", first, first + count);
    if (count == 0) return SGPU_OK;
    std::vector<unsigned long long> v((size_t)count);
    HIP_OK(ctx, hipStreamSynchronize(ctx->stream));
    HIP_OK(ctx, hipMemcpy(v.data(), fh.d_sec_evals + first, sizeof(unsigned long long) * (size_t)count,
                          hipMemcpyDeviceToHost));
    for (int32_t k = 0; k < count; ++k) out[k] = (int64_t)v[(size_t)k];
    return SGPU_OK;
}

}  // extern "C"

namespace sgpu {
int envelopes_device(sgpu_ctx *ctx, const double *d_hist, const std::vector<int64_t> &bases,
                     const std::vector<int32_t> &steps, int n_record, int n_drift, const int32_t *col_i,
                     const int32_t *col_j, const double *dx, double *disp_env, double *drift_env);
}

int sgpu_frame2d_envelopes(sgpu_ctx *ctx, int32_t frame_id, int n_drift, const int32_t *col_i, const int32_t *col_j,
                           const double *dx, double *disp_env, double *drift_env) {
    if (!ctx || !disp_env || (n_drift > 0 && !drift_env)) return SGPU_ERR_ARG;
    FrameHost &fh = frames_of(ctx);
    if (!fh.finalized) return fail(ctx, SGPU_ERR_STATE, "sgpu_frame2d_finalize has not run");
    if (frame_id < 0 || frame_id >= (int)fh.frames.size()) return fail(ctx, SGPU_ERR_ARG, "frame out of range");
    HIP_OK(ctx, hipStreamSynchronize(ctx->stream));
    std::vector<InstState> st(fh.insts.size());
    HIP_OK(ctx, hipMemcpy(st.data(), fh.d_state, sizeof(InstState) * st.size(), hipMemcpyDeviceToHost));
    std::vector<int64_t> bases;
    std::vector<int32_t> steps;
    for (size_t i = 0; i < fh.insts.size(); ++i) {
        if (fh.insts[i].frame != frame_id) continue;
        bases.push_back(fh.hist_base[i]);
        steps.push_back(std::min(st[i].step, (int32_t)fh.insts[i].n_steps));
    }
    return envelopes_device(ctx, fh.d_hist, bases, steps, fh.frames[frame_id].dev.n_record, n_drift, col_i, col_j, dx,
                            disp_env, drift_env);
}
