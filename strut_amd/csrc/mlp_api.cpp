// Host side of the learned Newton predictor's network (sgpu_mlp_*): FP32
// master weights with bf16 copies for the matrix cores, Adam training over
// minibatches of device-resident samples, batched inference, and the
// gradient read-out used to check the backward pass against the oracle.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "ctx.h"
#include "mlp_types.h"

namespace sgpu {
hipError_t launch_mlp_prep(hipStream_t s, const MlpPrepArgs &P);
hipError_t launch_mlp_fwdbwd(hipStream_t s, const MlpTrainArgs &T);
hipError_t launch_mlp_wgrad(hipStream_t s, const MlpTrainArgs &T, const MlpTiles &tiles);
hipError_t launch_mlp_adam(hipStream_t s, const MlpAdamArgs &P);
hipError_t launch_mlp_colstats(hipStream_t s, const float *x, int64_t rows, int cols, int chunks, double *partial);
hipError_t mlp_set_lds(int k0, int n5);
}  // namespace sgpu

using namespace sgpu;

struct sgpu_mlp {
    sgpu_ctx *ctx = nullptr;
    MlpDev dev{};
    int n_in = 0, n_out = 0, k0 = 0, n5 = 0;
    int N[kMlpLayers] = {}, K[kMlpLayers] = {};
    int64_t off_w[kMlpLayers] = {}, off_b[kMlpLayers] = {}, off_bf[kMlpLayers] = {};
    int64_t n_params = 0, n_wbf = 0;
    float *p = nullptr, *g = nullptr, *mom = nullptr, *vel = nullptr, *norm = nullptr;
    uint16_t *w = nullptr, *wt = nullptr;
    // training workspace, sized for cap_batch rows
    int cap_batch = 0;
    uint16_t *xb = nullptr, *xt = nullptr, *ht = nullptr, *dzt = nullptr;
    float *t = nullptr, *loss = nullptr;
    int32_t *perm = nullptr;
    int64_t perm_cap = 0;
    double *partial = nullptr;
    int64_t partial_cap = 0;
    int64_t adam_t = 0;
    std::vector<void *> bufs;
};

namespace {

constexpr int kStatChunks = 256;

int round16(int v) { return (v + 15) / 16 * 16; }

template <typename T>
int mlp_alloc(sgpu_mlp *m, T **p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void **)p, count * sizeof(T));
    if (e != hipSuccess)
        return fail(m->ctx, SGPU_ERR_HIP, "hipMalloc(%zu bytes): %s", count * sizeof(T), hipGetErrorString(e));
    m->bufs.push_back((void *)*p);
    return SGPU_OK;
}

template <typename T>
void mlp_release(sgpu_mlp *m, T *&p) {
    if (!p) return;
    for (auto &b : m->bufs)
        if (b == (void *)p) b = nullptr;
    (void)hipFree(p);
    p = nullptr;
}

uint64_t splitmix(uint64_t &s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

double uniform01(uint64_t &s) { return (double)(splitmix(s) >> 11) * (1.0 / 9007199254740992.0); }

// bf16 copies W [N][K] and W^T [K][N] of the master weights (host side).
void bf16_copies(const sgpu_mlp *m, const std::vector<float> &p, std::vector<uint16_t> &w,
                 std::vector<uint16_t> &wt) {
    w.assign((size_t)m->n_wbf, 0);
    wt.assign((size_t)m->n_wbf, 0);
    for (int l = 0; l < kMlpLayers; ++l)
        for (int n = 0; n < m->N[l]; ++n)
            for (int k = 0; k < m->K[l]; ++k) {
                const uint16_t h = f32_to_bf16(p[(size_t)(m->off_w[l] + (int64_t)n * m->K[l] + k)]);
                w[(size_t)(m->off_bf[l] + (int64_t)n * m->K[l] + k)] = h;
                wt[(size_t)(m->off_bf[l] + (int64_t)k * m->N[l] + n)] = h;
            }
}

int upload_params(sgpu_mlp *m, const std::vector<float> &p) {
    std::vector<uint16_t> w, wt;
    bf16_copies(m, p, w, wt);
    HIP_OK(m->ctx, hipMemcpy(m->p, p.data(), sizeof(float) * p.size(), hipMemcpyHostToDevice));
    HIP_OK(m->ctx, hipMemcpy(m->w, w.data(), sizeof(uint16_t) * w.size(), hipMemcpyHostToDevice));
    HIP_OK(m->ctx, hipMemcpy(m->wt, wt.data(), sizeof(uint16_t) * wt.size(), hipMemcpyHostToDevice));
    HIP_OK(m->ctx, hipMemset(m->mom, 0, sizeof(float) * (size_t)m->n_params));
    HIP_OK(m->ctx, hipMemset(m->vel, 0, sizeof(float) * (size_t)m->n_params));
    m->adam_t = 0;
    return SGPU_OK;
}

int ensure_workspace(sgpu_mlp *m, int batch) {
    if (batch <= m->cap_batch) return SGPU_OK;
    mlp_release(m, m->xb);
    mlp_release(m, m->xt);
    mlp_release(m, m->ht);
    mlp_release(m, m->dzt);
    mlp_release(m, m->t);
    mlp_release(m, m->loss);
    int rc;
    const size_t B = (size_t)batch;
    if ((rc = mlp_alloc(m, &m->xb, B * m->k0))) return rc;
    if ((rc = mlp_alloc(m, &m->xt, B * m->k0))) return rc;
    if ((rc = mlp_alloc(m, &m->ht, B * kMlpHidden * (kMlpLayers - 1)))) return rc;
    if ((rc = mlp_alloc(m, &m->dzt, B * (kMlpHidden * (kMlpLayers - 1) + m->n5)))) return rc;
    if ((rc = mlp_alloc(m, &m->t, B * m->n5))) return rc;
    if ((rc = mlp_alloc(m, &m->loss, B / 16))) return rc;
    m->cap_batch = batch;
    return SGPU_OK;
}

MlpTrainArgs train_args(sgpu_mlp *m, int batch, int rows, int backward) {
    MlpTrainArgs T{};
    T.m = m->dev;
    T.batch = batch;
    T.rows = rows;
    T.backward = backward;
    T.grad_scale = 2.0f / (float)((double)rows * m->n_out);
    T.xb = m->xb;
    T.xt = m->xt;
    T.t = m->t;
    for (int l = 0; l < kMlpLayers - 1; ++l) T.ht[l] = m->ht + (size_t)l * kMlpHidden * batch;
    for (int l = 0; l < kMlpLayers; ++l) T.dzt[l] = m->dzt + (size_t)l * kMlpHidden * batch;
    T.loss_acc = m->loss;
    T.g = m->g;
    return T;
}

MlpTiles tiles_of(const sgpu_mlp *m) {
    MlpTiles t{};
    t.base[0] = 0;
    for (int l = 0; l < kMlpLayers; ++l) {
        t.base[l + 1] = t.base[l] + (m->N[l] / 16) * (m->K[l] / 16);
        t.off_w[l] = m->off_w[l];
        t.off_b[l] = m->off_b[l];
    }
    return t;
}

MlpPrepArgs prep_args(sgpu_mlp *m, int batch, int rows, int64_t first, const int32_t *perm, const float *x,
                      const float *y) {
    MlpPrepArgs P{};
    P.m = m->dev;
    P.batch = batch;
    P.rows = rows;
    P.first = first;
    P.perm = perm;
    P.x = x;
    P.y = y;
    P.xb = m->xb;
    P.xt = m->xt;
    P.t = m->t;
    return P;
}

// Forward + backward of one minibatch: gradients into m->g, squared errors
// added to m->loss.
int fwdbwd(sgpu_mlp *m, int batch, int rows, int64_t first, const int32_t *perm, const float *x, const float *y) {
    sgpu_ctx *ctx = m->ctx;
    HIP_OK(ctx, launch_mlp_prep(ctx->stream, prep_args(m, batch, rows, first, perm, x, y)));
    const MlpTrainArgs T = train_args(m, batch, rows, 1);
    HIP_OK(ctx, launch_mlp_fwdbwd(ctx->stream, T));
    HIP_OK(ctx, launch_mlp_wgrad(ctx->stream, T, tiles_of(m)));
    return SGPU_OK;
}

}  // namespace

extern "C" {

int sgpu_mlp_create(sgpu_ctx *ctx, int32_t n_in, int32_t n_out, uint64_t seed, sgpu_mlp **out) {
    if (!ctx || !out) return SGPU_ERR_ARG;
    *out = nullptr;
    if (n_in <= 0 || n_out <= 0 || n_in > 4096 || n_out > 1024)
        return fail(ctx, SGPU_ERR_ARG, "network sizes out of range (n_in %d, n_out %d)", n_in, n_out);
    sgpu_mlp *m = new sgpu_mlp();
    m->ctx = ctx;
    m->n_in = n_in;
    m->n_out = n_out;
    m->k0 = round16(n_in);
    m->n5 = round16(n_out);
    int64_t off = 0, offbf = 0;
    for (int l = 0; l < kMlpLayers; ++l) {
        m->K[l] = l == 0 ? m->k0 : kMlpHidden;
        m->N[l] = l == kMlpLayers - 1 ? m->n5 : kMlpHidden;
        m->off_w[l] = off;
        off += (int64_t)m->N[l] * m->K[l];
        m->off_b[l] = off;
        off += m->N[l];
        m->off_bf[l] = offbf;
        offbf += (int64_t)m->N[l] * m->K[l];
    }
    m->n_params = off;
    m->n_wbf = offbf;
    int rc;
    if ((rc = mlp_alloc(m, &m->p, (size_t)off)) || (rc = mlp_alloc(m, &m->g, (size_t)off)) ||
        (rc = mlp_alloc(m, &m->mom, (size_t)off)) || (rc = mlp_alloc(m, &m->vel, (size_t)off)) ||
        (rc = mlp_alloc(m, &m->w, (size_t)offbf)) || (rc = mlp_alloc(m, &m->wt, (size_t)offbf)) ||
        (rc = mlp_alloc(m, &m->norm, 2 * (size_t)n_in + 2 * (size_t)n_out))) {
        sgpu_mlp_destroy(m);
        return rc;
    }
    // uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) on the valid entries; padding stays zero
    std::vector<float> p((size_t)off, 0.0f);
    uint64_t s = seed;
    for (int l = 0; l < kMlpLayers; ++l) {
        const int kv = l == 0 ? n_in : kMlpHidden, nv = l == kMlpLayers - 1 ? n_out : kMlpHidden;
        const double bound = 1.0 / std::sqrt((double)kv);
        for (int n = 0; n < nv; ++n)
            for (int k = 0; k < kv; ++k)
                p[(size_t)(m->off_w[l] + (int64_t)n * m->K[l] + k)] = (float)((2.0 * uniform01(s) - 1.0) * bound);
        for (int n = 0; n < nv; ++n) p[(size_t)(m->off_b[l] + n)] = (float)((2.0 * uniform01(s) - 1.0) * bound);
    }
    if ((rc = upload_params(m, p))) {
        sgpu_mlp_destroy(m);
        return rc;
    }
    std::vector<float> nrm(2 * (size_t)n_in + 2 * (size_t)n_out, 0.0f);
    for (int i = 0; i < n_in; ++i) nrm[(size_t)n_in + i] = 1.0f;
    for (int j = 0; j < n_out; ++j) nrm[2 * (size_t)n_in + n_out + j] = 1.0f;
    hipError_t e = hipMemcpy(m->norm, nrm.data(), sizeof(float) * nrm.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = mlp_set_lds(m->k0, m->n5);
    if (e != hipSuccess) {
        sgpu_mlp_destroy(m);
        return fail(ctx, SGPU_ERR_HIP, "network set-up: %s", hipGetErrorString(e));
    }
    MlpDev &d = m->dev;
    d.n_in = n_in;
    d.n_out = n_out;
    d.k0 = m->k0;
    d.n5 = m->n5;
    for (int l = 0; l < kMlpLayers; ++l) {
        d.w[l] = m->w + m->off_bf[l];
        d.wt[l] = m->wt + m->off_bf[l];
        d.b[l] = m->p + m->off_b[l];
    }
    d.x_mean = m->norm;
    d.x_scale = m->norm + n_in;
    d.y_mean = m->norm + 2 * n_in;
    d.y_scale = m->norm + 2 * n_in + n_out;
    *out = m;
    return SGPU_OK;
}

void sgpu_mlp_destroy(sgpu_mlp *m) {
    if (!m) return;
    for (void *b : m->bufs)
        if (b) (void)hipFree(b);
    delete m;
}

int64_t sgpu_mlp_param_count(const sgpu_mlp *m) { return m ? m->n_params : 0; }

int sgpu_mlp_layout(const sgpu_mlp *m, int32_t *out8) {
    if (!m || !out8) return SGPU_ERR_ARG;
    out8[0] = m->n_in;
    out8[1] = m->n_out;
    out8[2] = m->k0;
    out8[3] = m->n5;
    out8[4] = kMlpHidden;
    out8[5] = kMlpLayers;
    out8[6] = (int32_t)m->n_params;
    out8[7] = (int32_t)m->adam_t;
    return SGPU_OK;
}

int sgpu_mlp_get_params(sgpu_mlp *m, float *out) {
    if (!m || !out) return SGPU_ERR_ARG;
    HIP_OK(m->ctx, hipStreamSynchronize(m->ctx->stream));
    HIP_OK(m->ctx, hipMemcpy(out, m->p, sizeof(float) * (size_t)m->n_params, hipMemcpyDeviceToHost));
    return SGPU_OK;
}

int sgpu_mlp_set_params(sgpu_mlp *m, const float *in) {
    if (!m || !in) return SGPU_ERR_ARG;
    HIP_OK(m->ctx, hipStreamSynchronize(m->ctx->stream));
    std::vector<float> p(in, in + m->n_params);
    for (int l = 0; l < kMlpLayers; ++l) {  // padding entries must stay zero
        const int kv = l == 0 ? m->n_in : kMlpHidden, nv = l == kMlpLayers - 1 ? m->n_out : kMlpHidden;
        for (int n = 0; n < m->N[l]; ++n) {
            for (int k = 0; k < m->K[l]; ++k)
                if (n >= nv || k >= kv) p[(size_t)(m->off_w[l] + (int64_t)n * m->K[l] + k)] = 0.0f;
            if (n >= nv) p[(size_t)(m->off_b[l] + n)] = 0.0f;
        }
    }
    return upload_params(m, p);
}

int sgpu_mlp_set_normalization(sgpu_mlp *m, const float *x_mean, const float *x_scale, const float *y_mean,
                               const float *y_scale) {
    if (!m || !x_mean || !x_scale || !y_mean || !y_scale) return SGPU_ERR_ARG;
    for (int j = 0; j < m->n_out; ++j)
        if (!(y_scale[j] != 0.0f)) return fail(m->ctx, SGPU_ERR_ARG, "output scale %d is zero", j);
    std::vector<float> nrm;
    nrm.insert(nrm.end(), x_mean, x_mean + m->n_in);
    nrm.insert(nrm.end(), x_scale, x_scale + m->n_in);
    nrm.insert(nrm.end(), y_mean, y_mean + m->n_out);
    nrm.insert(nrm.end(), y_scale, y_scale + m->n_out);
    HIP_OK(m->ctx, hipStreamSynchronize(m->ctx->stream));
    HIP_OK(m->ctx, hipMemcpy(m->norm, nrm.data(), sizeof(float) * nrm.size(), hipMemcpyHostToDevice));
    return SGPU_OK;
}

int sgpu_mlp_get_normalization(sgpu_mlp *m, float *out) {
    if (!m || !out) return SGPU_ERR_ARG;
    HIP_OK(m->ctx, hipStreamSynchronize(m->ctx->stream));
    HIP_OK(m->ctx, hipMemcpy(out, m->norm, sizeof(float) * (2 * (size_t)m->n_in + 2 * (size_t)m->n_out),
                             hipMemcpyDeviceToHost));
    return SGPU_OK;
}

int sgpu_mlp_fit_normalization(sgpu_mlp *m, const float *x, const float *y, int64_t count) {
    if (!m || !x || !y || count <= 0) return SGPU_ERR_ARG;
    sgpu_ctx *ctx = m->ctx;
    const int64_t need = (int64_t)kStatChunks * (m->n_in > m->n_out ? m->n_in : m->n_out) * 2;
    if (need > m->partial_cap) {
        mlp_release(m, m->partial);
        int rc;
        if ((rc = mlp_alloc(m, &m->partial, (size_t)need))) return rc;
        m->partial_cap = need;
    }
    std::vector<float> nrm(2 * (size_t)m->n_in + 2 * (size_t)m->n_out);
    const float *src[2] = {x, y};
    const int cols[2] = {m->n_in, m->n_out};
    for (int which = 0; which < 2; ++which) {
        const int nc = cols[which];
        HIP_OK(ctx, launch_mlp_colstats(ctx->stream, src[which], count, nc, kStatChunks, m->partial));
        std::vector<double> part((size_t)kStatChunks * nc * 2);
        HIP_OK(ctx, hipMemcpyAsync(part.data(), m->partial, sizeof(double) * part.size(), hipMemcpyDeviceToHost,
                                   ctx->stream));
        HIP_OK(ctx, hipStreamSynchronize(ctx->stream));
        for (int c = 0; c < nc; ++c) {
            double s = 0.0, q = 0.0;
            for (int k = 0; k < kStatChunks; ++k) {
                s += part[((size_t)k * nc + c) * 2];
                q += part[((size_t)k * nc + c) * 2 + 1];
            }
            const double mean = s / (double)count;
            const double var = q / (double)count - mean * mean;
            const double sd = var > 0.0 ? std::sqrt(var) : 0.0;
            float *mean_out = nrm.data() + (which == 0 ? 0 : 2 * (size_t)m->n_in);
            float *scale_out = mean_out + nc;
            mean_out[c] = (float)mean;
            if (which == 0) scale_out[c] = sd > 0.0 && std::isfinite(1.0 / sd) ? (float)(1.0 / sd) : 0.0f;
            else scale_out[c] = sd > 0.0 && (float)sd > 0.0f ? (float)sd : 1.0f;
        }
    }
    return sgpu_mlp_set_normalization(m, nrm.data(), nrm.data() + m->n_in, nrm.data() + 2 * m->n_in,
                                      nrm.data() + 2 * m->n_in + m->n_out);
}

int sgpu_mlp_gradients(sgpu_mlp *m, const float *x, const float *y, int32_t rows, float *grad_out, double *loss) {
    if (!m || !x || !y || rows <= 0) return SGPU_ERR_ARG;
    sgpu_ctx *ctx = m->ctx;
    const int batch = round16(rows);
    int rc;
    if ((rc = ensure_workspace(m, batch))) return rc;
    HIP_OK(ctx, hipMemsetAsync(m->loss, 0, sizeof(float) * (size_t)(batch / 16), ctx->stream));
    if ((rc = fwdbwd(m, batch, rows, 0, nullptr, x, y))) return rc;
    std::vector<float> lp((size_t)batch / 16);
    HIP_OK(ctx, hipMemcpyAsync(lp.data(), m->loss, sizeof(float) * lp.size(), hipMemcpyDeviceToHost, ctx->stream));
    if (grad_out)
        HIP_OK(ctx, hipMemcpyAsync(grad_out, m->g, sizeof(float) * (size_t)m->n_params, hipMemcpyDeviceToHost,
                                   ctx->stream));
    HIP_OK(ctx, hipStreamSynchronize(ctx->stream));
    if (loss) {
        double s = 0.0;
        for (float v : lp) s += v;
        *loss = s / ((double)rows * m->n_out);
    }
    return SGPU_OK;
}

int sgpu_mlp_train(sgpu_mlp *m, const float *x, const float *y, int64_t count, int32_t batch, int32_t epochs,
                   float lr, uint64_t seed, double *epoch_loss) {
    if (!m || !x || !y || count <= 0 || batch <= 0 || batch % 16 || epochs <= 0 || !(lr > 0.0f))
        return fail(m ? m->ctx : nullptr, SGPU_ERR_ARG, "bad training arguments");
    sgpu_ctx *ctx = m->ctx;
    const int64_t nb = count / batch;
    if (nb <= 0) return fail(ctx, SGPU_ERR_ARG, "fewer samples (%lld) than one minibatch (%d)", (long long)count, batch);
    if (count > 0x7fffffff) return fail(ctx, SGPU_ERR_ARG, "too many samples");
    int rc;
    if ((rc = ensure_workspace(m, batch))) return rc;
    if (count > m->perm_cap) {
        mlp_release(m, m->perm);
        if ((rc = mlp_alloc(m, &m->perm, (size_t)count))) return rc;
        m->perm_cap = count;
    }
    std::vector<int32_t> perm((size_t)count);
    for (int64_t i = 0; i < count; ++i) perm[(size_t)i] = (int32_t)i;
    uint64_t s = seed;
    const int64_t total = nb * epochs, warm = total / 50 > 0 ? total / 50 : 1;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    MlpAdamArgs P{};
    P.p = m->p;
    P.g = m->g;
    P.mom = m->mom;
    P.vel = m->vel;
    P.w = m->w;
    P.wt = m->wt;
    P.n = m->n_params;
    P.b1 = b1;
    P.b2 = b2;
    for (int l = 0; l < kMlpLayers; ++l) {
        P.off_w[l] = m->off_w[l];
        P.off_bf[l] = m->off_bf[l];
        P.N[l] = m->N[l];
        P.K[l] = m->K[l];
    }
    int64_t stepi = 0;
    std::vector<float> lp((size_t)batch / 16);
    for (int ep = 0; ep < epochs; ++ep) {
        for (int64_t i = count - 1; i > 0; --i) {  // Fisher-Yates
            const int64_t j = (int64_t)(splitmix(s) % (uint64_t)(i + 1));
            std::swap(perm[(size_t)i], perm[(size_t)j]);
        }
        HIP_OK(ctx, hipMemcpyAsync(m->perm, perm.data(), sizeof(int32_t) * (size_t)count, hipMemcpyHostToDevice,
                                   ctx->stream));
        HIP_OK(ctx, hipMemsetAsync(m->loss, 0, sizeof(float) * (size_t)(batch / 16), ctx->stream));
        for (int64_t b = 0; b < nb; ++b, ++stepi) {
            if ((rc = fwdbwd(m, batch, batch, b * batch, m->perm, x, y))) return rc;
            // linear warm-up, then cosine decay to 1% of lr
            double f = stepi < warm ? (double)(stepi + 1) / warm
                                    : 0.01 + 0.99 * 0.5 * (1.0 + std::cos(M_PI * (double)(stepi - warm) /
                                                                        (double)(total - warm > 0 ? total - warm : 1)));
            m->adam_t += 1;
            const double c1 = 1.0 - std::pow((double)b1, (double)m->adam_t);
            const double c2 = 1.0 - std::pow((double)b2, (double)m->adam_t);
            P.lr_t = (float)(lr * f * std::sqrt(c2) / c1);
            P.eps_t = (float)(eps * std::sqrt(c2));
            HIP_OK(ctx, launch_mlp_adam(ctx->stream, P));
        }
        HIP_OK(ctx, hipMemcpyAsync(lp.data(), m->loss, sizeof(float) * lp.size(), hipMemcpyDeviceToHost, ctx->stream));
        HIP_OK(ctx, hipStreamSynchronize(ctx->stream));
        if (epoch_loss) {
            double sum = 0.0;
            for (float v : lp) sum += v;
            epoch_loss[ep] = sum / ((double)nb * batch * m->n_out);
        }
    }
    return SGPU_OK;
}

int sgpu_mlp_predict(sgpu_mlp *m, const float *x, int64_t count, float *y) {
    if (!m || !x || !y || count <= 0) return SGPU_ERR_ARG;
    sgpu_ctx *ctx = m->ctx;
    const int chunk = 1024;
    int rc;
    if ((rc = ensure_workspace(m, chunk))) return rc;
    for (int64_t first = 0; first < count; first += chunk) {
        const int rows = (int)(count - first < chunk ? count - first : chunk);
        const int batch = round16(rows);
        MlpPrepArgs P = prep_args(m, batch, rows, first, nullptr, x, nullptr);
        P.t = nullptr;
        HIP_OK(ctx, launch_mlp_prep(ctx->stream, P));
        MlpTrainArgs T = train_args(m, batch, rows, 0);
        T.y_raw = y + first * m->n_out;
        HIP_OK(ctx, launch_mlp_fwdbwd(ctx->stream, T));
    }
    HIP_OK(ctx, hipStreamSynchronize(ctx->stream));
    return SGPU_OK;
}

}  // extern "C"

namespace sgpu {
// Device view of a network (for the frame kernels' in-step inference).
const MlpDev *mlp_device_view(const sgpu_mlp *m) { return m ? &m->dev : nullptr; }
}  // namespace sgpu
