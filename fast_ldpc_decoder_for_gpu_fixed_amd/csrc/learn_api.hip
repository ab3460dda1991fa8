// learn_api.hip -- C ABI of the learned-table trainer: float shadow decoder
// forward/backward on the GPU, Adam on the host, training frames from the
// device channel with a per-frame Eb/N0.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ldpc_mi355x.h"
#include "code.h"
#include "kernels.h"

namespace ldpc {
void set_last_error(const std::string &m);  // api.hip: the error text of ldpc_last_error()
}

struct ldpc_trainer {
    ldpc_code code;
    ldpc::Plan plan;
    int mode = 0, T = 0, dcols = 0, FB = 4, lfb = 2;
    size_t frames = 0;
    hipStream_t stream = nullptr;
    uint2 *d_checks = nullptr;
    uint32_t *d_edge_var = nullptr;
    int *d_level_beg = nullptr;
    float *d_msg = nullptr, *d_r = nullptr, *d_pend = nullptr, *d_params = nullptr, *d_grad = nullptr,
          *d_loss = nullptr;
    int8_t *d_llr = nullptr;
};

#define LHIP(expr)                                                                          \
    do {                                                                                    \
        hipError_t _e = (expr);                                                             \
        if (_e != hipSuccess) {                                                             \
            ldpc::set_last_error(std::string(#expr) + ": " + hipGetErrorString(_e));        \
            return LDPC_ERR_HIP;                                                            \
        }                                                                                   \
    } while (0)

static void trainer_free(ldpc_trainer *t) {
    if (!t) return;
    (void)hipFree(t->d_checks);
    (void)hipFree(t->d_edge_var);
    (void)hipFree(t->d_level_beg);
    (void)hipFree(t->d_msg);
    (void)hipFree(t->d_r);
    (void)hipFree(t->d_pend);
    (void)hipFree(t->d_params);
    (void)hipFree(t->d_grad);
    (void)hipFree(t->d_loss);
    (void)hipFree(t->d_llr);
    if (t->stream) (void)hipStreamDestroy(t->stream);
    delete t;
}

extern "C" {

ldpc_trainer *ldpc_trainer_create(const ldpc_code *code, int mode, int iterations, int n_degs, size_t batch_frames) {
    if (!code || (mode != LDPC_LEARN_OFFSET && mode != LDPC_LEARN_SCALE) || iterations < 1 || n_degs < 1 ||
        n_degs > ldpc::MAX_DEG + 1 || batch_frames == 0) {
        ldpc::set_last_error("ldpc_trainer_create: bad arguments");
        return nullptr;
    }
    if (code->max_deg > 16) {
        ldpc::set_last_error("the float shadow decoder handles check degrees up to 16");
        return nullptr;
    }
    int FB = 4;
    while (FB > 1 && ((size_t)code->N * FB * 4 > ldpc::MAX_LDS_BYTES - 4096 || batch_frames % FB)) FB /= 2;
    if ((size_t)code->N * FB * 4 > ldpc::MAX_LDS_BYTES - 4096 || batch_frames % FB) {
        ldpc::set_last_error("frame too long for the shadow decoder's LDS block");
        return nullptr;
    }
    ldpc_trainer *t = new ldpc_trainer();
    t->code = *code;
    t->plan = ldpc::make_plan(t->code);
    t->mode = mode;
    t->T = iterations;
    t->dcols = n_degs;
    t->FB = FB;
    while ((1 << t->lfb) < FB) t->lfb++;
    t->frames = batch_frames;
    const size_t E = (size_t)code->E, N = (size_t)code->N, F = batch_frames, T = (size_t)iterations;
    std::vector<uint2> ck((size_t)code->C);
    for (int i = 0; i < code->C; i++) ck[i] = make_uint2(t->plan.check_info[i], t->plan.check_e0[i]);
    hipError_t e = hipSuccess;
    auto al = [&](auto **p, size_t n) {
        if (e == hipSuccess && n) e = hipMalloc((void **)p, n * sizeof(**p));
    };
    al(&t->d_checks, ck.size());
    al(&t->d_edge_var, E);
    al(&t->d_level_beg, t->plan.level_beg.size());
    al(&t->d_msg, E * F);
    al(&t->d_r, T * E * F);
    al(&t->d_pend, T * N * F);
    al(&t->d_params, T * (size_t)n_degs);
    al(&t->d_grad, T * (size_t)n_degs);
    al(&t->d_loss, 1);
    al(&t->d_llr, N * F);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&t->stream, hipStreamDefault);
    if (e == hipSuccess) e = hipMemcpy(t->d_checks, ck.data(), ck.size() * sizeof(uint2), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(t->d_edge_var, t->plan.edge_var.data(), E * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(t->d_level_beg, t->plan.level_beg.data(), t->plan.level_beg.size() * sizeof(int),
                      hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        ldpc::set_last_error(std::string("ldpc_trainer_create: ") + hipGetErrorString(e));
        trainer_free(t);
        return nullptr;
    }
    return t;
}

void ldpc_trainer_destroy(ldpc_trainer *tr) {
    if (tr) (void)hipStreamSynchronize(tr->stream);
    trainer_free(tr);
}

int ldpc_trainer_loss_grad(ldpc_trainer *tr, const float *params, const int8_t *d_llr, float *loss, float *grad,
                           float *d_post_out) {
    if (!tr || !params || !loss) {
        ldpc::set_last_error("ldpc_trainer_loss_grad: bad arguments");
        return LDPC_ERR_ARG;
    }
    const size_t np = (size_t)tr->T * tr->dcols;
    LHIP(hipMemcpyAsync(tr->d_params, params, np * sizeof(float), hipMemcpyHostToDevice, tr->stream));
    LHIP(hipMemsetAsync(tr->d_grad, 0, np * sizeof(float), tr->stream));
    LHIP(hipMemsetAsync(tr->d_loss, 0, sizeof(float), tr->stream));
    ldpc::ShadowArgs a;
    std::memset(&a, 0, sizeof a);
    a.checks = tr->d_checks;
    a.edge_var = tr->d_edge_var;
    a.level_beg = tr->d_level_beg;
    a.n_levels = tr->plan.n_levels;
    a.N = tr->code.N;
    a.C = tr->code.C;
    a.E = tr->code.E;
    a.FB = tr->FB;
    a.lfb = tr->lfb;
    a.frames = tr->frames;
    a.llr = d_llr ? d_llr : tr->d_llr;
    a.msg = tr->d_msg;
    a.rstore = tr->d_r;
    a.pend = tr->d_pend;
    a.params = tr->d_params;
    a.grad = tr->d_grad;
    a.loss = tr->d_loss;
    a.post_out = d_post_out;
    a.T = tr->T;
    a.dcols = tr->dcols;
    a.mode = tr->mode;
    LHIP(ldpc::launch_shadow(a, false, tr->stream));
    if (grad) LHIP(ldpc::launch_shadow(a, true, tr->stream));
    float l = 0.f;
    LHIP(hipMemcpyAsync(&l, tr->d_loss, sizeof(float), hipMemcpyDeviceToHost, tr->stream));
    if (grad) LHIP(hipMemcpyAsync(grad, tr->d_grad, np * sizeof(float), hipMemcpyDeviceToHost, tr->stream));
    LHIP(hipStreamSynchronize(tr->stream));
    *loss = l / (float)tr->frames;
    if (grad)
        for (size_t i = 0; i < np; i++) grad[i] /= (float)tr->frames;
    return LDPC_OK;
}

int ldpc_trainer_fit(ldpc_trainer *tr, const ldpc_learn_config *cfg, float *params, float *loss_history) {
    if (!tr || !cfg || !params || cfg->steps < 1 || !(cfg->lr > 0) || !(cfg->ebn0_hi >= cfg->ebn0_lo)) {
        ldpc::set_last_error("ldpc_trainer_fit: bad arguments");
        return LDPC_ERR_ARG;
    }
    const size_t np = (size_t)tr->T * tr->dcols;
    std::vector<double> m(np, 0.0), v(np, 0.0);
    std::vector<float> g(np);
    const double rate = cfg->rate > 0 ? cfg->rate : (double)(tr->code.N - tr->code.C) / tr->code.N;
    for (int step = 0; step < cfg->steps; step++) {
        const uint64_t f0 = cfg->first_frame + (uint64_t)step * tr->frames;
        LHIP(ldpc::launch_channel_range(tr->d_llr, tr->code.N, tr->frames, f0, cfg->seed, 0, cfg->ebn0_lo,
                                        cfg->ebn0_hi, rate, cfg->quant, tr->stream));
        float loss = 0.f;
        int rc = ldpc_trainer_loss_grad(tr, params, nullptr, &loss, g.data(), nullptr);
        if (rc) return rc;
        if (loss_history) loss_history[step] = loss;
        const double b1 = cfg->beta1 > 0 ? cfg->beta1 : 0.9, b2 = cfg->beta2 > 0 ? cfg->beta2 : 0.999;
        const double eps = cfg->eps > 0 ? cfg->eps : 1e-8;
        const double c1 = 1.0 - std::pow(b1, step + 1), c2 = 1.0 - std::pow(b2, step + 1);
        for (size_t i = 0; i < np; i++) {
            m[i] = b1 * m[i] + (1 - b1) * g[i];
            v[i] = b2 * v[i] + (1 - b2) * (double)g[i] * g[i];
            double p = params[i] - cfg->lr * (m[i] / c1) / (std::sqrt(v[i] / c2) + eps);
            if (p < cfg->pmin) p = cfg->pmin;
            if (p > cfg->pmax) p = cfg->pmax;
            params[i] = (float)p;
        }
    }
    return LDPC_OK;
}

int ldpc_channel_generate_range(void *d_out, int n_vars, size_t frames, uint64_t first_frame, uint64_t seed,
                                uint32_t snr_idx, double ebn0_lo, double ebn0_hi, double rate, int quant,
                                void *stream) {
    if (!d_out || n_vars <= 0 || frames == 0 || !(rate > 0) || !(ebn0_hi >= ebn0_lo) || (quant != 0 && quant != 1)) {
        ldpc::set_last_error("ldpc_channel_generate_range: bad arguments");
        return LDPC_ERR_ARG;
    }
    LHIP(ldpc::launch_channel_range((int8_t *)d_out, n_vars, frames, first_frame, seed, snr_idx, ebn0_lo, ebn0_hi,
                                    rate, quant, (hipStream_t)stream));
    return LDPC_OK;
}

}  // extern "C"
