// selector.cpp -- the learned precision selector's host side: the model object, its training and prediction (through
// the device kernels of selector_dev.hip), and the training logs of a factorization with precision 3.
#include "numeric_obj.h"
#include "selector_dev.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>
#include <vector>

using namespace fpa;

struct fplus_amd_selector {
    std::vector<float> w;          // SEL_NW floats
    float threshold = 0.5f;
    mutable std::mutex mu;
    mutable DevBuf<float> d_w;     // the device copy the factorization kernels read
    mutable bool dirty = true;
    const float* device_model() const
    {
        std::lock_guard<std::mutex> g(mu);
        if (dirty || !d_w.p) {
            d_w.alloc(SEL_NW);
            d_w.upload(w.data(), SEL_NW, devctx().stream);
            dirty = false;
        }
        return d_w.p;
    }
};

namespace fpa {
// for numeric.cpp: the device model of a selector and its threshold
SelectorArgs selector_args(const fplus_amd_selector* sel, const double* d_Dr, const double* d_Dc)
{
    SelectorArgs a;
    a.d_model = sel ? sel->device_model() : nullptr;
    a.threshold = sel ? sel->threshold : 0.5f;
    a.d_Dr = d_Dr;
    a.d_Dc = d_Dc;
    return a;
}
}  // namespace fpa

namespace {
void init_weights(std::vector<float>& w, uint64_t seed)
{
    w.assign(SEL_NW, 0.0f);
    std::mt19937_64 rng(seed);
    std::normal_distribution<float> nd(0.0f, 1.0f);
    float* W1 = w.data();
    float* W2 = W1 + SEL_H * SEL_NF + SEL_H;
    float* W3 = W2 + SEL_H * SEL_H + SEL_H;
    for (int i = 0; i < SEL_H * SEL_NF; i++) W1[i] = nd(rng) * std::sqrt(2.0f / SEL_NF);   // He initialisation
    for (int i = 0; i < SEL_H * SEL_H; i++) W2[i] = nd(rng) * std::sqrt(2.0f / SEL_H);
    for (int i = 0; i < 2 * SEL_H; i++) W3[i] = nd(rng) * std::sqrt(1.0f / SEL_H);
    float* mu = w.data() + SEL_NP;
    float* sd = mu + SEL_NF;
    for (int i = 0; i < SEL_NF; i++) { mu[i] = 0.0f; sd[i] = 1.0f; }
}
}  // namespace

extern "C" {

int fplus_amd_selector_create(uint64_t seed, fplus_amd_selector** sel)
{
    try {
        auto* s = new fplus_amd_selector();
        init_weights(s->w, seed);
        *sel = s;
        return FPLUS_OK;
    } catch (const std::bad_alloc&) {
        return FPLUS_ERR_NOMEM;
    }
}

void fplus_amd_selector_free(fplus_amd_selector* sel) { delete sel; }

int fplus_amd_selector_set_threshold(fplus_amd_selector* sel, float threshold)
{
    if (!sel || !(threshold >= 0.0f && threshold <= 1.0f)) return FPLUS_ERR_BADARG;
    sel->threshold = threshold;
    return FPLUS_OK;
}

int fplus_amd_selector_get_weights(const fplus_amd_selector* sel, float* w, int64_t cap)
{
    if (!sel || !w || cap < SEL_NW) return FPLUS_ERR_BADARG;
    std::memcpy(w, sel->w.data(), SEL_NW * sizeof(float));
    return FPLUS_OK;
}

int fplus_amd_selector_set_weights(fplus_amd_selector* sel, const float* w, int64_t n)
{
    if (!sel || !w || n != SEL_NW) return FPLUS_ERR_BADARG;
    for (int64_t i = 0; i < n; i++)
        if (!std::isfinite(w[i])) return FPLUS_ERR_BADARG;
    std::lock_guard<std::mutex> g(sel->mu);
    sel->w.assign(w, w + n);
    sel->dirty = true;
    return FPLUS_OK;
}

int fplus_amd_selector_train(fplus_amd_selector* sel, const float* features, const float* labels, int64_t nsamples, int epochs,
                             float lr, uint64_t seed, float* final_loss)
{
    if (!sel || !features || !labels || nsamples < 1 || epochs < 1 || !(lr > 0.0f)) return FPLUS_ERR_BADARG;
    if (nsamples >= ((int64_t)1 << 31)) return FPLUS_ERR_BADARG;
    try {
        hipStream_t s = devctx().stream;
        // the features' means and scales, from the samples
        std::vector<float> w(sel->w);
        float* mu = w.data() + SEL_NP;
        float* sd = mu + SEL_NF;
        for (int i = 0; i < SEL_NF; i++) {
            double m = 0, q = 0;
            for (int64_t t = 0; t < nsamples; t++) m += features[t * SEL_NF + i];
            m /= (double)nsamples;
            for (int64_t t = 0; t < nsamples; t++) { const double d = features[t * SEL_NF + i] - m; q += d * d; }
            mu[i] = (float)m;
            sd[i] = (float)std::sqrt(q / (double)nsamples) + 1e-3f;
        }
        DevBuf<float> d_f((size_t)nsamples * SEL_NF), d_l((size_t)nsamples), d_m(SEL_NW), d_mom(SEL_NP), d_loss(1);
        DevBuf<int32_t> d_ord((size_t)nsamples);
        d_f.upload(features, (size_t)nsamples * SEL_NF, s);
        d_l.upload(labels, (size_t)nsamples, s);
        d_m.upload(w.data(), SEL_NW, s);
        const std::vector<float> zero_mom(SEL_NP, 0.0f);
        d_mom.upload(zero_mom.data(), SEL_NP, s);
        std::vector<int32_t> ord((size_t)nsamples);
        std::iota(ord.begin(), ord.end(), 0);
        std::mt19937_64 rng(seed);
        float loss = 0;
        for (int e = 0; e < epochs; e++) {
            std::shuffle(ord.begin(), ord.end(), rng);
            d_ord.upload(ord.data(), ord.size(), s);
            const float zero = 0.0f;
            d_loss.upload(&zero, 1, s);
            const float lre = lr * (e < epochs / 2 ? 1.0f : (e < 3 * epochs / 4 ? 0.3f : 0.1f));   // step decay
            for (int64_t f0 = 0; f0 < nsamples; f0 += 256)
                k::sel_train_step(nsamples, d_ord.p, f0, d_f.p, d_l.p, d_m.p, d_mom.p, lre, 0.9f, 1e-4f, d_loss.p, s);
            d_loss.download(&loss, 1, s);
            loss /= (float)nsamples;
        }
        d_m.download(w.data(), SEL_NW, s);
        dev_sync(s);
        for (float v : w)
            if (!std::isfinite(v)) return FPLUS_ERR_INTERNAL;
        std::lock_guard<std::mutex> g(sel->mu);
        sel->w = w;
        sel->dirty = true;
        if (final_loss) *final_loss = loss;
    } catch (const GpuError&) {
        return FPLUS_ERR_GPU;
    } catch (const std::bad_alloc&) {
        return FPLUS_ERR_NOMEM;
    }
    return FPLUS_OK;
}

int fplus_amd_selector_predict(const fplus_amd_selector* sel, const float* features, int64_t nsamples, float* prob)
{
    if (!sel || !features || !prob || nsamples < 0) return FPLUS_ERR_BADARG;
    if (nsamples == 0) return FPLUS_OK;
    try {
        hipStream_t s = devctx().stream;
        DevBuf<float> d_f((size_t)nsamples * SEL_NF), d_p((size_t)nsamples);
        d_f.upload(features, (size_t)nsamples * SEL_NF, s);
        k::sel_infer(nsamples, nullptr, d_f.p, sel->device_model(), sel->threshold, d_p.p, nullptr, s);
        d_p.download(prob, (size_t)nsamples, s);
        dev_sync(s);
    } catch (const GpuError&) {
        return FPLUS_ERR_GPU;
    } catch (const std::bad_alloc&) {
        return FPLUS_ERR_NOMEM;
    }
    return FPLUS_OK;
}

int fplus_amd_selector_log(const fplus_numeric* num, float* features, double* diff, double* pmax, int64_t cap, int64_t* nblocks)
{
    if (!num || !nblocks || num->row) return FPLUS_ERR_BADARG;
    const DevFactor& F = num->F;
    const int64_t nb = (int64_t)F.blocks.size();
    *nblocks = nb;
    if (F.precision != 3 || !F.d_feat.p || !F.d_blk_err.p) return FPLUS_ERR_BADARG;
    if (!features || !diff || !pmax || cap < nb) return FPLUS_ERR_BADARG;
    try {
        hipStream_t s = devctx().stream;
        F.d_feat.download(features, (size_t)nb * SEL_NF, s);
        std::vector<unsigned long long> e((size_t)(2 * nb));
        F.d_blk_err.download(e.data(), e.size(), s);
        dev_sync(s);
        for (int64_t b = 0; b < nb; b++) {
            std::memcpy(&diff[b], &e[2 * b], 8);
            std::memcpy(&pmax[b], &e[2 * b + 1], 8);
        }
    } catch (const GpuError&) {
        return FPLUS_ERR_GPU;
    }
    return FPLUS_OK;
}

}  // extern "C"
