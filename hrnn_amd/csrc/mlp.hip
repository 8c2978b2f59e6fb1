// mlp.hip -- the learned proxy count and early stop: a 32-64-64-16 MLP (ReLU,
// softmax over 16 classes) with hand-written forward, backward and Adam
// kernels, and the two-stage query path it drives.
//
// No reference counterpart: the reference picks m from a fixed grid
// (src/hrnn_search.cpp:69) and runs every beam to ef_search.  Here the
// network predicts per query the cheapest class that reaches full recall:
// either a proxy count m of a 16-value grid (host-feature path) or an (ef, m)
// pair (adaptive path: a narrow stage-1 beam gives the features, and only the
// queries whose class needs a wider beam run phase A again).
//
// Parameter vector (FP32, HRNN_MLP_PARAMS floats):
//   W1[64][32] b1[64] W2[64][64] b2[64] W3[16][64] b3[16]
// Forward per sample: z1 = W1 x + b1, h1 = relu(z1), z2 = W2 h1 + b2,
// h2 = relu(z2), logits = W3 h2 + b3, p = softmax(logits).
// Loss: mean cross-entropy over the batch.  The batch gradient is the mean of
// per-sample gradients, reduced in a fixed order (deterministic).
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <random>
#include <vector>

#include "handle.h"
#include "mlp_core.h"

namespace hrnn {

// One 64-lane block per sample.  Writes probabilities (and, when grad != null,
// the per-sample gradient row grad[s][kParams] and loss[s]).
__global__ __launch_bounds__(64) void k_mlp_fwd_bwd(const float *__restrict__ P,
                                                    const float *__restrict__ X,
                                                    const uint32_t *__restrict__ idx,
                                                    const uint32_t *__restrict__ labels, uint32_t n,
                                                    float *__restrict__ probs,
                                                    float *__restrict__ grad,
                                                    float *__restrict__ loss) {
    __shared__ float x[kIn], z1[kH], h1[kH], z2[kH], h2[kH], lg[kOut], dl[kOut], dz2[kH], dz1[kH];
    const uint32_t s = blockIdx.x;
    if (s >= n) return;
    const uint32_t j = threadIdx.x;
    const uint32_t row = idx ? idx[s] : s;
    if (j < kIn) x[j] = X[(uint64_t)row * kIn + j];
    __syncthreads();
    mlp_forward_wave(P, x, z1, h1, z2, h2, lg, dl, j);   // dl holds p for now
    if (j == 0 && loss && labels) loss[s] = -logf(fmaxf(dl[labels[row]], 1e-30f));
    __syncthreads();
    if (probs && j < kOut) probs[(uint64_t)s * kOut + j] = dl[j];
    if (!grad) return;
    float *g = grad + (uint64_t)s * kParams;
    __syncthreads();
    if (j < kOut) dl[j] -= (labels[row] == j) ? 1.f : 0.f;   // dlogits = p - onehot
    __syncthreads();
    // layer 3
    for (int t = j; t < kOut * kH; t += 64) g[oW3 + t] = dl[t / kH] * h2[t % kH];
    if (j < kOut) g[oB3 + j] = dl[j];
    {
        float a = 0.f;
        for (int c = 0; c < kOut; c++) a = __builtin_fmaf(P[oW3 + c * kH + j], dl[c], a);
        dz2[j] = z2[j] > 0.f ? a : 0.f;
    }
    __syncthreads();
    for (int t = j; t < kH * kH; t += 64) g[oW2 + t] = dz2[t / kH] * h1[t % kH];
    g[oB2 + j] = dz2[j];
    {
        float a = 0.f;
        for (int c = 0; c < kH; c++) a = __builtin_fmaf(P[oW2 + c * kH + j], dz2[c], a);
        dz1[j] = z1[j] > 0.f ? a : 0.f;
    }
    __syncthreads();
    for (int t = j; t < kH * kIn; t += 64) g[oW1 + t] = dz1[t / kIn] * x[t % kIn];
    g[oB1 + j] = dz1[j];
}

// mean over the batch rows, fixed order: out[p] = (1/B) sum_s grad[s][p]
__global__ void k_mlp_reduce(const float *__restrict__ grad, uint32_t B, float *__restrict__ out) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (uint32_t)kParams) return;
    float a = 0.f;
    for (uint32_t s = 0; s < B; s++) a += grad[(uint64_t)s * kParams + p];
    out[p] = a / (float)B;
}

__global__ void k_mlp_adam(float *__restrict__ P, const float *__restrict__ g, float *__restrict__ m1,
                           float *__restrict__ m2, float lr, float b1, float b2, float eps,
                           float c1, float c2, float wd) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (uint32_t)kParams) return;
    float gv = g[p] + wd * P[p];
    float a = b1 * m1[p] + (1.f - b1) * gv;
    float b = b2 * m2[p] + (1.f - b2) * gv * gv;
    m1[p] = a;
    m2[p] = b;
    P[p] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

// In-place feature standardisation f = (f - mu) / sd over n rows of kIn.
__global__ void k_standardize(float *__restrict__ f, uint32_t n, const float *__restrict__ mu,
                              const float *__restrict__ sd) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint64_t)n * kIn) return;
    uint32_t j = (uint32_t)(i % kIn);
    f[i] = (f[i] - mu[j]) / sd[j];
}

// Per-query m: the smallest grid class whose cumulative probability reaches
// tau (argmax when tau <= 0); proxies beyond that m are masked out.
__global__ void k_pick_m(const float *__restrict__ probs, uint32_t nq, const uint32_t *__restrict__ grid,
                         float tau, uint32_t *__restrict__ prox, uint32_t mstride,
                         uint32_t *__restrict__ m_out) {
    uint32_t qi = blockIdx.x * blockDim.x + threadIdx.x;
    if (qi >= nq) return;
    const float *p = probs + (uint64_t)qi * kOut;
    int c = kOut - 1;
    if (tau <= 0.f) {
        c = 0;
        for (int j = 1; j < kOut; j++)
            if (p[j] > p[c]) c = j;
    } else {
        float acc = 0.f;
        for (int j = 0; j < kOut; j++) {
            acc += p[j];
            if (acc >= tau) {
                c = j;
                break;
            }
        }
    }
    uint32_t m = grid[c] < mstride ? grid[c] : mstride;
    for (uint32_t j = m; j < mstride; j++) prox[(uint64_t)qi * mstride + j] = kInvalidId;
    if (m_out) m_out[qi] = m;
}

// Learned early stop: per query the cheapest (ef, m) class.  Classes are
// ordered by cost, each dominating the previous one (ef and m both
// non-decreasing).  With probs: the first class whose cumulative probability
// reaches tau (argmax when tau <= 0); without: the class given in cls.  Writes
// the class, the stage-2 beam width (0 when the class's ef <= ef1: the
// stage-1 proxies serve) and the class's m (at most mstride), and masks the
// stage-1 proxies beyond that m.  classes = [ef of 16 classes, m of 16 classes].
// exp2 (optional): the stage-2 expansion counters, zeroed here for the
// queries that have no stage 2.  lists (optional, 2 nq u32): the escalated
// queries are appended to lists[0..), the others to lists[nq..), counted in
// list_cnt[0] and list_cnt[1] (zero before the launch; list order is
// arbitrary).  zero_a / zero_b (optional): words set to zero (the work queues
// of the searches launched after this kernel).
__global__ void k_pick_class(const float *__restrict__ probs, uint32_t nq,
                             const uint32_t *__restrict__ classes, float tau, uint32_t ef1,
                             uint32_t *__restrict__ prox, uint32_t mstride, uint32_t *__restrict__ cls,
                             uint32_t *__restrict__ ef_q, uint32_t *__restrict__ m_q,
                             uint32_t *__restrict__ exp2, uint32_t *__restrict__ lists,
                             uint32_t *__restrict__ list_cnt, uint32_t *__restrict__ zero_a,
                             uint32_t *__restrict__ zero_b) {
    const uint32_t qi = blockIdx.x * blockDim.x + threadIdx.x;
    if (qi == 0) {   // the work queues of the two searches that follow
        if (zero_a) *zero_a = 0;
        if (zero_b) *zero_b = 0;
    }
    if (qi >= nq) return;
    uint32_t c = kOut - 1;
    if (probs) {
        c = pick_class_of(probs + (uint64_t)qi * kOut, tau);
        cls[qi] = c;
    } else {
        c = min(cls[qi], (uint32_t)kOut - 1u);
    }
    const uint32_t ef = classes[c];
    const uint32_t m = min(max(classes[kOut + c], 1u), mstride);
    ef_q[qi] = ef > ef1 ? ef : 0u;
    m_q[qi] = m;
    if (lists) {   // escalated queries to lists[0..), the others to lists[nq..)
        // one atomic per wave and list (10k single-lane atomics on two words
        // would serialise)
        const bool esc = ef > ef1;
        const uint32_t lane = threadIdx.x & 63u;
        const uint64_t me = __ballot(esc), mr = __ballot(!esc);
        const uint64_t mine = esc ? me : mr;
        const uint32_t leader = (uint32_t)__ffsll((long long)mine) - 1u;
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(&list_cnt[esc ? 0 : 1], (uint32_t)__popcll(mine));
        base = (uint32_t)__shfl((int)base, (int)leader, 64);
        lists[(uint64_t)(esc ? 0u : 1u) * nq + base + (uint32_t)__popcll(mine & lanemask_lt())] = qi;
    }
    if (exp2 && ef <= ef1) {   // no stage-2 beam for this query
        exp2[2 * (uint64_t)qi] = 0;
        exp2[2 * (uint64_t)qi + 1] = 0;
    }
    for (uint32_t j = m; j < mstride; j++) prox[(uint64_t)qi * mstride + j] = kInvalidId;
}

}  // namespace hrnn

using namespace hrnn;

extern "C" {

int hrnn_dev_rknn_search_learned_dptr(hrnn_dev_t h, const float *q_dev, uint32_t nq, float *feats_dev,
                                      const float *model_dev, const uint32_t *grid_dev, float tau,
                                      uint32_t *proxies_dev, uint32_t m_max, uint32_t *m_dev, uint32_t k,
                                      uint32_t K_prime, float threshold, uint32_t *res_cnt_dev,
                                      uint32_t *res_ids_dev, uint32_t cap_per_q, uint32_t *work_dev) {
    if (!h) return HRNN_E_ARG;
    return guarded(h, [&] {
        REQUIRE(h->csr_valid && h->K > 0, HRNN_E_STATE, "search_learned: build the CSR first");
        REQUIRE(k >= 1 && k <= h->K, HRNN_E_ARG, "search_learned: k out of [1, K]");
        REQUIRE(m_max >= 1 && m_max <= 1024, HRNN_E_ARG, "search_learned: m_max must be in [1, 1024]");
        REQUIRE(q_dev && feats_dev && model_dev && grid_dev && proxies_dev && res_cnt_dev && res_ids_dev,
                HRNN_E_ARG, "search_learned: null argument");
        if (!nq) return;
        h->s8.reserve((uint64_t)nq * kOut * 4);
        float *probs = h->s8.as<float>();
        hipLaunchKernelGGL(k_standardize, dim3(ceil_div((uint64_t)nq * kIn, 256)), dim3(256), 0, h->stream,
                           feats_dev, nq, model_dev + kParams, model_dev + kParams + kIn);
        hipLaunchKernelGGL(k_mlp_fwd_bwd, dim3(nq), dim3(64), 0, h->stream, model_dev, feats_dev, nullptr,
                           nullptr, nq, probs, nullptr, nullptr);
        hipLaunchKernelGGL(k_pick_m, dim3(ceil_div(nq, 128)), dim3(128), 0, h->stream, probs, nq, grid_dev,
                           tau, proxies_dev, m_max, m_dev);
        HIPC(hipGetLastError());
        rknn_search_resident(h, q_dev, nq, proxies_dev, m_max, nullptr, k, K_prime, threshold, res_cnt_dev,
                             res_ids_dev, cap_per_q, work_dev);
    });
}

int hrnn_dev_stage1_features_dptr(hrnn_dev_t h, const float *q_dev, uint32_t nq, uint32_t ef1,
                                  uint32_t m, uint32_t k, uint32_t *proxies_dev, float *feats_dev,
                                  uint32_t *expansions_dev) {
    if (!h) return HRNN_E_ARG;
    return guarded(h, [&] {
        REQUIRE(q_dev && proxies_dev && feats_dev, HRNN_E_ARG, "stage1_features: null argument");
        REQUIRE(ef1 >= 1 && m >= 1 && m <= 1024, HRNN_E_ARG, "stage1_features: bad ef1 or m");
        BeamOpt o;
        o.feats = feats_dev;
        o.feat_k = k;
        beam_launch(h, q_dev, nq, m, ef1, proxies_dev, expansions_dev, nullptr, &o);
    });
}

int hrnn_dev_rknn_query_adaptive_dptr(hrnn_dev_t h, const float *q_dev, uint32_t nq,
                                      const float *model_dev, const uint32_t *classes_dev,
                                      uint32_t ef_max, float tau, uint32_t ef1, uint32_t m_max,
                                      uint32_t k, uint32_t K_prime, float threshold,
                                      uint32_t *proxies_dev, uint32_t *cls_dev,
                                      uint32_t *res_cnt_dev, uint32_t *res_ids_dev,
                                      uint32_t cap_per_q, uint32_t *work_dev,
                                      uint32_t *expansions_dev) {
    if (!h) return HRNN_E_ARG;
    return guarded(h, [&] {
        REQUIRE(h->csr_valid && h->K > 0, HRNN_E_STATE, "query_adaptive: build the CSR first");
        REQUIRE(k >= 1 && k <= h->K, HRNN_E_ARG, "query_adaptive: k out of [1, K]");
        REQUIRE(m_max >= 1 && m_max <= 1024, HRNN_E_ARG, "query_adaptive: m_max must be in [1, 1024]");
        REQUIRE(ef1 >= 1 && ef_max >= ef1 && ef_max <= 512, HRNN_E_ARG,
                "query_adaptive: need 1 <= ef1 <= ef_max <= 512");
        REQUIRE(q_dev && classes_dev && proxies_dev && cls_dev && (!res_cnt_dev || res_ids_dev), HRNN_E_ARG,
                "query_adaptive: null argument");
        const bool search = res_cnt_dev != nullptr;   // NULL: phase A only
        if (!nq) return;
        h->s8.reserve((uint64_t)nq * (kIn + kOut + 4) * 4);
        float *feats = h->s8.as<float>();
        float *probs = feats + (uint64_t)nq * kIn;
        uint32_t *ef_q = reinterpret_cast<uint32_t *>(probs + (uint64_t)nq * kOut);
        uint32_t *m_q = ef_q + nq;
        uint32_t *lists = m_q + nq;   // [escalated | the rest], nq each
        uint32_t *list_cnt = h->counters.as<uint32_t>() + kListCountWord;
        ensure_stream2(h);
        // the r_k column of the two concurrent searches, before they fork
        if (search) rk_column_ready(h, k, h->stream);
        HIPC(hipMemsetAsync(list_cnt, 0, 8, h->stream));
        // stage 1: every query at ef1; with a model, the features, the MLP and
        // the class decision run inside the beam kernel's epilogue
        uint32_t *exp1 = expansions_dev, *exp2 = expansions_dev ? expansions_dev + 2ull * nq : nullptr;
        BeamOpt o1;
        if (model_dev) {
            o1.feat_k = k;
            o1.mu = model_dev + kParams;
            o1.sd = model_dev + kParams + kIn;
            o1.mlp = model_dev;
            o1.classes = classes_dev;
            o1.tau = tau;
            o1.ef1 = ef1;
            o1.nq = nq;
            o1.cls = cls_dev;
            o1.ef_q_out = ef_q;
            o1.m_q_out = m_q;
            o1.exp2 = exp2;
            o1.lists = lists;
            o1.list_cnt = list_cnt;
            o1.zero_a = h->counters.as<uint32_t>() + kQueueWord;
            o1.zero_b = h->counters.as<uint32_t>() + kQueue2Word;
            o1.max_exp = (uint32_t)h->tune_stage1_cap;
        }
        // queries in page-locked host memory: read over the link once, by
        // stage 1, which leaves a device copy for everything after it
        const float *qs = q_dev;
        if (host_memory(q_dev)) {
            h->q_stage.reserve((uint64_t)nq * h->d * 4);
            o1.q_copy = h->q_stage.as<float>();
            qs = o1.q_copy;
        }
        beam_launch(h, q_dev, nq, m_max, ef1, proxies_dev, exp1, nullptr, &o1);
        if (!model_dev)   // forced classes
            hipLaunchKernelGGL(k_pick_class, dim3(ceil_div(nq, 128)), dim3(128), 0, h->stream, nullptr, nq,
                               classes_dev, tau, ef1, proxies_dev, m_max, cls_dev, ef_q, m_q, exp2, lists,
                               list_cnt, h->counters.as<uint32_t>() + kQueueWord,
                               h->counters.as<uint32_t>() + kQueue2Word);
        HIPC(hipGetLastError());
        // Phases B+C of the queries that stay at stage 1 run on the second
        // stream while stage 2 and then the escalated queries' phases B+C run
        // on the first; the call ends when both are done.
        if (search) {
            HIPC(hipEventRecord(h->ev_a, h->stream));
            HIPC(hipStreamWaitEvent(h->stream2, h->ev_a, 0));
            SearchSel rest;
            rest.qlist = lists + nq;
            rest.qcount = list_cnt + 1;
            rest.stream = h->stream2;
            rest.queue_word = kQueue2Word;
            rest.queue_zeroed = true;
            rknn_search_resident(h, qs, nq, proxies_dev, m_max, nullptr, k, K_prime, threshold, res_cnt_dev,
                                 res_ids_dev, cap_per_q, work_dev, &rest);
            HIPC(hipEventRecord(h->ev_b, h->stream2));
        }
        if (ef_max > ef1) {   // stage 2: one workgroup per escalated query
            BeamOpt o2;
            o2.ef_q = ef_q;
            o2.m_q = m_q;
            o2.qlist = lists;
            o2.qcount = list_cnt;
            beam_launch(h, qs, nq, m_max, ef_max, proxies_dev, exp2, nullptr, &o2);
        }
        // the escalated queries' search overlaps the other one too: its own
        // fallback buffers and a few workgroups
        if (!search) return;
        SearchSel esc;
        esc.qlist = lists;
        esc.qcount = list_cnt;
        esc.queue_zeroed = true;
        esc.buffer_set = 1;
        esc.max_slots = (uint32_t)std::max(1, h->n_cu);
        rknn_search_resident(h, qs, nq, proxies_dev, m_max, nullptr, k, K_prime, threshold, res_cnt_dev,
                             res_ids_dev, cap_per_q, work_dev, &esc);
        HIPC(hipStreamWaitEvent(h->stream, h->ev_b, 0));
    });
}

int hrnn_dev_mlp_forward(hrnn_dev_t h, const float *params, const float *feats, uint32_t n,
                         float *probs) {
    if (!h) return HRNN_E_ARG;
    return guarded(h, [&] {
        REQUIRE(params && feats && probs, HRNN_E_ARG, "mlp_forward: null argument");
        if (!n) return;
        DevBuf dp, dx, dpr;
        dp.reserve(kParams * 4);
        dx.reserve((uint64_t)n * kIn * 4);
        dpr.reserve((uint64_t)n * kOut * 4);
        h2d(h, dp.p, params, kParams * 4);
        h2d(h, dx.p, feats, (uint64_t)n * kIn * 4);
        hipLaunchKernelGGL(k_mlp_fwd_bwd, dim3(n), dim3(64), 0, h->stream, dp.as<float>(), dx.as<float>(),
                           nullptr, nullptr, n, dpr.as<float>(), nullptr, nullptr);
        HIPC(hipGetLastError());
        d2h(h, probs, dpr.p, (uint64_t)n * kOut * 4);
        dp.release(); dx.release(); dpr.release();
    });
}

int hrnn_dev_mlp_grad(hrnn_dev_t h, const float *params, const float *feats, const uint32_t *labels,
                      uint32_t n, float *grad, float *loss) {
    if (!h) return HRNN_E_ARG;
    return guarded(h, [&] {
        REQUIRE(params && feats && labels && grad, HRNN_E_ARG, "mlp_grad: null argument");
        REQUIRE(n > 0, HRNN_E_ARG, "mlp_grad: empty batch");
        for (uint32_t i = 0; i < n; i++) REQUIRE(labels[i] < (uint32_t)kOut, HRNN_E_ARG, "mlp_grad: label >= 16");
        DevBuf dp, dx, dl, dg, dr, dloss;
        dp.reserve(kParams * 4);
        dx.reserve((uint64_t)n * kIn * 4);
        dl.reserve((uint64_t)n * 4);
        dg.reserve((uint64_t)n * kParams * 4);
        dr.reserve(kParams * 4);
        dloss.reserve((uint64_t)n * 4);
        h2d(h, dp.p, params, kParams * 4);
        h2d(h, dx.p, feats, (uint64_t)n * kIn * 4);
        h2d(h, dl.p, labels, (uint64_t)n * 4);
        hipLaunchKernelGGL(k_mlp_fwd_bwd, dim3(n), dim3(64), 0, h->stream, dp.as<float>(), dx.as<float>(),
                           nullptr, dl.as<uint32_t>(), n, nullptr, dg.as<float>(), dloss.as<float>());
        hipLaunchKernelGGL(k_mlp_reduce, dim3(ceil_div(kParams, 256)), dim3(256), 0, h->stream,
                           dg.as<float>(), n, dr.as<float>());
        HIPC(hipGetLastError());
        d2h(h, grad, dr.p, kParams * 4);
        if (loss) {
            std::vector<float> ls(n);
            d2h(h, ls.data(), dloss.p, (uint64_t)n * 4);
            double a = 0;
            for (float v : ls) a += v;
            *loss = (float)(a / n);
        }
        dp.release(); dx.release(); dl.release(); dg.release(); dr.release(); dloss.release();
    });
}

int hrnn_dev_mlp_train(hrnn_dev_t h, float *params, const float *feats, const uint32_t *labels,
                       uint32_t n, int epochs, uint32_t batch, float lr, float weight_decay,
                       uint64_t seed, float *epoch_loss) {
    if (!h) return HRNN_E_ARG;
    return guarded(h, [&] {
        REQUIRE(params && feats && labels, HRNN_E_ARG, "mlp_train: null argument");
        REQUIRE(n > 0 && batch > 0 && epochs >= 0, HRNN_E_ARG, "mlp_train: bad sizes");
        for (uint32_t i = 0; i < n; i++) REQUIRE(labels[i] < (uint32_t)kOut, HRNN_E_ARG, "mlp_train: label >= 16");
        batch = std::min(batch, n);
        DevBuf dp, dx, dl, dg, dr, dm1, dm2, didx, dloss;
        dp.reserve(kParams * 4);
        dx.reserve((uint64_t)n * kIn * 4);
        dl.reserve((uint64_t)n * 4);
        dg.reserve((uint64_t)batch * kParams * 4);
        dr.reserve(kParams * 4);
        dm1.reserve(kParams * 4);
        dm2.reserve(kParams * 4);
        didx.reserve((uint64_t)n * 4);
        dloss.reserve((uint64_t)n * 4);
        h2d(h, dp.p, params, kParams * 4);
        h2d(h, dx.p, feats, (uint64_t)n * kIn * 4);
        h2d(h, dl.p, labels, (uint64_t)n * 4);
        HIPC(hipMemsetAsync(dm1.p, 0, kParams * 4, h->stream));
        HIPC(hipMemsetAsync(dm2.p, 0, kParams * 4, h->stream));
        std::vector<uint32_t> order(n);
        for (uint32_t i = 0; i < n; i++) order[i] = i;
        std::mt19937_64 rng(seed);
        const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
        uint64_t t = 0;
        std::vector<float> ls(n);
        for (int ep = 0; ep < epochs; ep++) {
            std::shuffle(order.begin(), order.end(), rng);
            h2d(h, didx.p, order.data(), (uint64_t)n * 4);
            for (uint32_t b0 = 0; b0 + batch <= n; b0 += batch) {
                t++;
                float c1 = 1.f - std::pow(b1, (float)t), c2 = 1.f - std::pow(b2, (float)t);
                hipLaunchKernelGGL(k_mlp_fwd_bwd, dim3(batch), dim3(64), 0, h->stream, dp.as<float>(),
                                   dx.as<float>(), didx.as<uint32_t>() + b0, dl.as<uint32_t>(), batch,
                                   nullptr, dg.as<float>(), dloss.as<float>() + b0);
                hipLaunchKernelGGL(k_mlp_reduce, dim3(ceil_div(kParams, 256)), dim3(256), 0, h->stream,
                                   dg.as<float>(), batch, dr.as<float>());
                hipLaunchKernelGGL(k_mlp_adam, dim3(ceil_div(kParams, 256)), dim3(256), 0, h->stream,
                                   dp.as<float>(), dr.as<float>(), dm1.as<float>(), dm2.as<float>(), lr,
                                   b1, b2, eps, c1, c2, weight_decay);
            }
            HIPC(hipGetLastError());
            if (epoch_loss) {
                uint32_t used = (n / batch) * batch;
                d2h(h, ls.data(), dloss.p, (uint64_t)used * 4);
                double a = 0;
                for (uint32_t i = 0; i < used; i++) a += ls[i];
                epoch_loss[ep] = used ? (float)(a / used) : 0.f;
            }
        }
        d2h(h, params, dp.p, kParams * 4);
        DevBuf *all[] = {&dp, &dx, &dl, &dg, &dr, &dm1, &dm2, &didx, &dloss};
        for (DevBuf *b : all) b->release();
    });
}

}  // extern "C"
