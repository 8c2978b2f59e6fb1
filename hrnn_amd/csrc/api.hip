// api.hip -- C-ABI entry points of libhrnn_hip.so (declared in include/hrnn_dev.h).
#include <algorithm>
#include <cfloat>
#include <cstring>
#include <thread>
#include <vector>

#include "handle.h"

using namespace hrnn;

hrnn_dev::~hrnn_dev() {
    (void)hipSetDevice(device);
    DevBuf *bufs[] = {&x, &xb, &xh, &xu8, &norms, &bf_norm16, &knn_ids, &knn_d, &knn_cnt, &knn_new, &rk_col, &csr_off,
                      &csr_ent, &search_mark, &search_cand, &search_mark2, &search_cand2, &fetch_off, &fetch_ids, &l0, &up_off, &up_pool, &beam_vis, &levels_dev, &hb_prop, &hb_inc_cnt, &hb_inc,
                      &hb_touched, &hb_ctr,
                      &s0, &s1, &s2, &s3, &s4, &s5, &s6, &s7, &s8, &nnd_aux, &counters};
    for (DevBuf *b : bufs) b->release();
    if (pin_small) (void)hipHostFree(pin_small);
    if (ev_a) (void)hipEventDestroy(ev_a);
    if (ev_b) (void)hipEventDestroy(ev_b);
    for (int b = 0; b < 2; b++) {
        if (pin_ev[b]) (void)hipEventDestroy(pin_ev[b]);
        if (pin[b]) (void)hipHostFree(pin[b]);
    }
    if (stream2) (void)hipStreamDestroy(stream2);
    if (stream) (void)hipStreamDestroy(stream);
}

namespace hrnn {

// Large copies between caller (pageable) memory and the device go through
// two pinned staging buffers: the host-side copy of one chunk (split over a
// few threads) overlaps the DMA of the other, and the DMA runs from pinned
// memory at the link's rate.
static constexpr size_t kStageBytes = 64ull << 20;

static void par_memcpy(void *dst, const void *src, size_t n) {
    const unsigned T = std::min(8u, std::max(1u, std::thread::hardware_concurrency()));
    if (n < (8ull << 20) || T == 1) {
        std::memcpy(dst, src, n);
        return;
    }
    const size_t per = (n + T - 1) / T;
    std::vector<std::thread> th;
    th.reserve(T);
    for (unsigned t = 0; t < T; t++) {
        const size_t o = (size_t)t * per;
        if (o >= n) break;
        const size_t l = std::min(per, n - o);
        auto part = [=] { std::memcpy(static_cast<char *>(dst) + o, static_cast<const char *>(src) + o, l); };
        try {
            th.emplace_back(part);
        } catch (...) {   // no thread to spare: copy this part here
            part();
        }
    }
    for (auto &x : th) x.join();
}

static void staged_copy(hrnn_dev_t h, void *dst, const void *src, size_t bytes, bool to_dev) {
    for (int b = 0; b < 2; b++)
        if (!h->pin[b]) {
            HIPC(hipHostMalloc(&h->pin[b], kStageBytes, hipHostMallocDefault));
            HIPC(hipEventCreateWithFlags(&h->pin_ev[b], hipEventDisableTiming));
        }
    const size_t nch = (bytes + kStageBytes - 1) / kStageBytes;
    char *d = static_cast<char *>(dst);
    const char *s = static_cast<const char *>(src);
    if (to_dev) {
        for (size_t c = 0; c < nch; c++) {
            const int b = (int)(c & 1);
            const size_t off = c * kStageBytes, len = std::min(kStageBytes, bytes - off);
            if (c >= 2) HIPC(hipEventSynchronize(h->pin_ev[b]));   // chunk c - 2 has left this buffer
            par_memcpy(h->pin[b], s + off, len);
            HIPC(hipMemcpyAsync(d + off, h->pin[b], len, hipMemcpyHostToDevice, h->stream));
            HIPC(hipEventRecord(h->pin_ev[b], h->stream));
        }
        HIPC(hipStreamSynchronize(h->stream));
        return;
    }
    for (size_t c = 0; c <= nch; c++) {   // DMA chunk c while chunk c - 1 is copied out
        if (c < nch) {
            const size_t off = c * kStageBytes, len = std::min(kStageBytes, bytes - off);
            HIPC(hipMemcpyAsync(h->pin[c & 1], s + off, len, hipMemcpyDeviceToHost, h->stream));
            HIPC(hipEventRecord(h->pin_ev[c & 1], h->stream));
        }
        if (c >= 1) {
            const size_t off = (c - 1) * kStageBytes, len = std::min(kStageBytes, bytes - off);
            HIPC(hipEventSynchronize(h->pin_ev[(c - 1) & 1]));
            par_memcpy(d + off, h->pin[(c - 1) & 1], len);
        }
    }
}

void h2d(hrnn_dev_t h, void *dst, const void *src, size_t bytes) {
    if (!bytes) return;
    if (bytes >= 2 * kStageBytes) {
        staged_copy(h, dst, src, bytes, true);
        return;
    }
    HIPC(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, h->stream));
    HIPC(hipStreamSynchronize(h->stream));
}

void d2h(hrnn_dev_t h, void *dst, const void *src, size_t bytes) {
    if (!bytes) return;
    if (bytes >= 2 * kStageBytes) {
        staged_copy(h, dst, src, bytes, false);
        return;
    }
    HIPC(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, h->stream));
    HIPC(hipStreamSynchronize(h->stream));
}

void check_device_error(hrnn_dev_t h) {
    uint32_t *w = h->counters.as<uint32_t>() + kErrWord;
    uint32_t e = 0;
    d2h(h, &e, w, 4);
    if (e) {
        HIPC(hipMemsetAsync(w, 0, 4, h->stream));
        HIPC(hipStreamSynchronize(h->stream));
        throw Error{HRNN_E_STATE, "a kernel met invalid graph data (ids or counts out of range); "
                                  "results of the last launches are not valid"};
    }
}

// BF16/FP16 mirrors, FP32 row norms (sum of squares, sequential FMA order) and
// the value-range flags for rows [row0, row0 + rows): flags bit 0 = some value
// is not an integer in [-256, 256], bit 1 = some value is not one in [0, 255].
__global__ void k_base_mirrors(const float *__restrict__ x, __hip_bfloat16 *__restrict__ xb,
                               __half *__restrict__ xh, uint8_t *__restrict__ xu8,
                               float *__restrict__ norms, uint64_t row0, uint64_t rows, uint32_t d,
                               unsigned int *__restrict__ flags) {
    uint64_t r = row0 + blockIdx.x * (uint64_t)(blockDim.x / 64) + (threadIdx.x / 64);
    if (r >= row0 + rows) return;
    uint32_t lane = threadIdx.x & 63u;
    const float *xr = x + r * d;
    float acc = 0.f;
    bool bad = false, not_u8 = false;
    for (uint32_t e = lane; e < d; e += 64) {
        float v = xr[e];
        xb[r * d + e] = __float2bfloat16(v);
        xh[r * d + e] = __float2half(v);
        xu8[r * d + e] = (uint8_t)fminf(fmaxf(rintf(v), 0.f), 255.f);   // exact when base_u8
        acc = __builtin_fmaf(v, v, acc);
        const bool integral = v == rintf(v);
        bad |= !(integral && fabsf(v) <= 256.f);
        not_u8 |= !(integral && v >= 0.f && v <= 255.f);
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
    if (lane == 0) norms[r] = acc;
    const unsigned f = (__ballot(bad) ? 1u : 0u) | (__ballot(not_u8) ? 2u : 0u);
    if (lane == 0 && f) atomicOr(flags, f);
}

void launch_base_mirrors(hrnn_dev_t h, uint64_t row0, uint64_t rows) {
    if (!rows) return;
    h->counters.reserve(4096);
    unsigned int *flag = h->counters.as<unsigned int>();
    HIPC(hipMemsetAsync(flag, 0, sizeof(unsigned int), h->stream));
    unsigned blocks = ceil_div(rows, 4);
    hipLaunchKernelGGL(k_base_mirrors, dim3(blocks), dim3(256), 0, h->stream, h->x.as<float>(),
                       h->xb.as<__hip_bfloat16>(), h->xh.as<__half>(), h->xu8.as<uint8_t>(),
                       h->norms.as<float>(),
                       row0, rows, h->d, flag);
    HIPC(hipGetLastError());
    unsigned int f = 0;
    d2h(h, &f, flag, sizeof(f));
    h->small_int = (row0 == 0) ? ((f & 1u) == 0) : (h->small_int && (f & 1u) == 0);
    h->base_u8 = (row0 == 0) ? ((f & 2u) == 0) : (h->base_u8 && (f & 2u) == 0);
}

// Flags (*bad = 1) any listed neighbour id >= n (entries j < cnt[row]).
__global__ void k_check_knng(const uint32_t *__restrict__ ids, const uint32_t *__restrict__ cnt, uint64_t n,
                             uint32_t K, uint32_t *__restrict__ bad) {
    const uint64_t row = (uint64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
    if (row >= n) return;
    const uint32_t c = cnt[row];
    bool b = false;
    for (uint32_t j = threadIdx.x & 63u; j < c; j += 64) b |= ids[row * K + j] >= n;
    if (__ballot(b) && (threadIdx.x & 63u) == 0) atomicOr(bad, 1u);
}

__global__ void k_l2_pairs(const float *__restrict__ a, const float *__restrict__ x,
                           const uint32_t *__restrict__ ia, const uint32_t *__restrict__ ib,
                           uint64_t npairs, uint32_t d, float *__restrict__ out) {
    uint64_t grp = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / 16;
    int g = threadIdx.x & 15;
    if (grp >= npairs) return;   // uniform across the 16-lane group
    const float *pa = a + (uint64_t)ia[grp] * d;
    const float *pb = x + (uint64_t)ib[grp] * d;
    float r;
    if ((d & 15u) == 0) {
        r = l2_ref16_group(pa, pb, d, g);
    } else {
        r = l2_seq(pa, pb, d);
    }
    if (g == 0) out[grp] = r;
}

void l2_pairs(hrnn_dev_t h, const float *a_dev, const uint32_t *ia, const uint32_t *ib,
              uint64_t npairs, float *out) {
    if (!npairs) return;
    hipLaunchKernelGGL(k_l2_pairs, dim3(ceil_div(npairs * 16, 256)), dim3(256), 0, h->stream,
                       a_dev, h->x.as<float>(), ia, ib, npairs, h->d, out);
    HIPC(hipGetLastError());
}

}  // namespace hrnn

extern "C" {

int hrnn_dev_device_count(int *count) {
    if (!count) return HRNN_E_ARG;
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    if (e != hipSuccess) {
        *count = 0;
        return HRNN_E_HIP;
    }
    *count = c;
    return HRNN_OK;
}

int hrnn_dev_open(int device, hrnn_dev_t *out) {
    if (!out) return HRNN_E_ARG;
    *out = nullptr;
    int cnt = 0;
    if (hipGetDeviceCount(&cnt) != hipSuccess || cnt <= 0) return HRNN_E_HIP;
    if (device < 0 || device >= cnt) return HRNN_E_ARG;
    hrnn_dev_t h = new hrnn_dev();
    h->device = device;
    int rc = guarded(h, [&] {
        HIPC(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
        hipDeviceProp_t prop;
        HIPC(hipGetDeviceProperties(&prop, device));
        h->n_cu = prop.multiProcessorCount;
        h->lds_per_block = prop.sharedMemPerBlock;
        h->counters.reserve(4096);
        HIPC(hipMemsetAsync(h->counters.p, 0, 4096, h->stream));
        HIPC(hipStreamSynchronize(h->stream));
        noalloc_scratch_init(h);
    });
    if (rc != HRNN_OK) {
        delete h;
        return rc;
    }
    *out = h;
    return HRNN_OK;
}

int hrnn_dev_close(hrnn_dev_t h) {
    if (!h) return HRNN_E_ARG;
    delete h;
    return HRNN_OK;
}

const char *hrnn_dev_last_error(hrnn_dev_t h) { return h ? h->err.c_str() : "null handle"; }

int hrnn_dev_sync(hrnn_dev_t h) {
    if (!h) return HRNN_E_ARG;
    return guarded(h, [&] {
        HIPC(hipStreamSynchronize(h->stream));
        check_device_error(h);
    });
}

void *hrnn_dev_stream(hrnn_dev_t h) { return h ? (void *)h->stream : nullptr; }

static void set_base_common(hrnn_dev_t h, const float *x, uint64_t n, uint32_t d, bool device) {
    REQUIRE(x && n > 0 && d > 0, HRNN_E_ARG, "set_base: empty input");
    REQUIRE(n < 0xFFFFFFFFull, HRNN_E_ARG, "set_base: n must fit in 32-bit ids");
    h->n = n;
    h->n_cap = n;
    h->d = d;
    h->x.reserve(n * d * sizeof(float));
    h->xb.reserve(n * d * 2);
    h->xh.reserve(n * d * 2);
    h->xu8.reserve(n * d);
    h->norms.reserve(n * sizeof(float));
    if (device)
        HIPC(hipMemcpyAsync(h->x.p, x, n * d * sizeof(float), hipMemcpyDeviceToDevice, h->stream));
    else
        h2d(h, h->x.p, x, n * d * sizeof(float));
    launch_base_mirrors(h, 0, n);
    h->base_version++;
    h->K = 0;
    h->csr_valid = false;
    h->hnsw_valid = false;
    h->hnsw_rows = 0;
    h->beam_vis.release();     // sized by n
    h->search_mark.release();
    h->search_mark2.release();
}

int hrnn_dev_set_base(hrnn_dev_t h, const float *x, uint64_t n, uint32_t d) {
    if (!h) return HRNN_E_ARG;
    return guarded(h, [&] { set_base_common(h, x, n, d, false); });
}

int hrnn_dev_set_base_dptr(hrnn_dev_t h, const float *x, uint64_t n, uint32_t d) {
    if (!h) return HRNN_E_ARG;
    return guarded(h, [&] { set_base_common(h, x, n, d, true); });
}

int hrnn_dev_append_base(hrnn_dev_t h, const float *x, uint64_t n_add) {
    if (!h) return HRNN_E_ARG;
    return guarded(h, [&] {
        REQUIRE(h->n > 0 && x, HRNN_E_STATE, "append_base: no base");
        uint64_t n_new = h->n + n_add;
        REQUIRE(n_new < 0xFFFFFFFFull, HRNN_E_ARG, "append_base: n must fit in 32-bit ids");
        if (n_new > h->n_cap) {
            // grow by copying into larger buffers (geometric)
            uint64_t cap = std::max<uint64_t>(n_new, h->n_cap * 3 / 2);
            DevBuf nx, nb, nh, nu, nn;
            nx.reserve(cap * h->d * 4);
            nb.reserve(cap * h->d * 2);
            nh.reserve(cap * h->d * 2);
            nu.reserve(cap * h->d);
            nn.reserve(cap * 4);
            HIPC(hipMemcpyAsync(nx.p, h->x.p, h->n * h->d * 4, hipMemcpyDeviceToDevice, h->stream));
            HIPC(hipMemcpyAsync(nb.p, h->xb.p, h->n * h->d * 2, hipMemcpyDeviceToDevice, h->stream));
            HIPC(hipMemcpyAsync(nh.p, h->xh.p, h->n * h->d * 2, hipMemcpyDeviceToDevice, h->stream));
            HIPC(hipMemcpyAsync(nu.p, h->xu8.p, h->n * h->d, hipMemcpyDeviceToDevice, h->stream));
            HIPC(hipMemcpyAsync(nn.p, h->norms.p, h->n * 4, hipMemcpyDeviceToDevice, h->stream));
            HIPC(hipStreamSynchronize(h->stream));
            h->x.release(); h->xb.release(); h->xh.release(); h->xu8.release(); h->norms.release();
            h->x = nx; h->xb = nb; h->xh = nh; h->xu8 = nu; h->norms = nn;
            nx.p = nb.p = nh.p = nu.p = nn.p = nullptr;
            h->n_cap = cap;
        }
        HIPC(hipMemcpyAsync(h->x.as<float>() + h->n * h->d, x, n_add * h->d * 4,
                            hipMemcpyHostToDevice, h->stream));
        launch_base_mirrors(h, h->n, n_add);
        h->n = n_new;
        h->base_version++;
        h->csr_valid = false;
        h->hnsw_valid = false;
        h->beam_vis.release();
        h->search_mark.release();
        h->search_mark2.release();
    });
}

int hrnn_dev_base_is_small_int(hrnn_dev_t h, int *is_int) {
    if (!h || !is_int) return HRNN_E_ARG;
    *is_int = h->small_int;
    return HRNN_OK;
}

int hrnn_dev_base_is_u8(hrnn_dev_t h, int *is_u8) {
    if (!h || !is_u8) return HRNN_E_ARG;
    *is_u8 = h->base_u8 ? 1 : 0;
    return HRNN_OK;
}

const float *hrnn_dev_base_ptr(hrnn_dev_t h) { return h ? h->x.as<float>() : nullptr; }

static int *tuning_slot(hrnn_dev_t h, const char *key) {
    std::string k = key ? key : "";
    if (k == "search_wg_per_cu") return &h->tune_search_wg_per_cu;
    if (k == "beam_wg_per_cu") return &h->tune_beam_wg_per_cu;
    if (k == "search_lds_kib") return &h->tune_search_lds_kib;
    if (k == "stage1_cap") return &h->tune_stage1_cap;
    if (k == "hnsw_batch_permille") return &h->tune_hnsw_batch_permille;
    if (k == "hnsw_seed_exp") return &h->tune_hnsw_seed_exp;
    if (k == "insert_offer_cap") return &h->tune_insert_offer_cap;
    if (k == "insert_lds_bitmap") return &h->tune_insert_lds_bitmap;
    if (k == "beam_vis_shift") return &h->tune_beam_vis_shift;
    if (k == "bf_row_frags") return &h->tune_bf_frags;
    if (k == "nnd_join") return &h->tune_nnd_join;
    if (k == "query_streams") return &h->tune_query_streams;
    if (k == "query_split") return &h->tune_query_split;
    if (k == "csr_wide") return &h->tune_csr_wide;
    if (k == "bf_layout") return &h->tune_bf_layout;
    return nullptr;
}

int hrnn_dev_set_tuning(hrnn_dev_t h, const char *key, int value) {
    if (!h) return HRNN_E_ARG;
    int *slot = tuning_slot(h, key);
    const bool cap = slot == &h->tune_stage1_cap;
    const bool permille = slot == &h->tune_hnsw_batch_permille;
    const bool seedx = slot == &h->tune_hnsw_seed_exp;
    const bool offers = slot == &h->tune_insert_offer_cap;
    const bool flag = slot == &h->tune_insert_lds_bitmap;
    const bool vshift = slot == &h->tune_beam_vis_shift;
    const bool frags = slot == &h->tune_bf_frags;
    const bool join = slot == &h->tune_nnd_join;
    const bool qstreams = slot == &h->tune_query_streams;
    const bool qsplit = slot == &h->tune_query_split;
    const bool cwide = slot == &h->tune_csr_wide;
    const bool blay = slot == &h->tune_bf_layout;
    if (!slot || value < (vshift ? -1 : cap || seedx || offers || flag || join || cwide || blay ? 0 : 1) ||
        value > (offers ? (1 << 30) : flag || cwide || blay ? 1 : vshift || frags || qstreams ? 2 : join ? 3 : cap ? 1024
                                                                   : permille || seedx || qsplit ? 1000 : 64) ||
        (qsplit && (value < 100 || value > 900))) {
        h->err = std::string("set_tuning: unknown key or value out of range: ") + (key ? key : "");
        return HRNN_E_ARG;
    }
    *slot = value;
    // the per-workgroup private state is sized by the workgroup count
    h->search_mark.release();
    h->search_mark2.release();
    h->beam_vis.release();
    return HRNN_OK;
}

int hrnn_dev_get_tuning(hrnn_dev_t h, const char *key, int *value) {
    if (!h || !value) return HRNN_E_ARG;
    int *slot = tuning_slot(h, key);
    if (!slot) return HRNN_E_ARG;
    *value = *slot;
    return HRNN_OK;
}

int hrnn_dev_set_vector_precision(hrnn_dev_t h, int prec) {
    if (!h || (prec != HRNN_PREC_F32 && prec != HRNN_PREC_BF16 && prec != HRNN_PREC_U8)) return HRNN_E_ARG;
    h->vec_bf16 = (prec != HRNN_PREC_F32);
    h->vec_u8 = (prec == HRNN_PREC_U8);
    return HRNN_OK;
}

int hrnn_dev_set_knng(hrnn_dev_t h, uint32_t K, const uint32_t *ids, const float *dists,
                      const uint16_t *cnt) {
    if (!h) return HRNN_E_ARG;
    return guarded(h, [&] {
        REQUIRE(h->n > 0, HRNN_E_STATE, "set_knng: set_base first");
        REQUIRE(K > 0 && K <= 65535 && ids && dists && cnt, HRNN_E_ARG, "set_knng: bad args");
        uint64_t n = h->n;
        h->K = K;
        h->knn_ids.reserve(n * K * 4);
        h->knn_d.reserve(n * K * 4);
        h->knn_cnt.reserve(n * 4);
        h->knn_new.reserve(n * K);
        std::vector<uint32_t> c32(n);
        for (uint64_t i = 0; i < n; i++) {
            REQUIRE(cnt[i] <= K, HRNN_E_ARG, "set_knng: cnt > K");
            c32[i] = cnt[i];
        }
        h->csr_valid = false;
        h->knn_version++;
        h2d(h, h->knn_ids.p, ids, n * K * 4);
        h2d(h, h->knn_d.p, dists, n * K * 4);
        h2d(h, h->knn_cnt.p, c32.data(), n * 4);
        HIPC(hipMemsetAsync(h->knn_new.p, 1, n * K, h->stream));
        // every listed id < n, checked on the device (the rows are there now)
        uint32_t *bad = h->counters.as<uint32_t>() + kKnngCheckWord;
        HIPC(hipMemsetAsync(bad, 0, 4, h->stream));
        hipLaunchKernelGGL(k_check_knng, dim3(ceil_div(n, 4)), dim3(256), 0, h->stream, h->knn_ids.as<uint32_t>(),
                           h->knn_cnt.as<uint32_t>(), n, K, bad);
        HIPC(hipGetLastError());
        uint32_t flag = 0;
        d2h(h, &flag, bad, 4);
        if (flag) h->K = 0;   // no usable graph on the device
        REQUIRE(!flag, HRNN_E_ARG, "set_knng: neighbour id >= n");
    });
}

int hrnn_dev_get_knng(hrnn_dev_t h, uint32_t *ids, float *dists, uint16_t *cnt) {
    if (!h) return HRNN_E_ARG;
    return guarded(h, [&] {
        REQUIRE(h->K > 0, HRNN_E_STATE, "get_knng: no graph");
        uint64_t n = h->n;
        uint32_t K = h->K;
        std::vector<uint32_t> c32(n);
        if (ids) d2h(h, ids, h->knn_ids.p, n * K * 4);
        if (dists) d2h(h, dists, h->knn_d.p, n * K * 4);
        d2h(h, c32.data(), h->knn_cnt.p, n * 4);
        if (cnt)
            for (uint64_t i = 0; i < n; i++) cnt[i] = (uint16_t)c32[i];
    });
}

int hrnn_dev_get_knng_flags(hrnn_dev_t h, uint8_t *flags) {
    if (!h) return HRNN_E_ARG;
    return guarded(h, [&] {
        REQUIRE(h->K > 0, HRNN_E_STATE, "get_knng_flags: no graph");
        REQUIRE(flags, HRNN_E_ARG, "get_knng_flags: null argument");
        d2h(h, flags, h->knn_new.p, h->n * h->K);
    });
}

int hrnn_dev_set_knng_flags(hrnn_dev_t h, const uint8_t *flags) {
    if (!h) return HRNN_E_ARG;
    return guarded(h, [&] {
        REQUIRE(h->K > 0, HRNN_E_STATE, "set_knng_flags: no graph");
        if (flags) {
            h2d(h, h->knn_new.p, flags, h->n * h->K);
        } else {
            HIPC(hipMemsetAsync(h->knn_new.p, 1, h->n * h->K, h->stream));
            HIPC(hipStreamSynchronize(h->stream));
        }
    });
}

int hrnn_dev_knng_ptrs(hrnn_dev_t h, uint32_t **ids, float **dists, uint32_t **cnt) {
    if (!h) return HRNN_E_ARG;
    h->knn_version++;   // the caller may write through these pointers
    if (ids) *ids = h->knn_ids.as<uint32_t>();
    if (dists) *dists = h->knn_d.as<float>();
    if (cnt) *cnt = h->knn_cnt.as<uint32_t>();
    return h->K ? HRNN_OK : HRNN_E_STATE;
}

int hrnn_dev_nndescent(hrnn_dev_t h, int num_iters, int sample_size, uint32_t bucket,
                       hrnn_nnd_iter_stats *stats) {
    if (!h) return HRNN_E_ARG;
    h->knn_version++;
    return guarded(h, [&] {
        REQUIRE(h->K > 0, HRNN_E_STATE, "nndescent: set_knng first");
        REQUIRE(num_iters >= 0 && sample_size > 0 && sample_size <= 64, HRNN_E_ARG,
                "nndescent: sample_size must be in [1, 64]");
        nndescent_run(h, num_iters, sample_size, bucket, stats);
        h->csr_valid = false;
    });
}

int hrnn_dev_build_rknn_csr(hrnn_dev_t h, uint32_t *rank_bits, int *wide, uint64_t *n_entries) {
    if (!h) return HRNN_E_ARG;
    return guarded(h, [&] {
        REQUIRE(h->K > 0, HRNN_E_STATE, "build_rknn_csr: no KNN graph");
        csr_build(h);
        if (rank_bits) *rank_bits = h->rank_bits;
        if (wide) *wide = h->wide;
        if (n_entries) *n_entries = h->n_ent;
    });
}

int hrnn_dev_get_rknn_csr(hrnn_dev_t h, uint64_t *offsets, void *entries) {
    if (!h) return HRNN_E_ARG;
    return guarded(h, [&] {
        REQUIRE(h->csr_valid, HRNN_E_STATE, "get_rknn_csr: no CSR");
        if (offsets) d2h(h, offsets, h->csr_off.p, (h->n + 1) * 8);
        if (entries) d2h(h, entries, h->csr_ent.p, h->n_ent * (h->wide ? 8 : 4));
    });
}

int hrnn_dev_set_rknn_csr(hrnn_dev_t h, uint32_t rank_bits, int wide, uint64_t n,
                          const uint64_t *offsets, const void *entries, uint64_t n_entries) {
    if (!h) return HRNN_E_ARG;
    return guarded(h, [&] {
        REQUIRE(n == h->n && offsets, HRNN_E_ARG, "set_rknn_csr: n mismatch");
        REQUIRE(rank_bits < 32, HRNN_E_ARG, "set_rknn_csr: rank_bits");
        h->csr_off.reserve((n + 1) * 8);
        h->csr_ent.reserve(std::max<uint64_t>(1, n_entries) * (wide ? 8 : 4));
        h2d(h, h->csr_off.p, offsets, (n + 1) * 8);
        if (n_entries) h2d(h, h->csr_ent.p, entries, n_entries * (wide ? 8 : 4));
        h->rank_bits = rank_bits;
        h->wide = wide ? 1 : 0;
        h->n_ent = n_entries;
        h->max_deg = max_list_len(h);
        h->csr_valid = true;
    });
}

int hrnn_dev_rknn_search_dptr(hrnn_dev_t h, const float *q_dev, uint32_t nq,
                              const uint32_t *proxies_dev, uint32_t m,
                              const uint32_t *k_per_q_dev, uint32_t k, uint32_t K_prime,
                              float threshold, uint32_t *res_cnt_dev, uint32_t *res_ids_dev,
                              uint32_t cap_per_q, uint32_t *work_dev) {
    if (!h) return HRNN_E_ARG;
    return guarded(h, [&] {
        REQUIRE(h->csr_valid, HRNN_E_STATE, "rknn_search: build the CSR first");
        REQUIRE(h->K > 0, HRNN_E_STATE, "rknn_search: no KNN graph");
        REQUIRE(k_per_q_dev || (k >= 1 && k <= h->K), HRNN_E_ARG, "rknn_search: k out of [1, K]");
        REQUIRE(m >= 1 && m <= 1024, HRNN_E_ARG, "rknn_search: m out of [1, 1024]");
        if (nq == 0) return;
        rknn_search_resident(h, q_dev, nq, proxies_dev, m, k_per_q_dev, k, K_prime, threshold,
                             res_cnt_dev, res_ids_dev, cap_per_q, work_dev);
    });
}

int hrnn_dev_rknn_search(hrnn_dev_t h, const float *q, uint32_t nq, const uint32_t *proxies,
                         uint32_t m, const uint32_t *k_per_q, uint32_t k, uint32_t K_prime,
                         float threshold, uint64_t *res_off, uint32_t *res_ids, uint64_t cap,
                         uint64_t *needed) {
    if (!h) return HRNN_E_ARG;
    return guarded(h, [&] {
        REQUIRE(h->csr_valid, HRNN_E_STATE, "rknn_search: build the CSR first");
        REQUIRE(res_off && q && proxies, HRNN_E_ARG, "rknn_search: null argument");
        REQUIRE(m >= 1 && m <= 1024, HRNN_E_ARG, "rknn_search: m out of [1, 1024]");
        if (k_per_q) {
            for (uint32_t i = 0; i < nq; i++)
                REQUIRE(k_per_q[i] >= 1 && k_per_q[i] <= h->K, HRNN_E_ARG, "rknn_search: k out of [1, K]");
        } else {
            REQUIRE(k >= 1 && k <= h->K, HRNN_E_ARG, "rknn_search: k out of [1, K]");
        }
        for (uint64_t i = 0; i < (uint64_t)nq * m; i++)
            REQUIRE(proxies[i] < h->n || proxies[i] == kInvalidId, HRNN_E_ARG,
                    "rknn_search: proxy id >= n");
        res_off[0] = 0;
        if (nq == 0) { if (needed) *needed = 0; return; }
        // a query's results are distinct candidates: at most m reverse-list
        // prefixes, each no longer than the longest reverse list (a proxy's
        // in-degree, which K' does not bound)
        const uint64_t worst = std::min<uint64_t>((uint64_t)m * std::max<uint64_t>(h->max_deg, 1), h->n);
        uint32_t cap_q = (uint32_t)std::min<uint64_t>(worst, 4096);
        const uint32_t chunk = 16384;
        uint32_t qc = std::min(nq, chunk);
        // device staging: queries, proxies, k, counts, ids
        h->s5.reserve((uint64_t)qc * h->d * 4);
        h->s6.reserve((uint64_t)qc * m * 4 + (uint64_t)qc * 4);
        h->s7.reserve((uint64_t)qc * 4);
        h->s8.reserve((uint64_t)qc * cap_q * 4);
        std::vector<uint32_t> cnt(nq);
        std::vector<std::vector<uint32_t>> big;   // overflowed queries: full results
        std::vector<uint32_t> ids_all;            // chunk results (cap_q per query)
        std::vector<uint32_t> chunk_ids;
        uint64_t total = 0;
        std::vector<uint64_t> per_q_off(nq + 1, 0);
        std::vector<uint32_t> out;
        out.reserve(1024);
        for (uint32_t q0 = 0; q0 < nq; q0 += chunk) {
            uint32_t nqc = std::min(chunk, nq - q0);
            float *qd = h->s5.as<float>();
            uint32_t *pd = h->s6.as<uint32_t>();
            uint32_t *kd = pd + (uint64_t)nqc * m;
            uint32_t *cd = h->s7.as<uint32_t>();
            uint32_t *id = h->s8.as<uint32_t>();
            HIPC(hipMemcpyAsync(qd, q + (uint64_t)q0 * h->d, (uint64_t)nqc * h->d * 4,
                                hipMemcpyHostToDevice, h->stream));
            HIPC(hipMemcpyAsync(pd, proxies + (uint64_t)q0 * m, (uint64_t)nqc * m * 4,
                                hipMemcpyHostToDevice, h->stream));
            if (k_per_q)
                HIPC(hipMemcpyAsync(kd, k_per_q + q0, (uint64_t)nqc * 4, hipMemcpyHostToDevice, h->stream));
            rknn_search_resident(h, qd, nqc, pd, m, k_per_q ? kd : nullptr, k, K_prime, threshold,
                                 cd, id, cap_q);
            chunk_ids.resize((uint64_t)nqc * cap_q);
            HIPC(hipMemcpyAsync(cnt.data() + q0, cd, (uint64_t)nqc * 4, hipMemcpyDeviceToHost, h->stream));
            HIPC(hipMemcpyAsync(chunk_ids.data(), id, (uint64_t)nqc * cap_q * 4, hipMemcpyDeviceToHost,
                                h->stream));
            HIPC(hipStreamSynchronize(h->stream));
            for (uint32_t i = 0; i < nqc; i++) {
                uint32_t c = cnt[q0 + i];
                if (c <= cap_q) {
                    out.insert(out.end(), chunk_ids.begin() + (uint64_t)i * cap_q,
                               chunk_ids.begin() + (uint64_t)i * cap_q + c);
                } else {
                    // rerun this query alone with the worst-case capacity
                    DevBuf tmp;
                    tmp.reserve(worst * 4 + 64);
                    uint32_t *tc = tmp.as<uint32_t>();
                    uint32_t *ti = tc + 16;
                    rknn_search_resident(h, qd + (uint64_t)i * h->d, 1, pd + (uint64_t)i * m, m,
                                         k_per_q ? kd + i : nullptr, k, K_prime, threshold, tc, ti,
                                         (uint32_t)worst);
                    uint32_t c2 = 0;
                    d2h(h, &c2, tc, 4);
                    REQUIRE(c2 <= worst, HRNN_E_STATE, "rknn_search: result count above its bound");
                    std::vector<uint32_t> v(c2);
                    d2h(h, v.data(), ti, (uint64_t)c2 * 4);
                    out.insert(out.end(), v.begin(), v.end());
                    tmp.release();
                }
                per_q_off[q0 + i + 1] = out.size();
            }
        }
        total = out.size();
        for (uint32_t i = 0; i <= nq; i++) res_off[i] = per_q_off[i];
        if (needed) *needed = total;
        if (total > cap || (total && !res_ids)) throw Error{HRNN_E_CAP, "rknn_search: result capacity too small"};
        if (total) std::memcpy(res_ids, out.data(), total * 4);
    });
}

int hrnn_dev_l2sqr_pairs(hrnn_dev_t h, const float *a, uint64_t na, const uint32_t *ia,
                         const uint32_t *ib, uint64_t npairs, float *out) {
    if (!h) return HRNN_E_ARG;
    return guarded(h, [&] {
        REQUIRE(h->n > 0, HRNN_E_STATE, "l2sqr_pairs: set_base first");
        REQUIRE(a && ia && ib && out, HRNN_E_ARG, "l2sqr_pairs: null argument");
        if (!npairs) return;
        for (uint64_t i = 0; i < npairs; i++) {
            REQUIRE(ia[i] < na && ib[i] < h->n, HRNN_E_ARG, "l2sqr_pairs: index out of range");
        }
        DevBuf da, di, dj, dout;
        da.reserve(na * h->d * 4);
        di.reserve(npairs * 4);
        dj.reserve(npairs * 4);
        dout.reserve(npairs * 4);
        h2d(h, da.p, a, na * h->d * 4);
        h2d(h, di.p, ia, npairs * 4);
        h2d(h, dj.p, ib, npairs * 4);
        l2_pairs(h, da.as<float>(), di.as<uint32_t>(), dj.as<uint32_t>(), npairs, dout.as<float>());
        d2h(h, out, dout.p, npairs * 4);
        da.release(); di.release(); dj.release(); dout.release();
    });
}

}  // extern "C"
