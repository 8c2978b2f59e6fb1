// handle.h -- the per-index device state behind an hrnn_dev_t.
#pragma once
#include <functional>

#include "common.h"

struct hrnn_dev {
    int device = 0;
    hipStream_t stream = nullptr;
    // second stream (default priority) and two events for work that overlaps
    // within one call; created on first use
    hipStream_t stream2 = nullptr;
    hipEvent_t ev_a = nullptr, ev_b = nullptr;
    // two pinned host staging buffers (and their events) for large host <->
    // device copies; allocated on first use
    void *pin[2] = {nullptr, nullptr};
    hipEvent_t pin_ev[2] = {nullptr, nullptr};
    std::string err;
    int n_cu = 0;
    size_t lds_per_block = 0;

    // base vectors (row-major n x d): FP32, BF16 and FP16 mirrors, FP32 norms
    uint64_t n = 0;
    uint64_t n_cap = 0;
    uint32_t d = 0;
    hrnn::DevBuf x, xb, xh, xu8, norms;
    int small_int = 0;   // every value an integer in [-256, 256] (exact in BF16/FP16)
    // every value an integer in [0, 255]: with d <= 128 every sum formed by the
    // distance kernels (direct difference or Gram form, any order) is an
    // integer below 2^24, so all of them are exact
    int base_u8 = 0;
    // read vectors from the BF16 mirror where that is exact (small_int), or
    // from the u8 mirror where that is (base_u8; HRNN_PREC_U8)
    bool vec_bf16 = true;
    bool vec_u8 = false;
    // changes whenever the base vectors change (set_base, append_base)
    uint64_t base_version = 1;
    // brute force: squared norms of the BF16- and FP16-rounded base rows
    // [2n floats], then the largest rounded norm and the largest rounding
    // error norm ||round(x) - x|| over the base for each format (4 floats, as
    // the bits of non-negative floats); valid for bf_prep_version
    hrnn::DevBuf bf_norm16;
    uint64_t bf_prep_version = 0;

    // launch shapes chosen by the autotuner (hrnn_dev_set_tuning)
    int tune_search_wg_per_cu = 8;   // persistent RkNN search workgroups per CU
    int tune_search_lds_kib = 16;    // LDS dedup table of the RkNN search, KiB
    int tune_beam_wg_per_cu = 16;    // persistent beam workgroups (one wave each) per CU
    int tune_beam_vis_shift = 0;     // phase A visited-set slots: log2 shift from the ef rule (-1 .. 2)
    int tune_bf_frags = 1;           // brute force: 16-row fragments per wave (2: 32-wide lists, d <= 256)
    int tune_bf_layout = 0;          // brute force: waves split a tile's rows (0) or its columns (1, d <= 128)
    int tune_query_streams = 1;      // resident query batches: halves on two streams (2) or one (1)
    int tune_query_split = 500;      // with two streams: the first part's share of the batch, permille
    int tune_csr_wide = 0;           // reverse CSR: u64 entries even where u32 would do (1)
    int tune_nnd_join = 0;           // NNDescent join staging (0: chosen from d and data; 1 chunked BF16,
                                     // 2 whole-sample BF16, 3 whole-sample int8)
    int tune_stage1_cap = 0;         // learned early stop: stage-1 expansion cap (0: none)
    int tune_hnsw_batch_permille = 100;  // HNSW construction: batch size / graph size, per mille
    int tune_hnsw_seed_exp = 32;         // HNSW construction: expansions of the seed completion
    int tune_insert_offer_cap = 0;       // incremental insertion: first offer buffer, entries (0: max(2^22, 24n))
    int tune_insert_lds_bitmap = 1;      // incremental insertion: visited bitmap in LDS when it fits (0: in HBM)

    // ranked KNN graph (SoA); knn_version changes whenever it may have changed
    uint32_t K = 0;
    hrnn::DevBuf knn_ids, knn_d, knn_cnt, knn_new;
    uint64_t knn_version = 1;
    // r_k column of the graph for one k (verification reads one float per
    // candidate from a 4n-byte array instead of a line of the n x K matrix)
    hrnn::DevBuf rk_col;
    hrnn::DevBuf q_stage;   // device copy of a query batch read from page-locked host memory
    uint32_t rk_col_k = 0;
    uint64_t rk_col_version = 0, rk_col_n = 0;

    // reverse CSR
    bool csr_valid = false;
    uint32_t rank_bits = 0;
    int wide = 0;
    uint64_t n_ent = 0;
    uint64_t max_deg = 0;   // longest reverse list
    hrnn::DevBuf csr_off, csr_ent;

    // RkNN search private state: per-workgroup visited bitmaps / u16 counts
    // (all-zero between launches) and candidate buffers
    hrnn::DevBuf search_mark, search_cand;
    hrnn::DevBuf search_mark2, search_cand2;   // a second set for an overlapping search
    // result download (hrnn_dev_fetch_results): offsets (+ total and largest
    // count), the densely packed ids, and a small pinned word for the count
    hrnn::DevBuf fetch_off, fetch_ids;
    void *pin_small = nullptr;

    // HNSW graph for phase A on the GPU: layer-0 links [n x (maxM0+1)] (count
    // word + ids), per-node offset into the upper-layer pool (blocks of maxM+1
    // words per level, 0xFFFFFFFF for level-0-only nodes)
    bool hnsw_valid = false;
    // rows the resident graph links (appending base rows leaves the graph of
    // the earlier rows in place for an incremental insertion)
    uint64_t hnsw_rows = 0;
    uint32_t l0_words = 0, up_words = 0, entry = 0;
    uint64_t up_pool_words = 0;
    int maxlevel = 0;
    hrnn::DevBuf l0, up_off, up_pool, beam_vis;
    // HNSW construction: node levels, and per batch the proposed links back,
    // each node's incoming links, the touched nodes and counters
    hrnn::DevBuf levels_dev, hb_prop, hb_inc_cnt, hb_inc, hb_touched, hb_ctr;

    // scratch shared by the build kernels
    hrnn::DevBuf s0, s1, s2, s3, s4, s5, s6, s7, s8;
    hrnn::DevBuf nnd_aux;   // NNDescent proposal offsets, scan partials, total
    // small device counters, by byte offset: 0 base-int flag, 32 long-list
    // count, 128 CSR total, 192 longest list, 256 search queue, 384 device
    // error word (set by a kernel that met invalid graph data; read and
    // cleared by hrnn_dev_sync), 512 second search queue, 640 the two list
    // counts of the learned query path
    hrnn::DevBuf counters;

    ~hrnn_dev();
};

namespace hrnn {

template <class F>
int guarded(hrnn_dev_t h, F &&f) {
    try {
        if (h) HIPC(hipSetDevice(h->device));
        f();
        return HRNN_OK;
    } catch (Error &e) {
        if (h) h->err = e.msg;
        (void)hipGetLastError();   // a failed HIP call must not surface in a later one
        return e.code;
    } catch (std::exception &e) {
        if (h) h->err = e.what();
        (void)hipGetLastError();
        return HRNN_E_ARG;
    }
}

// Scoped host <-> device staging helpers (synchronous on the handle stream).
void h2d(hrnn_dev_t h, void *dst, const void *src, size_t bytes);
void d2h(hrnn_dev_t h, void *dst, const void *src, size_t bytes);
// Read the device error word (synchronous); throws HRNN_E_STATE and clears
// it when a kernel reported invalid data.
void check_device_error(hrnn_dev_t h);

// Launchers defined in the kernel files.
void launch_base_mirrors(hrnn_dev_t h, uint64_t row0, uint64_t rows);
void csr_build(hrnn_dev_t h);
// Exclusive scan of n u32 counts into off[n + 1] (u64; off[n] = *total = the
// sum); bsum holds scan_block_count(n) u64.  Kernels only (capturable).
uint32_t scan_block_count(uint64_t n);
void scan_u32(hrnn_dev_t h, const uint32_t *cnt, uint64_t n, uint64_t *off, uint64_t *bsum,
              uint64_t *total);
uint64_t max_list_len(hrnn_dev_t h);
constexpr uint32_t kQueueWord = 64;        // search work queue (u32 index in counters)
constexpr uint32_t kQueue2Word = 128;      // second search queue (overlapping launches)
constexpr uint32_t kListCountWord = 160;   // two list counts of the learned query path
constexpr uint32_t kKnngCheckWord = 168;   // set_knng's device-side id check flag
constexpr uint32_t kNndActiveWord = 172;   // NNDescent: the current iteration has new entries
// Optional query subset and stream of a search launch: queries qlist[0 ..
// *qcount) (device memory) instead of 0 .. nq-1, on `stream` with its work
// queue at counters[queue_word].
struct SearchSel {
    const uint32_t *qlist = nullptr;
    const uint32_t *qcount = nullptr;
    hipStream_t stream = nullptr;
    uint32_t queue_word = kQueueWord;
    bool queue_zeroed = false;   // an earlier kernel on the stream already zeroed the queue
    int buffer_set = 0;          // 1: the second fallback buffers (a search overlapping another)
    uint32_t max_slots = 0;      // cap on the persistent workgroups (0: the tuned count)
    // caller-free form: the queue word and the fallback bitmaps / candidate
    // buffers carved from this device memory instead of the handle's buffers
    // (nothing is allocated, nothing is cleared with a memset; the r_k column
    // is used only when it is already built for this k)
    uint8_t *scratch = nullptr;
    size_t scratch_bytes = 0;
    // bytes from the start of the bitmap area known to be zero: the launch
    // clears the part of its bitmaps beyond it (earlier calls may have left
    // candidate ids there) and sets it to its own bitmaps' size (the kernel
    // leaves its bitmaps zero; the candidate buffers after them are dirty)
    size_t *scratch_clean = nullptr;
};

void rknn_search_resident(hrnn_dev_t h, const float *q, uint32_t nq, const uint32_t *proxies,
                          uint32_t m, const uint32_t *k_per_q, uint32_t k, uint32_t K_prime,
                          float threshold, uint32_t *res_cnt, uint32_t *res_ids,
                          uint32_t cap_per_q, uint32_t *work = nullptr,
                          const SearchSel *sel = nullptr);
void nndescent_run(hrnn_dev_t h, int iters, int S, uint32_t bucket, hrnn_nnd_iter_stats *st);
// Top-K merge of per-node proposals into the resident ranked lists (the
// NNDescent merge): node u takes prop_cnt[u] keys (order-preserving bits of
// d2 << 32 | id) from prop_key[prop_off[u] ..), keys at or past cap dropped;
// stats: the NNDescent counter block (updates counted).
void nnd_merge_launch(hrnn_dev_t h, uint64_t n, const uint32_t *prop_cnt, const uint64_t *prop_off,
                      const unsigned long long *prop_key, uint64_t cap, unsigned long long *stats);
// NNDescent counter block size (u64 words) and the updates it counted.
uint64_t nnd_stat_words();
uint64_t nnd_stat_updates(const unsigned long long *host_block);
// Batched HNSW insertion with KNN seeding of the resident rows [b_first, n):
// b_first = 0 builds the graph from scratch; otherwise the resident graph
// must link exactly the rows [0, b_first) and its buffers grow (the KNN lists
// too, new rows empty).  levels[n] for every row.  after(b, b1) runs on the
// stream after each batch's links and seeds are in place.
void hnsw_insert_rows(hrnn_dev_t h, const int *levels, uint64_t b_first, uint32_t M, uint32_t ef_construction,
                      uint32_t K, hrnn_hnsw_build_stats *stats, const std::function<void(uint32_t, uint32_t)> &after);
// Look up, with the handle's device current, this device's copy of the
// no-allocation search's scratch array (once per device; called by
// hrnn_dev_open) and clear it.
void noalloc_scratch_init(hrnn_dev_t h);
// Per-query beam statistics (raw words): 0 d(q, entry), 1 d(q, layer-0
// start), 2..17 the 16 closest beam distances, 18 expansions, 19 list size.
constexpr int kRawWords = 20;
constexpr uint32_t kErrWord = 96;   // u32 index of the device error word in counters
constexpr int kFeatWords = 32;       // learned-m features per query
// Optional per-query controls and outputs of a phase-A launch (device
// pointers; null = unused).
struct BeamOpt {
    hipStream_t stream = nullptr;     // the launch's stream (null: the handle's)
    const uint32_t *ef_q = nullptr;   // per-query ef, <= the launch's ef; 0 skips the query
    const uint32_t *m_q = nullptr;    // per-query proxy count; later slots get 0xFFFFFFFF
    float *feats = nullptr;           // kFeatWords features per query
    const float *knn_d = nullptr;     // ranked KNN distances [n x K] (features' r_k)
    const uint32_t *knn_cnt = nullptr;
    uint32_t K = 0, feat_k = 0;       // r_k = knn_d[p][feat_k - 1]
    const float *mu = nullptr;        // standardisation (f - mu) / sd, both or neither
    const float *sd = nullptr;
    const uint32_t *qlist = nullptr;  // workgroup i takes query qlist[i], i < *qcount
    const uint32_t *qcount = nullptr;
    // learned early stop fused into the stage-1 beam (all of these, or mlp
    // null): MLP parameters (features standardised with mu/sd above), the 16
    // classes' ef then m, tau, the stage-1 width; per query the class, the
    // stage-2 ef (0: none) and m, stage-2 expansion counters zeroed for the
    // rest, and the escalated / other query lists (lists[0..), lists[nq..))
    // with their counts; zero_a / zero_b (also without mlp) are set to zero
    // by the launch's first workgroup (the queues of the searches after it)
    const float *mlp = nullptr;
    const uint32_t *classes = nullptr;
    float tau = 0.f;
    uint32_t ef1 = 0, nq = 0;
    uint32_t *cls = nullptr, *ef_q_out = nullptr, *m_q_out = nullptr, *exp2 = nullptr;
    uint32_t *lists = nullptr, *list_cnt = nullptr, *zero_a = nullptr, *zero_b = nullptr;
    // with mlp: stop the beam after max_exp expansions (0: no cap); a query
    // stopped early whose class keeps it at stage 1 takes the first class with
    // a wider beam instead
    uint32_t max_exp = 0;
    // the query rows written here as the beam reads them (nq x d floats, the
    // layout of q), so the phases that follow read device memory
    float *q_copy = nullptr;
};
// The r_k column for k (n floats, rk[i] = k-th distance of i's list) built on
// stream st unless it is current; searches that run concurrently on several
// streams call this first, on the stream they all follow.
void rk_column_ready(hrnn_dev_t h, uint32_t k, hipStream_t st);
// True when p points into page-locked host memory (mapped for the device).
bool host_memory(const void *p);
// The handle's second stream and its two events (created on first use).
void ensure_stream2(hrnn_dev_t h);
void beam_launch(hrnn_dev_t h, const float *q_dev, uint32_t nq, uint32_t m, uint32_t ef,
                 uint32_t *proxies_dev, uint32_t *expansions_dev, float *raw_dev = nullptr,
                 const BeamOpt *opt = nullptr);
// Optional outputs of bf_topk_base (device pointers; null = unused).
struct BfExtra {
    double *r64 = nullptr;         // nq: the kk-th smallest FP64 distance of each row
    float *raw_kth = nullptr;      // nq: the first candidate pass's own kk-th distance
    uint32_t *raw_ids = nullptr;   // nq * kk: the first candidate pass's first kk ids
    hrnn_bf_stats *st = nullptr;   // host: certificate statistics
};
void bf_topk_base(hrnn_dev_t h, const float *q_dev, uint32_t nq, bool self_base, uint32_t kk,
                  int prec, uint32_t *ids_dev, float *d_dev, const BfExtra *ex = nullptr);
// RkNN hits of nq device queries; with kth64_dev, pairs within FP32 rounding
// of r_k are decided in FP64 and chk_dev (2 u64) counts them and the changed
// decisions.
void bf_rknn_count_fill(hrnn_dev_t h, const float *q_dev, uint32_t nq, const float *kth_dev,
                        int prec, uint32_t *cnt_dev, uint32_t *ids_dev, float *d_dev,
                        uint32_t cap_per_q, const double *kth64_dev = nullptr,
                        unsigned long long *chk_dev = nullptr);
void l2_pairs(hrnn_dev_t h, const float *a_dev, const uint32_t *ia, const uint32_t *ib,
              uint64_t npairs, float *out);

}  // namespace hrnn
