// hnsw_build.hip -- HNSW construction with KNN seeding on the GPU.
//
// Reference: HRNN::addPoint (include/hnswlib/hrnn.h:1378-1517) with
// searchBaseLayer (:401-481), getNeighborsByHeuristic2 (:619-666) and
// mutuallyConnectNewElement (:682-806).  Per inserted point: greedy descent
// over the layers above its level, then at each of its layers a beam of width
// ef_construction, the layer-0 beam sorted by (dist, id) seeding the point's
// ranked KNN list (self skipped, the first K; :1479-1501), a diverse subset of
// at most M beam entries as the point's links (a candidate is dropped when a
// kept one is closer to it than the point is), and a link back from each of
// them, a full list (2M at layer 0, M above) being pruned by the same rule
// over its entries and the new link.
//
// The reference inserts with many threads, each insertion seeing the graph as
// the others have left it.  Here points are inserted in batches, each batch a
// few percent of the graph it is inserted into: one wave per point searches
// the graph as it was when the batch started (nothing a search reads is
// written during the search kernel; a point writes only its own lists), then
// per layer the links back are gathered per target node (k_hnsw_scatter) and
// each target's list is rebuilt once from its entries and all its new links
// (k_hnsw_link: appended when they fit, else pruned by the diversity rule
// over (dist, id)-sorted candidates).  Levels are drawn on the host with the
// reference's generator; the entry point and top level change between
// batches exactly as sequential insertion changes them.
//
// The beams use the phase-A machinery of beam.hip (LDS list of (dist, id)
// keys, visited set, wave-parallel merge) with the construction stop rule:
// the search ends when every list entry has been expanded, the list holding
// at most ef nodes and admitting a node while it is not full or the node is
// below the worst entry (searchBaseLayer's candidate set, without the nodes it
// would never expand).  Distances are the reference's FP32 16-partial order
// (exact integer sums for 8-bit data), so accept/reject decisions agree with
// a CPU insertion of the same point into the same graph.
#include <algorithm>
#include <cfloat>
#include <chrono>
#include <vector>

#include "beam_core.h"
#include "handle.h"

namespace hrnn {

constexpr uint32_t kIncMax = 64;       // new links one node takes per batch and layer
constexpr uint32_t kLinkCand = 128;    // candidates of one list rebuild (<= 64 entries + kIncMax)
constexpr uint32_t kMaxSelect = 64;    // links one point selects per layer

struct BuildGraph {
    const float *xf;            // FP32 base rows [n x d]
    uint32_t d, n;
    uint32_t *l0;               // [n x l0_words]: count word, then ids
    uint32_t l0_words;
    const uint32_t *up_off;     // [n] word offset of the node's upper blocks, or 0xFFFFFFFF
    uint32_t *up_pool;          // blocks of up_words words per upper layer
    uint64_t pool_words;
    uint32_t up_words;
    const int *levels;          // [n]
};

// The link list of node c at layer L (count word then ids), or null when the
// node has no such layer (the caller flags the graph as bad).
__device__ __forceinline__ uint32_t *layer_list(const BuildGraph &G, int L, uint32_t c, uint32_t *words) {
    if (L == 0) {
        *words = G.l0_words;
        return G.l0 + (uint64_t)c * G.l0_words;
    }
    const uint32_t uo = G.up_off[c];
    const uint64_t base = (uint64_t)uo + (uint64_t)(L - 1) * G.up_words;
    *words = G.up_words;
    if (uo == kInvalidId || G.levels[c] < L || base + G.up_words > G.pool_words) return nullptr;
    return G.up_pool + base;
}

// Row of node c at layer L: the count (uniform) and, in lane j, the j-th id.
__device__ __forceinline__ void layer_row(const BuildGraph &G, int L, uint32_t c, uint32_t lane,
                                          uint32_t *cnt, uint32_t *nb, uint32_t *bad) {
    uint32_t words = 0;
    const uint32_t *ll = layer_list(G, L, c, &words);
    if (!ll) {
        *cnt = 0;
        *nb = 0;
        if (lane == 0) *bad = 1u;
        return;
    }
    const uint32_t w0 = ll[0];
    *nb = (lane + 1u < words) ? ll[1 + lane] : 0u;
    *cnt = uni(min(w0 & 0xFFFFu, words - 1u));
}

// Squared L2 between base rows a and b by a 16-lane group (the reference's
// order; sequential when d % 16 != 0).
__device__ __forceinline__ float pair_dist(const float *__restrict__ xf, uint32_t a, uint32_t b, uint32_t d,
                                           uint32_t g) {
    const float *pa = xf + (uint64_t)a * d, *pb = xf + (uint64_t)b * d;
    return ((d & 15u) == 0) ? l2_ref16_group(pa, pb, d, (int)g) : l2_seq(pa, pb, d);
}

// Diversity selection (getNeighborsByHeuristic2): cand[0 .. nc) ascending
// (dist, id) keys of nodes at distance dist from a base node; keep at most M
// of them in order, dropping a candidate c when a kept r has d(r, c) < d(c).
// Fewer than M candidates are all kept.  Kept keys go to sel[0 .. ns), ns is
// returned (the same in every lane).  Candidates are taken four at a time,
// one per 16-lane group: each group measures its candidate against the kept
// ones and the earlier candidates of its four; the four are then decided in
// order.  cand and sel are LDS arrays; the whole wave calls this.
__device__ uint32_t select_diverse_wave(const BuildGraph &G, const unsigned long long *cand, uint32_t nc,
                                        uint32_t M, unsigned long long *sel, uint32_t lane) {
    const uint32_t g = lane & 15u, grp = lane >> 4;
    if (nc < M) {
        for (uint32_t t = lane; t < nc; t += 64) sel[t] = cand[t];
        __syncthreads();
        return nc;
    }
    uint32_t ns = 0;
    for (uint32_t t0 = 0; t0 < nc && ns < M; t0 += 4) {
        const uint32_t ci = t0 + grp;
        const bool valid = ci < nc;
        const unsigned long long kc = valid ? cand[ci] : cand[t0];
        const uint32_t idc = (uint32_t)kc;
        const float dc = key_float((uint32_t)(kc >> 32));
        bool dom = false;
        for (uint32_t j = 0; j < ns; j++) {   // ns is uniform
            const float dd = pair_dist(G.xf, (uint32_t)sel[j], idc, G.d, g);
            dom |= dd < dc;
        }
        // distances to the earlier candidates of this four (group a < grp)
        float dp[3];
#pragma unroll
        for (uint32_t a = 0; a < 3; a++) {
            const uint32_t ida = (uint32_t)__shfl((int)idc, (int)(a * 16u), 64);
            const float dd = pair_dist(G.xf, ida, idc, G.d, g);
            dp[a] = (a < grp) ? dd : FLT_MAX;
        }
        // decide the four in order; every lane computes the same decisions
        bool kept[4] = {false, false, false, false};
#pragma unroll
        for (uint32_t a = 0; a < 4; a++) {
            const bool va = __shfl((int)valid, (int)(a * 16u), 64) != 0;
            const bool doma = __shfl((int)dom, (int)(a * 16u), 64) != 0;
            const float da = __shfl(dc, (int)(a * 16u), 64);
            bool ok = va && !doma && ns < M;
#pragma unroll
            for (uint32_t b = 0; b < 3; b++) {
                const float dab = __shfl(dp[b], (int)(a * 16u), 64);   // d(candidate b, candidate a)
                if (b < a && kept[b] && dab < da) ok = false;
            }
            kept[a] = ok;
            if (ok) {
                const unsigned long long ka = ((unsigned long long)(uint32_t)__shfl((int)(uint32_t)(kc >> 32), (int)(a * 16u), 64) << 32) |
                                              (uint32_t)__shfl((int)idc, (int)(a * 16u), 64);
                if (lane == 0) sel[ns] = ka;
                ns++;
            }
        }
        __syncthreads();   // sel visible to every lane before the next four read it
    }
    return ns;
}

// LDS state and per-beam constants of one construction beam.
struct BeamCtx {
    float *qv;
    uint32_t *q8, *vis;
    unsigned long long *kl;
    uint8_t *xl;
    float *nd;
    uint32_t *ni;
    unsigned long long *ck;
    uint32_t *lost, *bad;
    uint32_t ef, ef_cap, log_vis, qq;
    bool exact;
    __amdgpu_buffer_rsrc_t xrs;
};

// The beam's expansion loop at layer L over the list kl[buf] of Ln entries
// (with expanded flags; the visited set already holds every listed node):
// expand the first unexpanded entry, admit its unvisited neighbours with id
// below hi while the list holds fewer than ef entries or the (dist, id) key
// is below the worst, keep the list sorted; stop when every entry is
// expanded or after max_exp expansions.  Ln and buf are updated.
template <class VT>
__device__ void beam_expand(const BuildGraph &G, const VT *__restrict__ x, BeamCtx &cx, int L, uint32_t &Ln,
                            uint32_t &buf, uint32_t n, uint32_t hi, unsigned long long &n_exp,
                            unsigned long long &n_eval, uint32_t max_exp = 0xFFFFFFFFu) {
    const uint32_t lane = threadIdx.x, g = lane & 15u, grp = lane >> 4;
    const unsigned long long exp_end = max_exp == 0xFFFFFFFFu ? ~0ull : n_exp + max_exp;
    // the first unexpanded entry
    uint32_t pick = kInvalidId;
    for (uint32_t q0 = 0; q0 < Ln; q0 += 64) {
        const uint32_t t = q0 + lane;
        const uint64_t mask = __ballot(t < Ln && cx.xl[buf * cx.ef_cap + t] == 0);
        if (mask) {
            pick = q0 + (uint32_t)__ffsll((long long)mask) - 1;
            break;
        }
    }
    for (uint32_t iter = 0; iter <= n && pick != kInvalidId && n_exp < exp_end; iter++) {
        const unsigned long long *lk = cx.kl + buf * cx.ef_cap;
        uint8_t *lx = cx.xl + buf * cx.ef_cap;
        const uint32_t c = uni((uint32_t)lk[pick]);
        uint32_t cnt, nbw;
        layer_row(G, L, c, lane, &cnt, &nbw, cx.bad);
        __syncthreads();
        if (lane == 0) lx[pick] = 1;
        n_exp++;
        uint32_t nb = 0;
        bool fresh = false;
        if (lane < cnt) {
            nb = nbw;
            if (nb < n) fresh = vis_insert(cx.vis, cx.log_vis, nb, cx.lost);
            else *cx.bad = 1u;
        }
        const uint64_t fm = __ballot(fresh);
        const uint32_t nf = (uint32_t)__popcll(fm);
        n_eval += nf;
        if (fresh) cx.ni[__popcll(fm & lanemask_lt())] = nb;
        __syncthreads();
        uint32_t o = kInvalidId;   // the first unexpanded entry after pick
        for (uint32_t q0 = pick + 1; q0 < Ln; q0 += 64) {
            const uint32_t t = q0 + lane;
            const uint64_t mask = __ballot(t < Ln && lx[t] == 0);
            if (mask) {
                o = q0 + (uint32_t)__ffsll((long long)mask) - 1;
                break;
            }
        }
        if (nf == 0) {
            pick = o;
            continue;
        }
        constexpr int kRows = std::is_same<VT, uint8_t>::value ? 8 : 4;
        for (uint32_t b = 0; b < nf; b += 4 * kRows) {
            uint32_t idr[kRows];
            bool vr[kRows];
#pragma unroll
            for (int r = 0; r < kRows; r++) {
                const uint32_t t = b + grp + 4u * (uint32_t)r;
                vr[r] = t < nf;
                idr[r] = cx.ni[vr[r] ? t : b];
            }
            if constexpr (std::is_same<VT, float>::value) {
                if (!cx.exact && (G.d & 15u) == 0) {   // float data: quad r of each group takes row r
                    const float s4 = dist4_quad_f32(cx.qv, x, idr, vr, G.d, g);
                    const uint32_t t = b + grp + 4u * (g >> 2);
                    if ((g & 3u) == 0 && t < nf) cx.nd[t] = s4;
                    continue;
                }
            }
            float orr[kRows];
            beam_dist4<VT, kRows>(cx.exact, cx.qv, cx.q8, cx.qq, x, cx.xrs, idr, vr, G.d, g, orr);
            if (g == 0) {
#pragma unroll
                for (int r = 0; r < kRows; r++)
                    if (vr[r]) cx.nd[b + grp + 4u * (uint32_t)r] = orr[r];
            }
        }
        __syncthreads();
        unsigned long long key = ~0ull;
        if (lane < nf) {
            const unsigned long long kc = beam_key(cx.nd[lane], cx.ni[lane]);
            bool ok = (uint32_t)kc < hi && (Ln < cx.ef || kc < lk[Ln - 1]);
            if (ok && *cx.lost) {
                for (uint32_t j = 0; j < Ln && ok; j++) ok = (uint32_t)lk[j] != (uint32_t)kc;
                for (uint32_t j = 0; j < lane && ok; j++) ok = cx.ni[j] != (uint32_t)kc;
            }
            if (ok) key = kc;
        }
        const uint64_t am = __ballot(key != ~0ull);
        const uint32_t nv = (uint32_t)__popcll(am);
        if (nv == 0) {
            pick = o;
            continue;
        }
        if (key != ~0ull) cx.ck[__popcll(am & lanemask_lt())] = key;
        __syncthreads();
        key = (lane < nv) ? cx.ck[lane] : ~0ull;
        key = wave_sort64(key, lane, next_pow2_dev(nv));
        const unsigned long long kbest = read_lane64(key, 0);
        const unsigned long long ko = (o != kInvalidId) ? uni64(lk[o]) : ~0ull;
        const uint32_t nbuf = buf ^ 1u;
        unsigned long long *nk = cx.kl + nbuf * cx.ef_cap;
        uint8_t *nx = cx.xl + nbuf * cx.ef_cap;
        uint32_t below = 0;
        for (uint32_t q0 = 0; q0 < Ln; q0 += 64) {
            const uint32_t e = q0 + lane;
            const bool valid = e < Ln;
            const unsigned long long ke = valid ? lk[e] : ~0ull;
            const uint8_t xe = valid ? lx[e] : (uint8_t)1;
            uint32_t less = 0;
            for (uint32_t j = 0; j < nv; j++) {
                const unsigned long long kj = read_lane64(key, j);
                const bool lt = kj < ke;
                less += lt ? 1u : 0u;
                const uint32_t cb = (uint32_t)__popcll(__ballot(valid && !lt));
                if (lane == j) below += cb;
            }
            const uint32_t pos = e + less;
            if (valid && pos < cx.ef) {
                nk[pos] = ke;
                nx[pos] = xe;
            }
        }
        if (lane < nv) {
            const uint32_t pos = lane + below;
            if (pos < cx.ef) {
                nk[pos] = key;
                nx[pos] = 0;
            }
        }
        pick = (kbest < ko) ? (uint32_t)__builtin_amdgcn_readlane((int)below, 0) : o;
        __syncthreads();
        Ln = min(Ln + nv, cx.ef);
        buf = nbuf;
    }
}

// One inserted point per workgroup of one wave.
template <class VT>
__global__ __launch_bounds__(64) void k_hnsw_insert(
    const VT *__restrict__ x, BuildGraph G, uint32_t b0, uint32_t b1, uint32_t entry, int maxlevel,
    uint32_t ef, uint32_t ef_cap, uint32_t log_vis, uint32_t M, int base_exact, uint32_t K,
    uint32_t *__restrict__ knn_ids, float *__restrict__ knn_d, uint32_t *__restrict__ knn_cnt,
    unsigned long long *__restrict__ prop, uint32_t prop_stride, uint32_t *__restrict__ err,
    unsigned long long *__restrict__ stats) {
    __shared__ float s_nd[kBeamMaxDeg];
    __shared__ uint32_t s_ni[kBeamMaxDeg];
    __shared__ unsigned long long s_ck[kBeamMaxDeg];
    __shared__ unsigned long long s_sel[kMaxSelect];
    __shared__ uint32_t s_lost, s_bad;
    extern __shared__ __align__(16) unsigned char s_dyn[];
    const BeamLds lay(G.d, ef_cap, log_vis);
    float *s_qv = reinterpret_cast<float *>(s_dyn);
    uint32_t *s_q8 = reinterpret_cast<uint32_t *>(s_dyn + lay.q8_off);
    uint32_t *s_vis = reinterpret_cast<uint32_t *>(s_dyn + lay.vis_off);
    unsigned long long *s_kl = reinterpret_cast<unsigned long long *>(s_dyn + lay.k_off);
    uint8_t *s_xl = s_dyn + lay.x_off;

    const uint32_t lane = threadIdx.x;
    const uint32_t g = lane & 15u, grp = lane >> 4;
    const uint32_t p = uni(b0 + blockIdx.x);
    if (p >= b1) return;
    const uint32_t d = G.d, n = G.n;
    bool q_u8 = true;
    for (uint32_t e = lane; e < d; e += 64) {
        const float v = G.xf[(uint64_t)p * d + e];
        s_qv[e] = v;
        q_u8 &= (v == rintf(v)) && v >= 0.f && v <= 255.f;
    }
    const bool exact = base_exact && __ballot(!q_u8) == 0;
    uint32_t qq = 0;
    if constexpr (std::is_same<VT, uint8_t>::value) {
        if (exact) {
            uint32_t part = 0;
            for (uint32_t w = lane; 4 * w < d; w += 64) {
                const float *qw = G.xf + (uint64_t)p * d + 4 * w;
                const uint32_t c0 = (uint32_t)qw[0], c1 = (uint32_t)qw[1], c2 = (uint32_t)qw[2], c3 = (uint32_t)qw[3];
                s_q8[w] = c0 | (c1 << 8) | (c2 << 16) | (c3 << 24);
                part += c0 * c0 + c1 * c1 + c2 * c2 + c3 * c3;
            }
            for (int s2 = 32; s2 > 0; s2 >>= 1) part += (uint32_t)__shfl_xor((int)part, s2, 64);
            qq = uni(part);
        }
    }
    const __amdgpu_buffer_rsrc_t xrs = make_rsrc(x, exact ? (uint32_t)((uint64_t)n * d * sizeof(VT)) : 0u);
    if (lane == 0) {
        s_lost = 0;
        s_bad = 0;
    }
    __syncthreads();
    const int lp = G.levels[p];
    // this point's proposal slots start empty at every layer (lane t owns
    // slot t of each layer, so the later writes of the same lane follow)
    unsigned long long *prow = prop + (uint64_t)blockIdx.x * prop_stride;
    for (uint32_t L = 0; L < prop_stride / M; L++)
        if (lane < M) prow[(uint64_t)L * M + lane] = ~0ull;

    // ---- greedy descent over the layers above the point's level ----
    uint32_t cur = entry < n ? entry : 0;
    float cd;
    {
        const uint32_t id4[4] = {cur, cur, cur, cur};
        const bool v4[4] = {true, false, false, false};
        float o[4];
        beam_dist4<VT>(exact, s_qv, s_q8, qq, x, xrs, id4, v4, d, g, o);
        cd = uni_f(__shfl(o[0], 0, 64));
    }
    for (int level = maxlevel; level > lp; level--) {
        for (uint32_t step = 0; step < n; step++) {   // each move strictly lowers cd
            uint32_t cnt, w;
            layer_row(G, level, cur, lane, &cnt, &w, &s_bad);
            float best = cd;
            uint32_t best_id = cur;
            for (uint32_t b = 0; b < cnt; b += 16) {
                uint32_t id4[4];
                bool v4[4];
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const uint32_t t = b + grp + 4u * (uint32_t)r;
                    uint32_t v = (uint32_t)__shfl((int)w, (int)min(t, cnt - 1u), 64);
                    v4[r] = t < cnt;
                    if (v >= n) {
                        if (v4[r] && g == 0) s_bad = 1u;
                        v4[r] = false;
                        v = cur;
                    }
                    id4[r] = v;
                }
                float o[4];
                beam_dist4<VT>(exact, s_qv, s_q8, qq, x, xrs, id4, v4, d, g, o);
                unsigned long long mk = ~0ull;
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const unsigned long long kk =
                        ((unsigned long long)float_key(o[r]) << 32) | (b + grp + 4u * (uint32_t)r);
                    if (v4[r] && kk < mk) mk = kk;
                }
#pragma unroll
                for (int s2 = 16; s2 <= 32; s2 <<= 1) {
                    const unsigned long long other = shfl_xor64(mk, s2);
                    mk = other < mk ? other : mk;
                }
                if (mk != ~0ull) {
                    const float ds = key_float((uint32_t)(mk >> 32));
                    if (ds < best) {
                        best = ds;
                        best_id = (uint32_t)__shfl((int)w, (int)(uint32_t)mk, 64);
                    }
                }
            }
            if (best_id == cur) break;
            cur = uni(best_id);
            cd = uni_f(best);
        }
    }

    unsigned long long n_exp = 0, n_eval = 1;
    BeamCtx bs{s_qv, s_q8, s_vis, s_kl, s_xl, s_nd, s_ni, s_ck, &s_lost, &s_bad, ef, ef_cap, log_vis, qq, exact, xrs};
    for (int L = min(lp, maxlevel); L >= 0; L--) {
        // ---- beam at layer L from (cur, cd) ----
        for (uint32_t t = lane; t < (1u << log_vis) / 4; t += 64)
            reinterpret_cast<uint4 *>(s_vis)[t] = make_uint4(kInvalidId, kInvalidId, kInvalidId, kInvalidId);
        if (lane == 0) s_lost = 0;
        __syncthreads();
        if (lane == 0) {
            s_kl[0] = beam_key(cd, cur);
            s_xl[0] = 0;
            vis_insert(s_vis, log_vis, cur, &s_lost);
        }
        __syncthreads();
        uint32_t Ln = 1, buf = 0;
        beam_expand<VT>(G, x, bs, L, Ln, buf, n, 0xFFFFFFFFu, n_exp, n_eval);
        const unsigned long long *lk = s_kl + buf * ef_cap;

        // ---- layer 0: the KNN seed, the beam in (dist, id) order without p ----
        if (L == 0 && K > 0) {
            uint32_t written = 0;
            for (uint32_t q0 = 0; q0 < Ln && written < K; q0 += 64) {
                const uint32_t t = q0 + lane;
                const unsigned long long kt = t < Ln ? lk[t] : ~0ull;
                const bool keep = t < Ln && (uint32_t)kt != p;
                const uint64_t bal = __ballot(keep);
                const uint32_t pos = written + (uint32_t)__popcll(bal & lanemask_lt());
                if (keep && pos < K) {
                    knn_ids[(uint64_t)p * K + pos] = (uint32_t)kt;
                    knn_d[(uint64_t)p * K + pos] = key_float((uint32_t)(kt >> 32));
                }
                written += (uint32_t)__popcll(bal);
            }
            if (lane == 0) knn_cnt[p] = min(written, K);
        }

        // ---- the point's links at L and the links back to it ----
        const uint32_t ns = select_diverse_wave(G, lk, Ln, M, s_sel, lane);
        uint32_t words = 0;
        uint32_t *ll = layer_list(G, L, p, &words);
        if (!ll || ns + 1 > words) {
            if (lane == 0) s_bad = 1u;
        } else {
            if (lane < ns) ll[1 + lane] = (uint32_t)s_sel[lane];
            if (lane == 0) ll[0] = ns;
        }
        if (lane < M) prow[(uint64_t)L * M + lane] = (lane < ns && ll) ? s_sel[lane] : ~0ull;
        if (ns > 0) {
            cur = uni((uint32_t)s_sel[0]);
            cd = uni_f(key_float((uint32_t)(s_sel[0] >> 32)));
        }
        __syncthreads();
    }
    if (lane == 0) {
        atomicAdd(&stats[0], n_exp);
        atomicAdd(&stats[1], n_eval);
        if (s_bad) atomicOr(err, 1u);
    }
}

// Seed completion after a batch's links are merged: the points of one batch
// searched the graph as it was before the batch, so a point's seed list
// misses the points of its own batch that sequential insertion would have
// let it see (those before it).  Starting from its seed list (all entries
// unexpanded), each point continues its layer-0 beam over the graph that now
// links the batch for at most max_exp expansions, admitting only ids below
// its own, and keeps the first K entries.  One wave per point.
template <class VT>
__global__ __launch_bounds__(64) void k_hnsw_seed_fix(
    const VT *__restrict__ x, BuildGraph G, uint32_t b0, uint32_t b1, uint32_t ef, uint32_t ef_cap,
    uint32_t log_vis, int base_exact, uint32_t K, uint32_t max_exp, uint32_t *__restrict__ knn_ids,
    float *__restrict__ knn_d, uint32_t *__restrict__ knn_cnt, uint32_t *__restrict__ err,
    unsigned long long *__restrict__ stats) {
    __shared__ float s_nd[kBeamMaxDeg];
    __shared__ uint32_t s_ni[kBeamMaxDeg];
    __shared__ unsigned long long s_ck[kBeamMaxDeg];
    __shared__ uint32_t s_lost, s_bad;
    extern __shared__ __align__(16) unsigned char s_dyn[];
    const BeamLds lay(G.d, ef_cap, log_vis);
    float *s_qv = reinterpret_cast<float *>(s_dyn);
    uint32_t *s_q8 = reinterpret_cast<uint32_t *>(s_dyn + lay.q8_off);
    uint32_t *s_vis = reinterpret_cast<uint32_t *>(s_dyn + lay.vis_off);
    unsigned long long *s_kl = reinterpret_cast<unsigned long long *>(s_dyn + lay.k_off);
    uint8_t *s_xl = s_dyn + lay.x_off;
    const uint32_t lane = threadIdx.x;
    const uint32_t p = uni(b0 + blockIdx.x);
    if (p >= b1) return;
    const uint32_t d = G.d, n = G.n;
    const uint32_t c0 = uni(min(knn_cnt[p], min(ef, K)));
    if (c0 == 0) return;
    bool q_u8 = true;
    for (uint32_t e = lane; e < d; e += 64) {
        const float v = G.xf[(uint64_t)p * d + e];
        s_qv[e] = v;
        q_u8 &= (v == rintf(v)) && v >= 0.f && v <= 255.f;
    }
    const bool exact = base_exact && __ballot(!q_u8) == 0;
    uint32_t qq = 0;
    if constexpr (std::is_same<VT, uint8_t>::value) {
        if (exact) {
            uint32_t part = 0;
            for (uint32_t w = lane; 4 * w < d; w += 64) {
                const float *qw = G.xf + (uint64_t)p * d + 4 * w;
                const uint32_t c0b = (uint32_t)qw[0], c1 = (uint32_t)qw[1], c2 = (uint32_t)qw[2], c3 = (uint32_t)qw[3];
                s_q8[w] = c0b | (c1 << 8) | (c2 << 16) | (c3 << 24);
                part += c0b * c0b + c1 * c1 + c2 * c2 + c3 * c3;
            }
            for (int s2 = 32; s2 > 0; s2 >>= 1) part += (uint32_t)__shfl_xor((int)part, s2, 64);
            qq = uni(part);
        }
    }
    const __amdgpu_buffer_rsrc_t xrs = make_rsrc(x, exact ? (uint32_t)((uint64_t)n * d * sizeof(VT)) : 0u);
    for (uint32_t t = lane; t < (1u << log_vis) / 4; t += 64)
        reinterpret_cast<uint4 *>(s_vis)[t] = make_uint4(kInvalidId, kInvalidId, kInvalidId, kInvalidId);
    if (lane == 0) {
        s_lost = 0;
        s_bad = 0;
    }
    __syncthreads();
    // the seed list, ascending (dist, id), as the start list; p and its seeds visited
    for (uint32_t t = lane; t < c0; t += 64) {
        const uint32_t id = knn_ids[(uint64_t)p * K + t];
        s_kl[t] = beam_key(knn_d[(uint64_t)p * K + t], id);
        s_xl[t] = 0;
        vis_insert(s_vis, log_vis, id, &s_lost);
    }
    if (lane == 0) vis_insert(s_vis, log_vis, p, &s_lost);
    __syncthreads();
    unsigned long long n_exp = 0, n_eval = 0;
    BeamCtx bs{s_qv, s_q8, s_vis, s_kl, s_xl, s_nd, s_ni, s_ck, &s_lost, &s_bad, ef, ef_cap, log_vis, qq, exact, xrs};
    uint32_t Ln = c0, buf = 0;
    beam_expand<VT>(G, x, bs, 0, Ln, buf, n, p, n_exp, n_eval, max_exp);
    const unsigned long long *lk = s_kl + buf * ef_cap;
    const uint32_t keep = min(Ln, K);
    for (uint32_t t = lane; t < keep; t += 64) {
        const unsigned long long kt = lk[t];
        knn_ids[(uint64_t)p * K + t] = (uint32_t)kt;
        knn_d[(uint64_t)p * K + t] = key_float((uint32_t)(kt >> 32));
    }
    if (lane == 0) {
        knn_cnt[p] = keep;
        atomicAdd(&stats[3], n_exp);
        atomicAdd(&stats[4], n_eval);
        if (s_bad) atomicOr(err, 1u);
    }
}

// Links back at layer L: each proposal (target, dist) of the batch's points
// joins its target's incoming list; the first proposal for a target records
// the target as touched.  Beyond kIncMax a proposal is counted and dropped.
__global__ void k_hnsw_scatter(const unsigned long long *__restrict__ prop, uint32_t nb, uint32_t prop_stride,
                               uint32_t M, int L, uint32_t b0, uint32_t n, uint32_t *__restrict__ inc_cnt,
                               unsigned long long *__restrict__ inc, uint32_t *__restrict__ touched,
                               uint32_t *__restrict__ n_touched, unsigned long long *__restrict__ dropped,
                               uint32_t *__restrict__ err) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint64_t)nb * M) return;
    const uint32_t pi = (uint32_t)(i / M), j = (uint32_t)(i % M);
    const unsigned long long key = prop[(uint64_t)pi * prop_stride + (uint64_t)L * M + j];
    if (key == ~0ull) return;
    const uint32_t t = (uint32_t)key;
    if (t >= n) {
        atomicOr(err, 1u);
        return;
    }
    const uint32_t s = atomicAdd(&inc_cnt[t], 1u);
    if (s == 0) touched[atomicAdd(n_touched, 1u)] = t;
    if (s < kIncMax) inc[(uint64_t)t * kIncMax + s] = (key & 0xFFFFFFFF00000000ull) | (uint64_t)(b0 + pi);
    else atomicAdd(dropped, 1ull);
}

// Rebuild the list of every touched node at layer L from its entries and its
// incoming links: appended when they all fit in cap, else the diversity rule
// over all of them in (dist, id) order keeps at most cap.  One wave per node.
__global__ __launch_bounds__(64) void k_hnsw_link(BuildGraph G, int L, uint32_t cap,
                                                  const uint32_t *__restrict__ touched,
                                                  const uint32_t *__restrict__ n_touched,
                                                  uint32_t *__restrict__ inc_cnt,
                                                  const unsigned long long *__restrict__ inc,
                                                  uint32_t *__restrict__ err) {
    __shared__ unsigned long long s_c[kLinkCand];
    __shared__ unsigned long long s_sel[kLinkCand];
    const uint32_t lane = threadIdx.x, g = lane & 15u, grp = lane >> 4;
    if (blockIdx.x >= *n_touched) return;
    const uint32_t t = touched[blockIdx.x];
    const uint32_t ci = min(inc_cnt[t], kIncMax);
    uint32_t words = 0;
    uint32_t *ll = layer_list(G, L, t, &words);
    if (!ll || cap + 1 > words) {
        if (lane == 0) {
            atomicOr(err, 1u);
            inc_cnt[t] = 0;
        }
        return;
    }
    const uint32_t ce = min(ll[0] & 0xFFFFu, words - 1u);
    // incoming links sorted by (dist, id)
    unsigned long long kin = lane < ci ? inc[(uint64_t)t * kIncMax + lane] : ~0ull;
    kin = wave_sort64(kin, lane, next_pow2_dev(ci));
    if (ce + ci <= cap) {
        if (lane < ci) ll[1 + ce + lane] = (uint32_t)kin;
        __syncthreads();
        if (lane == 0) {
            ll[0] = ce + ci;
            inc_cnt[t] = 0;
        }
        return;
    }
    // entries' distances to t, four at a time per group
    for (uint32_t b = 0; b < ce; b += 4) {
        const uint32_t e = b + grp;
        uint32_t id = ll[1 + (e < ce ? e : b)];
        if (id >= G.n) {
            if (g == 0 && e < ce) atomicOr(err, 1u);
            id = t;
        }
        const float dd = pair_dist(G.xf, t, id, G.d, g);
        if (g == 0 && e < ce) s_c[e] = ((unsigned long long)float_key(dd) << 32) | id;
    }
    if (lane < ci) s_c[ce + lane] = kin;
    for (uint32_t e = ce + ci + lane; e < kLinkCand; e += 64) s_c[e] = ~0ull;
    bitonic_sort_lds(s_c, kLinkCand);
    const uint32_t ns = select_diverse_wave(G, s_c, ce + ci, cap, s_sel, lane);
    if (lane < ns) ll[1 + lane] = (uint32_t)s_sel[lane];
    __syncthreads();
    if (lane == 0) {
        ll[0] = ns;
        inc_cnt[t] = 0;
    }
}

__global__ void k_zero_u32(uint32_t *__restrict__ p, uint32_t count) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) p[i] = 0;
}

// Visited-set slots for a construction beam of width ef (as phase A's sizing).
static uint32_t build_log_vis(uint32_t ef) { return ef <= 16 ? 10u : ef <= 256 ? 11u : 12u; }

// Copy the first keep bytes of b into a buffer of at least want bytes.
static void grow_keep(hrnn_dev_t h, DevBuf &b, size_t keep, size_t want) {
    if (want <= b.bytes) return;
    DevBuf nb;
    nb.reserve(want);
    if (keep) HIPC(hipMemcpyAsync(nb.p, b.p, keep, hipMemcpyDeviceToDevice, h->stream));
    HIPC(hipStreamSynchronize(h->stream));
    b.release();
    b = nb;
}

void hnsw_insert_rows(hrnn_dev_t h, const int *levels, uint64_t b_first, uint32_t M, uint32_t ef_construction,
                      uint32_t K, hrnn_hnsw_build_stats *stats, const std::function<void(uint32_t, uint32_t)> &after) {
    const auto t_start = std::chrono::steady_clock::now();
    REQUIRE(h->n > 0, HRNN_E_STATE, "hnsw_build: set_base first");
    REQUIRE(levels, HRNN_E_ARG, "hnsw_build: null argument");
    REQUIRE(M >= 2 && 2 * M <= (uint32_t)kBeamMaxDeg && M <= kMaxSelect, HRNN_E_ARG,
            "hnsw_build: M must be in [2, 32]");
    const uint32_t ef = std::max(ef_construction, M);
    REQUIRE(ef <= (uint32_t)kBeamMaxEf, HRNN_E_ARG, "hnsw_build: ef_construction must be <= 512");
    REQUIRE(K <= 65535, HRNN_E_ARG, "hnsw_build: K must fit in 16 bits");
    REQUIRE(h->n < kInvalidId, HRNN_E_ARG, "hnsw_build: n must be below 2^32 - 1");
    REQUIRE(h->d >= 1 && h->d <= 4096, HRNN_E_ARG, "hnsw_build: d must be in [1, 4096]");
    const uint32_t n = (uint32_t)h->n, d = h->d;
    const uint32_t l0_words = 2 * M + 1, up_words = M + 1;
    const bool extend = b_first > 0;
    const uint32_t n_old = (uint32_t)b_first;
    if (extend) {
        REQUIRE(h->hnsw_rows == b_first && b_first < h->n, HRNN_E_STATE,
                "insert: the device graph must hold exactly the rows before the first new one");
        REQUIRE(h->l0_words == l0_words && h->up_words == up_words, HRNN_E_ARG,
                "insert: M differs from the resident graph's");
        REQUIRE(K == h->K && K > 0, HRNN_E_STATE, "insert: the resident KNN graph has another K");
    }
    // upper-layer blocks of the new rows from their levels, after the
    // resident pool; the top level over all rows
    std::vector<uint32_t> up_off(n - n_old, kInvalidId);
    uint64_t pool = extend ? h->up_pool_words : 0;
    const uint64_t pool_old = pool;
    int top = extend ? h->maxlevel : 0;
    for (uint32_t i = 0; i < n; i++)
        REQUIRE(levels[i] >= 0 && levels[i] <= 255, HRNN_E_ARG, "hnsw_build: level out of [0, 255]");
    for (uint32_t i = n_old; i < n; i++) {
        if (levels[i] > 0) {
            up_off[i - n_old] = (uint32_t)pool;
            pool += (uint64_t)levels[i] * up_words;
        }
        top = std::max(top, levels[i]);
    }
    REQUIRE(pool < kInvalidId, HRNN_E_ARG, "hnsw_build: upper layers exceed 2^32 words");
    REQUIRE(top < 100, HRNN_E_ARG, "hnsw_build: more than 100 layers");
    const uint32_t prop_stride = M * (uint32_t)(top + 1);
    uint32_t frac = (uint32_t)std::max(1, h->tune_hnsw_batch_permille);
    const uint32_t seed_exp = (uint32_t)std::max(0, h->tune_hnsw_seed_exp);
    const uint32_t bmax = 65536;

    hipStream_t st = h->stream;
    if (extend) {
        grow_keep(h, h->l0, (uint64_t)n_old * l0_words * 4, (uint64_t)n * l0_words * 4);
        grow_keep(h, h->up_off, (uint64_t)n_old * 4, (uint64_t)n * 4);
        grow_keep(h, h->up_pool, pool_old * 4, std::max<uint64_t>(pool, 1) * 4);
        grow_keep(h, h->knn_ids, (uint64_t)n_old * K * 4, (uint64_t)n * K * 4);
        grow_keep(h, h->knn_d, (uint64_t)n_old * K * 4, (uint64_t)n * K * 4);
        grow_keep(h, h->knn_cnt, (uint64_t)n_old * 4, (uint64_t)n * 4);
        grow_keep(h, h->knn_new, (uint64_t)n_old * K, (uint64_t)n * K);
        HIPC(hipMemsetAsync(h->l0.as<uint32_t>() + (uint64_t)n_old * l0_words, 0,
                            (uint64_t)(n - n_old) * l0_words * 4, st));
        if (pool > pool_old)
            HIPC(hipMemsetAsync(h->up_pool.as<uint32_t>() + pool_old, 0, (pool - pool_old) * 4, st));
        HIPC(hipMemsetAsync(h->knn_cnt.as<uint32_t>() + n_old, 0, (uint64_t)(n - n_old) * 4, st));
        HIPC(hipMemsetAsync(h->knn_new.as<uint8_t>() + (uint64_t)n_old * K, 1, (uint64_t)(n - n_old) * K, st));
        h->knn_version++;
    } else {
        h->l0.reserve((uint64_t)n * l0_words * 4);
        h->up_off.reserve((uint64_t)n * 4);
        h->up_pool.reserve(std::max<uint64_t>(pool, 1) * 4);
        HIPC(hipMemsetAsync(h->l0.p, 0, (uint64_t)n * l0_words * 4, st));
        if (pool) HIPC(hipMemsetAsync(h->up_pool.p, 0, pool * 4, st));
        if (K > 0) {
            h->K = K;
            h->knn_ids.reserve((uint64_t)n * K * 4);
            h->knn_d.reserve((uint64_t)n * K * 4);
            h->knn_cnt.reserve((uint64_t)n * 4);
            h->knn_new.reserve((uint64_t)n * K);
            HIPC(hipMemsetAsync(h->knn_cnt.p, 0, (uint64_t)n * 4, st));
            HIPC(hipMemsetAsync(h->knn_new.p, 1, (uint64_t)n * K, st));
            h->knn_version++;
        }
    }
    h->levels_dev.reserve((uint64_t)n * 4);
    h2d(h, h->up_off.as<uint32_t>() + n_old, up_off.data(), (uint64_t)(n - n_old) * 4);
    h2d(h, h->levels_dev.p, levels, (uint64_t)n * 4);
    h->csr_valid = false;
    h->hnsw_valid = false;
    h->hnsw_rows = 0;
    // batch buffers: proposals, incoming links per node, touched lists per layer
    h->hb_prop.reserve((uint64_t)bmax * prop_stride * 8);
    h->hb_inc_cnt.reserve((uint64_t)n * 4);
    h->hb_inc.reserve((uint64_t)n * kIncMax * 8);
    h->hb_touched.reserve((uint64_t)bmax * M * 4);
    h->hb_ctr.reserve(64 * 8);
    HIPC(hipMemsetAsync(h->hb_inc_cnt.p, 0, (uint64_t)n * 4, st));
    HIPC(hipMemsetAsync(h->hb_ctr.p, 0, 64 * 8, st));
    // [0] expansions [1] evaluations [2] dropped links [3] [4] seed completion expansions, evaluations
    unsigned long long *ctr = h->hb_ctr.as<unsigned long long>();
    uint32_t *n_touched = reinterpret_cast<uint32_t *>(ctr + 8);
    uint32_t *err = h->counters.as<uint32_t>() + kErrWord;

    BuildGraph G;
    G.xf = h->x.as<float>();
    G.d = d;
    G.n = n;
    G.l0 = h->l0.as<uint32_t>();
    G.l0_words = l0_words;
    G.up_off = h->up_off.as<uint32_t>();
    G.up_pool = h->up_pool.as<uint32_t>();
    G.pool_words = pool;
    G.up_words = up_words;
    G.levels = h->levels_dev.as<int>();

    const bool use_u8 = h->base_u8 != 0;
    const size_t vbytes = use_u8 ? 1 : h->small_int ? 2 : 4;
    const int base_exact =
        (h->base_u8 && d <= 258 && d % 8 == 0 && (uint64_t)n * d * vbytes < 0xFFFFFFF0ull) ? 1 : 0;
    const uint32_t ef_cap = (ef + 3u) & ~3u;
    const uint32_t log_vis = build_log_vis(ef);
    const BeamLds lay(d, ef_cap, log_vis);
    REQUIRE(lay.total <= 65536, HRNN_E_ARG, "hnsw_build: LDS need exceeds 64 KiB");

    uint32_t entry = extend ? h->entry : 0;
    int maxlevel = extend ? h->maxlevel : levels[0];
    uint64_t batches = 0;
    for (uint32_t b = extend ? n_old : 1; b < n;) {
        const uint32_t B = std::min<uint32_t>(bmax, std::max<uint32_t>(1, (uint32_t)((uint64_t)b * frac / 1000)));
        const uint32_t b1 = std::min<uint32_t>(n, b + B), nb = b1 - b;
        int lb = 0;
        for (uint32_t i = b; i < b1; i++) lb = std::max(lb, std::min(levels[i], maxlevel));
#define LAUNCH(VT, XP)                                                                                       \
    hipLaunchKernelGGL((k_hnsw_insert<VT>), dim3(nb), dim3(64), lay.total, st, XP, G, b, b1, entry, maxlevel, ef, \
                       ef_cap, log_vis, M, base_exact, K, h->knn_ids.as<uint32_t>(), h->knn_d.as<float>(),     \
                       h->knn_cnt.as<uint32_t>(), h->hb_prop.as<unsigned long long>(), prop_stride, err, ctr)
        if (use_u8) {
            LAUNCH(uint8_t, h->xu8.as<uint8_t>());
        } else if (h->small_int) {
            LAUNCH(__hip_bfloat16, h->xb.as<__hip_bfloat16>());
        } else {
            LAUNCH(float, h->x.as<float>());
        }
#undef LAUNCH
        HIPC(hipGetLastError());
        for (int L = 0; L <= lb; L++) {
            const uint32_t cap = L ? M : 2 * M;
            uint32_t *nt = n_touched + L;
            hipLaunchKernelGGL(k_hnsw_scatter, dim3(ceil_div((uint64_t)nb * M, 256)), dim3(256), 0, st,
                               h->hb_prop.as<unsigned long long>(), nb, prop_stride, M, L, b, n,
                               h->hb_inc_cnt.as<uint32_t>(), h->hb_inc.as<unsigned long long>(),
                               h->hb_touched.as<uint32_t>(), nt, ctr + 2, err);
            hipLaunchKernelGGL(k_hnsw_link, dim3(std::min<uint64_t>((uint64_t)nb * M, n)), dim3(64), 0, st, G, L,
                               cap, h->hb_touched.as<uint32_t>(), nt, h->hb_inc_cnt.as<uint32_t>(),
                               h->hb_inc.as<unsigned long long>(), err);
            HIPC(hipGetLastError());
        }
        if (K > 0 && seed_exp > 0) {
#define LAUNCH(VT, XP)                                                                                       \
    hipLaunchKernelGGL((k_hnsw_seed_fix<VT>), dim3(nb), dim3(64), lay.total, st, XP, G, b, b1, ef, ef_cap,   \
                       log_vis, base_exact, K, seed_exp, h->knn_ids.as<uint32_t>(), h->knn_d.as<float>(),    \
                       h->knn_cnt.as<uint32_t>(), err, ctr)
            if (use_u8) {
                LAUNCH(uint8_t, h->xu8.as<uint8_t>());
            } else if (h->small_int) {
                LAUNCH(__hip_bfloat16, h->xb.as<__hip_bfloat16>());
            } else {
                LAUNCH(float, h->x.as<float>());
            }
#undef LAUNCH
            HIPC(hipGetLastError());
        }
        hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, st, n_touched, (uint32_t)(lb + 1));
        HIPC(hipGetLastError());
        for (uint32_t i = b; i < b1; i++)
            if (levels[i] > maxlevel) {
                maxlevel = levels[i];
                entry = i;
            }
        if (after) after(b, b1);
        b = b1;
        batches++;
    }
    HIPC(hipStreamSynchronize(st));
    unsigned long long hc[5] = {0, 0, 0, 0, 0};
    d2h(h, hc, ctr, sizeof(hc));
    check_device_error(h);
    h->l0_words = l0_words;
    h->up_words = up_words;
    h->up_pool_words = pool;
    h->entry = entry;
    h->maxlevel = maxlevel;
    h->hnsw_valid = true;
    h->hnsw_rows = n;
    if (stats) {
        stats->batches = batches;
        stats->expansions = hc[0];
        stats->evaluations = hc[1];
        stats->dropped_links = hc[2];
        stats->seed_expansions = hc[3];
        stats->seed_evaluations = hc[4];
        stats->ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    }
}

}  // namespace hrnn

using namespace hrnn;

extern "C" {

int hrnn_dev_hnsw_build(hrnn_dev_t h, const int *levels, uint32_t M, uint32_t ef_construction, uint32_t K,
                        uint32_t *entry_out, int *maxlevel_out, hrnn_hnsw_build_stats *stats) {
    if (!h) return HRNN_E_ARG;
    return guarded(h, [&] {
        REQUIRE(entry_out && maxlevel_out, HRNN_E_ARG, "hnsw_build: null argument");
        hnsw_insert_rows(h, levels, 0, M, ef_construction, K, stats, nullptr);
        *entry_out = h->entry;
        *maxlevel_out = h->maxlevel;
    });
}

int hrnn_dev_get_hnsw(hrnn_dev_t h, uint32_t *l0, uint32_t *up_pool, uint64_t pool_words) {
    if (!h) return HRNN_E_ARG;
    return guarded(h, [&] {
        REQUIRE(h->hnsw_valid, HRNN_E_STATE, "get_hnsw: no graph on the device");
        REQUIRE(l0, HRNN_E_ARG, "get_hnsw: null argument");
        REQUIRE(pool_words == h->up_pool_words && (pool_words == 0 || up_pool), HRNN_E_ARG,
                "get_hnsw: pool size differs from the device's");
        d2h(h, l0, h->l0.p, h->n * h->l0_words * 4);
        if (pool_words) d2h(h, up_pool, h->up_pool.p, pool_words * 4);
    });
}

}  // extern "C"
