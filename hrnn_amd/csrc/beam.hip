// beam.hip -- RkNN phase A on the GPU: HNSW greedy descent over the upper
// layers, then the layer-0 beam of width ef = max(ef_search, m), keeping the m
// closest nodes as proxies.
//
// Reference: searchRknn_impl phase A (include/hnswlib/hrnn.h:1992-2019) with
// searchBaseLayerST<true> (hrnn.h:485-616).  The result list is kept sorted
// by (dist, id); a node enters it when the list holds fewer than ef nodes or
// its distance is strictly below the worst; the closest unexpanded entry is
// expanded next; the search ends when every entry is expanded (= the closest
// open candidate is farther than the worst result).  Distances use the
// reference's 16-partial FMA order, so GPU and host beams agree up to the
// order of equal distances.
//
// k_phase_a: one query per workgroup of one wave (64 lanes), no state kept
// between workgroups.  The sorted list lives in LDS as 64-bit (dist, id) keys,
// double buffered, with expanded flags; visited nodes go into an LDS hash set
// of 1024-4096 ids sized by ef (triangular probing).  When a probe run is exhausted
// the node is treated as unseen and admission checks the list for it, so the
// result does not depend on the table size.  Per expansion the fresh
// neighbours' distances are computed by 16-lane groups, four rows per group
// in flight; the admissible ones are sorted across next_pow2(count) lanes and
// merged into the list with ranks counted by wave ballots (the first ef of the
// merged (dist, id) order: the same set a sequential insertion keeps).  The
// next node to expand is known before the merge is written (the smaller of
// the best candidate and the first open entry), so its layer-0 row is loaded
// while the merge runs.  When every base and query value is an integer in
// [0, 255] and d <= 258, every distance sum is an exact integer, so rows are
// read with 16-byte buffer loads (32-bit offsets, base below 4 GiB) and summed
// in any order; otherwise the reference's 16-partial order is used.  Every id
// or count read from the graph is range-checked; a bad one is skipped and
// raises the handle's device error word (reported by hrnn_dev_sync).  The
// descent and the beam run bounded loops.
#include <algorithm>
#include <cfloat>
#include <type_traits>

#include "beam_core.h"
#include "handle.h"
#include "mlp_core.h"

namespace hrnn {

// kPerQueryEf: the stage-2 form of the two-stage learned query (opt.ef_q
// gives each query's beam width; 0 skips the query).  A separate
// instantiation so profiles tell the two stages apart.
// Occupancy: 5 waves per SIMD (the register budget at which the compiler
// needs no spills).
template <class VT, bool kPerQueryEf>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(5))) void k_phase_a(
    const VT *__restrict__ x, uint32_t d, uint32_t n, const uint32_t *__restrict__ l0,
    uint32_t l0_words, const uint32_t *__restrict__ up_off, const uint32_t *__restrict__ up_pool,
    uint64_t pool_words, uint32_t up_words, uint32_t entry, int maxlevel, const float *__restrict__ q,
    uint32_t ef, uint32_t ef_cap, uint32_t log_vis, uint32_t m, int base_exact,
    uint32_t *__restrict__ proxies, uint32_t *__restrict__ expansions, float *__restrict__ raw,
    uint32_t *__restrict__ err, BeamOpt opt) {
    __shared__ float s_nd[kBeamMaxDeg];
    __shared__ uint32_t s_ni[kBeamMaxDeg];
    __shared__ unsigned long long s_ck[kBeamMaxDeg];   // admitted keys of one expansion
    __shared__ uint32_t s_lost, s_bad;
    extern __shared__ __align__(16) unsigned char s_dyn[];
    const BeamLds lay(d, ef_cap, log_vis);
    float *s_qv = reinterpret_cast<float *>(s_dyn);
    uint32_t *s_q8 = reinterpret_cast<uint32_t *>(s_dyn + lay.q8_off);
    uint32_t *s_vis = reinterpret_cast<uint32_t *>(s_dyn + lay.vis_off);
    // the sorted list as (dist, id) keys, double buffered, with expanded flags
    unsigned long long *s_kl = reinterpret_cast<unsigned long long *>(s_dyn + lay.k_off);
    uint8_t *s_xl = s_dyn + lay.x_off;

    const uint32_t lane = threadIdx.x;
    const uint32_t g = lane & 15u, grp = lane >> 4;
    uint32_t qi = blockIdx.x;   // uniform: kept in a scalar register
    if (blockIdx.x == 0 && threadIdx.x == 0) {   // the queues of the searches that follow
        if (opt.zero_a) *opt.zero_a = 0;
        if (opt.zero_b) *opt.zero_b = 0;
    }
    if (opt.qlist) {   // a query subset: workgroup i takes qlist[i]
        if (blockIdx.x >= *opt.qcount) return;
        qi = uni(opt.qlist[blockIdx.x]);
    }
    if (kPerQueryEf) {   // per-query beam width; 0 leaves this query to another launch
        const uint32_t e = opt.ef_q[qi];
        if (e == 0) return;
        ef = min(e, ef);
    }
    // The query is read once (16-byte reads where its row allows: it may sit
    // in page-locked host memory, read over the link): to LDS as FP32 and,
    // for u8 storage, packed 4 bytes per word with its sum of squares.
    bool q_u8 = true;
    uint32_t part = 0;
    const float *qrow = q + (uint64_t)qi * d;
    if ((d & 3u) == 0 && (reinterpret_cast<uintptr_t>(qrow) & 15u) == 0) {
        for (uint32_t w = lane; 4 * w < d; w += 64) {
            const float4 v = reinterpret_cast<const float4 *>(qrow)[w];
            reinterpret_cast<float4 *>(s_qv)[w] = v;
            if (opt.q_copy) reinterpret_cast<float4 *>(opt.q_copy + (uint64_t)qi * d)[w] = v;
            q_u8 &= is_u8_value(v.x) && is_u8_value(v.y) && is_u8_value(v.z) && is_u8_value(v.w);
            if constexpr (std::is_same<VT, uint8_t>::value) {
                const uint32_t b0 = clamp_u8(v.x), b1 = clamp_u8(v.y), b2 = clamp_u8(v.z), b3 = clamp_u8(v.w);
                s_q8[w] = b0 | (b1 << 8) | (b2 << 16) | (b3 << 24);
                part += b0 * b0 + b1 * b1 + b2 * b2 + b3 * b3;
            }
        }
    } else {
        for (uint32_t e = lane; e < d; e += 64) {
            const float v = qrow[e];
            s_qv[e] = v;
            if (opt.q_copy) opt.q_copy[(uint64_t)qi * d + e] = v;
            q_u8 &= is_u8_value(v);
        }
        if constexpr (std::is_same<VT, uint8_t>::value) {
            __syncthreads();   // s_qv complete (one wave)
            for (uint32_t w = lane; 4 * w < d; w += 64) {
                const uint32_t b0 = clamp_u8(s_qv[4 * w]), b1 = 4 * w + 1 < d ? clamp_u8(s_qv[4 * w + 1]) : 0u;
                const uint32_t b2 = 4 * w + 2 < d ? clamp_u8(s_qv[4 * w + 2]) : 0u;
                const uint32_t b3 = 4 * w + 3 < d ? clamp_u8(s_qv[4 * w + 3]) : 0u;
                s_q8[w] = b0 | (b1 << 8) | (b2 << 16) | (b3 << 24);
                part += b0 * b0 + b1 * b1 + b2 * b2 + b3 * b3;
            }
        }
    }
    // exact sums for this query: every base and query value an integer in [0, 255]
    const bool exact = base_exact && __ballot(!q_u8) == 0;
    uint32_t qq = 0;   // u8 storage: the query's sum q^2 (used only when exact)
    if constexpr (std::is_same<VT, uint8_t>::value) {
        for (int s2 = 32; s2 > 0; s2 >>= 1) part += (uint32_t)__shfl_xor((int)part, s2, 64);
        qq = uni(part);
    }
    // the base as a buffer (exact path): loads past its end return zeros
    const __amdgpu_buffer_rsrc_t xrs =
        make_rsrc(x, exact ? (uint32_t)((uint64_t)n * d * sizeof(VT)) : 0u);
    for (uint32_t t = lane; t < (1u << log_vis) / 4; t += 64)   // 16-byte stores
        reinterpret_cast<uint4 *>(s_vis)[t] = make_uint4(kInvalidId, kInvalidId, kInvalidId, kInvalidId);
    if (lane == 0) {
        s_lost = 0;
        s_bad = 0;
    }
    __syncthreads();

    // ---- greedy descent over the upper layers ----
    uint32_t cur = entry < n ? entry : 0;
    float cd;
    {
        const uint32_t id4[4] = {cur, cur, cur, cur};
        const bool v4[4] = {true, false, false, false};
        float o[4];
        beam_dist4<VT>(exact, s_qv, s_q8, qq, x, xrs, id4, v4, d, g, o);
        cd = uni_f(__shfl(o[0], 0, 64));
    }
    const float d_entry = cd;
    for (int level = maxlevel; level > 0; level--) {
        for (uint32_t step = 0; step < n; step++) {   // each move strictly lowers cd
            const uint32_t uo = up_off[cur];
            const uint64_t base = (uint64_t)uo + (uint64_t)(level - 1) * up_words;
            if (uo == kInvalidId || base + up_words > pool_words) {
                if (lane == 0) s_bad = 1u;
                break;
            }
            const uint32_t *ll = up_pool + base;
            const uint32_t w = (lane < up_words) ? ll[lane] : 0u;
            const uint32_t cnt = min((uint32_t)__shfl((int)w, 0, 64) & 0xFFFFu, up_words - 1u);
            float best = cd;
            uint32_t best_id = cur;
            for (uint32_t b = 0; b < cnt; b += 16) {
                uint32_t id4[4];
                bool v4[4];
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const uint32_t t = b + grp + 4u * (uint32_t)r;
                    uint32_t v = (uint32_t)__shfl((int)w, (int)(1u + min(t, cnt - 1u)), 64);
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
                // the host's sequential scan keeps the first strictly smaller
                // entry: the minimum of (dist, list position) over the chunk,
                // taken if it is below the running best
                unsigned long long mk = ~0ull;
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const unsigned long long kk =
                        ((unsigned long long)float_key(o[r]) << 32) | (b + grp + 4u * (uint32_t)r);
                    if (v4[r] && kk < mk) mk = kk;
                }
#pragma unroll
                for (int s2 = 16; s2 <= 32; s2 <<= 1) {   // across the four groups
                    const unsigned long long other = shfl_xor64(mk, s2);
                    mk = other < mk ? other : mk;
                }
                if (mk != ~0ull) {
                    const float ds = key_float((uint32_t)(mk >> 32));
                    if (ds < best) {
                        best = ds;
                        best_id = (uint32_t)__shfl((int)w, (int)(1u + (uint32_t)mk), 64);
                    }
                }
            }
            if (best_id == cur) break;
            cur = uni(best_id);
            cd = uni_f(best);
        }
    }

    // ---- layer-0 beam ----
    const float d_start = cd;
    if (lane == 0) {
        s_kl[0] = beam_key(cd, cur);
        s_xl[0] = 0;
        vis_insert(s_vis, log_vis, cur, &s_lost);
    }
    __syncthreads();
    uint32_t L = 1, buf = 0, n_exp = 0, n_eval = 1;
    // pick: position of the first unexpanded list entry (every entry before it
    // is expanded), kInvalidId when none is left
    uint32_t pick = 0;
    // layer-0 row of the node to be expanded next, loaded while the current
    // expansion merges its candidates
    uint32_t pf_node = kInvalidId, pf_w = 0, pf_nb = 0;
    const uint32_t max_exp = opt.mlp && opt.max_exp ? opt.max_exp : 0xFFFFFFFFu;
    for (uint32_t iter = 0; iter <= n && pick != kInvalidId && n_exp < max_exp; iter++) {   // one expansion per pass
        const unsigned long long *lk = s_kl + buf * ef_cap;
        uint8_t *lx = s_xl + buf * ef_cap;
        const uint32_t c = uni((uint32_t)lk[pick]);
        uint32_t w, nbw;
        if (c == pf_node) {
            w = pf_w;
            nbw = pf_nb;
        } else {
            l0_row_issue(l0, l0_words, c, lane, &w, &nbw);
        }
        const uint32_t cnt = l0_count(w, l0_words);
        __syncthreads();
        if (lane == 0) lx[pick] = 1;
        n_exp++;
        uint32_t nb = 0;
        bool fresh = false;
        if (lane < cnt) {
            nb = nbw;
            if (nb < n) fresh = vis_insert(s_vis, log_vis, nb, &s_lost);
            else s_bad = 1u;
        }
        const uint64_t fm = __ballot(fresh);
        const uint32_t nf = (uint32_t)__popcll(fm);
        n_eval += nf;
        if (fresh) s_ni[__popcll(fm & lanemask_lt())] = nb;
        __syncthreads();
        // the first unexpanded entry after pick
        uint32_t o = kInvalidId;
        for (uint32_t b0 = pick + 1; b0 < L; b0 += 64) {
            const uint32_t t = b0 + lane;
            const uint64_t mask = __ballot(t < L && lx[t] == 0);
            if (mask) {
                o = b0 + (uint32_t)__ffsll((long long)mask) - 1;
                break;
            }
        }
        if (nf == 0) {
            pick = o;
            continue;
        }
        // distances of the fresh neighbours, 4 kRows per pass (kRows per
        // 16-lane group; u8 rows take 8 at once, so a full layer-0 row of 32
        // needs one round trip)
        constexpr int kRows = std::is_same<VT, uint8_t>::value ? 8 : 4;
        for (uint32_t b = 0; b < nf; b += 4 * kRows) {
            uint32_t idr[kRows];
            bool vr[kRows];
#pragma unroll
            for (int r = 0; r < kRows; r++) {
                const uint32_t t = b + grp + 4u * (uint32_t)r;
                vr[r] = t < nf;
                idr[r] = s_ni[vr[r] ? t : b];
            }
            if constexpr (std::is_same<VT, float>::value) {
                if (!exact && (d & 15u) == 0) {   // float data: quad r of each group takes row r
                    const float s4 = dist4_quad_f32(s_qv, x, idr, vr, d, g);
                    const uint32_t t = b + grp + 4u * (g >> 2);
                    if ((g & 3u) == 0 && t < nf) s_nd[t] = s4;
                    continue;
                }
            }
            float orr[kRows];
            beam_dist4<VT, kRows>(exact, s_qv, s_q8, qq, x, xrs, idr, vr, d, g, orr);
            if (g == 0) {
#pragma unroll
                for (int r = 0; r < kRows; r++)
                    if (vr[r]) s_nd[b + grp + 4u * (uint32_t)r] = orr[r];
            }
        }
        __syncthreads();
        // candidates that can enter: the list is not full, or (dist, id) below
        // its worst entry; not already listed (checked only when the visited
        // set lost an insertion)
        unsigned long long key = ~0ull;
        if (lane < nf) {
            const unsigned long long kc = beam_key(s_nd[lane], s_ni[lane]);
            bool ok = L < ef || kc < lk[L - 1];
            if (ok && s_lost) {   // nor twice among this expansion's candidates
                for (uint32_t j = 0; j < L && ok; j++) ok = (uint32_t)lk[j] != (uint32_t)kc;
                for (uint32_t j = 0; j < lane && ok; j++) ok = s_ni[j] != (uint32_t)kc;
            }
            if (ok) key = kc;
        }
        const uint64_t am = __ballot(key != ~0ull);
        const uint32_t nv = (uint32_t)__popcll(am);
        if (nv == 0) {   // uniform
            pick = o;
            continue;
        }
        // admitted keys to lanes [0, nv), then sorted over next_pow2(nv) lanes
        if (key != ~0ull) s_ck[__popcll(am & lanemask_lt())] = key;
        __syncthreads();
        key = (lane < nv) ? s_ck[lane] : ~0ull;
        key = wave_sort64(key, lane, next_pow2_dev(nv));
        // the next expansion takes the smaller of the best admitted candidate
        // and the first unexpanded entry after pick; start its row load
        const unsigned long long kbest = read_lane64(key, 0);
        const unsigned long long ko = (o != kInvalidId) ? uni64(lk[o]) : ~0ull;
        pf_node = (uint32_t)(kbest < ko ? kbest : ko);
        l0_row_issue(l0, l0_words, pf_node, lane, &pf_w, &pf_nb);
        // merge the sorted candidates into the sorted list, keeping the first
        // ef: entry e moves up by the candidates below it; candidate j lands
        // after the entries below it
        const uint32_t nbuf = buf ^ 1u;
        unsigned long long *nk = s_kl + nbuf * ef_cap;
        uint8_t *nx = s_xl + nbuf * ef_cap;
        uint32_t below = 0;   // lane j < nv: list entries below candidate j
        for (uint32_t b0 = 0; b0 < L; b0 += 64) {
            const uint32_t e = b0 + lane;
            const bool valid = e < L;
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
            if (valid && pos < ef) {
                nk[pos] = ke;
                nx[pos] = xe;
            }
        }
        if (lane < nv) {
            const uint32_t pos = lane + below;
            if (pos < ef) {
                nk[pos] = key;
                nx[pos] = 0;
            }
        }
        // position of the next expansion in the merged list: candidate 0's, or
        // entry o's, which no candidate precedes
        pick = (kbest < ko) ? (uint32_t)__builtin_amdgcn_readlane((int)below, 0) : o;
        __syncthreads();
        L = min(L + nv, ef);
        buf = nbuf;
    }
    // ---- output ----
    const unsigned long long *lk = s_kl + buf * ef_cap;
    uint32_t mq = opt.m_q ? min(opt.m_q[qi], m) : m;
    if (expansions && lane == 0) {
        expansions[2 * (uint64_t)qi] = n_exp;
        expansions[2 * (uint64_t)qi + 1] = n_eval;
    }
    if (raw) {   // per-query statistics for the learned m (kRawWords floats)
        float *r = raw + (uint64_t)qi * kRawWords;
        if (lane < 16) r[2 + lane] = key_float((uint32_t)(lk[lane < L ? lane : L - 1] >> 32));
        if (lane == 0) {
            r[0] = d_entry;
            r[1] = d_start;
            r[18] = (float)n_exp;
            r[19] = (float)L;
        }
    }
    if (opt.feats || opt.mlp) {
        // The learned-m features (the same definitions as the host's
        // HRNN::queryFeatures over this beam's statistics):
        //   0 log1p d(q,entry)   1 log1p d(q,start)   2..17 log1p of the 16
        //   closest list distances (the last one repeated past the list end)
        //   18..25 log1p r_k of list entries 1..8   26..28 mean / min / max of
        //   log1p r_k over entries 1..16   29 fraction of them with
        //   d(q,p) <= r_k(p)   30 log1p d1 - log1p r_k(p1)   31 log1p expansions
        // r_k(p) = the feat_k-th distance of p's ranked KNN list (3e38 when
        // the list is shorter).  Lanes 0..15 take one list entry each.
        const uint32_t j = lane & 15u;
        const float dj = key_float((uint32_t)(lk[j < L ? j : L - 1] >> 32));
        const float ld = log1pf(fminf(dj, 3.0e38f));
        float rk = 0.f, inside = 0.f, cntf = 0.f;
        if (j < L) {
            const uint32_t p = (uint32_t)lk[j];
            if (p < n) {
                const float v = (opt.feat_k - 1u < opt.knn_cnt[p]) ? opt.knn_d[(uint64_t)p * opt.K + opt.feat_k - 1u]
                                                                   : 3.0e38f;
                rk = log1pf(v);
                inside = (dj <= v) ? 1.f : 0.f;
                cntf = 1.f;
            } else {
                s_bad = 1u;
            }
        }
        // reductions over lanes 0..15 (each 16-lane group holds the same values)
        float sum = rk * cntf, ins = inside, cnt = cntf;
        float mn = cntf > 0.f ? rk : FLT_MAX, mx = cntf > 0.f ? rk : -FLT_MAX;
#pragma unroll
        for (int s2 = 8; s2 > 0; s2 >>= 1) {
            sum += __shfl_xor(sum, s2, 16);
            ins += __shfl_xor(ins, s2, 16);
            cnt += __shfl_xor(cnt, s2, 16);
            mn = fminf(mn, __shfl_xor(mn, s2, 16));
            mx = fmaxf(mx, __shfl_xor(mx, s2, 16));
        }
        const float rk1 = __shfl(rk, 0, 64);
        const float ld0 = __shfl(ld, 0, 64);
        const uint32_t t = lane & 31u;
        const float from_ld = __shfl(ld, (int)((t - 2u) & 15u), 64);
        const float from_rk = __shfl(rk, (int)((t - 18u) & 15u), 64);
        float f;
        if (t == 0) f = log1pf(d_entry);
        else if (t == 1) f = log1pf(d_start);
        else if (t < 18) f = from_ld;
        else if (t < 26) f = from_rk;
        else if (t == 26) f = cnt > 0.f ? sum / cnt : 0.f;
        else if (t == 27) f = cnt > 0.f ? mn : 0.f;
        else if (t == 28) f = cnt > 0.f ? mx : 0.f;
        else if (t == 29) f = cnt > 0.f ? ins / cnt : 0.f;
        else if (t == 30) f = ld0 - rk1;
        else f = log1pf((float)n_exp);
        if (opt.mu) f = (f - opt.mu[t]) / opt.sd[t];
        if (opt.feats && lane < (uint32_t)kFeatWords) opt.feats[(uint64_t)qi * kFeatWords + lane] = f;
        if (opt.mlp) {
            // The learned early stop's decision for this query: the MLP over
            // the (standardised) features, the cheapest class whose cumulative
            // probability reaches tau, its beam width for stage 2 (0: the
            // stage-1 proxies serve) and proxy count; the query joins the
            // escalated list or the list of the rest.
            // scratch in the visited set's LDS (the beam is done with it; it
            // holds >= 512 words, the network needs 320)
            float *m_x = reinterpret_cast<float *>(s_vis), *m_z1 = m_x + kIn, *m_h1 = m_z1 + kH,
                  *m_z2 = m_h1 + kH, *m_h2 = m_z2 + kH, *m_lg = m_h2 + kH, *m_p = m_lg + kOut;
            static_assert(kIn + 4 * kH + 2 * kOut <= 512, "MLP scratch exceeds the smallest visited set");
            __syncthreads();
            if (lane < (uint32_t)kIn) m_x[lane] = f;
            __syncthreads();
            mlp_forward_wave(opt.mlp, m_x, m_z1, m_h1, m_z2, m_h2, m_lg, m_p, lane);
            uint32_t c = uni(pick_class_of(m_p, opt.tau));
            if (pick != kInvalidId && opt.classes[c] <= opt.ef1) {   // stopped by the cap
                uint32_t e = c;
                while (e + 1 < (uint32_t)kOut && opt.classes[e] <= opt.ef1) e++;
                c = uni(e);
            }
            const uint32_t cef = opt.classes[c];
            mq = min(max(opt.classes[kOut + c], 1u), m);
            if (lane == 0) {
                const bool esc = cef > opt.ef1;
                opt.cls[qi] = c;
                opt.ef_q_out[qi] = esc ? cef : 0u;
                opt.m_q_out[qi] = mq;
                if (opt.exp2 && !esc) {
                    opt.exp2[2 * (uint64_t)qi] = 0;
                    opt.exp2[2 * (uint64_t)qi + 1] = 0;
                }
                opt.lists[(uint64_t)(esc ? 0u : 1u) * opt.nq + atomicAdd(&opt.list_cnt[esc ? 0 : 1], 1u)] =
                    qi;
            }
        }
    }
    for (uint32_t t = lane; t < m; t += 64)
        proxies[(uint64_t)qi * m + t] = (t < L && t < mq) ? (uint32_t)lk[t] : kInvalidId;
    __syncthreads();
    if (lane == 0 && s_bad) atomicOr(err, 1u);
}

// Visited-set slots for a beam of width ef.  The ids evaluated per query grow
// slower than ef (about 370, 560, 900, 1,020 and 1,110 at ef = 16, 32, 64, 128
// and 200 on a 1M-point, M = 16 graph), so 1,024 slots up to ef = 16, 2,048 up
// to ef = 256 and 4,096 beyond keep the load near or below 1/2.  An
// overflowing probe run stays correct, only slower.  Harder data evaluate
// more (about 4,300 at ef 200 on 96-dim unit vectors), where a larger table
// trades occupancy (LDS per wave) against probe runs: the "beam_vis_shift"
// tuning key moves the rule's choice by a power of two, within 2^10 .. 2^13.
static uint32_t beam_log_vis(uint32_t ef) { return ef <= 16 ? 10u : ef <= 256 ? 11u : 12u; }

bool host_memory(const void *p) {
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return a.type == hipMemoryTypeHost;
}

void beam_launch(hrnn_dev_t h, const float *q_dev, uint32_t nq, uint32_t m, uint32_t ef,
                 uint32_t *proxies_dev, uint32_t *expansions_dev, float *raw_dev, const BeamOpt *opt) {
    REQUIRE(h->hnsw_valid, HRNN_E_STATE, "proxies: upload the HNSW graph first");
    REQUIRE(ef >= 1 && ef <= (uint32_t)kBeamMaxEf, HRNN_E_ARG, "proxies: ef must be in [1, 512]");
    REQUIRE(m >= 1 && m <= 1024, HRNN_E_ARG, "proxies: m must be in [1, 1024]");
    REQUIRE(h->l0_words >= 1 && h->l0_words - 1 <= (uint32_t)kBeamMaxDeg, HRNN_E_ARG,
            "proxies: layer-0 degree must be in [0, 64]");
    REQUIRE(h->up_words >= 1 && h->up_words <= 64, HRNN_E_ARG,
            "proxies: upper-layer degree must be in [0, 63]");
    REQUIRE(h->d >= 1 && h->d <= 4096, HRNN_E_ARG, "proxies: d must be in [1, 4096]");
    REQUIRE(q_dev && proxies_dev, HRNN_E_ARG, "proxies: null argument");
    BeamOpt o = opt ? *opt : BeamOpt{};
    REQUIRE(!o.q_copy || (!o.ef_q && !o.qlist), HRNN_E_ARG, "proxies: the query copy needs every query");
    REQUIRE(!o.mlp || (o.classes && o.cls && o.ef_q_out && o.m_q_out && o.lists && o.list_cnt && o.nq == nq &&
                       o.mu && o.sd),
            HRNN_E_ARG, "proxies: incomplete learned-stop arguments");
    if (o.feats || o.mlp) {
        REQUIRE(h->K > 0 && o.feat_k >= 1 && o.feat_k <= h->K, HRNN_E_ARG,
                "proxies: feature k must be in [1, K]");
        REQUIRE((o.mu == nullptr) == (o.sd == nullptr), HRNN_E_ARG, "proxies: give mu and sd together");
        o.knn_d = h->knn_d.as<float>();
        o.knn_cnt = h->knn_cnt.as<uint32_t>();
        o.K = h->K;
    }
    if (nq == 0) return;
    uint32_t *err = h->counters.as<uint32_t>() + kErrWord;
    // every sum exact in any order: values in [0, 255] and d <= 258, with the
    // row width the wide loads need (8 BF16 or 4 FP32 elements)
    const bool use_u8 = h->vec_u8 && h->base_u8;
    const size_t vbytes = use_u8 ? 1 : (h->small_int && h->vec_bf16) ? 2 : 4;
    const int base_exact = (h->base_u8 && h->d <= 258 && h->d % 8 == 0 &&
                            (uint64_t)h->n * h->d * vbytes < 0xFFFFFFF0ull) ? 1 : 0;
    const uint32_t ef_cap = (ef + 3u) & ~3u;
    const uint32_t log_vis =
        (uint32_t)std::min(13, std::max(10, (int)beam_log_vis(ef) + h->tune_beam_vis_shift));
    const BeamLds lay(h->d, ef_cap, log_vis);
    REQUIRE(lay.total <= 65536, HRNN_E_ARG, "proxies: LDS need exceeds 64 KiB");
    hipStream_t st = o.stream ? o.stream : h->stream;
#define LAUNCH1(VT, PQ, XP)                                                                       \
    hipLaunchKernelGGL((k_phase_a<VT, PQ>), dim3(nq), dim3(64), lay.total, st, XP, h->d,         \
                       (uint32_t)h->n, h->l0.as<uint32_t>(), h->l0_words, h->up_off.as<uint32_t>(), \
                       h->up_pool.as<uint32_t>(), h->up_pool_words, h->up_words, h->entry,        \
                       h->maxlevel, q_dev, ef, ef_cap, log_vis, m, base_exact, proxies_dev,       \
                       expansions_dev, raw_dev, err, o)
#define LAUNCH(VT, XP)                  \
    if (o.ef_q) LAUNCH1(VT, true, XP);  \
    else LAUNCH1(VT, false, XP)
    // u8 or BF16 mirror where it is exact (same distances, a quarter or half
    // of the bytes)
    if (use_u8) {
        LAUNCH(uint8_t, h->xu8.as<uint8_t>());
    } else if (h->small_int && h->vec_bf16) {
        LAUNCH(__hip_bfloat16, h->xb.as<__hip_bfloat16>());
    } else {
        LAUNCH(float, h->x.as<float>());
    }
#undef LAUNCH
#undef LAUNCH1
    HIPC(hipGetLastError());
}

}  // namespace hrnn

using namespace hrnn;

extern "C" {

int hrnn_dev_set_hnsw(hrnn_dev_t h, uint32_t maxM0, uint32_t maxM, const uint32_t *l0,
                      const int *levels, const uint32_t *up_off, const uint32_t *up_pool,
                      uint64_t pool_words, uint32_t entry, int maxlevel) {
    if (!h) return HRNN_E_ARG;
    return guarded(h, [&] {
        REQUIRE(h->n > 0, HRNN_E_STATE, "set_hnsw: set_base first");
        REQUIRE(l0 && levels && up_off && maxM0 >= 1 && maxM0 <= (uint32_t)kBeamMaxDeg && maxM >= 1,
                HRNN_E_ARG, "set_hnsw: bad arguments");
        REQUIRE(entry < h->n && maxlevel >= 0 && levels[entry] == maxlevel, HRNN_E_ARG,
                "set_hnsw: bad entry point");
        const uint64_t n = h->n;
        const uint32_t w0 = maxM0 + 1, wu = maxM + 1;
        // Validate every access the kernel can make: layer-0 ids < n; a node of
        // level L >= 1 owns L blocks inside the pool; a link at layer l points
        // to a node whose level is >= l.
        for (uint64_t i = 0; i < n; i++) {
            const uint32_t *ll = l0 + i * w0;
            uint32_t c = ll[0] & 0xFFFFu;
            REQUIRE(c <= maxM0, HRNN_E_ARG, "set_hnsw: layer-0 count > maxM0");
            for (uint32_t t = 0; t < c; t++) REQUIRE(ll[1 + t] < n, HRNN_E_ARG, "set_hnsw: link id >= n");
            int lv = levels[i];
            REQUIRE(lv >= 0 && lv <= maxlevel, HRNN_E_ARG, "set_hnsw: level out of range");
            if (lv == 0) continue;
            REQUIRE(up_pool && up_off[i] != kInvalidId && (uint64_t)up_off[i] + (uint64_t)lv * wu <= pool_words,
                    HRNN_E_ARG, "set_hnsw: upper block out of pool");
            for (int l = 1; l <= lv; l++) {
                const uint32_t *ul = up_pool + up_off[i] + (uint64_t)(l - 1) * wu;
                uint32_t uc = ul[0] & 0xFFFFu;
                REQUIRE(uc <= maxM, HRNN_E_ARG, "set_hnsw: upper count > maxM");
                for (uint32_t t = 0; t < uc; t++)
                    REQUIRE(ul[1 + t] < n && levels[ul[1 + t]] >= l, HRNN_E_ARG,
                            "set_hnsw: upper link to a node below that layer");
            }
        }
        h->l0.reserve(n * w0 * 4);
        h->up_off.reserve(n * 4);
        h->up_pool.reserve(std::max<uint64_t>(pool_words, 1) * 4);
        h2d(h, h->l0.p, l0, n * w0 * 4);
        h2d(h, h->up_off.p, up_off, n * 4);
        if (pool_words) h2d(h, h->up_pool.p, up_pool, pool_words * 4);
        h->l0_words = w0;
        h->up_words = wu;
        h->up_pool_words = pool_words;
        h->entry = entry;
        h->maxlevel = maxlevel;
        h->hnsw_valid = true;
        h->hnsw_rows = n;
    });
}

int hrnn_dev_proxies(hrnn_dev_t h, const float *q, uint32_t nq, uint32_t m, uint32_t ef,
                     uint32_t *proxies) {
    if (!h) return HRNN_E_ARG;
    return guarded(h, [&] {
        REQUIRE(q && proxies, HRNN_E_ARG, "proxies: null argument");
        REQUIRE(m >= 1 && m <= ef, HRNN_E_ARG, "proxies: m must be in [1, ef]");
        if (!nq) return;
        DevBuf dq, dp;
        dq.reserve((uint64_t)nq * h->d * 4);
        dp.reserve((uint64_t)nq * m * 4);
        h2d(h, dq.p, q, (uint64_t)nq * h->d * 4);
        beam_launch(h, dq.as<float>(), nq, m, ef, dp.as<uint32_t>(), nullptr);
        d2h(h, proxies, dp.p, (uint64_t)nq * m * 4);
        dq.release();
        dp.release();
        check_device_error(h);
    });
}

int hrnn_dev_proxies_dptr(hrnn_dev_t h, const float *q_dev, uint32_t nq, uint32_t m, uint32_t ef,
                          uint32_t *proxies_dev, uint32_t *expansions_dev) {
    if (!h) return HRNN_E_ARG;
    return guarded(h, [&] {
        REQUIRE(m >= 1 && m <= ef, HRNN_E_ARG, "proxies: m must be in [1, ef]");
        beam_launch(h, q_dev, nq, m, ef, proxies_dev, expansions_dev);
    });
}

int hrnn_dev_rknn_query_dptr(hrnn_dev_t h, const float *q_dev, uint32_t nq, uint32_t m,
                             uint32_t ef, const uint32_t *k_per_q_dev, uint32_t k,
                             uint32_t K_prime, float threshold, uint32_t *proxies_dev,
                             uint32_t *res_cnt_dev, uint32_t *res_ids_dev, uint32_t cap_per_q,
                             uint32_t *work_dev) {
    if (!h) return HRNN_E_ARG;
    return guarded(h, [&] {
        REQUIRE(h->csr_valid && h->K > 0, HRNN_E_STATE, "rknn_query: build the CSR first");
        REQUIRE(k_per_q_dev || (k >= 1 && k <= h->K), HRNN_E_ARG, "rknn_query: k out of [1, K]");
        if (nq == 0) return;
        const uint32_t efm = std::max(ef, m);
        // Queries in page-locked host memory are read over the link once: the
        // beam writes each row to a device copy, which the search reads.
        BeamOpt o1;
        const float *qs = q_dev;
        if (host_memory(q_dev)) {
            h->q_stage.reserve((uint64_t)nq * h->d * 4);
            o1.q_copy = h->q_stage.as<float>();
            qs = o1.q_copy;
        }
        // Each search's work queue is zeroed by the phase-A launch before it
        // on its stream (no separate fill command between the two kernels).
        SearchSel s1;
        s1.queue_zeroed = true;
        o1.zero_a = h->counters.as<uint32_t>() + kQueueWord;
        if (h->tune_query_streams != 2 || nq < 2048) {
            beam_launch(h, q_dev, nq, m, efm, proxies_dev, nullptr, nullptr, &o1);
            rknn_search_resident(h, qs, nq, proxies_dev, m, k_per_q_dev, k, K_prime, threshold,
                                 res_cnt_dev, res_ids_dev, cap_per_q, work_dev, &s1);
            return;
        }
        // Two parts (the first query_split permille of the batch, then the
        // rest), each phase A then phases B+C, on the handle's two streams:
        // the parts' kernels fill each other's idle slots, and with a smaller
        // first part its search runs beside the second part's phase A (the
        // same answers; the second part's search has its own queue word and
        // fallback buffers).  The main stream waits for both.
        ensure_stream2(h);
        if (!k_per_q_dev) rk_column_ready(h, k, h->stream);
        HIPC(hipEventRecord(h->ev_a, h->stream));
        HIPC(hipStreamWaitEvent(h->stream2, h->ev_a, 0));
        const uint32_t n0 = (uint32_t)((uint64_t)nq * (uint32_t)h->tune_query_split / 1000), n1 = nq - n0;
        const uint64_t d = h->d;
        beam_launch(h, q_dev, n0, m, efm, proxies_dev, nullptr, nullptr, &o1);
        BeamOpt o2;
        o2.stream = h->stream2;
        o2.zero_a = h->counters.as<uint32_t>() + kQueue2Word;
        if (o1.q_copy) o2.q_copy = o1.q_copy + n0 * d;
        beam_launch(h, q_dev + n0 * d, n1, m, efm, proxies_dev + (uint64_t)n0 * m, nullptr, nullptr, &o2);
        rknn_search_resident(h, qs, n0, proxies_dev, m, k_per_q_dev, k, K_prime, threshold, res_cnt_dev,
                             res_ids_dev, cap_per_q, work_dev, &s1);
        SearchSel s2;
        s2.stream = h->stream2;
        s2.queue_word = kQueue2Word;
        s2.queue_zeroed = true;
        s2.buffer_set = 1;
        rknn_search_resident(h, qs + n0 * d, n1, proxies_dev + (uint64_t)n0 * m,
                             m, k_per_q_dev ? k_per_q_dev + n0 : nullptr, k, K_prime, threshold,
                             res_cnt_dev + n0, res_ids_dev + (uint64_t)n0 * cap_per_q, cap_per_q,
                             work_dev ? work_dev + 2ull * n0 : nullptr, &s2);
        HIPC(hipEventRecord(h->ev_b, h->stream2));
        HIPC(hipStreamWaitEvent(h->stream, h->ev_b, 0));
    });
}

}  // extern "C"
