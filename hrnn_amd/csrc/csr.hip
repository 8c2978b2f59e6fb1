// csr.hip -- (b) reverse-KNN CSR: integer histogram, exclusive scan and stable
// scatter, bit-exact with the canonical (rank, src) order.
//
// Reference: HRNN::buildRKNNG (include/hnswlib/hrnn.h:1869-1909) and
// buildRKNNG_fill (hrnn.h:1912-1935).  The reference fills lists in (src, rank)
// order and std::sort-s each list by rank; this build produces each list in
// (rank, src) order, which is what a stable sort by rank gives.
//
// HBM layout: knn_ids u32[n*K] (row i = node i's ranked list), knn_cnt u32[n];
// csr_off u64[n+1]; csr_ent u32 or u64[E] with E = sum(cnt).
//
// u32 entries (node_bits + rank_bits <= 32): the stable radix sort below
// (k_rs_hist, k_rs_scatter, k_rs_offsets).  u64 entries, the fallback:
// (one wave of 64 lanes per source row for the row passes):
//   k_degree    deg[dst] += 1 for every (src, rank)            reads 4*E B
//   k_scan_*    offsets = exclusive scan of deg (u64)           reads/writes 12*n B
//   k_scatter   tmp[off[dst] + atomic slot] = (rank<<nb)|src   reads 4*E, writes 4*E
//   k_seg_sort  per list: sort keys, repack (src<<rb)|rank      reads 4*E, writes 4*E
// Lists longer than the LDS capacity are sorted by k_seg_sort_global.
#include "handle.h"

namespace hrnn {

__global__ void k_degree(const uint32_t *__restrict__ ids, const uint32_t *__restrict__ cnt,
                         uint64_t n, uint32_t K, uint32_t *__restrict__ deg) {
    uint64_t row = (uint64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x / 64);
    if (row >= n) return;
    uint32_t lane = threadIdx.x & 63u;
    uint32_t c = cnt[row];
    const uint32_t *r = ids + row * K;
    for (uint32_t j = lane; j < c; j += 64) atomicAdd(&deg[r[j]], 1u);
}

// Block-level exclusive scan of u32 degrees into u64 partial offsets.
constexpr int kScanBlock = 1024;
constexpr int kScanItems = 4;   // per thread -> 4096 per block

__global__ void k_scan_blocks(const uint32_t *__restrict__ deg, uint64_t n,
                              uint64_t *__restrict__ out, uint64_t *__restrict__ block_sums) {
    __shared__ uint64_t warp_tot[kScanBlock / 64];
    uint64_t base = (uint64_t)blockIdx.x * kScanBlock * kScanItems;
    uint64_t v[kScanItems];
    uint64_t local = 0;
#pragma unroll
    for (int t = 0; t < kScanItems; t++) {
        uint64_t i = base + (uint64_t)threadIdx.x * kScanItems + t;
        v[t] = (i < n) ? deg[i] : 0;
        local += v[t];
    }
    // inclusive scan of 'local' within the wave
    uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x / 64;
    uint64_t incl = local;
    for (int off = 1; off < 64; off <<= 1) {
        uint64_t y = __shfl_up(incl, off, 64);
        if (lane >= (uint32_t)off) incl += y;
    }
    if (lane == 63) warp_tot[wid] = incl;
    __syncthreads();
    if (wid == 0) {
        uint64_t w = (lane < kScanBlock / 64) ? warp_tot[lane] : 0;
        uint64_t wi = w;
        for (int off = 1; off < 64; off <<= 1) {
            uint64_t y = __shfl_up(wi, off, 64);
            if (lane >= (uint32_t)off) wi += y;
        }
        if (lane < kScanBlock / 64) warp_tot[lane] = wi - w;   // exclusive
        if (lane == kScanBlock / 64 - 1) block_sums[blockIdx.x] = wi;
    }
    __syncthreads();
    uint64_t run = warp_tot[wid] + incl - local;
#pragma unroll
    for (int t = 0; t < kScanItems; t++) {
        uint64_t i = base + (uint64_t)threadIdx.x * kScanItems + t;
        if (i < n) out[i] = run;
        run += v[t];
    }
}

// Serial-over-blocks exclusive scan of the block sums by one block.
__global__ void k_scan_sums(uint64_t *__restrict__ block_sums, uint32_t nb, uint64_t *__restrict__ total) {
    __shared__ uint64_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    uint32_t lane = threadIdx.x & 63u;
    for (uint32_t base = 0; base < nb; base += 64) {
        if (threadIdx.x < 64) {
            uint32_t i = base + lane;
            uint64_t v = (i < nb) ? block_sums[i] : 0;
            uint64_t incl = v;
            for (int off = 1; off < 64; off <<= 1) {
                uint64_t y = __shfl_up(incl, off, 64);
                if (lane >= (uint32_t)off) incl += y;
            }
            uint64_t c = carry;
            if (i < nb) block_sums[i] = c + incl - v;
            uint64_t last = __shfl(incl, 63, 64);
            if (lane == 0) carry = c + last;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = carry;
}

__global__ void k_scan_add(uint64_t *__restrict__ out, uint64_t n,
                           const uint64_t *__restrict__ block_sums, const uint64_t *__restrict__ total) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] += block_sums[i / (kScanBlock * kScanItems)];
    if (i == 0) out[n] = *total;
}

uint32_t scan_block_count(uint64_t n) { return ceil_div(n, kScanBlock * kScanItems); }

void scan_u32(hrnn_dev_t h, const uint32_t *cnt, uint64_t n, uint64_t *off, uint64_t *bsum,
              uint64_t *total) {
    uint32_t nblk = scan_block_count(n);
    hipLaunchKernelGGL(k_scan_blocks, dim3(nblk), dim3(kScanBlock), 0, h->stream, cnt, n, off, bsum);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(64), 0, h->stream, bsum, nblk, total);
    hipLaunchKernelGGL(k_scan_add, dim3(ceil_div(n, 256)), dim3(256), 0, h->stream, off, n, bsum, total);
}

template <class KeyT>
__global__ void k_scatter(const uint32_t *__restrict__ ids, const uint32_t *__restrict__ cnt,
                          uint64_t n, uint32_t K, const uint64_t *__restrict__ off,
                          uint32_t *__restrict__ fill, uint32_t node_bits, KeyT *__restrict__ tmp) {
    uint64_t row = (uint64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x / 64);
    if (row >= n) return;
    uint32_t lane = threadIdx.x & 63u;
    uint32_t c = cnt[row];
    const uint32_t *r = ids + row * K;
    for (uint32_t j = lane; j < c; j += 64) {
        uint32_t dst = r[j];
        uint32_t slot = atomicAdd(&fill[dst], 1u);
        tmp[off[dst] + slot] = ((KeyT)j << node_bits) | (KeyT)row;
    }
}

// One block per list; lists of length <= cap are sorted in LDS.
template <class KeyT>
__global__ void k_seg_sort(const uint64_t *__restrict__ off, uint64_t n, const KeyT *__restrict__ tmp,
                           KeyT *__restrict__ out, uint32_t node_bits, uint32_t rank_bits,
                           uint32_t cap, uint32_t *__restrict__ big_list, uint32_t *__restrict__ n_big) {
    extern __shared__ unsigned char smem_raw[];
    KeyT *keys = reinterpret_cast<KeyT *>(smem_raw);
    for (uint64_t list = blockIdx.x; list < n; list += gridDim.x) {
        uint64_t b = off[list], e = off[list + 1];
        uint32_t L = (uint32_t)(e - b);
        if (L == 0) continue;
        if (L > cap) {
            if (threadIdx.x == 0) {
                uint32_t s = atomicAdd(n_big, 1u);
                big_list[s] = (uint32_t)list;
            }
            continue;
        }
        uint32_t P = 1;
        while (P < L) P <<= 1;
        for (uint32_t t = threadIdx.x; t < P; t += blockDim.x) keys[t] = (t < L) ? tmp[b + t] : (KeyT)~(KeyT)0;
        bitonic_sort_lds(keys, P);
        const KeyT nmask = ((KeyT)1 << node_bits) - 1;
        for (uint32_t t = threadIdx.x; t < L; t += blockDim.x) {
            KeyT k = keys[t];
            KeyT src = k & nmask, rank = k >> node_bits;
            out[b + t] = (src << rank_bits) | rank;
        }
        __syncthreads();
    }
}

// Long lists: bitonic sort in global scratch (padded to a power of two) by one block.
template <class KeyT>
__global__ void k_seg_sort_global(const uint64_t *__restrict__ off, const uint32_t *__restrict__ big_list,
                                  const KeyT *__restrict__ tmp, KeyT *__restrict__ out, KeyT *scratch,
                                  uint32_t node_bits, uint32_t rank_bits, uint64_t scratch_stride) {
    uint32_t list = big_list[blockIdx.x];
    uint64_t b = off[list], e = off[list + 1];
    uint32_t L = (uint32_t)(e - b);
    uint32_t P = 1;
    while (P < L) P <<= 1;
    KeyT *keys = scratch + (uint64_t)blockIdx.x * scratch_stride;
    for (uint32_t t = threadIdx.x; t < P; t += blockDim.x) keys[t] = (t < L) ? tmp[b + t] : (KeyT)~(KeyT)0;
    for (uint32_t size = 2; size <= P; size <<= 1) {
        for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
            __syncthreads();
            for (uint32_t t = threadIdx.x; t < (P >> 1); t += blockDim.x) {
                uint32_t lo = 2 * t - (t & (stride - 1));
                uint32_t hi = lo + stride;
                bool up = ((lo & size) == 0);
                KeyT a = keys[lo], c = keys[hi];
                if ((a > c) == up) {
                    keys[lo] = c;
                    keys[hi] = a;
                }
            }
        }
    }
    __syncthreads();
    const KeyT nmask = ((KeyT)1 << node_bits) - 1;
    for (uint32_t t = threadIdx.x; t < L; t += blockDim.x) {
        KeyT k = keys[t];
        out[b + t] = ((k & nmask) << rank_bits) | (k >> node_bits);
    }
}

template <class KeyT>
static void csr_fill(hrnn_dev_t h, uint32_t node_bits, uint32_t rank_bits) {
    uint64_t n = h->n, E = h->n_ent;
    uint32_t K = h->K;
    h->csr_ent.reserve(std::max<uint64_t>(E, 1) * sizeof(KeyT));
    h->s1.reserve(std::max<uint64_t>(E, 1) * sizeof(KeyT));   // unsorted keys
    h->s0.reserve(n * 4 + 64);                                // fill counters
    uint32_t *fill = h->s0.as<uint32_t>();
    HIPC(hipMemsetAsync(fill, 0, n * 4, h->stream));
    unsigned rows_blocks = ceil_div(n, 4);
    hipLaunchKernelGGL(k_scatter<KeyT>, dim3(rows_blocks), dim3(256), 0, h->stream,
                       h->knn_ids.as<uint32_t>(), h->knn_cnt.as<uint32_t>(), n, K,
                       h->csr_off.as<uint64_t>(), fill, node_bits, h->s1.as<KeyT>());
    HIPC(hipGetLastError());

    uint32_t cap = (uint32_t)(std::min<size_t>(h->lds_per_block, 65536) / sizeof(KeyT));
    uint32_t capp = 1;
    while (capp * 2 <= cap) capp <<= 1;
    cap = capp;
    uint32_t *nbig = h->counters.as<uint32_t>() + 8;
    HIPC(hipMemsetAsync(nbig, 0, 4, h->stream));
    h->s2.reserve(n * 4 + 64);
    uint32_t *big = h->s2.as<uint32_t>();
    unsigned grid = (unsigned)std::min<uint64_t>(n, (uint64_t)h->n_cu * 64);
    hipLaunchKernelGGL(k_seg_sort<KeyT>, dim3(grid), dim3(256), cap * sizeof(KeyT), h->stream,
                       h->csr_off.as<uint64_t>(), n, h->s1.as<KeyT>(), h->csr_ent.as<KeyT>(),
                       node_bits, rank_bits, cap, big, nbig);
    HIPC(hipGetLastError());
    uint32_t n_big = 0;
    d2h(h, &n_big, nbig, 4);
    if (n_big) {
        // process long lists in batches bounded by scratch size
        std::vector<uint32_t> lists(n_big);
        d2h(h, lists.data(), big, (uint64_t)n_big * 4);
        std::vector<uint64_t> offs(n + 1);
        d2h(h, offs.data(), h->csr_off.p, (n + 1) * 8);
        uint64_t maxP = 1;
        for (uint32_t l : lists) {
            uint64_t L = offs[l + 1] - offs[l];
            while (maxP < L) maxP <<= 1;
        }
        uint32_t batch = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(n_big, (1ull << 31) / (maxP * sizeof(KeyT))));
        h->s3.reserve((uint64_t)batch * maxP * sizeof(KeyT));
        for (uint32_t s = 0; s < n_big; s += batch) {
            uint32_t nb = std::min(batch, n_big - s);
            hipLaunchKernelGGL(k_seg_sort_global<KeyT>, dim3(nb), dim3(1024), 0, h->stream,
                               h->csr_off.as<uint64_t>(), big + s, h->s1.as<KeyT>(),
                               h->csr_ent.as<KeyT>(), h->s3.as<KeyT>(), node_bits, rank_bits, maxP);
            HIPC(hipGetLastError());
        }
    }
}

// ---- u32 entries: stable LSD radix sort on (dst, rank) ------------------------
//
// Edges enumerated row-major (src ascending, then rank) are sorted stably by
// key = (dst << rank_bits) | rank, carrying src: each list then holds its
// entries in (rank, src) order with no per-list sort and no global atomics.
// 8-bit digits; each pass = a per-tile LDS histogram, one exclusive scan over
// [digit][tile] counts, and a scatter that sorts the tile by digit in LDS
// (stable: wave ballots over the 8 digit bits, per-wave counts) and writes
// it out digit run by digit run.  The
// first pass reads the KNN rows (entries past a row's count are dropped), the
// last writes the CSR entries (src << rank_bits) | rank and the sorted keys,
// from which the offsets follow.
constexpr int kRsThreads = 256;
constexpr int kRsRounds = 16;                         // rounds of 256 elements per tile
constexpr int kRsTile = kRsThreads * kRsRounds;       // 4096 elements
constexpr uint32_t kRsBins = 256;

// Element t of the row-major enumeration of rows [row0, ...): its key and
// value; false for an entry past its row's count (or past the last row).
__device__ __forceinline__ bool rs_edge(const uint32_t *__restrict__ ids, const uint32_t *__restrict__ cnt,
                                        uint64_t n, uint32_t K, uint32_t rb, uint64_t row0, uint32_t t,
                                        uint32_t *key, uint32_t *val) {
    const uint64_t row = row0 + t / K;
    const uint32_t j = t % K;
    if (row >= n || j >= cnt[row]) return false;
    *key = (ids[row * K + j] << rb) | j;
    *val = (uint32_t)row;
    return true;
}

// Per-tile digit histogram, written digit-major: hist[digit * tiles + tile].
// kRows: tiles of `rows` KNN rows (first pass); otherwise tiles of kRsTile
// consecutive elements of keys[0..E).
template <bool kRows>
__global__ __launch_bounds__(kRsThreads) void k_rs_hist(const uint32_t *__restrict__ ids,
                                                        const uint32_t *__restrict__ cnt, uint64_t n,
                                                        uint32_t K, uint32_t rb, uint32_t rows,
                                                        const uint32_t *__restrict__ keys, uint64_t E,
                                                        uint32_t shift, uint32_t tiles,
                                                        uint32_t *__restrict__ hist) {
    __shared__ uint32_t h[kRsBins];
    h[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t tile = blockIdx.x;
    if (kRows) {
        const uint64_t row0 = (uint64_t)tile * rows;
        const uint32_t span = rows * K;
        for (uint32_t t = threadIdx.x; t < span; t += kRsThreads) {
            uint32_t key, val;
            if (rs_edge(ids, cnt, n, K, rb, row0, t, &key, &val)) atomicAdd(&h[(key >> shift) & 255u], 1u);
        }
    } else {
        const uint64_t b = (uint64_t)tile * kRsTile;
        for (uint32_t t = threadIdx.x; t < (uint32_t)kRsTile; t += kRsThreads) {
            const uint64_t i = b + t;
            if (i < E) atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
        }
    }
    __syncthreads();
    hist[(uint64_t)threadIdx.x * tiles + tile] = h[threadIdx.x];
}

// Stable scatter of one pass.  goff: exclusive scan of hist (digit-major).
// kLast: write CSR entries to out_ent and sorted keys to out_key; otherwise
// keys and values to out_key / out_val.
//
// The tile is first sorted by digit in LDS, then written out digit run by
// digit run, so consecutive lanes store to consecutive addresses (runs of
// ~16 elements per digit and tile instead of 256 scattered stores per
// round).  Local ranks: wave w ranks its contiguous quarter of the tile in
// rounds of 64 (lanes with the same digit found by 8 ballots), keeping its
// running per-digit counts in LDS without barriers (one wave's LDS accesses
// are performed in order); one barrier later the tile's digit offsets and
// the counts of the earlier waves give every element its place.
template <bool kRows, bool kLast>
__global__ __launch_bounds__(kRsThreads) void k_rs_scatter(
    const uint32_t *__restrict__ ids, const uint32_t *__restrict__ cnt, uint64_t n, uint32_t K, uint32_t rb,
    uint32_t rows, const uint32_t *__restrict__ in_key, const uint32_t *__restrict__ in_val, uint64_t E,
    uint32_t shift, uint32_t tiles, const uint64_t *__restrict__ goff, uint32_t *__restrict__ out_key,
    uint32_t *__restrict__ out_val, uint32_t *__restrict__ out_ent) {
    constexpr int kWaves = kRsThreads / 64;
    constexpr int kPerWave = kRsTile / kWaves;   // 1024 consecutive elements per wave
    constexpr int kR = kPerWave / 64;            // rounds of 64
    __shared__ uint32_t s_key[kRsTile], s_val[kRsTile];
    __shared__ uint32_t wc[kWaves][kRsBins];     // per-wave running digit counts
    __shared__ uint32_t loff[kRsBins + 1];       // tile-local start of each digit; [256] = tile total
    __shared__ uint64_t gbase[kRsBins];          // global start of each digit's run of this tile
    __shared__ uint32_t wsum[kWaves];
    const uint32_t tile = blockIdx.x, lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
#pragma unroll
    for (int w = 0; w < kWaves; w++) wc[w][threadIdx.x] = 0;
    gbase[threadIdx.x] = goff[(uint64_t)threadIdx.x * tiles + tile];
    __syncthreads();
    const uint32_t rmask = (1u << rb) - 1u;
    const uint32_t span = kRows ? rows * K : (uint32_t)kRsTile;
    const uint64_t row0 = (uint64_t)tile * rows, b = (uint64_t)tile * kRsTile;
    uint32_t key[kR], val[kR], pos[kR];
    uint32_t vmask = 0;   // bit r: element of round r is valid
#pragma unroll
    for (int r = 0; r < kR; r++) {   // every load of the quarter in flight at once
        const uint32_t t = wave * kPerWave + r * 64 + lane;
        bool valid = false;
        key[r] = 0;
        val[r] = 0;
        if (t < span) {
            if (kRows) {
                valid = rs_edge(ids, cnt, n, K, rb, row0, t, &key[r], &val[r]);
            } else if (b + t < E) {
                valid = true;
                key[r] = in_key[b + t];
                val[r] = in_val[b + t];
            }
        }
        vmask |= valid ? (1u << r) : 0u;
    }
#pragma unroll
    for (int r = 0; r < kR; r++) {
        const bool valid = (vmask >> r) & 1u;
        const uint32_t dg = (key[r] >> shift) & 255u;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < 8; bit++) {
            const uint64_t bb = __ballot((dg >> bit) & 1u);
            peers &= ((dg >> bit) & 1u) ? bb : ~bb;
        }
        const uint32_t below = (uint32_t)__popcll(peers & lanemask_lt());
        uint32_t base = 0;
        if (valid) base = wc[wave][dg];
        pos[r] = base + below;
        if (valid && below == 0) wc[wave][dg] = base + (uint32_t)__popcll(peers);
    }
    __syncthreads();
    // tile-local digit offsets: exclusive scan of the per-digit totals
    {
        uint32_t tot = 0;
#pragma unroll
        for (int w = 0; w < kWaves; w++) tot += wc[w][threadIdx.x];
        uint32_t incl = tot;
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(incl, o, 64);
            if (lane >= (uint32_t)o) incl += y;
        }
        if (lane == 63) wsum[wave] = incl;
        __syncthreads();
        uint32_t woff = 0;
        for (uint32_t w = 0; w < wave; w++) woff += wsum[w];
        loff[threadIdx.x] = woff + incl - tot;
        if (threadIdx.x == kRsThreads - 1) loff[kRsBins] = woff + incl;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kR; r++) {
        if ((vmask >> r) & 1u) {
            const uint32_t dg = (key[r] >> shift) & 255u;
            uint32_t p = loff[dg] + pos[r];
            for (uint32_t w = 0; w < wave; w++) p += wc[w][dg];
            s_key[p] = key[r];
            s_val[p] = val[r];
        }
    }
    __syncthreads();
    const uint32_t total = loff[kRsBins];
    for (uint32_t i = threadIdx.x; i < total; i += kRsThreads) {
        const uint32_t k = s_key[i];
        const uint32_t dg = (k >> shift) & 255u;
        const uint64_t gp = gbase[dg] + (i - loff[dg]);
        if (kLast) {
            out_ent[gp] = (s_val[i] << rb) | (k & rmask);
            out_key[gp] = k;
        } else {
            out_key[gp] = k;
            out_val[gp] = s_val[i];
        }
    }
}

// Offsets from the sorted keys: off[v] = first position whose dst >= v.
__global__ void k_rs_offsets(const uint32_t *__restrict__ keys, uint64_t E, uint32_t rb, uint64_t n,
                             uint64_t *__restrict__ off) {
    const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p > E) return;
    const uint64_t hi = (p < E) ? (uint64_t)(keys[p] >> rb) : n;          // dst at p (n past the end)
    const uint64_t lo = (p > 0) ? (uint64_t)(keys[p - 1] >> rb) + 1 : 0;  // first dst not before p
    for (uint64_t v = lo; v <= hi && v <= n; v++) off[v] = p;
}

static void csr_radix(hrnn_dev_t h, uint32_t nb, uint32_t rb) {
    const uint64_t n = h->n;
    const uint32_t K = h->K;
    const uint32_t rows = std::max<uint32_t>(1, kRsTile / K);
    const uint32_t tiles0 = ceil_div(n, rows);
    const uint64_t Emax = n * K;
    // key/value ping-pong buffers and the histogram
    h->s1.reserve(Emax * 4 * 2 + 64);
    h->s2.reserve(Emax * 4 * 2 + 64);
    uint32_t *ka = h->s1.as<uint32_t>(), *va = ka + Emax;
    uint32_t *kb = h->s2.as<uint32_t>(), *vb = kb + Emax;
    const uint32_t tiles_max = std::max<uint32_t>(tiles0, ceil_div(Emax, kRsTile));
    h->s3.reserve((uint64_t)kRsBins * tiles_max * 4 + 64);
    h->s5.reserve((uint64_t)kRsBins * tiles_max * 8 + 64);
    h->s4.reserve((uint64_t)scan_block_count((uint64_t)kRsBins * tiles_max) * 8 + 64);
    uint32_t *hist = h->s3.as<uint32_t>();
    uint64_t *goff = h->s5.as<uint64_t>();
    uint64_t *bsum = h->s4.as<uint64_t>();
    uint64_t *total = h->counters.as<uint64_t>() + 16;
    const uint32_t key_bits = nb + rb;
    const uint32_t passes = std::max<uint32_t>(1, (key_bits + 7) / 8);
    h->csr_ent.reserve(std::max<uint64_t>(Emax, 1) * 4);
    const uint32_t *ids = h->knn_ids.as<uint32_t>(), *cn = h->knn_cnt.as<uint32_t>();
    uint64_t E = 0;
    const uint32_t *ik = nullptr, *iv = nullptr;
    for (uint32_t pass = 0; pass < passes; pass++) {
        const uint32_t shift = 8 * pass;
        const bool first = pass == 0, last = pass + 1 == passes;
        const uint32_t tiles = first ? tiles0 : ceil_div(E, kRsTile);
        if (first)
            hipLaunchKernelGGL(k_rs_hist<true>, dim3(tiles), dim3(kRsThreads), 0, h->stream, ids, cn, n, K, rb,
                               rows, nullptr, 0, shift, tiles, hist);
        else
            hipLaunchKernelGGL(k_rs_hist<false>, dim3(tiles), dim3(kRsThreads), 0, h->stream, nullptr, nullptr, n,
                               K, rb, rows, ik, E, shift, tiles, hist);
        HIPC(hipGetLastError());
        scan_u32(h, hist, (uint64_t)kRsBins * tiles, goff, bsum, total);
        if (first) {   // the number of edges
            d2h(h, &E, total, 8);
            h->n_ent = E;
            if (E == 0) {
                HIPC(hipMemsetAsync(h->csr_off.p, 0, (n + 1) * 8, h->stream));
                return;
            }
        }
        uint32_t *okey = (pass & 1) ? kb : ka, *oval = (pass & 1) ? vb : va;
#define RS_SCATTER(ROWS, LAST)                                                                          \
    hipLaunchKernelGGL((k_rs_scatter<ROWS, LAST>), dim3(tiles), dim3(kRsThreads), 0, h->stream, ids, cn, n, \
                       K, rb, rows, ik, iv, E, shift, tiles, goff, okey, oval, h->csr_ent.as<uint32_t>())
        if (first && last) RS_SCATTER(true, true);
        else if (first) RS_SCATTER(true, false);
        else if (last) RS_SCATTER(false, true);
        else RS_SCATTER(false, false);
#undef RS_SCATTER
        HIPC(hipGetLastError());
        ik = okey;
        iv = oval;
    }
    // offsets from the sorted keys (left in ik by the last pass)
    hipLaunchKernelGGL(k_rs_offsets, dim3(ceil_div(E + 1, 256)), dim3(256), 0, h->stream, ik, E, rb, n,
                       h->csr_off.as<uint64_t>());
    HIPC(hipGetLastError());
}

__global__ void k_max_len(const uint64_t *__restrict__ off, uint64_t n, unsigned long long *__restrict__ out) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long v = (i < n) ? (unsigned long long)(off[i + 1] - off[i]) : 0ull;
    for (int o = 32; o > 0; o >>= 1) {
        unsigned long long y = __shfl_down(v, o, 64);
        v = y > v ? y : v;
    }
    if ((threadIdx.x & 63u) == 0 && v) atomicMax(out, v);
}

uint64_t max_list_len(hrnn_dev_t h) {
    unsigned long long *mx = reinterpret_cast<unsigned long long *>(h->counters.as<uint64_t>() + 24);
    HIPC(hipMemsetAsync(mx, 0, 8, h->stream));
    hipLaunchKernelGGL(k_max_len, dim3(ceil_div(h->n, 256)), dim3(256), 0, h->stream,
                       h->csr_off.as<uint64_t>(), h->n, mx);
    HIPC(hipGetLastError());
    unsigned long long v = 0;
    d2h(h, &v, mx, 8);
    return (uint64_t)v;
}

void csr_build(hrnn_dev_t h) {
    uint64_t n = h->n;
    uint32_t K = h->K;
    uint32_t rb = 0, nb = 0;
    while ((1u << rb) < K) rb++;
    while ((1ull << nb) < n) nb++;
    REQUIRE(nb + rb <= 64, HRNN_E_RANGE, "buildRKNNG: exceeds 64-bit packing");
    // u64 entries when node and rank do not fit 32 bits, or when the
    // "csr_wide" tuning key asks for them (the same values, wider words)
    h->wide = (nb + rb > 32 || h->tune_csr_wide) ? 1 : 0;
    h->rank_bits = rb;

    h->csr_off.reserve((n + 1) * 8);
    if (!h->wide && K <= (uint32_t)kRsTile) {   // u32 entries, rows within a radix tile: the radix-sort build
        csr_radix(h, nb, rb);
        HIPC(hipStreamSynchronize(h->stream));
        h->max_deg = max_list_len(h);
        h->csr_valid = true;
        return;
    }
    h->s0.reserve(n * 4 + 64);
    uint32_t *deg = h->s0.as<uint32_t>();
    HIPC(hipMemsetAsync(deg, 0, n * 4, h->stream));
    unsigned rows_blocks = ceil_div(n, 4);
    hipLaunchKernelGGL(k_degree, dim3(rows_blocks), dim3(256), 0, h->stream,
                       h->knn_ids.as<uint32_t>(), h->knn_cnt.as<uint32_t>(), n, K, deg);
    HIPC(hipGetLastError());

    uint32_t nblk = ceil_div(n, kScanBlock * kScanItems);
    h->s4.reserve((uint64_t)nblk * 8 + 64);
    uint64_t *bsum = h->s4.as<uint64_t>();
    uint64_t *total = h->counters.as<uint64_t>() + 16;
    hipLaunchKernelGGL(k_scan_blocks, dim3(nblk), dim3(kScanBlock), 0, h->stream, deg, n,
                       h->csr_off.as<uint64_t>(), bsum);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(64), 0, h->stream, bsum, nblk, total);
    hipLaunchKernelGGL(k_scan_add, dim3(ceil_div(n, 256)), dim3(256), 0, h->stream,
                       h->csr_off.as<uint64_t>(), n, bsum, total);
    HIPC(hipGetLastError());
    uint64_t E = 0;
    d2h(h, &E, total, 8);
    h->n_ent = E;
    if (h->wide)
        csr_fill<uint64_t>(h, nb, rb);
    else
        csr_fill<uint32_t>(h, nb, rb);
    HIPC(hipStreamSynchronize(h->stream));
    h->max_deg = max_list_len(h);
    h->csr_valid = true;
}

}  // namespace hrnn
