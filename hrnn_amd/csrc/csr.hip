// csr.hip -- (b) reverse-KNN CSR: integer histogram, exclusive scan, scatter and
// per-list ordering, bit-exact with the canonical (rank, src) order.
//
// Reference: HRNN::buildRKNNG (include/hnswlib/hrnn.h:1869-1909) and
// buildRKNNG_fill (hrnn.h:1912-1935).  The reference fills lists in (src, rank)
// order and std::sort-s each list by rank; this build produces each list in
// (rank, src) order, which is what a stable sort by rank gives.
//
// HBM layout: knn_ids u32[n*K] (row i = node i's ranked list), knn_cnt u32[n];
// csr_off u64[n+1]; csr_ent u32 or u64[E] with E = sum(cnt).
//
// Kernels (one wave of 64 lanes per source row for the row passes):
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
    h->wide = (nb + rb > 32) ? 1 : 0;
    h->rank_bits = rb;

    h->csr_off.reserve((n + 1) * 8);
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
