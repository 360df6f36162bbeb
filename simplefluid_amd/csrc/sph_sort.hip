// Stable LSD radix sort of (cell key, particle index) for gfx950.
//
// One histogram kernel computes the 256-bin digit counts of every pass at
// once; then one kernel per 8-bit digit ranks a tile of keys in LDS and finds
// its global offsets by decoupled look-back over earlier tiles (tiles are
// taken in order from an atomic ticket, so every tile a block waits on is
// already resident and running).  Ranking inside a wave uses ballot-matching
// of the eight digit bits across the 64 lanes, so equal digits are ranked in
// lane order without LDS atomics, which keeps the sort stable.
//
// Traffic per pass: read key + value (8 B), write key + value (8 B); the
// histogram kernel reads each key once for all passes.
#include "sph_internal.hpp"

namespace sph {

namespace {

constexpr int kBins = 256;
constexpr int kSortBlock = 256;
constexpr int kSortWaves = kSortBlock / 64;
constexpr uint32_t kFlagAgg = 1u << 30;
constexpr uint32_t kFlagInc = 2u << 30;
constexpr uint32_t kValMask = (1u << 30) - 1u;
constexpr uint32_t kSpinLimit = 1u << 22;
constexpr int kHistItems = 16;

__device__ __forceinline__ uint32_t ld_agent(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Exclusive scan of one value per thread over a 256-thread block.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* s_w) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) s_w[wave] = x;
    __syncthreads();
    uint32_t pre = 0;
#pragma unroll
    for (int w = 0; w < kSortWaves; ++w)
        if (w < wave) pre += s_w[w];
    __syncthreads();
    return pre + x - v;
}

// Digit histograms of all passes; zeroes the look-back status words and the
// tile tickets of this sort; clears the cell entries the previous step wrote
// (prev_sorted holds that step's sorted keys, prev_n of them).
__global__ __launch_bounds__(kSortBlock) void radix_histogram_kernel(
    const uint32_t* __restrict__ keys, uint32_t n, int passes, uint32_t* __restrict__ hist,
    uint32_t* __restrict__ status, uint32_t status_words, uint32_t* __restrict__ tickets,
    const uint32_t* __restrict__ prev_sorted, uint32_t prev_n, uint2* __restrict__ cells) {
    __shared__ uint32_t s_hist[kMaxPasses * kBins];
    for (int i = threadIdx.x; i < kMaxPasses * kBins; i += kSortBlock) s_hist[i] = 0;
    __syncthreads();

    // Each thread counts a contiguous run of keys and only flushes a digit to
    // LDS when it changes: sorted-ish input keeps the high digits constant.
    const uint32_t first = (blockIdx.x * kSortBlock + threadIdx.x) * kHistItems;
    uint32_t cur[kMaxPasses], cnt[kMaxPasses];
#pragma unroll
    for (int p = 0; p < kMaxPasses; ++p) { cur[p] = 0xffffffffu; cnt[p] = 0; }
    for (int q = 0; q < kHistItems; ++q) {
        const uint32_t i = first + q;
        if (i >= n) break;
        const uint32_t k = keys[i];
#pragma unroll
        for (int p = 0; p < kMaxPasses; ++p) {
            if (p >= passes) break;
            const uint32_t d = (k >> (8 * p)) & 255u;
            if (d != cur[p]) {
                if (cnt[p]) atomicAdd(&s_hist[p * kBins + cur[p]], cnt[p]);
                cur[p] = d;
                cnt[p] = 0;
            }
            cnt[p]++;
        }
    }
#pragma unroll
    for (int p = 0; p < kMaxPasses; ++p)
        if (p < passes && cnt[p]) atomicAdd(&s_hist[p * kBins + cur[p]], cnt[p]);

    const uint32_t stride = gridDim.x * kSortBlock;
    const uint32_t tid = blockIdx.x * kSortBlock + threadIdx.x;
    for (uint32_t i = tid; i < status_words; i += stride) status[i] = 0u;
    if (tid < (uint32_t)kMaxPasses) tickets[tid] = 0u;
    for (uint32_t i = tid; i < prev_n; i += stride) {
        const uint32_t k = prev_sorted[i];
        if (i == 0 || prev_sorted[i - 1] != k) cells[k] = make_uint2(0xffffffffu, 0u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < passes * kBins; i += kSortBlock) {
        const uint32_t v = s_hist[i];
        if (v) atomicAdd(&hist[i], v);
    }
}

template <int ITEMS>
__global__ __launch_bounds__(kSortBlock) void radix_pass_kernel(
    const uint32_t* __restrict__ keys_in, const uint32_t* __restrict__ vals_in,
    uint32_t* __restrict__ keys_out, uint32_t* __restrict__ vals_out,
    const uint32_t* __restrict__ hist, uint32_t* status, uint32_t* ticket, uint32_t n, int shift,
    uint32_t* err) {
    constexpr int TILE = kSortBlock * ITEMS;
    __shared__ uint32_t s_keys[TILE];
    __shared__ uint32_t s_vals[TILE];
    __shared__ uint32_t s_wcnt[kSortWaves][kBins];
    __shared__ uint32_t s_gbase[kBins];
    __shared__ uint32_t s_lstart[kBins];
    __shared__ uint32_t s_w[kSortWaves];
    __shared__ uint32_t s_tile;

    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    if (tid == 0) s_tile = atomicAdd(ticket, 1u);
#pragma unroll
    for (int w = 0; w < kSortWaves; ++w) s_wcnt[w][tid] = 0u;
    __syncthreads();
    const uint32_t tile = s_tile;
    const uint32_t tbase = tile * (uint32_t)TILE;

    uint32_t key[ITEMS], val[ITEMS], rank[ITEMS];
    const uint32_t wbase = tbase + (uint32_t)(wave * 64 * ITEMS) + (uint32_t)lane;
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
        const uint32_t idx = wbase + (uint32_t)(i * 64);
        const bool valid = idx < n;
        key[i] = valid ? keys_in[idx] : 0xffffffffu;
        val[i] = valid ? (vals_in ? vals_in[idx] : idx) : 0u;
    }
    const uint64_t lt_mask = (1ull << lane) - 1ull;
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
        const bool valid = wbase + (uint32_t)(i * 64) < n;
        const uint32_t d = (key[i] >> shift) & 255u;
        uint64_t m = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bb = __ballot(bit);
            m &= bit ? bb : ~bb;
        }
        const uint32_t pre = (uint32_t)__popcll(m & lt_mask);
        const uint32_t cnt = (uint32_t)__popcll(m);
        uint32_t base = 0;
        if (valid) base = s_wcnt[wave][d];
        __builtin_amdgcn_wave_barrier();
        rank[i] = base + pre;
        if (valid && pre == 0) s_wcnt[wave][d] = base + cnt;
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();

    // Thread tid owns digit tid from here on.
    uint32_t tot = 0;
#pragma unroll
    for (int w = 0; w < kSortWaves; ++w) {
        const uint32_t c = s_wcnt[w][tid];
        s_wcnt[w][tid] = tot;
        tot += c;
    }
    uint32_t* my = status + (size_t)tile * kBins + tid;
    st_agent(my, (tile == 0 ? kFlagInc : kFlagAgg) | tot);

    const uint32_t lstart = block_exclusive_scan(tot, s_w);
    const uint32_t hbase = block_exclusive_scan(hist[tid], s_w);
    s_lstart[tid] = lstart;

    uint32_t prefix = 0;
    if (tile > 0) {
        int t = (int)tile - 1;
        uint32_t spins = 0;
        while (true) {
            const uint32_t s = ld_agent(status + (size_t)t * kBins + tid);
            const uint32_t f = s & ~kValMask;
            if (f == 0u) {
                if (++spins > kSpinLimit) {
                    atomicOr(err, 1u);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
                continue;
            }
            prefix += s & kValMask;
            if (f == kFlagInc || t == 0) break;
            --t;
        }
        st_agent(my, kFlagInc | (prefix + tot));
    }
    s_gbase[tid] = hbase + prefix;
    __syncthreads();

#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
        if (wbase + (uint32_t)(i * 64) < n) {
            const uint32_t d = (key[i] >> shift) & 255u;
            const uint32_t pos = s_lstart[d] + s_wcnt[wave][d] + rank[i];
            s_keys[pos] = key[i];
            s_vals[pos] = val[i];
        }
    }
    __syncthreads();
    const uint32_t count = n - tbase < (uint32_t)TILE ? n - tbase : (uint32_t)TILE;
#pragma unroll
    for (int q = 0; q < ITEMS; ++q) {
        const uint32_t idx = (uint32_t)(q * kSortBlock + tid);
        if (idx < count) {
            const uint32_t k = s_keys[idx];
            const uint32_t d = (k >> shift) & 255u;
            const uint32_t out = s_gbase[d] + idx - s_lstart[d];
            if (out < n) {
                keys_out[out] = k;
                vals_out[out] = s_vals[idx];
            } else {
                atomicOr(err, 2u);  // inconsistent histogram: never write out of bounds
            }
        }
    }
}

}  // namespace

uint32_t sort_tiles(uint32_t n, int items) {
    const uint32_t tile = (uint32_t)(kSortBlock * items);
    return (n + tile - 1) / tile;
}

int sort_key_passes(uint32_t ncells) {
    int bits = 1;
    while (bits < 32 && (1ull << bits) < (uint64_t)ncells) ++bits;
    return (bits + 7) / 8;
}

hipError_t launch_radix_sort(hipStream_t st, const SortPlan& sp) {
    if (sp.n == 0) return hipSuccess;
    const uint32_t tiles = sort_tiles(sp.n, sp.items);
    const uint32_t hist_blocks = (sp.n + kSortBlock * kHistItems - 1) / (kSortBlock * kHistItems);
    const uint32_t status_words = (uint32_t)sp.passes * tiles * kBins;
    hipLaunchKernelGGL(radix_histogram_kernel, dim3(hist_blocks), dim3(kSortBlock), 0, st,
                       sp.keys_in, sp.n, sp.passes, sp.hist, sp.status, status_words, sp.tickets,
                       sp.prev_sorted, sp.prev_n, sp.cells);
    for (int p = 0; p < sp.passes; ++p) {
        const uint32_t* kin = p == 0 ? sp.keys_in : sp.kbuf[(p - 1) & 1];
        const uint32_t* vin = p == 0 ? nullptr : sp.vbuf[(p - 1) & 1];
        uint32_t* kout = sp.kbuf[p & 1];
        uint32_t* vout = sp.vbuf[p & 1];
        uint32_t* status = sp.status + (size_t)p * tiles * kBins;
        const uint32_t* hist = sp.hist + p * kBins;
        switch (sp.items) {
            case 4:
                hipLaunchKernelGGL(radix_pass_kernel<4>, dim3(tiles), dim3(kSortBlock), 0, st, kin, vin,
                                   kout, vout, hist, status, sp.tickets + p, sp.n, 8 * p, sp.err);
                break;
            case 12:
                hipLaunchKernelGGL(radix_pass_kernel<12>, dim3(tiles), dim3(kSortBlock), 0, st, kin, vin,
                                   kout, vout, hist, status, sp.tickets + p, sp.n, 8 * p, sp.err);
                break;
            case 16:
                hipLaunchKernelGGL(radix_pass_kernel<16>, dim3(tiles), dim3(kSortBlock), 0, st, kin, vin,
                                   kout, vout, hist, status, sp.tickets + p, sp.n, 8 * p, sp.err);
                break;
            default:
                hipLaunchKernelGGL(radix_pass_kernel<8>, dim3(tiles), dim3(kSortBlock), 0, st, kin, vin,
                                   kout, vout, hist, status, sp.tickets + p, sp.n, 8 * p, sp.err);
                break;
        }
    }
    return hipGetLastError();
}

}  // namespace sph
