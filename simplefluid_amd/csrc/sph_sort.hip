// Sorting particles by cell key on gfx950, and the cell-start table.
//
// Both sort modes produce the same stable permutation (ties keep the input
// order) and the same lower-bound table cstart[c] = #particles with key < c.
//
// Cell-count sort (default): a one-digit radix sort whose digit is the whole
// cell key.  The integrate pass already counted particles per cell with
// atomics (count[key]++, keeping the returned slot in rank0).  Then
//   cell_scan        exclusive scan of count -> cstart; each block of 4096
//                    cells takes its prefix from the per-tile sums the
//                    integrate pass also counted (zeroes count for the next step)
//   cell_scatter     tmp[cstart[key] + rank0] = k   (slot order within a
//                    cell is whatever order the atomics ran in)
//   cell_rank        the deterministic rank of k in its cell is the number of
//                    members with a smaller input index; k is gathered
//                    straight into its sorted place (pos, vel, id, key, perm)
// Radix sort: one histogram kernel for all 8-bit digits, then one kernel per
// digit that ranks a tile of keys in LDS (wave64 ballot-matching of the
// digit bits, so equal digits rank in lane order without LDS atomics) and
// finds its global offsets by decoupled look-back; a gather kernel then
// reorders the state.  It is kept for key spaces too large to count.
//
// Look-back waits are bounded: a wait that exceeds its bound raises the
// device error flag instead of hanging, and every data-dependent store is
// range-checked.
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

// Exclusive scan of one value per thread over a 256-thread block; *total
// receives the block sum.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* s_w, uint32_t* total = nullptr) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) s_w[wave] = x;
    __syncthreads();
    uint32_t pre = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < kSortWaves; ++w) {
        const uint32_t sw = s_w[w];
        if (w < wave) pre += sw;
        tot += sw;
    }
    __syncthreads();
    if (total) *total = tot;
    return pre + x - v;
}

// Inclusive prefix of the tiles before `tile` in a look-back status array
// (one word per tile and digit: 2-bit flag, 30-bit value).
__device__ __forceinline__ uint32_t look_back(const uint32_t* status, uint32_t stride, uint32_t tile, uint32_t* err) {
    uint32_t prefix = 0;
    int t = (int)tile - 1;
    uint32_t spins = 0;
    while (t >= 0) {
        const uint32_t s = ld_agent(status + (size_t)t * stride);
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
        if (f == kFlagInc) break;
        --t;
    }
    return prefix;
}

// ---------------------------------------------------------------- cell count

// Exclusive scan of the T per-tile sums (count[C .. C+T)) into
// count[C+T .. C+2T), one block of 1024 threads, each over a contiguous run.
constexpr int kTileScanBlock = 1024;
constexpr uint32_t kInlineTiles = 1024;

__global__ __launch_bounds__(kTileScanBlock) void tile_scan_kernel(uint32_t* __restrict__ count, uint32_t C,
                                                                   uint32_t T, const uint32_t* __restrict__ halt) {
    if (halt && *halt) return;
    // Rounds of 4096 tile sums: thread t takes 4 consecutive ones (16-byte
    // access when aligned), a block scan adds the round's carry.
    __shared__ uint32_t s_w[kTileScanBlock / 64];
    __shared__ uint32_t s_carry;
    const uint32_t* tsum = count + C;
    uint32_t* tpre = count + C + T;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0) s_carry = 0u;
    __syncthreads();
    for (uint32_t r0 = 0; r0 < T; r0 += 4u * kTileScanBlock) {
        const uint32_t i0 = r0 + 4u * (uint32_t)tid;
        uint32_t v[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) v[e] = i0 + e < T ? tsum[i0 + e] : 0u;
        const uint32_t own = v[0] + v[1] + v[2] + v[3];
        uint32_t x = own;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(x, o, 64);
            if (lane >= o) x += y;
        }
        if (lane == 63) s_w[wave] = x;
        __syncthreads();
        uint32_t run = s_carry + x - own;
        uint32_t tot = 0;
        for (int w = 0; w < kTileScanBlock / 64; ++w) {
            if (w < wave) run += s_w[w];
            tot += s_w[w];
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            if (i0 + e < T) tpre[i0 + e] = run;
            run += v[e];
        }
        __syncthreads();
        if (tid == 0) s_carry += tot;
        __syncthreads();
    }
}

// Block b scans the counts of cells [4096 b, 4096 (b + 1)).  Its prefix is
// the exclusive scan of the tile sums (counted by the integrate pass next to
// the cell counts): summed here when there are at most kInlineTiles tiles,
// else read from tile_scan_kernel's output.  No block waits for another; a
// tile with no particles only writes its prefix.  Wave w takes 1024 cells in 4 rounds of 64 lanes x
// 4 consecutive cells (16-byte loads and stores); the counts are zeroed for
// the next step as they are read.
__global__ __launch_bounds__(kSortBlock) void cell_scan_kernel(uint32_t* __restrict__ count,
                                                               uint32_t* __restrict__ cstart, uint32_t C,
                                                               uint32_t* __restrict__ queue,
                                                               const uint32_t* __restrict__ halt,
                                                               uint32_t* __restrict__ tile_pre) {
    if (halt && *halt) return;
    __shared__ uint32_t s_tot[kSortWaves];
    const uint32_t b = blockIdx.x;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // Tile counters of this step's persistent density and force launches.
    if (b == 0 && tid < kQueueWords && queue) queue[tid] = 0u;
    const uint32_t T = (C + (1u << kScanTileShift) - 1u) >> kScanTileShift;
    uint32_t pre;
    if (T <= kInlineTiles) {
        __shared__ uint32_t s_pre[kSortWaves];
        uint32_t x = 0;
        for (uint32_t i = (uint32_t)tid; i < b; i += kSortBlock) x += count[C + i];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
        if (lane == 0) s_pre[wave] = x;
        __syncthreads();
        pre = s_pre[0] + s_pre[1] + s_pre[2] + s_pre[3];
    } else {
        pre = count[C + T + b];
    }
    // What the table holds for this tile, read by every thread before any
    // thread of the block rewrites it.
    const uint32_t held = tile_pre[b];
    __syncthreads();
    if (count[C + b] == 0u) {
        // An empty tile's cells all hold its prefix; skip the rewrite when
        // this table already holds that prefix there (most empty tiles of
        // a sparse grid keep theirs from step to step).
        if (held == pre) return;
        if (tid == 0) tile_pre[b] = pre;
        const uint32_t t0 = b << kScanTileShift;
        const uint32_t t1 = t0 + (1u << kScanTileShift) < C ? t0 + (1u << kScanTileShift) : C;
        const uint4 v4 = make_uint4(pre, pre, pre, pre);
        for (uint32_t i = t0 + 4u * (uint32_t)tid; i < t1; i += 4u * kSortBlock) {
            if (i + 4u <= t1) {
                *reinterpret_cast<uint4*>(cstart + i) = v4;
            } else {
                for (uint32_t e = i; e < t1; ++e) cstart[e] = pre;
            }
        }
        if (t1 == C && tid == 0) cstart[C] = pre;
        return;
    }
    if (tid == 0) tile_pre[b] = 0xffffffffu;
    const uint32_t base = (b << kScanTileShift) + (uint32_t)wave * 1024u;
    uint4 v[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const uint32_t i0 = base + 256u * (uint32_t)r + 4u * (uint32_t)lane;
        if (i0 + 4u <= C) {
            v[r] = *reinterpret_cast<const uint4*>(count + i0);
            *reinterpret_cast<uint4*>(count + i0) = make_uint4(0u, 0u, 0u, 0u);
        } else {
            v[r].x = i0 < C ? count[i0] : 0u;
            v[r].y = i0 + 1u < C ? count[i0 + 1u] : 0u;
            v[r].z = i0 + 2u < C ? count[i0 + 2u] : 0u;
            v[r].w = i0 + 3u < C ? count[i0 + 3u] : 0u;
            for (uint32_t e = 0; e < 4u; ++e)
                if (i0 + e < C) count[i0 + e] = 0u;
        }
    }
    // Wave-local exclusive prefix of each lane's four cells, round by round.
    uint32_t ex[4];
    uint32_t carry = 0;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const uint32_t own = v[r].x + v[r].y + v[r].z + v[r].w;
        uint32_t x = own;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(x, o, 64);
            if (lane >= o) x += y;
        }
        ex[r] = carry + x - own;
        carry += __shfl(x, 63, 64);
    }
    if (lane == 0) s_tot[wave] = carry;
    __syncthreads();
    uint32_t off = pre;
#pragma unroll
    for (int w = 0; w < kSortWaves; ++w)
        if (w < wave) off += s_tot[w];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const uint32_t i0 = base + 256u * (uint32_t)r + 4u * (uint32_t)lane;
        uint4 o;
        o.x = off + ex[r];
        o.y = o.x + v[r].x;
        o.z = o.y + v[r].y;
        o.w = o.z + v[r].z;
        const uint32_t end = o.w + v[r].w;
        if (i0 + 4u <= C) {
            *reinterpret_cast<uint4*>(cstart + i0) = o;
        } else {
            const uint32_t val[4] = {o.x, o.y, o.z, o.w};
            for (uint32_t e = 0; e < 4u; ++e)
                if (i0 + e < C) cstart[i0 + e] = val[e];
        }
        if (i0 < C && C <= i0 + 4u) {
            const uint32_t val[5] = {o.x, o.y, o.z, o.w, end};
            cstart[C] = val[C - i0];
        }
    }
}

__global__ void cell_scatter_kernel(uint32_t n, uint32_t C, const uint32_t* __restrict__ key0,
                                    const uint32_t* __restrict__ rank0, const uint32_t* __restrict__ cstart,
                                    uint32_t* __restrict__ tmp, uint32_t* __restrict__ err,
                                    const uint32_t* __restrict__ halt) {
    if (halt && *halt) return;
    const uint32_t k = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint32_t key = key0[k];
    const uint32_t slot = key < C ? cstart[key] + rank0[k] : 0xffffffffu;
    if (slot < n)
        tmp[slot] = k;
    else
        atomicOr(err, 2u);
}

// Members a cell may hold for the stable in-cell rank (cell_rank_kernel
// scans its cell once per member).  Rest-density fluid holds a few per cell.
constexpr uint32_t kMaxRankCell = 1024;

__global__ void cell_rank_kernel(uint32_t n, uint32_t C, const uint32_t* __restrict__ key0,
                                 const uint32_t* __restrict__ rank0,
                                 const uint32_t* __restrict__ cstart, const uint32_t* __restrict__ tmp,
                                 const float4* __restrict__ pos_in, const float4* __restrict__ vel_in,
                                 const uint32_t* __restrict__ id_in, float4* __restrict__ pos_out,
                                 float4* __restrict__ vel_out, uint32_t* __restrict__ id_out,
                                 uint32_t* __restrict__ skey, uint32_t* __restrict__ perm,
                                 uint32_t* __restrict__ tsum, uint32_t tsum_words,
                                 uint32_t* __restrict__ err, float* __restrict__ soa,
                                 const uint32_t* __restrict__ halt) {
    if (halt && *halt) return;
    // The scan has read the tile sums; clear them for the integrate pass.
    if (blockIdx.x == 0)
        for (uint32_t i = threadIdx.x; i < tsum_words; i += blockDim.x) tsum[i] = 0u;
    const uint32_t k = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint32_t key = key0[k];
    if (key >= C) {
        atomicOr(err, 4u);
        return;
    }
    const uint32_t lo = cstart[key], hi = cstart[key + 1];
    uint32_t r;
    if (hi - lo <= kMaxRankCell) {
        r = 0;
        for (uint32_t t = lo; t < hi; ++t) r += tmp[t] < k ? 1u : 0u;
    } else {
        // A cell this full only arises from collapsed or non-finite
        // positions (clamped into one cell).  Ranking it costs its size per
        // member, so it keeps the arrival order (a valid permutation, not
        // the stable one) and raises the error flag instead.
        r = rank0[k];
        atomicOr(err, 8u);
    }
    const uint32_t dst = lo + r;
    if (dst >= hi || hi > n) {
        atomicOr(err, 4u);
        return;
    }
    perm[dst] = k;
    skey[dst] = key;
    const float4 pk = pos_in[k];
    pos_out[dst] = pk;
    if (soa) {  // the planes, for the gathered density pass
        soa[dst] = pk.x;
        soa[n + dst] = pk.y;
        soa[2 * (size_t)n + dst] = pk.z;
    }
    vel_out[dst] = vel_in[k];
    id_out[dst] = id_in[k];
}

// --------------------------------------------------------------------- radix

// Digit histograms of all passes; zeroes the look-back status words and the
// tile tickets of this sort.
__global__ __launch_bounds__(kSortBlock) void radix_histogram_kernel(
    const uint32_t* __restrict__ keys, uint32_t n, int passes, uint32_t* __restrict__ hist,
    uint32_t* __restrict__ status, uint32_t status_words, uint32_t* __restrict__ tickets) {
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
        prefix = look_back(status + tid, kBins, tile, err);
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

__global__ void radix_reorder_kernel(uint32_t n, const uint32_t* __restrict__ perm,
                                     const float4* __restrict__ pos_in, const float4* __restrict__ vel_in,
                                     const uint32_t* __restrict__ id_in, float4* __restrict__ pos_out,
                                     float4* __restrict__ vel_out, uint32_t* __restrict__ id_out,
                                     uint32_t* __restrict__ hist_zero, uint32_t* __restrict__ tsum,
                                     uint32_t tsum_words, uint32_t* __restrict__ err, float* __restrict__ soa) {
    if (blockIdx.x == 0) {
        for (uint32_t i = threadIdx.x; i < (uint32_t)(kMaxPasses * kBins); i += blockDim.x) hist_zero[i] = 0u;
        for (uint32_t i = threadIdx.x; i < tsum_words; i += blockDim.x) tsum[i] = 0u;
    }
    const uint32_t k = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint32_t s = perm[k];
    if (s >= n) {
        atomicOr(err, 4u);
        return;
    }
    const float4 ps = pos_in[s];
    pos_out[k] = ps;
    if (soa) {
        soa[k] = ps.x;
        soa[n + k] = ps.y;
        soa[2 * (size_t)n + k] = ps.z;
    }
    vel_out[k] = vel_in[s];
    id_out[k] = id_in[s];
}

inline uint32_t blocks_for(size_t n, int b) { return (uint32_t)((n + b - 1) / b); }

}  // namespace

uint32_t sort_tiles(uint32_t n, int items) {
    const uint32_t tile = (uint32_t)(kSortBlock * items);
    return (n + tile - 1) / tile;
}

uint32_t scan_tiles(uint32_t C) { return (C + (1u << kScanTileShift) - 1u) >> kScanTileShift; }

int sort_key_passes(uint32_t ncells) {
    int bits = 1;
    while (bits < 32 && (1ull << bits) < (uint64_t)ncells) ++bits;
    return (bits + 7) / 8;
}

hipError_t launch_sort_keys(hipStream_t st, const SortPlan& sp) {
    if (sp.n == 0) return hipSuccess;
    // The cell-start table (both modes).
    const uint32_t stiles = scan_tiles(sp.ncells);
    if (stiles > kInlineTiles)
        hipLaunchKernelGGL(tile_scan_kernel, dim3(1), dim3(kTileScanBlock), 0, st, sp.count, sp.ncells, stiles,
                           sp.halt);
    hipLaunchKernelGGL(cell_scan_kernel, dim3(stiles), dim3(kSortBlock), 0, st, sp.count, sp.cstart, sp.ncells,
                       sp.queue, sp.halt, sp.tile_pre);
    if (sp.mode == kSortCellCount) {
        hipLaunchKernelGGL(cell_scatter_kernel, dim3(blocks_for(sp.n, 256)), dim3(256), 0, st, sp.n, sp.ncells,
                           sp.key0, sp.rank0, sp.cstart, sp.tmp, sp.err, sp.halt);
        return hipGetLastError();
    }
    if (sp.halt) return hipErrorInvalidValue;  // the frame loop runs the cell-count sort only
    const uint32_t tiles = sort_tiles(sp.n, sp.items);
    const uint32_t hist_blocks = (sp.n + kSortBlock * kHistItems - 1) / (kSortBlock * kHistItems);
    const uint32_t status_words = (uint32_t)sp.passes * tiles * kBins;
    hipLaunchKernelGGL(radix_histogram_kernel, dim3(hist_blocks), dim3(kSortBlock), 0, st, sp.key0, sp.n,
                       sp.passes, sp.hist, sp.status, status_words, sp.tickets);
    for (int p = 0; p < sp.passes; ++p) {
        const bool last = p == sp.passes - 1;
        const uint32_t* kin = p == 0 ? sp.key0 : sp.kbuf[(p - 1) & 1];
        const uint32_t* vin = p == 0 ? nullptr : sp.vbuf[(p - 1) & 1];
        uint32_t* kout = last ? sp.skey : sp.kbuf[p & 1];
        uint32_t* vout = last ? sp.perm : sp.vbuf[p & 1];
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

hipError_t launch_sort_gather(hipStream_t st, const SortPlan& sp) {
    if (sp.n == 0) return hipSuccess;
    uint32_t* tsum = sp.count + sp.ncells;  // per-tile sums, cleared once the scan has read them
    const uint32_t tsum_words = scan_tiles(sp.ncells);
    if (sp.mode == kSortCellCount)
        hipLaunchKernelGGL(cell_rank_kernel, dim3(blocks_for(sp.n, sp.reorder_block)), dim3(sp.reorder_block), 0,
                           st, sp.n, sp.ncells, sp.key0, sp.rank0, sp.cstart, sp.tmp, sp.pos_in, sp.vel_in, sp.id_in,
                           sp.pos_out, sp.vel_out, sp.id_out, sp.skey, sp.perm, tsum,
                           tsum_words, sp.err, sp.soa_out, sp.halt);
    else
        hipLaunchKernelGGL(radix_reorder_kernel, dim3(blocks_for(sp.n, sp.reorder_block)), dim3(sp.reorder_block),
                           0, st, sp.n, sp.perm, sp.pos_in, sp.vel_in, sp.id_in, sp.pos_out, sp.vel_out,
                           sp.id_out, sp.hist, tsum, tsum_words, sp.err, sp.soa_out);
    return hipGetLastError();
}

}  // namespace sph
