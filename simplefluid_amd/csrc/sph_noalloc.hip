// One SPH step on caller-owned device arrays that allocates nothing: the
// intermediates of the step (sorted state, cell tables, neighbour masks,
// tile table, scene constants) live in one file-scope device array,
// `scratch`, of fixed size.  No counterpart in the reference: it is for
// callers that cannot allocate device memory, such as code under HIP graph
// capture or a serving loop that runs one fluid step per request.
//
// The step is the one sph_advance_steps runs as the first step after
// sph_set_particles (hash, cell-count sort, gather, density, force), with the
// same kernels, the same grid rule and the default launch shape, so its
// results equal that entry point's bit for bit.
//
// Limits: at most kMaxParticles particles and kMaxCells grid cells (the
// largest x subdivision of 4, 2, 1 whose grid fits is used, as a handle of
// that size would pick).  Every call on one GPU shares that GPU's scratch:
// calls that overlap on one GPU race, and each call rewrites the scratch
// from the start (counts, tile prefixes, error word and tile queues are
// cleared by kernels, not by memset, so the call may be graph-captured).
#include <cstring>
#include <mutex>

#include "sph_engine.hpp"

namespace sph {

namespace {

constexpr uint32_t kMaxParticles = 65536;
constexpr uint32_t kMaxCells = 1u << 20;
constexpr int kBlock = 256;
constexpr uint32_t kMaxTiles = kMaxParticles / kBlock;
constexpr uint32_t kScanTilesMax = (kMaxCells + 4095u) / 4096u;

// Byte offsets of the arrays inside the scratch, each 256-byte aligned.
constexpr size_t align256(size_t x) { return (x + 255u) & ~(size_t)255u; }
struct Layout {
    static constexpr size_t N = kMaxParticles;
    static constexpr size_t pos = 0;
    static constexpr size_t vel = align256(pos + 16 * N);
    static constexpr size_t pos_s = align256(vel + 16 * N);
    static constexpr size_t vel_s = align256(pos_s + 16 * N);
    static constexpr size_t soa = align256(vel_s + 16 * N);
    static constexpr size_t id = align256(soa + 12 * N);
    static constexpr size_t id_s = align256(id + 4 * N);
    static constexpr size_t key0 = align256(id_s + 4 * N);
    static constexpr size_t rank0 = align256(key0 + 4 * N);
    static constexpr size_t skey = align256(rank0 + 4 * N);
    static constexpr size_t perm = align256(skey + 4 * N);
    static constexpr size_t tmp = align256(perm + 4 * N);
    static constexpr size_t rho = align256(tmp + 4 * N);
    static constexpr size_t mask = align256(rho + 4 * N);
    static constexpr size_t count = align256(mask + 4 * (size_t)kRows * N);
    static constexpr size_t cstart = align256(count + 4 * ((size_t)kMaxCells + 2 * kScanTilesMax));
    static constexpr size_t tile_pre = align256(cstart + 4 * ((size_t)kMaxCells + 1));
    static constexpr size_t tiles = align256(tile_pre + 4 * (size_t)kScanTilesMax);
    static constexpr size_t scene = align256(tiles + 8 * (size_t)kMaxTiles);
    static constexpr size_t err = align256(scene + sizeof(SceneConst));
    static constexpr size_t queue = align256(err + 4);
    static constexpr size_t total = align256(queue + 4 * (size_t)kQueueWords);
};

// The step's scratch on each GPU the library runs on (about 17 MB), held
// until the process ends.  Host code reaches it only through the address
// hipGetSymbolAddress returns on that device.
static __device__ __align__(256) uint8_t scratch[Layout::total];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};

// Scene constants, the tile table (scene 0, first particle t * kBlock), the
// identity ids, and the error word and tile queues cleared.
__global__ void noalloc_setup_kernel(SceneConst sc, uint32_t n, uint8_t* __restrict__ base) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k == 0) *reinterpret_cast<SceneConst*>(base + Layout::scene) = sc;
    if (k < (uint32_t)kQueueWords) reinterpret_cast<uint32_t*>(base + Layout::queue)[k] = 0u;
    if (k == 0) *reinterpret_cast<uint32_t*>(base + Layout::err) = 0u;
    if (k < n) {
        reinterpret_cast<uint32_t*>(base + Layout::id)[k] = k;
        if (k % kBlock == 0) reinterpret_cast<uint2*>(base + Layout::tiles)[k / kBlock] = make_uint2(0u, k);
    }
}

// The scratch address on the current device, looked up once per device.
hipError_t scratch_ptr(uint8_t** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch[dev] = p;
    }
    *out = static_cast<uint8_t*>(g_scratch[dev]);
    return hipSuccess;
}

int fail_noalloc(int code, const char* msg) {
    set_thread_error(msg);
    return code;
}

}  // namespace

}  // namespace sph

using namespace sph;

extern "C" int sph_step_noalloc(const sph_params* p, const float* xyz_in, const float* vel_in, float* xyz_out,
                                float* vel_out, uint32_t n, void* stream) {
    if (!p || !xyz_in || !xyz_out || !vel_out) return fail_noalloc(SPH_E_INVALID, "sph_step_noalloc: null argument");
    if (n == 0 || n > kMaxParticles)
        return fail_noalloc(SPH_E_INVALID, "sph_step_noalloc: n must be 1 .. 65536 (the scratch's particle limit)");
    if (p->cfl_number != 0.0f) return fail_noalloc(SPH_E_INVALID, "sph_step_noalloc: fixed dt only (cfl_number = 0)");
    // The grid a handle of n <= 65536 particles picks: the largest x
    // subdivision whose cell count stays within 2^20.
    SceneConst sc;
    int rc = SPH_E_INVALID;
    for (int xsub : {4, 2, 1}) {
        rc = derive_scene_noalloc(*p, sc, xsub);
        if (rc != SPH_OK) return fail_noalloc(rc, "sph_step_noalloc: invalid parameters (as sph_create checks them)");
        if ((uint64_t)sc.G[0] * sc.G[1] * sc.G[2] <= kMaxCells) break;
        rc = SPH_E_CAPACITY;
    }
    if (rc != SPH_OK)
        return fail_noalloc(SPH_E_INVALID, "sph_step_noalloc: the grid exceeds 2^20 cells (the scratch's cell limit)");
    const uint32_t C = (uint32_t)((uint64_t)sc.G[0] * sc.G[1] * sc.G[2]);
    sc.cell_offset = 0;
    sc.p_begin = 0;
    sc.p_end = n;
    uint8_t* base = nullptr;
    hipError_t e = scratch_ptr(&base);
    if (e != hipSuccess) return fail_noalloc(SPH_E_HIP, "sph_step_noalloc: no scratch address on this device");
    hipStream_t st = static_cast<hipStream_t>(stream);
    auto at = [&](size_t off) { return base + off; };
    float4* pos = reinterpret_cast<float4*>(at(Layout::pos));
    float4* vel = reinterpret_cast<float4*>(at(Layout::vel));
    uint32_t* id = reinterpret_cast<uint32_t*>(at(Layout::id));
    uint32_t* count = reinterpret_cast<uint32_t*>(at(Layout::count));
    const uint32_t T = scan_tiles(C);
    const uint32_t ntiles = (n + kBlock - 1) / kBlock;

    hipLaunchKernelGGL(noalloc_setup_kernel, dim3((n + 255) / 256), dim3(256), 0, st, sc, n, base);
    e = hipGetLastError();
    if (e == hipSuccess) e = launch_fill_u32(st, count, 0u, (size_t)C + 2 * (size_t)T);
    if (e == hipSuccess) e = launch_fill_u32(st, reinterpret_cast<uint32_t*>(at(Layout::tile_pre)), 0xffffffffu, T);
    if (e == hipSuccess) e = launch_pack3(st, n, xyz_in, pos);
    if (e == hipSuccess) e = launch_pack3(st, n, vel_in, vel);
    const SceneConst* scd = reinterpret_cast<const SceneConst*>(at(Layout::scene));
    const uint2* tiles = reinterpret_cast<const uint2*>(at(Layout::tiles));
    uint32_t* key0 = reinterpret_cast<uint32_t*>(at(Layout::key0));
    uint32_t* rank0 = reinterpret_cast<uint32_t*>(at(Layout::rank0));
    if (e == hipSuccess) e = launch_hash(st, kBlock, ntiles, pos, vel, key0, rank0, count, C, nullptr, scd, tiles, ntiles);
    SortPlan sp;
    sp.n = n;
    sp.ncells = C;
    sp.mode = kSortCellCount;
    sp.reorder_block = kBlock;
    sp.key0 = key0;
    sp.rank0 = rank0;
    sp.count = count;
    sp.cstart = reinterpret_cast<uint32_t*>(at(Layout::cstart));
    sp.tile_pre = reinterpret_cast<uint32_t*>(at(Layout::tile_pre));
    sp.tmp = reinterpret_cast<uint32_t*>(at(Layout::tmp));
    sp.skey = reinterpret_cast<uint32_t*>(at(Layout::skey));
    sp.perm = reinterpret_cast<uint32_t*>(at(Layout::perm));
    sp.err = reinterpret_cast<uint32_t*>(at(Layout::err));
    sp.queue = reinterpret_cast<uint32_t*>(at(Layout::queue));
    sp.pos_in = pos;
    sp.vel_in = vel;
    sp.id_in = id;
    sp.pos_out = reinterpret_cast<float4*>(at(Layout::pos_s));
    sp.vel_out = reinterpret_cast<float4*>(at(Layout::vel_s));
    sp.id_out = reinterpret_cast<uint32_t*>(at(Layout::id_s));
    sp.soa_out = reinterpret_cast<float*>(at(Layout::soa));
    if (e == hipSuccess) e = launch_sort_keys(st, sp);
    if (e == hipSuccess) e = launch_sort_gather(st, sp);
    StepArgs a;
    std::memset(&a, 0, sizeof(a));
    a.pos_s = sp.pos_out;
    a.soa = sp.soa_out;
    a.vel_s = sp.vel_out;
    a.pos_o = pos;   // the new state, in sorted order, back into the input slots
    a.vel_o = vel;
    a.id_s = sp.id_out;
    a.id_o = id;
    a.key0 = key0;
    a.rank0 = rank0;
    a.count = count;
    a.cstart = sp.cstart;
    a.rho = reinterpret_cast<float*>(at(Layout::rho));
    a.mask = reinterpret_cast<uint32_t*>(at(Layout::mask));
    a.scenes = scd;
    a.tiles = tiles;
    a.ntiles = ntiles;
    a.n = n;
    a.ncells = C;
    if (e == hipSuccess) e = launch_density(st, kBlock, ntiles, a);
    if (e == hipSuccess) e = launch_force(st, kBlock, ntiles, a);
    if (e == hipSuccess) e = launch_unpermute3(st, n, pos, id, xyz_out);
    if (e == hipSuccess) e = launch_unpermute3(st, n, vel, id, vel_out);
    if (e != hipSuccess) return fail_noalloc(SPH_E_HIP, "sph_step_noalloc: kernel launch failed");
    return SPH_OK;
}
