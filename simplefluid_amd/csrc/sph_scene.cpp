// Reference scene generators: float lattices in the reference's loop order.
//
// Source/SceneManager.cpp
//   CubeDrop        :41-64   bMin -0.5, bMax 0.5, ppos = bMin + spacing*(i,j,k)
//   SphereDrop      :67-91   bounding-cube lattice of radius 0.5 at the origin,
//                            skip points with length(ppos - center) > radius
//   Dambreak        :94-119  bMin (-1,-1,-1) + r, bMax (0.4, 0.4, -0.5)
//   DoubleDambreak  :122-173 Dambreak's block, then a block laid out from its
//                            max corner: ppos = bMax - spacing*(i,j,k)
// The grid extent is the float quotient (bMax - bMin)/spacing truncated to
// int; i (x) is the outer loop, k (z) the inner one.  Radii follow
// Source/Controller.cpp:55 (kernelRadius = 2/res) with r = kernelRadius/4.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <utility>
#include <vector>

#include "sph_engine.h"

namespace {

struct V3 {
    float x, y, z;
};

void lattice(std::vector<V3>& out, V3 origin, float spacing, const int g[3], float sign) {
    for (int i = 0; i < g[0]; ++i)
        for (int j = 0; j < g[1]; ++j)
            for (int k = 0; k < g[2]; ++k) {
                const float ox = spacing * (float)i, oy = spacing * (float)j, oz = spacing * (float)k;
                if (sign > 0)
                    out.push_back({origin.x + ox, origin.y + oy, origin.z + oz});
                else
                    out.push_back({origin.x - ox, origin.y - oy, origin.z - oz});
            }
}

void extent(V3 bmin, V3 bmax, float spacing, int g[3]) {
    g[0] = (int)((bmax.x - bmin.x) / spacing);
    g[1] = (int)((bmax.y - bmin.y) / spacing);
    g[2] = (int)((bmax.z - bmin.z) / spacing);
}

int generate(int scene, float r, std::vector<V3>& out) {
    const float spacing = 2.0f * r;
    int g[3];
    switch (scene) {
        case SPH_SCENE_CUBE_DROP: {
            const V3 bmin{-0.5f, -0.5f, -0.5f}, bmax{0.5f, 0.5f, 0.5f};
            extent(bmin, bmax, spacing, g);
            lattice(out, bmin, spacing, g, 1.0f);
            return SPH_OK;
        }
        case SPH_SCENE_SPHERE_DROP: {
            const V3 c{0.0f, 0.0f, 0.0f};
            const float radius = 0.5f;
            const V3 bmin{c.x - radius, c.y - radius, c.z - radius};
            const int n = (int)(2.0f * radius / spacing);
            for (int i = 0; i < n; ++i)
                for (int j = 0; j < n; ++j)
                    for (int k = 0; k < n; ++k) {
                        const V3 p{bmin.x + spacing * (float)i, bmin.y + spacing * (float)j, bmin.z + spacing * (float)k};
                        const float dx = p.x - c.x, dy = p.y - c.y, dz = p.z - c.z;
                        const float len = std::sqrt((dx * dx + dy * dy) + dz * dz);
                        if (len > radius) continue;
                        out.push_back(p);
                    }
            return SPH_OK;
        }
        case SPH_SCENE_DAMBREAK:
        case SPH_SCENE_DOUBLE_DAMBREAK: {
            const V3 bmin{-1.0f + r, -1.0f + r, -1.0f + r}, bmax{0.4f, 0.4f, -0.5f};
            extent(bmin, bmax, spacing, g);
            lattice(out, bmin, spacing, g, 1.0f);
            if (scene == SPH_SCENE_DOUBLE_DAMBREAK) {
                const V3 bmin2{-0.4f + 0.0f, -1.0f + r, 0.5f + 0.0f};
                const V3 bmax2{1.0f - r, 0.4f - 0.0f, 1.0f - r};
                extent(bmin2, bmax2, spacing, g);
                lattice(out, bmax2, spacing, g, -1.0f);
            }
            return SPH_OK;
        }
        default:
            return SPH_E_INVALID;
    }
}

}  // namespace

extern "C" int sph_scene_generate(int scene, int resolution, float* xyz_out, uint32_t capacity, uint32_t* n_out) {
    if (resolution <= 0 || !n_out) return SPH_E_INVALID;
    const float kr = 2.0f / (float)resolution;
    const float r = kr / 4.0f;
    std::vector<V3> pts;
    const int rc = generate(scene, r, pts);
    if (rc != SPH_OK) return rc;
    *n_out = (uint32_t)pts.size();
    if (!xyz_out) return SPH_OK;
    if (pts.size() > capacity) return SPH_E_CAPACITY;
    std::memcpy(xyz_out, pts.data(), pts.size() * sizeof(V3));
    return SPH_OK;
}

// ---------------------------------------------------------------------------
// Seeded uniform sample of a ball with a minimum spacing: random sequential
// addition.  Darts are drawn uniformly in the ball (xoshiro256** seeded by
// splitmix64, 53-bit doubles, rejection from the bounding cube); a dart is
// kept when no kept point lies within min_spacing of it.  Kept points live in
// a grid of min_spacing cells (a dart checks the 27 around its own) as
// per-cell linked lists.
//
// Darts come in batches (4 count, within [2^16, 2^22]), each visited in cell order.  The grid is cut
// into slabs two cells thick in z; a dart only reads cells within one cell
// of its own, so darts of slabs two apart never see each other's points, and
// a batch runs in two phases (even slabs, then odd), each slab's darts in
// order on one thread.  The result depends on the seed only, not on the
// thread count.  Every dart is uniform and independent of the points kept,
// so the result is a random packing with exactly the spacing bound; when the
// last batch places more than `count`, a seeded random subset of that
// batch's points is dropped (a packing minus points is still one).
namespace {

struct Xoshiro256 {
    uint64_t s[4];
    static uint64_t splitmix(uint64_t& x) {
        uint64_t z = (x += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    explicit Xoshiro256(uint64_t seed) {
        for (int i = 0; i < 4; ++i) s[i] = splitmix(seed);
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        const uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }  // [0, 1)
};

// Points of one slab: coordinates and the next point of the same cell
// (a global reference: slab << 24 | index, -1 ends the list).
struct Slab {
    std::vector<double> p;
    std::vector<int32_t> next;
};

constexpr int kSlabShift = 22;

}  // namespace

extern "C" int sph_scene_sample_ball(uint64_t seed, const float* centre, float radius, float min_spacing,
                                     uint32_t count, float* xyz_out, uint32_t* n_out) {
    if (!centre || !xyz_out || !n_out) return SPH_E_INVALID;
    if (!(radius > 0.0f) || !(min_spacing > 0.0f) || !std::isfinite(radius) || !std::isfinite(min_spacing))
        return SPH_E_INVALID;
    *n_out = 0;
    const double R = radius, s = min_spacing, s2 = s * s;
    const double c[3] = {centre[0], centre[1], centre[2]};
    const int G = (int)std::ceil(2.0 * R / s) + 1;
    const int nslab = (G + 1) / 2;
    if ((double)G * G * G > 2.0e9 || nslab >= 512) return SPH_E_INVALID;
    const size_t plane = (size_t)G * G;
    std::vector<int32_t> head(plane * G, -1);
    std::vector<Slab> slabs(nslab);
    Xoshiro256 rng(seed);
    const uint32_t batch = (uint32_t)std::min<uint64_t>(1u << 22, std::max<uint64_t>(1u << 16, 4ull * count));
    const uint64_t max_darts = 64ull * count + batch;
    auto cell_of = [&](double v, int k) {
        const int q = (int)std::floor((v - (c[k] - R)) / s);
        return q < 0 ? 0 : (q >= G ? G - 1 : q);
    };
    std::vector<double> d(3 * (size_t)batch);
    std::vector<std::pair<uint64_t, uint32_t>> order(batch);
    std::vector<size_t> slab_begin(nslab + 1), before(nslab);
    int nthreads = (int)std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
    uint64_t darts = 0, kept = 0;
    auto run_slab = [&](int sl) {
        Slab& S = slabs[sl];
        for (size_t q = slab_begin[sl]; q < slab_begin[sl + 1]; ++q) {
            const double* p = &d[3 * (size_t)order[q].second];
            const int cx = cell_of(p[0], 0), cy = cell_of(p[1], 1), cz = cell_of(p[2], 2);
            bool ok = true;
            for (int z = std::max(cz - 1, 0); z <= std::min(cz + 1, G - 1) && ok; ++z)
                for (int y = std::max(cy - 1, 0); y <= std::min(cy + 1, G - 1) && ok; ++y)
                    for (int x = std::max(cx - 1, 0); x <= std::min(cx + 1, G - 1) && ok; ++x)
                        for (int32_t j = head[(size_t)z * plane + (size_t)y * G + x]; j >= 0;) {
                            const Slab& T = slabs[j >> kSlabShift];
                            const size_t l = (size_t)(j & ((1 << kSlabShift) - 1));
                            const double ex = p[0] - T.p[3 * l], ey = p[1] - T.p[3 * l + 1], ez = p[2] - T.p[3 * l + 2];
                            if (ex * ex + ey * ey + ez * ez < s2) {
                                ok = false;
                                break;
                            }
                            j = T.next[l];
                        }
            if (!ok) continue;
            const size_t cell = (size_t)cz * plane + (size_t)cy * G + cx;
            S.p.insert(S.p.end(), p, p + 3);
            S.next.push_back(head[cell]);
            head[cell] = (int32_t)(((uint32_t)sl << kSlabShift) | (uint32_t)(S.next.size() - 1));
        }
    };
    auto parallel = [&](int ntasks, auto&& task) {
        std::atomic<int> next_task{0};
        auto worker = [&]() {
            for (int t; (t = next_task.fetch_add(1)) < ntasks;) task(t);
        };
        std::vector<std::thread> pool;
        for (int t = 1; t < nthreads; ++t) pool.emplace_back(worker);
        worker();
        for (auto& t : pool) t.join();
    };
    constexpr int kChunks = 64;
    uint64_t nbatch = 0;
    std::vector<uint32_t> dslab(batch), cnt(nslab + 1);
    while (kept < count && darts < max_darts) {
        // Darts: 64 fixed chunks, each from its own stream of the seed.
        parallel(kChunks, [&](int ch) {
            Xoshiro256 r(seed ^ Xoshiro256(nbatch * kChunks + (uint64_t)ch + 1).next());
            const uint32_t i0 = (uint32_t)((uint64_t)batch * ch / kChunks);
            const uint32_t i1 = (uint32_t)((uint64_t)batch * (ch + 1) / kChunks);
            for (uint32_t i = i0; i < i1;) {
                const double x = 2.0 * r.uniform() - 1.0, y = 2.0 * r.uniform() - 1.0, z = 2.0 * r.uniform() - 1.0;
                if (x * x + y * y + z * z >= 1.0) continue;
                // Rounded to float first: the spacing is checked on the
                // coordinates the caller receives.
                d[3 * (size_t)i] = (double)(float)(c[0] + R * x);
                d[3 * (size_t)i + 1] = (double)(float)(c[1] + R * y);
                d[3 * (size_t)i + 2] = (double)(float)(c[2] + R * z);
                order[i].first = ((uint64_t)cell_of(d[3 * (size_t)i + 2], 2) * G +
                                  (uint64_t)cell_of(d[3 * (size_t)i + 1], 1)) * G +
                                 (uint64_t)cell_of(d[3 * (size_t)i], 0);
                dslab[i] = (uint32_t)(order[i].first / plane) / 2u;
                ++i;
            }
        });
        ++nbatch;
        darts += batch;
        // Group by slab (stable), then each slab's darts go in cell order.
        std::fill(cnt.begin(), cnt.end(), 0u);
        for (uint32_t i = 0; i < batch; ++i) cnt[dslab[i] + 1]++;
        for (int sl = 0; sl < nslab; ++sl) cnt[sl + 1] += cnt[sl];
        for (int sl = 0; sl <= nslab; ++sl) slab_begin[sl] = cnt[sl];
        {
            std::vector<std::pair<uint64_t, uint32_t>> tmp(batch);
            for (uint32_t i = 0; i < batch; ++i) tmp[cnt[dslab[i]]++] = {order[i].first, i};
            order.swap(tmp);
        }
        parallel(nslab, [&](int sl) {
            std::sort(order.begin() + (ptrdiff_t)slab_begin[sl], order.begin() + (ptrdiff_t)slab_begin[sl + 1]);
        });
        for (int sl = 0; sl < nslab; ++sl) before[sl] = slabs[sl].next.size();
        for (int phase = 0; phase < 2; ++phase)
            parallel((nslab - phase + 1) / 2, [&](int t) { run_slab(2 * t + phase); });
        kept = 0;
        for (const Slab& S : slabs) kept += S.next.size();
        for (const Slab& S : slabs)
            if (S.next.size() >= (1u << kSlabShift)) return SPH_E_INVALID;
    }
    // Points of the last batch beyond `count`: drop a seeded random subset of
    // that batch's points.
    std::vector<uint8_t> drop;
    std::vector<size_t> first(nslab + 1, 0);
    for (int sl = 0; sl < nslab; ++sl) first[sl + 1] = first[sl] + slabs[sl].next.size();
    if (kept > count) {
        std::vector<size_t> fresh;
        for (int sl = 0; sl < nslab; ++sl)
            for (size_t l = before[sl]; l < slabs[sl].next.size(); ++l) fresh.push_back(first[sl] + l);
        drop.assign(kept, 0);
        const uint64_t extra = kept - count;
        for (uint64_t i = 0; i < extra; ++i) {  // partial Fisher-Yates
            const size_t j = i + (size_t)(rng.next() % (uint64_t)(fresh.size() - i));
            std::swap(fresh[i], fresh[j]);
            drop[fresh[i]] = 1;
        }
    }
    uint64_t w = 0;
    for (int sl = 0; sl < nslab; ++sl)
        for (size_t l = 0; l < slabs[sl].next.size(); ++l) {
            if (!drop.empty() && drop[first[sl] + l]) continue;
            for (int k = 0; k < 3; ++k) xyz_out[3 * w + k] = (float)slabs[sl].p[3 * l + k];
            ++w;
        }
    *n_out = (uint32_t)w;
    return w == count ? SPH_OK : SPH_E_CAPACITY;
}
