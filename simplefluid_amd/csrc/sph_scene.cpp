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
#include <cmath>
#include <cstring>
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
