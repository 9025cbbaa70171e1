// Geometry tables: the reference's four parameter arrays repacked for the GPU.
//
// Input layout (PEGen_RayTracing_CircularHole/ReadMe.txt:1-61; read by
// PESysMatGen.cu:382-499 and scatter.cu:1077-1230):
//   Col[0] layers; layer m at (m+1)*10 + {0 holes,1 width,2 thickness,3 height,
//   4 offset,5 mu_t,6 mu_pe,7 mu_c}; hole h at 100+9h + {x,y1,y2,z,R,mu_t,mu_pe,mu_c,flag}
//   Det[0] N; crystal d at 12d + {1 x,2 y,3 z,4 width,5 thickness,6 height,
//   7 mu_t,8 mu_pe,9 mu_c,10 resolution,11 rotation,12 flag}
//
// Counts are taken from the headers (no fixed 80000-float limit).  All y
// coordinates in the tables already include the FOV-to-collimator distance
// (Img[11]) the way the kernels add it.
#pragma once
#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

namespace sysmat {

// One crystal, 5 x float4 = 80 B, read with wave-uniform (scalar) loads.
struct alignas(16) Crystal {
    float cx, cy, cz, mu_t;    // centre (y includes FOV2C, summed in FP32 as the reference does), mu_total
    float hx, hy, hz, rot;     // half extents (0.5*width, 0.5*thickness, 0.5*height), rotation about y
    float cr, sr, rbound, mu_pe;  // cos(-rot), sin(-rot), sqrt(w^2+h^2+t^2) (full diagonal), mu_pe
    float mu_c, res, slab, axis;  // mu_compton, energy resolution, y-slab id, 1 when rot == 0
    float y_raw, flag, hxz, pad1;   // y as stored (before FOV2C; FP64 paths add it in double), flag,
                                    // lateral half-diagonal sqrt(hx^2 + hz^2) (bound of a rotated crystal)
};

// One collimator hole, 3 x float4.
struct alignas(16) Hole {
    float x, y1, y2, z;        // axis x/z, faces y1/y2 (FOV2C added)
    float r, mu_t, slab, rbound;  // radius, total attenuation, y-slab id, sqrt(4R^2+dy^2)/2
    float mu_pe, mu_c, layer, flag;
    float y1_raw, y2_raw, pad0, pad1;  // faces as stored (before FOV2C)
};

// One collimator plate (layer), 2 x float4 + extras.
struct alignas(16) Plate {
    float x1, y1, z1, thick;   // box low corner, thickness
    float x2, y2, z2, mu_t;    // box high corner, total attenuation
    float mu_pe, mu_c, slab, n_holes;
    int hole_begin, hole_end, pad0, pad1;
};

// A distinct y-range shared by crystals / holes / plates.  The bundle filter
// bounds all sub-rays of an element laterally within each slab.
struct alignas(16) Slab {
    float y1, y2, pad0, pad1;
};

constexpr int kMaxLayers = 9;
constexpr int kMaxSlabs = 16;

struct Tables {
    std::vector<Crystal> crystals;
    std::vector<Hole> holes;
    std::vector<Plate> plates;
    std::vector<Slab> slabs;   // empty when more than kMaxSlabs distinct ranges (filter off)
    int n_det = 0, n_layers = 0, n_holes = 0;
    bool all_axis = true;      // every crystal has rotation exactly 0
};

inline int count_of(float v, float eps) { return (int)std::floor(v + eps); }

// Returns "" on success, else an error message.
inline std::string build_tables(const float* col, const float* det, float fov2c, Tables& T) {
    T = Tables();
    T.n_layers = count_of(col[0], 0.000001f);
    T.n_det = count_of(det[0], 0.000001f);
    if (T.n_layers < 0 || T.n_layers > kMaxLayers) return "collimator layer count out of range (0..9)";
    if (T.n_det <= 0) return "detector count must be positive";
    // the tracer keeps candidate crystal / hole indices in 16 bits; checked
    // before any crystal or hole record is read
    if (T.n_det > 65535) return "more than 65535 detector crystals";
    long long total_holes = 0;
    for (int m = 0; m < T.n_layers; ++m) {
        const int nh = (int)std::floor(col[(m + 1) * 10]);
        if (nh < 0) return "negative hole count";
        total_holes += nh;
    }
    if (total_holes > 65535) return "more than 65535 collimator holes";
    std::vector<float> ys;  // distinct y-ranges, in pairs
    auto slab_id = [&](float y1, float y2) -> int {
        for (size_t i = 0; i + 1 < ys.size(); i += 2)
            if (ys[i] == y1 && ys[i + 1] == y2) return (int)(i / 2);
        ys.push_back(y1);
        ys.push_back(y2);
        return (int)(ys.size() / 2 - 1);
    };
    int h = 0;
    for (int m = 0; m < T.n_layers; ++m) {
        const float* L = col + (m + 1) * 10;
        Plate p{};
        p.x1 = -L[1] / 2.0f;
        p.x2 = L[1] / 2.0f;
        p.y1 = -L[2] / 2.0f + fov2c + L[4];
        p.y2 = L[2] / 2.0f + fov2c + L[4];
        p.z1 = -L[3] / 2.0f;
        p.z2 = L[3] / 2.0f;
        p.thick = L[2];
        p.mu_t = L[5];
        p.mu_pe = L[6];
        p.mu_c = L[7];
        p.slab = (float)slab_id(p.y1, p.y2);
        const int nh = (int)std::floor(L[0]);
        p.n_holes = (float)nh;
        p.hole_begin = h;
        p.hole_end = h + nh;
        for (int q = 0; q < nh; ++q, ++h) {
            const float* H = col + 100 + 9 * h;
            Hole o{};
            o.x = H[0];
            o.y1 = H[1] + fov2c;
            o.y2 = H[2] + fov2c;
            o.z = H[3];
            o.r = H[4];
            o.mu_t = H[5];
            o.mu_pe = H[6];
            o.mu_c = H[7];
            o.flag = H[8];
            o.layer = (float)m;
            o.y1_raw = H[1];
            o.y2_raw = H[2];
            o.rbound = std::sqrt(o.r * o.r * 4 + (o.y1 - o.y2) * (o.y1 - o.y2)) / 2.0f;
            float lo = o.y1 < o.y2 ? o.y1 : o.y2, hi = o.y1 < o.y2 ? o.y2 : o.y1;
            o.slab = (float)slab_id(lo, hi);
            T.holes.push_back(o);
        }
        T.plates.push_back(p);
    }
    T.n_holes = h;
    for (int d = 0; d < T.n_det; ++d) {
        const float* D = det + 12 * d;
        Crystal c{};
        c.cx = D[1];
        c.cy = D[2] + fov2c;
        c.cz = D[3];
        c.mu_t = D[7];
        c.hx = 0.5f * D[4];
        c.hy = 0.5f * D[5];
        c.hz = 0.5f * D[6];
        c.rot = D[11];
        c.cr = std::cos(-D[11]);
        c.sr = std::sin(-D[11]);
        c.rbound = std::sqrt(D[4] * D[4] + D[6] * D[6] + D[5] * D[5]);
        c.mu_pe = D[8];
        c.mu_c = D[9];
        c.res = D[10];
        c.axis = (D[11] == 0.0f) ? 1.0f : 0.0f;
        if (D[11] != 0.0f) T.all_axis = false;
        c.y_raw = D[2];
        c.flag = D[12];
        c.hxz = std::sqrt(c.hx * c.hx + c.hz * c.hz);
        c.slab = (float)slab_id(c.cy - c.hy, c.cy + c.hy);
        T.crystals.push_back(c);
    }
    if (ys.size() / 2 <= (size_t)kMaxSlabs) {
        for (size_t i = 0; i + 1 < ys.size(); i += 2) T.slabs.push_back(Slab{ys[i], ys[i + 1], 0.f, 0.f});
    } else {
        for (auto& c : T.crystals) c.slab = -1.0f;
        for (auto& o : T.holes) o.slab = -1.0f;
        for (auto& p : T.plates) p.slab = -1.0f;
    }
    return "";
}

}  // namespace sysmat
