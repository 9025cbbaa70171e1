// C ABI and host orchestration of the Compton-scatter matrices.
//
// Replaces scatter() (ScatterGen_RayTracing_CircularHole/scatter.cu:2013-2378 and
// the ScatterGen_Collimator / ScatterGen_Crystal variants): bitmaps are built
// on the GPU (or read from the GeometryRelationShip_* files), per-pair
// constants are prepared once, and the per-voxel kernel runs over detector
// blocks.  Host-memory entry points mirror the reference signatures; plans keep
// everything resident in HBM for repeated launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "../../include/sysmat.h"
#include "common.hpp"
#include "scatter.hpp"
#include "surrogate.hpp"

namespace sysmat {

// Reference KN normalisation: h * sum_{i=0..n} f(i h), n = (int)(pi_f / h), in
// double with the reference's float constants (scatter.cu:1169).
static double kn_total_ref(double e0) {
    const double h = (double)0.01f, top = (double)(float)M_PI;
    const int n = (int)((top - 0.0) / h);
    double s = 0.0;
    for (int i = 0; i <= n; ++i) {
        const double th = 0.0 + (double)(float)i * h;
        const double c = std::cos(th), a = e0 / 511.0;
        const double f1 = a * (1.0 - c), f2 = 1.0 + c * c;
        s += f2 / (1.0 + f1) / (1.0 + f1) * (1.0 + f1 * f1 / f2 / (1.0 + f1)) * std::sin(th) * h;
    }
    return s;
}

// Scatter angles (theta_lo, theta_hi) where the 2-sigma energy-window test
// of scatter.cu:1263 can pass (E' is monotone in theta); a small margin keeps
// the pruning conservative.
static void window_band(double e0, double res, double lo, double hi, double& th_lo, double& th_hi) {
    auto ep = [&](double th) { return e0 / (1.0 + (e0 / 511.0) * (1.0 - std::cos(th))); };
    auto fail_low = [&](double th) {
        const double e = ep(th), k = 2.0 * res * std::sqrt(e / e0) / (double)2.35482f;
        return e * (1.0 + k) <= lo;
    };
    auto fail_high = [&](double th) {
        const double e = ep(th), k = 2.0 * res * std::sqrt(e / e0) / (double)2.35482f;
        return e * (1.0 - k) >= hi;
    };
    // fail_low holds for large theta, fail_high for small theta (E' decreases with theta)
    double a = 0.0, b = M_PI;
    if (fail_low(0.0)) { th_lo = 1.0; th_hi = -1.0; return; }  // nothing passes
    if (!fail_low(M_PI)) {
        th_hi = M_PI;
    } else {
        for (int i = 0; i < 80; ++i) { const double m = 0.5 * (a + b); (fail_low(m) ? b : a) = m; }
        th_hi = b;
    }
    if (!fail_high(0.0)) {
        th_lo = 0.0;
    } else {
        a = 0.0; b = M_PI;
        for (int i = 0; i < 80; ++i) { const double m = 0.5 * (a + b); (fail_high(m) ? a : b) = m; }
        th_lo = a;
    }
    th_lo = std::max(0.0, th_lo - 1e-6);
    th_hi = std::min(M_PI, th_hi + 1e-6);
}

struct ScatterGeom {
    int device = 0;
    Tables T;
    DeviceBuf<Crystal> cry;
    DeviceBuf<Hole> hol;
    DeviceBuf<ScRow> rows;
    std::vector<ScRow> hrows;
    DeviceBuf<SlabGrid> grids;
    DeviceBuf<int> grid_cell, grid_item;
    DeviceBuf<double2> kn_tab;
    ScParams base{};
    int n_holes0 = 0;
};

// Crystals of each y-slab binned by centre on an x-z grid whose cell is the
// slab's largest crystal footprint (doubled until the grid has at most
// 4 x members + 65536 cells).  Members of a cell stay in index order.
static int build_grids(ScatterGeom& G, double fov2c) {
    const std::vector<Crystal>& C = G.T.crystals;
    if (G.T.slabs.empty()) return SM_OK;  // too many distinct y-ranges: scan every crystal
    int max_slab = -1;
    for (const Crystal& c : C) max_slab = std::max(max_slab, (int)c.slab);
    std::vector<SlabGrid> grids;
    std::vector<int> cell, item;
    for (int sid = 0; sid <= max_slab; ++sid) {
        std::vector<int> mem;
        for (int k = 0; k < (int)C.size(); ++k)
            if ((int)C[k].slab == sid) mem.push_back(k);
        if (mem.empty()) continue;
        SlabGrid g{};
        double xl = INFINITY, xh = -INFINITY, zl = INFINITY, zh = -INFINITY, cs = 0.0, m = 0.0;
        g.y1 = INFINITY;
        g.y2 = -INFINITY;
        for (int k : mem) {
            const Crystal& c = C[k];
            const double yc = (double)c.y_raw + fov2c;
            g.y1 = std::min(g.y1, yc - (double)c.hy);
            g.y2 = std::max(g.y2, yc + (double)c.hy);
            xl = std::min(xl, (double)c.cx);
            xh = std::max(xh, (double)c.cx);
            zl = std::min(zl, (double)c.cz);
            zh = std::max(zh, (double)c.cz);
            const double e = (c.rot == 0.0f) ? std::max((double)c.hx, (double)c.hz)
                                             : std::sqrt((double)c.hx * c.hx + (double)c.hz * c.hz);
            m = std::max(m, e);
            cs = std::max(cs, 2.0 * e);
        }
        if (!(cs > 0.0)) cs = 1.0;
        long long nx, nz;
        for (;;) {
            nx = (long long)std::floor((xh - xl) / cs) + 2;
            nz = (long long)std::floor((zh - zl) / cs) + 2;
            if (nx * nz <= 4LL * (long long)mem.size() + 65536) break;
            cs *= 2.0;
        }
        g.cx = g.cz = cs;
        g.x0 = xl - 0.5 * cs;
        g.z0 = zl - 0.5 * cs;
        g.nx = (int)nx;
        g.nz = (int)nz;
        g.m = m + 0.01;
        g.cell_base = (int)cell.size();
        std::vector<int> cnt((size_t)(nx * nz) + 1, 0);
        std::vector<int> where(mem.size());
        for (size_t i = 0; i < mem.size(); ++i) {
            const Crystal& c = C[mem[i]];
            const long long ix = std::min(nx - 1, std::max(0LL, (long long)std::floor(((double)c.cx - g.x0) / cs)));
            const long long iz = std::min(nz - 1, std::max(0LL, (long long)std::floor(((double)c.cz - g.z0) / cs)));
            where[i] = (int)(iz * nx + ix);
            ++cnt[(size_t)where[i] + 1];
        }
        for (size_t c = 1; c < cnt.size(); ++c) cnt[c] += cnt[c - 1];
        const int item0 = (int)item.size();
        item.resize(item.size() + mem.size());
        std::vector<int> fill(cnt.begin(), cnt.end() - 1);
        for (size_t i = 0; i < mem.size(); ++i) item[(size_t)item0 + fill[(size_t)where[i]]++] = mem[i];
        for (int& c : cnt) c += item0;
        cell.insert(cell.end(), cnt.begin(), cnt.end());
        grids.push_back(g);
    }
    int rc;
    if ((rc = G.grids.upload(grids)) < 0) return rc;
    if ((rc = G.grid_cell.upload(cell)) < 0) return rc;
    if ((rc = G.grid_item.upload(item)) < 0) return rc;
    G.base.grids = G.grids.ptr;
    G.base.grid_cell = G.grid_cell.ptr;
    G.base.grid_item = G.grid_item.ptr;
    G.base.n_grids = (int)grids.size();
    return SM_OK;
}

static int build_geom(const float* col, const float* det, const float* img, const float* phy, int device,
                      int rotation, ScatterGeom& G) {
    G.device = device;
    std::string err = build_tables(col, det, img[11], G.T);
    if (!err.empty()) return set_error(SM_E_ARG, err);
    int rc;
    if ((rc = G.cry.upload(G.T.crystals)) < 0) return rc;
    if ((rc = G.hol.upload(G.T.holes)) < 0) return rc;
    ScParams& p = G.base;
    p = ScParams{};
    p.X = (int)std::floor(img[0] + 0.001f);
    p.Y = (int)std::floor(img[1] + 0.001f);
    p.Z = (int)std::floor(img[2] + 0.001f);
    if (p.X <= 0 || p.Y <= 0 || p.Z <= 0) return set_error(SM_E_ARG, "image voxel counts must be positive");
    p.wx = img[3]; p.wy = img[4]; p.wz = img[5];
    p.sx = img[8]; p.sy = img[9]; p.sz = img[10];
    const double ang = (double)(float)rotation * (double)img[7];
    p.rc = std::cos(ang);
    p.rs = std::sin(ang);
    p.e0 = phy[7];
    if (!(p.e0 > 0.0)) return set_error(SM_E_ARG, "photopeak energy Phys[7] must be positive");
    p.kn_total = kn_total_ref(p.e0);
    p.ikn_total = 1.0 / p.kn_total;
    p.h = (double)0.01f;
    p.ihf = (float)(1.0 / p.h);
    {
        std::vector<double2> tab(kKnSteps + 1);
        for (int n = 0; n <= kKnSteps; ++n) {
            const double a = (double)n * p.h;
            tab[n] = make_double2(std::cos(a), std::sin(a));
        }
        int rc0 = G.kn_tab.upload(tab);
        if (rc0 < 0) return rc0;
        p.kn_tab = G.kn_tab.ptr;
    }
    p.n_det = G.T.n_det;
    p.V = (long long)p.X * p.Y * p.Z;
    p.fov2c = img[11];
    G.n_holes0 = (int)std::floor(col[10] + 0.000001f);
    if (G.n_holes0 > G.T.n_holes) G.n_holes0 = G.T.n_holes;
    p.w_col = (col[15] != 0.0f) ? (double)col[17] / (double)col[15] : 0.0;
    const bool use_compton = (int)std::floor(phy[0] + 0.5f) == 1;
    const bool same_window = (int)std::floor(phy[4] + 0.5f) > 0;
    G.hrows.resize(G.T.n_det);
    for (int r = 0; r < G.T.n_det; ++r) {
        const float* D = det + 12 * r;
        ScRow& R = G.hrows[r];
        R = ScRow{};
        R.cx = D[1];
        R.cy = (double)D[2] + (double)img[11];
        R.cz = D[3];
        R.crn = std::cos(-(double)D[11]);
        R.srn = std::sin(-(double)D[11]);
        R.hx = 0.5 * (double)D[4];
        R.hy = 0.5 * (double)D[5];
        R.hz = 0.5 * (double)D[6];
        R.res = D[10];
        if (same_window) {
            R.lo = phy[5];
            R.hi = phy[6];
        } else {
            R.lo = (1.0 - R.res / 2.0) * p.e0;
            R.hi = (1.0 + R.res / 2.0) * p.e0;
        }
        R.mu_pe = D[8];
        R.mu_c = D[9];
        const double w = D[4], h = D[6], t = D[5];
        R.rdet = std::sqrt(w * w / 1.0 / 1.0 + h * h / 1.0 / 1.0 + t * t / 1.0 / 1.0) / 2.0;
        window_band(p.e0, R.res, R.lo, R.hi, R.th_lo, R.th_hi);
        R.use = (use_compton && D[7] > 0.01f && D[8] != 0.0f) ? 1.0f : 0.0f;
    }
    int rc2 = G.rows.upload(G.hrows);
    if (rc2 < 0) return rc2;
    return build_grids(G, (double)img[11]);
}

long long crystal_words(int n) { return ((long long)n * (n + 1) / 2 * n + 31) / 32; }
long long collimator_words(int nh, int n) { return ((long long)n * n * nh + 31) / 32; }

// Run prep + accumulation for detector rows [det_begin, det_begin + count) in
// blocks that keep the pair table under `pair_budget` bytes.
static int run_blocks(const ScatterGeom& G, ScParams p, const unsigned* bits_dev, const float* pe_dev, int det_begin,
                      int det_count, float* dst_dev, long long row_stride, hipStream_t s);

static int run_scatter(ScatterGeom& G, int mode, const unsigned* bits_dev, const float* pe_dev, int det_begin,
                       int det_count, float* dst_dev, long long row_stride, hipStream_t s) {
    ScParams p = G.base;
    p.mode = mode;
    p.max_sites = (mode == 0) ? G.T.n_det : std::max(1, G.n_holes0);
    if (mode != 0 && G.n_holes0 <= 0) {  // no layer-0 holes: the collimator term is zero
        for (int r = 0; r < det_count; ++r)
            HIP_TRY(hipMemsetAsync(dst_dev + (long long)r * row_stride, 0, sizeof(float) * p.V, s));
        return SM_OK;
    }
    return run_blocks(G, p, bits_dev, pe_dev, det_begin, det_count, dst_dev, row_stride, s);
}

// Collimator plate cells resident on the device for one (cell size, keep distance).
struct CellModel {
    double h = -1.0, keep = 0.0;
    std::vector<Cell> cells;
    DeviceBuf<Cell> dcells;
    DeviceBuf<CellPlate> dplates;
    DeviceBuf<CellHole> dholes;
    int n_plates = 0;
};

static int prepare_cells(const float* col, float fov2c, double h, double keep, CellModel& M) {
    if (M.h == h && M.keep == keep && M.dcells.ptr) return SM_OK;
    std::vector<CellPlate> plates;
    std::vector<CellHole> holes;
    const std::string err = build_cells(col, fov2c, h, keep, M.cells, plates, holes);
    if (!err.empty()) return set_error(SM_E_ARG, err);
    int rc;
    if ((rc = M.dcells.upload(M.cells)) < 0) return rc;
    if ((rc = M.dplates.upload(plates)) < 0) return rc;
    if ((rc = M.dholes.upload(holes)) < 0) return rc;
    M.n_plates = (int)plates.size();
    M.h = h;
    M.keep = keep;
    return SM_OK;
}

// Cell-model collimator scatter: cells in chunks whose weight table stays under
// 4 GiB (or `per_chunk` cells); the first chunk writes, later chunks accumulate.
static int run_scatter_cells(ScatterGeom& G, CellModel& M, int det_begin, int det_count, float* dst_dev,
                             long long row_stride, int per_chunk, hipStream_t s) {
    ScParams p = G.base;
    p.mode = 3;
    p.cplates = M.dplates.ptr;
    p.choles = M.dholes.ptr;
    p.n_cplates = M.n_plates;
    const int n_cells = (int)M.cells.size();
    if (n_cells == 0) {
        for (int r = 0; r < det_count; ++r)
            HIP_TRY(hipMemsetAsync(dst_dev + (long long)r * row_stride, 0, sizeof(float) * p.V, s));
        return SM_OK;
    }
    const long long table_budget = 4096ll << 20;
    int chunk = per_chunk > 0 ? per_chunk
                              : (int)std::max(1ll, std::min<long long>(n_cells, table_budget / (p.V * (long long)sizeof(float))));
    chunk = std::min(chunk, n_cells);
    DeviceBuf<float> W;
    int rc;
    if ((rc = W.alloc((size_t)chunk * p.V)) < 0) return rc;
    for (int c0 = 0; c0 < n_cells; c0 += chunk) {
        const int cnt = std::min(chunk, n_cells - c0);
        p.cells = M.dcells.ptr + c0;
        p.max_sites = cnt;
        p.accumulate = c0 > 0 ? 1 : 0;
        HIP_TRY(launch_cell_weights(p, cnt, W.ptr, s));
        if ((rc = run_blocks(G, p, nullptr, W.ptr, det_begin, det_count, dst_dev, row_stride, s)) < 0) return rc;
    }
    return SM_OK;
}

// Prep + accumulation of rows [det_begin, det_begin + det_count) for the mode and
// sites in `p`, in row blocks whose pair table stays under 2 GiB.
static int run_blocks(const ScatterGeom& G, ScParams p, const unsigned* bits_dev, const float* pe_dev, int det_begin,
                      int det_count, float* dst_dev, long long row_stride, hipStream_t s) {
    const long long pair_budget = 2048ll << 20;  // 2 GiB of pair records per block
    int block = (int)std::max(4ll, std::min<long long>(det_count, pair_budget / ((long long)p.max_sites * sizeof(ScPair))));
    block = (block + 3) / 4 * 4;
    p.site_chunk = kPrepChunk;
    DeviceBuf<ScPair> pairs;
    DeviceBuf<int> counts;
    int rc;
    if ((rc = pairs.alloc((size_t)block * p.max_sites)) < 0) return rc;
    if ((rc = counts.alloc((size_t)block * site_chunks(p))) < 0) return rc;
    for (int b0 = 0; b0 < det_count; b0 += block) {
        const int cnt = std::min(block, det_count - b0);
        p.det_begin = det_begin + b0;
        p.det_count = cnt;
        p.row_stride = row_stride;
        set_tiles(p, true);
        p.n_rgroups = (cnt + 3) / 4;
        p.n_blocks = (long long)p.n_tiles * p.n_rgroups;
        HIP_TRY(launch_scatter_prep(p, G.rows.ptr, G.cry.ptr, G.hol.ptr, bits_dev, pairs.ptr, counts.ptr, s));
        HIP_TRY(launch_scatter(p, G.rows.ptr, pairs.ptr, counts.ptr, pe_dev, dst_dev + (long long)b0 * row_stride, s));
    }
    HIP_TRY(hipStreamSynchronize(s));
    return SM_OK;
}

// Exact scatter of an explicit detector-row list.  Output row of rows[i] is
// rows[i] - det_begin (dst_by_row) or i.  With `items` ((index into rows, voxel
// tile) pairs) only those 64-voxel tiles are computed.  Rows are processed in
// chunks whose pair tables stay under the pair budget.
static int run_scatter_list(ScatterGeom& G, int mode, const unsigned* bits, const float* pe_dev,
                            const std::vector<int>& rows, int det_begin, bool dst_by_row,
                            const std::vector<int2>* items, float* dst_dev, long long row_stride, hipStream_t s) {
    if (rows.empty()) return SM_OK;
    ScParams p = G.base;
    p.mode = mode;
    p.max_sites = (mode == 0) ? G.T.n_det : std::max(1, G.n_holes0);
    p.row_stride = row_stride;
    p.det_begin = det_begin;
    p.dst_by_row = dst_by_row ? 1 : 0;
    set_tiles(p, true);  // listed tiles are 4 x 4 x 4 bricks too
    const long long pair_budget = 2048ll << 20;
    const int chunk = (int)std::max(1ll, std::min<long long>((long long)rows.size(),
                                                             pair_budget / ((long long)p.max_sites * sizeof(ScPair))));
    p.site_chunk = kPrepChunk;
    DeviceBuf<ScPair> pairs;
    DeviceBuf<int> counts, drows;
    DeviceBuf<int2> ditems;
    int rc;
    if ((rc = pairs.alloc((size_t)chunk * p.max_sites)) < 0) return rc;
    if ((rc = counts.alloc((size_t)chunk * site_chunks(p))) < 0) return rc;
    if ((rc = drows.alloc(chunk)) < 0) return rc;
    std::vector<int2> local;
    for (size_t c0 = 0; c0 < rows.size(); c0 += chunk) {
        const int cnt = (int)std::min<size_t>(chunk, rows.size() - c0);
        HIP_TRY(hipMemcpyAsync(drows.ptr, rows.data() + c0, sizeof(int) * cnt, hipMemcpyHostToDevice, s));
        p.row_list = drows.ptr;
        p.det_count = cnt;
        if (items) {
            local.clear();
            for (const int2& it : *items)
                if (it.x >= (int)c0 && it.x < (int)c0 + cnt) local.push_back(make_int2(it.x - (int)c0, it.y));
            if (local.empty()) continue;
        }
        HIP_TRY(launch_scatter_prep(p, G.rows.ptr, G.cry.ptr, G.hol.ptr, bits, pairs.ptr, counts.ptr, s));
        float* out = dst_by_row ? dst_dev : dst_dev + (long long)c0 * row_stride;
        if (items) {
            if ((rc = ditems.alloc(local.size())) < 0) return rc;
            HIP_TRY(hipMemcpyAsync(ditems.ptr, local.data(), sizeof(int2) * local.size(), hipMemcpyHostToDevice, s));
            p.items = ditems.ptr;
            p.n_items = (long long)local.size();
        } else {
            p.items = nullptr;
            p.n_items = 0;
            p.n_rgroups = (cnt + 3) / 4;
            p.n_blocks = (long long)p.n_tiles * p.n_rgroups;
        }
        HIP_TRY(launch_scatter(p, G.rows.ptr, pairs.ptr, counts.ptr, pe_dev, out, s));
        HIP_TRY(hipStreamSynchronize(s));  // host vectors / buffers are reused by the next chunk
    }
    HIP_TRY(hipStreamSynchronize(s));
    return SM_OK;
}

static int read_words(const char* fname, std::vector<unsigned>& w) {
    FILE* f = std::fopen(fname, "rb");
    if (!f) return set_error(SM_E_IO, std::string("cannot open ") + fname);
    size_t got = std::fread(w.data(), sizeof(unsigned), w.size(), f);
    std::fclose(f);
    if (got != w.size()) return set_error(SM_E_IO, std::string("short geometry relationship file ") + fname);
    return SM_OK;
}

static int write_words(const char* fname, const std::vector<unsigned>& w) {
    FILE* f = std::fopen(fname, "wb+");
    if (!f) return set_error(SM_E_IO, std::string("cannot create ") + fname);
    size_t put = std::fwrite(w.data(), sizeof(unsigned), w.size(), f);
    std::fclose(f);
    if (put != w.size()) return set_error(SM_E_IO, std::string("short write of ") + fname);
    return SM_OK;
}

// Device bitmap for `mode` (0 crystal, else collimator): computed, or loaded from
// the file when `compute` is false; optionally written to the file after computing.
static int make_bitmap(ScatterGeom& G, int mode, bool compute, const char* fname, bool write_file,
                       DeviceBuf<unsigned>& bits, hipStream_t s) {
    const int n = G.T.n_det;
    const long long nw = (mode == 0) ? crystal_words(n) : collimator_words(G.n_holes0, n);
    int rc;
    if ((rc = bits.alloc((size_t)std::max(1ll, nw))) < 0) return rc;
    if (compute) {
        HIP_TRY(hipMemsetAsync(bits.ptr, 0, sizeof(unsigned) * (size_t)std::max(1ll, nw), s));
        if (mode == 0)
            HIP_TRY(launch_bitmap_crystal(G.cry.ptr, n, bits.ptr, s));
        else
            HIP_TRY(launch_bitmap_collimator(G.hol.ptr, G.n_holes0, G.cry.ptr, n, bits.ptr, s));
        HIP_TRY(hipStreamSynchronize(s));
        if (write_file && fname && nw > 0) {
            std::vector<unsigned> h((size_t)nw);
            HIP_TRY(hipMemcpy(h.data(), bits.ptr, sizeof(unsigned) * (size_t)nw, hipMemcpyDeviceToHost));
            if ((rc = write_words(fname, h)) < 0) return rc;
        }
    } else {
        std::vector<unsigned> h((size_t)std::max(1ll, nw), 0u);
        if (nw > 0 && (rc = read_words(fname, h)) < 0) return rc;
        HIP_TRY(hipMemcpy(bits.ptr, h.data(), sizeof(unsigned) * h.size(), hipMemcpyHostToDevice));
    }
    return SM_OK;
}

}  // namespace sysmat

using namespace sysmat;

struct SmScatterPlan {
    ScatterGeom G;
    DeviceBuf<unsigned> bits_crystal, bits_collimator;  // empty: criterion evaluated inline
    int variant = 0;
    bool has_collimator = false;
    std::vector<float> col;   // collimator record (cell model)
    float fov2c = 0.0f;
    CellModel cells;
};

extern "C" {

long long sm_crystal_bitmap_words(int n_det) { return crystal_words(n_det); }
long long sm_collimator_bitmap_words(int n_holes0, int n_det) { return collimator_words(n_holes0, n_det); }

int sm_bitmap_crystal(const float* det, unsigned* words, long long n_words, int device) {
    clear_error();
    if (!det || !words) return set_error(SM_E_ARG, "null argument");
    int rc = use_device(device);
    if (rc < 0) return rc;
    float col0[110] = {0};
    float img0[32] = {0};
    float phy0[10] = {0};
    phy0[7] = 140.0f;
    img0[0] = img0[1] = img0[2] = 1.0f;
    ScatterGeom G;
    if ((rc = build_geom(col0, det, img0, phy0, device, 0, G)) < 0) return rc;
    if (n_words < crystal_words(G.T.n_det)) return set_error(SM_E_ARG, "bitmap buffer too small");
    DeviceBuf<unsigned> bits;
    if ((rc = make_bitmap(G, 0, true, nullptr, false, bits, nullptr)) < 0) return rc;
    HIP_TRY(hipMemcpy(words, bits.ptr, sizeof(unsigned) * crystal_words(G.T.n_det), hipMemcpyDeviceToHost));
    return SM_OK;
}

int sm_bitmap_collimator(const float* col, const float* det, unsigned* words, long long n_words, int device) {
    clear_error();
    if (!col || !det || !words) return set_error(SM_E_ARG, "null argument");
    int rc = use_device(device);
    if (rc < 0) return rc;
    float img0[32] = {0};
    float phy0[10] = {0};
    phy0[7] = 140.0f;
    img0[0] = img0[1] = img0[2] = 1.0f;
    ScatterGeom G;
    if ((rc = build_geom(col, det, img0, phy0, device, 0, G)) < 0) return rc;
    const long long nw = collimator_words(G.n_holes0, G.T.n_det);
    if (n_words < nw) return set_error(SM_E_ARG, "bitmap buffer too small");
    DeviceBuf<unsigned> bits;
    if ((rc = make_bitmap(G, 1, true, nullptr, false, bits, nullptr)) < 0) return rc;
    if (nw > 0) HIP_TRY(hipMemcpy(words, bits.ptr, sizeof(unsigned) * nw, hipMemcpyDeviceToHost));
    return SM_OK;
}

int sm_scatter_plan_create_ex(const float* col, const float* det, const float* img, const float* phy, int rotation,
                              int collimator_variant, int flags, const unsigned* crystal_words_host,
                              const unsigned* collimator_words_host, int device, SmScatterPlan** plan) {
    clear_error();
    if (!det || !img || !phy || !plan) return set_error(SM_E_ARG, "null argument");
    int rc = use_device(device);
    if (rc < 0) return rc;
    float col0[110] = {0};
    auto* P = new SmScatterPlan();
    P->variant = collimator_variant;
    rc = build_geom(col ? col : col0, det, img, phy, device, rotation, P->G);
    if (rc >= 0 && !(flags & SM_PLAN_NO_CRYSTAL_BITMAP)) {
        const int n = P->G.T.n_det;
        if (crystal_words_host) {
            const long long nw = crystal_words(n);
            if ((rc = P->bits_crystal.alloc(nw)) >= 0) {
                hipError_t e = hipMemcpy(P->bits_crystal.ptr, crystal_words_host, sizeof(unsigned) * nw,
                                         hipMemcpyHostToDevice);
                if (e != hipSuccess) rc = set_error(SM_E_DEVICE, "bitmap upload failed");
            }
        } else {
            rc = make_bitmap(P->G, 0, true, nullptr, false, P->bits_crystal, nullptr);
        }
    }
    if (rc >= 0 && col) {
        P->has_collimator = true;
        const int n_layers = std::max(0, std::min(9, (int)std::floor(col[0] + 0.000001f)));
        int n_h = 0;
        for (int m = 0; m < n_layers; ++m) n_h += std::max(0, (int)std::floor(col[(m + 1) * 10]));
        P->col.assign(col, col + 100 + 9 * (size_t)n_h);
        P->fov2c = img[11];
        const long long nw = collimator_words(P->G.n_holes0, P->G.T.n_det);
        if (flags & SM_PLAN_NO_COLLIMATOR_BITMAP) {
            // criterion evaluated inline by the pair preparation
        } else if (collimator_words_host) {
            if ((rc = P->bits_collimator.alloc(std::max(1ll, nw))) >= 0 && nw > 0) {
                hipError_t e = hipMemcpy(P->bits_collimator.ptr, collimator_words_host, sizeof(unsigned) * nw,
                                         hipMemcpyHostToDevice);
                if (e != hipSuccess) rc = set_error(SM_E_DEVICE, "bitmap upload failed");
            }
        } else {
            rc = make_bitmap(P->G, 1, true, nullptr, false, P->bits_collimator, nullptr);
        }
    }
    if (rc < 0) {
        delete P;
        return rc;
    }
    *plan = P;
    return SM_OK;
}

int sm_scatter_plan_create(const float* col, const float* det, const float* img, const float* phy, int rotation,
                           int collimator_variant, const unsigned* crystal_words_host,
                           const unsigned* collimator_words_host, int device, SmScatterPlan** plan) {
    return sm_scatter_plan_create_ex(col, det, img, phy, rotation, collimator_variant, 0, crystal_words_host,
                                     collimator_words_host, device, plan);
}

int sm_scatter_plan_crystal(SmScatterPlan* plan, const float* pe_dev, int det_begin, int det_count, float* dst_dev,
                            long long row_stride, void* stream) {
    clear_error();
    if (!plan || !pe_dev || !dst_dev) return set_error(SM_E_ARG, "null argument");
    const int n = plan->G.T.n_det;
    if (det_begin < 0 || det_count <= 0 || det_begin + det_count > n)
        return set_error(SM_E_ARG, "detector range outside the detector table");
    int rc = use_device(plan->G.device);
    if (rc < 0) return rc;
    return run_scatter(plan->G, 0, plan->bits_crystal.ptr, pe_dev, det_begin, det_count, dst_dev,
                       row_stride > 0 ? row_stride : plan->G.base.V, (hipStream_t)stream);
}

int sm_scatter_plan_crystal_rows(SmScatterPlan* plan, const float* pe_dev, const int* rows, int n_rows,
                                 float* dst_dev, long long row_stride, void* stream) {
    clear_error();
    if (!plan || !pe_dev || !rows || !dst_dev || n_rows <= 0) return set_error(SM_E_ARG, "bad argument");
    const int n = plan->G.T.n_det;
    for (int i = 0; i < n_rows; ++i)
        if (rows[i] < 0 || rows[i] >= n) return set_error(SM_E_ARG, "row outside the detector table");
    int rc = use_device(plan->G.device);
    if (rc < 0) return rc;
    std::vector<int> list(rows, rows + n_rows);
    return run_scatter_list(plan->G, 0, plan->bits_crystal.ptr, pe_dev, list, 0, false, nullptr, dst_dev,
                            row_stride > 0 ? row_stride : plan->G.base.V, (hipStream_t)stream);
}

int sm_scatter_plan_collimator(SmScatterPlan* plan, int det_begin, int det_count, float* dst_dev,
                               long long row_stride, void* stream) {
    clear_error();
    if (!plan || !dst_dev) return set_error(SM_E_ARG, "null argument");
    if (!plan->has_collimator) return set_error(SM_E_ARG, "plan created without a collimator");
    const int n = plan->G.T.n_det;
    if (det_begin < 0 || det_count <= 0 || det_begin + det_count > n)
        return set_error(SM_E_ARG, "detector range outside the detector table");
    int rc = use_device(plan->G.device);
    if (rc < 0) return rc;
    return run_scatter(plan->G, plan->variant == 1 ? 2 : 1, plan->bits_collimator.ptr, nullptr, det_begin, det_count,
                       dst_dev, row_stride > 0 ? row_stride : plan->G.base.V, (hipStream_t)stream);
}

int sm_scatter_plan_cells(SmScatterPlan* plan, float cell_size, float keep_dist, int det_begin, int det_count,
                          float* dst_dev, long long row_stride, int cells_per_chunk, void* stream) {
    clear_error();
    if (!plan || !dst_dev) return set_error(SM_E_ARG, "null argument");
    if (!plan->has_collimator) return set_error(SM_E_ARG, "plan created without a collimator");
    if (!(cell_size > 0.0f)) return set_error(SM_E_ARG, "cell size must be positive");
    const int n = plan->G.T.n_det;
    if (det_begin < 0 || det_count <= 0 || det_begin + det_count > n)
        return set_error(SM_E_ARG, "detector range outside the detector table");
    int rc = use_device(plan->G.device);
    if (rc < 0) return rc;
    if ((rc = prepare_cells(plan->col.data(), plan->fov2c, cell_size, keep_dist, plan->cells)) < 0) return rc;
    rc = run_scatter_cells(plan->G, plan->cells, det_begin, det_count, dst_dev,
                           row_stride > 0 ? row_stride : plan->G.base.V, cells_per_chunk, (hipStream_t)stream);
    if (rc < 0) return rc;
    HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
    return SM_OK;
}

long long sm_scatter_cell_table(SmScatterPlan* plan, float cell_size, float keep_dist, double* out, long long cap) {
    clear_error();
    if (!plan) return set_error(SM_E_ARG, "null argument");
    if (!plan->has_collimator) return set_error(SM_E_ARG, "plan created without a collimator");
    if (!(cell_size > 0.0f)) return set_error(SM_E_ARG, "cell size must be positive");
    int rc = use_device(plan->G.device);
    if (rc < 0) return rc;
    if ((rc = prepare_cells(plan->col.data(), plan->fov2c, cell_size, keep_dist, plan->cells)) < 0) return rc;
    const long long nc = (long long)plan->cells.cells.size();
    if (out) {
        for (long long i = 0; i < std::min(nc, cap); ++i) {
            const Cell& c = plan->cells.cells[(size_t)i];
            out[4 * i + 0] = c.x;
            out[4 * i + 1] = c.y;
            out[4 * i + 2] = c.z;
            out[4 * i + 3] = c.w;
        }
    }
    return nc;
}

int sm_scatter_plan_destroy(SmScatterPlan* plan) {
    if (plan) {
        use_device(plan->G.device);
        delete plan;
    }
    return SM_OK;
}

// Host-memory convenience: crystal scatter rows of the N x V PE matrix `pe`.
int sm_scatter_crystal(const float* det, const float* img, const float* phy, const float* pe,
                       const unsigned* crystal_words_host, int rotation, int det_begin, int det_count, float* dst,
                       int device) {
    clear_error();
    if (!pe || !dst) return set_error(SM_E_ARG, "null argument");
    SmScatterPlan* plan = nullptr;
    int rc = sm_scatter_plan_create(nullptr, det, img, phy, rotation, 0, crystal_words_host, nullptr, device, &plan);
    if (rc < 0) return rc;
    const long long V = plan->G.base.V, N = plan->G.T.n_det;
    if (det_count <= 0) det_count = (int)N - det_begin;
    if (det_begin < 0 || det_count <= 0 || det_begin + (long long)det_count > N) {
        sm_scatter_plan_destroy(plan);
        return set_error(SM_E_ARG, "detector range outside the detector table");
    }
    DeviceBuf<float> dpe, dout;
    if ((rc = dpe.alloc((size_t)(N * V))) >= 0 && (rc = dout.alloc((size_t)det_count * V)) >= 0) {
        hipError_t e = hipMemcpy(dpe.ptr, pe, sizeof(float) * N * V, hipMemcpyHostToDevice);
        if (e != hipSuccess) rc = set_error(SM_E_DEVICE, "PE upload failed");
        if (rc >= 0) rc = sm_scatter_plan_crystal(plan, dpe.ptr, det_begin, det_count, dout.ptr, V, nullptr);
        if (rc >= 0) {
            e = hipMemcpy(dst, dout.ptr, sizeof(float) * det_count * V, hipMemcpyDeviceToHost);
            if (e != hipSuccess) rc = set_error(SM_E_DEVICE, "result download failed");
        }
    }
    sm_scatter_plan_destroy(plan);
    return rc;
}

int sm_scatter_collimator(const float* col, const float* det, const float* img, const float* phy,
                          const unsigned* collimator_words_host, int rotation, int variant, int det_begin,
                          int det_count, float* dst, int device) {
    clear_error();
    if (!dst) return set_error(SM_E_ARG, "null argument");
    SmScatterPlan* plan = nullptr;
    int rc = sm_scatter_plan_create(col, det, img, phy, rotation, variant, nullptr, collimator_words_host, device, &plan);
    if (rc < 0) return rc;
    const long long V = plan->G.base.V, N = plan->G.T.n_det;
    if (det_count <= 0) det_count = (int)N - det_begin;
    if (det_begin < 0 || det_count <= 0 || det_begin + (long long)det_count > N) {
        sm_scatter_plan_destroy(plan);
        return set_error(SM_E_ARG, "detector range outside the detector table");
    }
    DeviceBuf<float> dout;
    if ((rc = dout.alloc((size_t)det_count * V)) >= 0) {
        rc = sm_scatter_plan_collimator(plan, det_begin, det_count, dout.ptr, V, nullptr);
        if (rc >= 0) {
            hipError_t e = hipMemcpy(dst, dout.ptr, sizeof(float) * det_count * V, hipMemcpyDeviceToHost);
            if (e != hipSuccess) rc = set_error(SM_E_DEVICE, "result download failed");
        }
    }
    sm_scatter_plan_destroy(plan);
    return rc;
}

// Host-memory convenience: plate-cell collimator scatter rows.
int sm_scatter_cells(const float* col, const float* det, const float* img, const float* phy, int rotation,
                     float cell_size, float keep_dist, int det_begin, int det_count, float* dst, int device) {
    clear_error();
    if (!col || !dst) return set_error(SM_E_ARG, "null argument");
    SmScatterPlan* plan = nullptr;
    int rc = sm_scatter_plan_create_ex(col, det, img, phy, rotation, 0,
                                       SM_PLAN_NO_CRYSTAL_BITMAP | SM_PLAN_NO_COLLIMATOR_BITMAP, nullptr, nullptr,
                                       device, &plan);
    if (rc < 0) return rc;
    const long long V = plan->G.base.V, N = plan->G.T.n_det;
    if (det_count <= 0) det_count = (int)N - det_begin;
    if (det_begin < 0 || det_count <= 0 || det_begin + (long long)det_count > N) {
        sm_scatter_plan_destroy(plan);
        return set_error(SM_E_ARG, "detector range outside the detector table");
    }
    DeviceBuf<float> dout;
    if ((rc = dout.alloc((size_t)det_count * V)) >= 0) {
        rc = sm_scatter_plan_cells(plan, cell_size, keep_dist, det_begin, det_count, dout.ptr, V, 0, nullptr);
        if (rc >= 0) {
            hipError_t e = hipMemcpy(dst, dout.ptr, sizeof(float) * det_count * V, hipMemcpyDeviceToHost);
            if (e != hipSuccess) rc = set_error(SM_E_DEVICE, "result download failed");
        }
    }
    sm_scatter_plan_destroy(plan);
    return rc;
}

// Drop-in for the combined scatter() (scatter.cu:2013-2378): crystal and
// collimator terms of one rotation (Img[20]), summed into dst (N x V).  The
// bitmaps are computed (and written to the files) when Phys[8] / Phys[9] are 1,
// otherwise read from the files, as the reference does.
int sm_scatter(const float* col, const float* det, const float* img, const float* phy, const float* pe,
               const char* fname_geo_crystal, const char* fname_geo_collimator, float* dst, int cuda_id) {
    clear_error();
    if (!col || !det || !img || !phy || !pe || !dst) return set_error(SM_E_ARG, "null argument");
    int rc = use_device(cuda_id);
    if (rc < 0) return rc;
    const int rot = (int)std::floor(img[20] + 0.001f);
    auto* P = new SmScatterPlan();
    std::unique_ptr<SmScatterPlan> guard(P);
    if ((rc = build_geom(col, det, img, phy, cuda_id, rot, P->G)) < 0) return rc;
    const bool calc_c = (int)std::floor(phy[8]) == 1, calc_h = (int)std::floor(phy[9] + 0.001f) == 1;
    if ((rc = make_bitmap(P->G, 0, calc_c, fname_geo_crystal, true, P->bits_crystal, nullptr)) < 0) return rc;
    if ((rc = make_bitmap(P->G, 1, calc_h, fname_geo_collimator, true, P->bits_collimator, nullptr)) < 0) return rc;
    const long long V = P->G.base.V, N = P->G.T.n_det;
    DeviceBuf<float> dpe, dc, dh;
    if ((rc = dpe.alloc((size_t)(N * V))) < 0 || (rc = dc.alloc((size_t)(N * V))) < 0 ||
        (rc = dh.alloc((size_t)(N * V))) < 0)
        return rc;
    HIP_TRY(hipMemcpy(dpe.ptr, pe, sizeof(float) * N * V, hipMemcpyHostToDevice));
    if ((rc = run_scatter(P->G, 0, P->bits_crystal.ptr, dpe.ptr, 0, (int)N, dc.ptr, V, nullptr)) < 0) return rc;
    if ((rc = run_scatter(P->G, 1, P->bits_collimator.ptr, nullptr, 0, (int)N, dh.ptr, V, nullptr)) < 0) return rc;
    std::vector<float> a((size_t)(N * V)), b((size_t)(N * V));
    HIP_TRY(hipMemcpy(a.data(), dc.ptr, sizeof(float) * N * V, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(b.data(), dh.ptr, sizeof(float) * N * V, hipMemcpyDeviceToHost));
    for (long long i = 0; i < N * V; ++i) dst[i] = b[i] + a[i];  // collimator + crystal (scatter.cu:2354-2357)
    return (int)V;
}

// Drop-in for ScatterGen_Collimator's scatter() (ScatterGen_Collimator/scatter.cu:1511-1747).
int sm_scatter_collimator_only(const float* col, const float* det, const float* img, const float* phy,
                               const float* pe, const char* fname_geo, float* dst, int cuda_id) {
    clear_error();
    (void)pe;
    if (!col || !det || !img || !phy || !dst) return set_error(SM_E_ARG, "null argument");
    int rc = use_device(cuda_id);
    if (rc < 0) return rc;
    const int rot = (int)std::floor(img[20] + 0.001f);
    auto* P = new SmScatterPlan();
    std::unique_ptr<SmScatterPlan> guard(P);
    if ((rc = build_geom(col, det, img, phy, cuda_id, rot, P->G)) < 0) return rc;
    const bool calc_h = (int)std::floor(phy[9]) == 1;
    if ((rc = make_bitmap(P->G, 1, calc_h, fname_geo, true, P->bits_collimator, nullptr)) < 0) return rc;
    const long long V = P->G.base.V, N = P->G.T.n_det;
    DeviceBuf<float> dh;
    if ((rc = dh.alloc((size_t)(N * V))) < 0) return rc;
    if ((rc = run_scatter(P->G, 2, P->bits_collimator.ptr, nullptr, 0, (int)N, dh.ptr, V, nullptr)) < 0) return rc;
    HIP_TRY(hipMemcpy(dst, dh.ptr, sizeof(float) * N * V, hipMemcpyDeviceToHost));
    return (int)V;
}

// Drop-in for ScatterGen_Crystal's scatter() (ScatterGen_Crystal/scatter.cu:1528-1738).
int sm_scatter_crystal_only(const float* det, const float* img, const float* phy, const float* pe,
                            const char* fname_geo, float* dst, int cuda_id) {
    clear_error();
    if (!det || !img || !phy || !pe || !dst) return set_error(SM_E_ARG, "null argument");
    int rc = use_device(cuda_id);
    if (rc < 0) return rc;
    const int rot = (int)std::floor(img[20] + 0.001f);
    float col0[110] = {0};
    auto* P = new SmScatterPlan();
    std::unique_ptr<SmScatterPlan> guard(P);
    if ((rc = build_geom(col0, det, img, phy, cuda_id, rot, P->G)) < 0) return rc;
    const bool calc_c = (int)std::floor(phy[8]) == 1;
    if ((rc = make_bitmap(P->G, 0, calc_c, fname_geo, true, P->bits_crystal, nullptr)) < 0) return rc;
    const long long V = P->G.base.V, N = P->G.T.n_det;
    DeviceBuf<float> dpe, dc;
    if ((rc = dpe.alloc((size_t)(N * V))) < 0 || (rc = dc.alloc((size_t)(N * V))) < 0) return rc;
    HIP_TRY(hipMemcpy(dpe.ptr, pe, sizeof(float) * N * V, hipMemcpyHostToDevice));
    if ((rc = run_scatter(P->G, 0, P->bits_crystal.ptr, dpe.ptr, 0, (int)N, dc.ptr, V, nullptr)) < 0) return rc;
    HIP_TRY(hipMemcpy(dst, dc.ptr, sizeof(float) * N * V, hipMemcpyDeviceToHost));
    return (int)V;
}

}  // extern "C"

// C++-linkage drop-ins with the reference's exact signatures.
int scatter(float* parameter_Collimator, float* parameter_Detector, float* parameter_Image, float* parameter_Physics,
            float* PE_SysMat, const char* FnameGeoCrystal, const char* FnameGeoCollimator, float* dst, int cuda_id) {
    int rc = sm_scatter(parameter_Collimator, parameter_Detector, parameter_Image, parameter_Physics, PE_SysMat,
                        FnameGeoCrystal, FnameGeoCollimator, dst, cuda_id);
    if (rc < 0) std::fprintf(stderr, "scatter: %s\n", sm_last_error());
    return rc;
}

int scatter(float* parameter_Collimator, float* parameter_Detector, float* parameter_Image, float* parameter_Physics,
            float* PE_SysMat, const char* FnameGeo, float* dst, int cuda_id) {
    int rc = sm_scatter_collimator_only(parameter_Collimator, parameter_Detector, parameter_Image, parameter_Physics,
                                        PE_SysMat, FnameGeo, dst, cuda_id);
    if (rc < 0) std::fprintf(stderr, "scatter: %s\n", sm_last_error());
    return rc;
}

int scatter(float* parameter_Detector, float* parameter_Image, float* parameter_Physics, float* PE_SysMat,
            const char* FnameGeo, float* dst, int cuda_id) {
    int rc = sm_scatter_crystal_only(parameter_Detector, parameter_Image, parameter_Physics, PE_SysMat, FnameGeo, dst,
                                     cuda_id);
    if (rc < 0) std::fprintf(stderr, "scatter: %s\n", sm_last_error());
    return rc;
}

// ---------------------------------------------------------------------------
// Scatter surrogate (crystal scatter): exact near field + learned far field
// ---------------------------------------------------------------------------

struct SmSurrogate {
    ScatterGeom G;
    DeviceBuf<unsigned> bits;  // crystal bitmap (empty: criterion inline)
    DeviceBuf<ScPair> keep;    // [N][n_keep] ranked sites
    DeviceBuf<int> keep_cnt;   // [N]
    DeviceBuf<SgRow> srow;     // [N]
    DeviceBuf<float> ftab;     // [N][kSgTab + 1] angular factor of far sites
    DeviceBuf<float4> ring;    // [N][n_keep][2] compact ranked sites for the ring prior
    DeviceBuf<int> ring_cnt;   // [N] the ring prior's loop bound
    SgParams q{};
    int n_near = 0, n_ring = 0;
};

namespace {

// the n_cand crystals nearest to each crystal (centre distance, index order on ties)
void nearest_candidates(const Tables& T, int n_cand, std::vector<int>& out) {
    const int n = T.n_det;
    out.assign((size_t)n * n_cand, -1);
    auto work = [&](int r0, int r1) {
        std::vector<std::pair<double, int>> d;
        d.reserve(n);
        for (int r = r0; r < r1; ++r) {
            d.clear();
            const Crystal& R = T.crystals[r];
            for (int k = 0; k < n; ++k) {
                if (k == r) continue;
                const Crystal& K = T.crystals[k];
                const double dx = (double)K.cx - R.cx, dy = (double)K.y_raw - R.y_raw, dz = (double)K.cz - R.cz;
                d.emplace_back(dx * dx + dy * dy + dz * dz, k);
            }
            const int m = std::min<int>(n_cand, (int)d.size());
            std::partial_sort(d.begin(), d.begin() + m, d.end());
            for (int i = 0; i < m; ++i) out[(size_t)r * n_cand + i] = d[i].second;
        }
    };
    const int nt = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    const int per = (n + nt - 1) / nt;
    for (int t = 0; t < nt; ++t) {
        const int a = t * per, b = std::min(n, a + per);
        if (a < b) th.emplace_back(work, a, b);
    }
    for (auto& x : th) x.join();
}

int check_rows(const SmSurrogate* S, int det_begin, int det_count) {
    if (det_begin < 0 || det_count <= 0 || det_begin + det_count > S->G.T.n_det)
        return set_error(SM_E_ARG, "detector range outside the detector table");
    return SM_OK;
}

}  // namespace

extern "C" {

int sm_surrogate_create(const float* det, const float* img, const float* phy, int rotation,
                        const unsigned* crystal_words_host, int use_bitmap, int n_near, int n_ring, int device,
                        SmSurrogate** out) {
    clear_error();
    if (!det || !img || !phy || !out) return set_error(SM_E_ARG, "null argument");
    if (n_near < 1 || n_ring < 0 || n_near + n_ring > 256) return set_error(SM_E_ARG, "need 1 <= n_near, n_near + n_ring <= 256");
    int rc = use_device(device);
    if (rc < 0) return rc;
    float col0[110] = {0};
    auto* S = new SmSurrogate();
    auto fail = [&](int code) {
        delete S;
        return code;
    };
    if ((rc = build_geom(col0, det, img, phy, device, rotation, S->G)) < 0) return fail(rc);
    const int n = S->G.T.n_det;
    if (use_bitmap) {
        if (crystal_words_host) {
            const long long nw = crystal_words(n);
            if ((rc = S->bits.alloc(nw)) < 0) return fail(rc);
            if (hipMemcpy(S->bits.ptr, crystal_words_host, sizeof(unsigned) * nw, hipMemcpyHostToDevice) != hipSuccess)
                return fail(set_error(SM_E_DEVICE, "bitmap upload failed"));
        } else if ((rc = make_bitmap(S->G, 0, true, nullptr, false, S->bits, nullptr)) < 0) {
            return fail(rc);
        }
    }
    S->n_near = n_near;
    S->n_ring = n_ring;
    const int n_keep = n_near + n_ring;
    const int n_cand = std::min(256, n - 1);
    std::vector<int> cand;
    nearest_candidates(S->G.T, std::max(1, n_cand), cand);
    DeviceBuf<int> dcand;
    if ((rc = dcand.upload(cand)) < 0) return fail(rc);
    if ((rc = S->keep.alloc((size_t)n * n_keep)) < 0) return fail(rc);
    if ((rc = S->keep_cnt.alloc(n)) < 0) return fail(rc);
    ScParams p = S->G.base;
    p.mode = 0;
    p.det_begin = 0;
    p.det_count = n;
    // angular factor per detector: KN shape x window probability at the scattered energy
    {
        std::vector<float> tab((size_t)n * (kSgTab + 1));
        const double e0 = S->G.base.e0, ea = e0 / 511.0;
        for (int r = 0; r < n; ++r) {
            const ScRow& R = S->G.hrows[r];
            for (int i = 0; i <= kSgTab; ++i) {
                const double c = -1.0 + 2.0 * i / kSgTab;
                const double a = ea * (1.0 - c), c2 = 1.0 + c * c;
                const double kn = c2 / ((1.0 + a) * (1.0 + a)) * (1.0 + a * a / (c2 * (1.0 + a)));
                const double es = e0 / (1.0 + a);
                const double res_s = R.res * std::sqrt(es / e0);
                const double k2 = 2.0 * res_s / 2.35482;
                double pw = 0.0;
                if (!(es * (1.0 + k2) <= R.lo || es * (1.0 - k2) >= R.hi)) {
                    const double s2 = res_s / 2.35482 * es * std::sqrt(2.0);
                    pw = 0.5 * (std::erf((R.hi - es) / s2) - std::erf((R.lo - es) / s2));
                }
                tab[(size_t)r * (kSgTab + 1) + i] = (float)(kn * pw);
            }
        }
        if ((rc = S->ftab.upload(tab)) < 0) return fail(rc);
    }
    SurrogateRank k{dcand.ptr, std::max(1, n_cand), n_keep, S->ftab.ptr};
    if (n_cand >= 1) {
        HIP_TRY(launch_surrogate_rank(p, k, S->G.rows.ptr, S->G.cry.ptr, S->bits.ptr, S->keep.ptr, S->keep_cnt.ptr,
                                      nullptr));
    } else {
        HIP_TRY(hipMemset(S->keep_cnt.ptr, 0, sizeof(int) * n));
    }
    if ((rc = S->ring.alloc((size_t)n * n_keep * 2)) < 0) return fail(rc);
    if ((rc = S->ring_cnt.alloc((size_t)n)) < 0) return fail(rc);
    HIP_TRY(launch_surrogate_ring_table(S->keep.ptr, S->keep_cnt.ptr, n, n_keep, n_near, S->ring.ptr,
                                        S->ring_cnt.ptr, nullptr));
    HIP_TRY(hipDeviceSynchronize());
    // per-row feature constants
    float bmin[3] = {1e30f, 1e30f, 1e30f}, bmax[3] = {-1e30f, -1e30f, -1e30f};
    double cen[3] = {0, 0, 0};
    for (const Crystal& c : S->G.T.crystals) {
        const float v[3] = {c.cx, c.cy, c.cz};
        for (int a = 0; a < 3; ++a) {
            bmin[a] = std::min(bmin[a], v[a]);
            bmax[a] = std::max(bmax[a], v[a]);
            cen[a] += v[a];
        }
    }
    for (double& c : cen) c /= n;
    std::vector<SgRow> rows(n);
    for (int r = 0; r < n; ++r) {
        const Crystal& c = S->G.T.crystals[r];
        SgRow& R = rows[r];
        R.cx = c.cx;
        R.cy = c.cy;
        R.cz = c.cz;
        R.use = S->G.hrows[r].use;
        const float v[3] = {c.cx, c.cy, c.cz};
        float f[3];
        for (int a = 0; a < 3; ++a) {
            const float mid = 0.5f * (bmin[a] + bmax[a]), half = std::max(0.5f * (bmax[a] - bmin[a]), 1.0f);
            f[a] = (v[a] - mid) / half;
        }
        R.f3 = f[0];
        R.f4 = f[1];
        R.f5 = f[2];
    }
    if ((rc = S->srow.upload(rows)) < 0) return fail(rc);
    SgParams& q = S->q;
    const ScParams& b = S->G.base;
    q.X = b.X; q.Y = b.Y; q.Z = b.Z;
    q.wx = (float)b.wx; q.wy = (float)b.wy; q.wz = (float)b.wz;
    q.sx = (float)b.sx; q.sy = (float)b.sy; q.sz = (float)b.sz;
    q.rc = (float)b.rc; q.rs = (float)b.rs;
    q.icx = (float)(b.sx * b.rc - b.sy * b.rs);
    q.icy = (float)(b.sx * b.rs + b.sy * b.rc);
    q.icz = (float)b.sz;
    q.irho = (float)(0.5 * std::sqrt((double)b.X * b.X * b.wx * b.wx + (double)b.Y * b.Y * b.wy * b.wy +
                                     (double)b.Z * b.Z * b.wz * b.wz));
    const double ddx = cen[0] - q.icx, ddy = cen[1] - q.icy, ddz = cen[2] - q.icz;
    q.d_ref = (float)std::max(1.0, std::sqrt(ddx * ddx + ddy * ddy + ddz * ddz));
    q.V = b.V;
    q.n_det = n;
    q.n_near = n_near;
    q.n_keep = n_keep;
    q.pe_scale = 1.0f;
    q.s_scale = 1.0f;
    q.mu_y = 0.0f;
    q.sig_y = 1.0f;
    q.phi = 0.0f;
    q.n_tiles = (long long)((b.X + 3) / 4) * ((b.Y + 3) / 4) * ((b.Z + 3) / 4);
    q.ftab = S->ftab.ptr;
    q.ring = S->ring.ptr;
    q.ring_cnt = S->ring_cnt.ptr;
    *out = S;
    return SM_OK;
}

int sm_surrogate_set_norm(SmSurrogate* S, float pe_scale, float s_scale, float mu_y, float sig_y, float phi) {
    clear_error();
    if (!S || !(pe_scale > 0.0f) || !(s_scale > 0.0f) || !(sig_y > 0.0f) || !(phi >= 0.0f))
        return set_error(SM_E_ARG, "bad normalisation");
    S->q.pe_scale = pe_scale;
    S->q.s_scale = s_scale;
    S->q.mu_y = mu_y;
    S->q.sig_y = sig_y;
    S->q.phi = phi;
    return SM_OK;
}

int sm_surrogate_row_sites(const SmSurrogate* S, int row, int* sites, float* weights, int cap) {
    clear_error();
    if (!S || row < 0 || row >= S->G.T.n_det) return set_error(SM_E_ARG, "bad row");
    const int nk = S->q.n_keep;
    int cnt = 0;
    HIP_TRY(hipMemcpy(&cnt, S->keep_cnt.ptr + row, sizeof(int), hipMemcpyDeviceToHost));
    std::vector<ScPair> P(std::max(1, cnt));
    if (cnt > 0)
        HIP_TRY(hipMemcpy(P.data(), S->keep.ptr + (long long)row * nk, sizeof(ScPair) * cnt, hipMemcpyDeviceToHost));
    for (int i = 0; i < cnt && i < cap; ++i) {
        if (sites) sites[i] = P[i].site;
        if (weights) weights[i] = P[i].pad;
    }
    return cnt;
}

int sm_surrogate_near(SmSurrogate* S, const float* pe_dev, int det_begin, int det_count, float* dst_dev,
                      long long row_stride, void* stream) {
    clear_error();
    if (!S || !pe_dev || !dst_dev) return set_error(SM_E_ARG, "null argument");
    int rc = check_rows(S, det_begin, det_count);
    if (rc < 0) return rc;
    if ((rc = use_device(S->G.device)) < 0) return rc;
    ScParams p = S->G.base;
    p.mode = 0;
    p.det_begin = det_begin;
    p.det_count = det_count;
    p.max_sites = S->q.n_keep;
    p.site_cap = S->n_near;
    p.row_stride = row_stride > 0 ? row_stride : p.V;
    set_tiles(p, true);
    p.n_rgroups = (det_count + 3) / 4;
    p.n_blocks = (long long)p.n_tiles * p.n_rgroups;
    HIP_TRY(launch_scatter(p, S->G.rows.ptr, S->keep.ptr + (long long)det_begin * S->q.n_keep,
                           S->keep_cnt.ptr + det_begin, pe_dev, dst_dev, (hipStream_t)stream));
    return SM_OK;
}

int sm_surrogate_exact_rows(SmSurrogate* S, const float* pe_dev, const int* rows, int n_rows, float* dst_dev,
                            long long row_stride, void* stream) {
    clear_error();
    if (!S || !pe_dev || !rows || !dst_dev || n_rows <= 0) return set_error(SM_E_ARG, "bad argument");
    for (int i = 0; i < n_rows; ++i)
        if (rows[i] < 0 || rows[i] >= S->G.T.n_det) return set_error(SM_E_ARG, "row outside the detector table");
    int rc = use_device(S->G.device);
    if (rc < 0) return rc;
    std::vector<int> list(rows, rows + n_rows);
    return run_scatter_list(S->G, 0, S->bits.ptr, pe_dev, list, 0, false, nullptr, dst_dev,
                            row_stride > 0 ? row_stride : S->G.base.V, (hipStream_t)stream);
}

int sm_surrogate_features(SmSurrogate* S, const float* pe_dev, const float* near_dev, long long near_stride,
                          int det_begin, int det_count, const int* vox_dev, int n_vox, float* x_dev, void* stream) {
    clear_error();
    if (!S || !pe_dev || !near_dev || !x_dev || n_vox <= 0) return set_error(SM_E_ARG, "bad argument");
    int rc = check_rows(S, det_begin, det_count);
    if (rc < 0) return rc;
    if ((rc = use_device(S->G.device)) < 0) return rc;
    SgParams q = S->q;
    q.det_begin = det_begin;
    q.det_count = det_count;
    HIP_TRY(launch_surrogate_features(q, S->srow.ptr, S->keep.ptr, S->keep_cnt.ptr, pe_dev, near_dev,
                                      near_stride > 0 ? near_stride : q.V, vox_dev, n_vox, x_dev,
                                      (hipStream_t)stream));
    return SM_OK;
}

int sm_surrogate_apply(SmSurrogate* S, const void* packed_dev, int n_members, const float* pe_dev, int det_begin,
                       int det_count, float* dst_dev, long long row_stride, float tol, float floor_abs,
                       unsigned char* tile_flags_dev, void* stream) {
    clear_error();
    if (!S || !packed_dev || !pe_dev || !dst_dev || n_members < 1) return set_error(SM_E_ARG, "bad argument");
    int rc = check_rows(S, det_begin, det_count);
    if (rc < 0) return rc;
    if ((rc = use_device(S->G.device)) < 0) return rc;
    SgParams q = S->q;
    q.det_begin = det_begin;
    q.det_count = det_count;
    q.row_stride = row_stride > 0 ? row_stride : q.V;
    q.n_members = n_members;
    q.tol = tol;
    q.floor_abs = floor_abs;
    HIP_TRY(launch_surrogate_apply(q, S->srow.ptr, S->keep.ptr, S->keep_cnt.ptr, (const uint8_t*)packed_dev, pe_dev,
                                   dst_dev, tile_flags_dev, (hipStream_t)stream));
    return SM_OK;
}

long long sm_surrogate_fallback(SmSurrogate* S, const float* pe_dev, int det_begin, int det_count,
                                const unsigned char* tile_flags_dev, float* dst_dev, long long row_stride,
                                void* stream) {
    clear_error();
    if (!S || !pe_dev || !tile_flags_dev || !dst_dev) return set_error(SM_E_ARG, "null argument");
    int rc = check_rows(S, det_begin, det_count);
    if (rc < 0) return rc;
    if ((rc = use_device(S->G.device)) < 0) return rc;
    hipStream_t s = (hipStream_t)stream;
    const long long nt = S->q.n_tiles;
    std::vector<unsigned char> f((size_t)det_count * nt);
    HIP_TRY(hipMemcpyAsync(f.data(), tile_flags_dev, f.size(), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    std::vector<int> rows;
    std::vector<int2> items;
    for (int rr = 0; rr < det_count; ++rr) {
        bool any = false;
        for (long long t = 0; t < nt; ++t) {
            if (!f[(size_t)rr * nt + t]) continue;
            if (!any) {
                rows.push_back(det_begin + rr);
                any = true;
            }
            items.push_back(make_int2((int)rows.size() - 1, (int)t));
        }
    }
    if (items.empty()) return 0;
    rc = run_scatter_list(S->G, 0, S->bits.ptr, pe_dev, rows, det_begin, true, &items, dst_dev,
                          row_stride > 0 ? row_stride : S->q.V, s);
    if (rc < 0) return rc;
    return (long long)items.size();
}

long long sm_surrogate_tile_count(const SmSurrogate* S) { return S ? S->q.n_tiles : (long long)SM_E_ARG; }

int sm_surrogate_destroy(SmSurrogate* S) {
    if (S) {
        use_device(S->G.device);
        delete S;
    }
    return SM_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batched scatter over collimator variants (shared detector and image): every
// variant's collimator term then crystal term onto the same rows, all kernels
// on one stream with preallocated workspace, optionally captured once into a
// hipGraph and replayed.
// ---------------------------------------------------------------------------
struct SmScatterBatch {
    std::vector<std::unique_ptr<ScatterGeom>> G;
    DeviceBuf<ScPair> pairs;
    DeviceBuf<int> counts;
    int block = 0;  // detector rows per prep block
    int device = 0;
    hipStream_t own_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<const float*> cap_pe;
    std::vector<float*> cap_dst;
    hipStream_t cap_stream = nullptr;
    ~SmScatterBatch() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (own_stream) (void)hipStreamDestroy(own_stream);
    }
};

namespace {

// Kernel launches only (capturable): for every variant and row block, collimator
// prep + accumulate (write), crystal prep + accumulate (add).
int batch_enqueue(SmScatterBatch& B, const float* const* pe, float* const* dst, hipStream_t s) {
    for (size_t vi = 0; vi < B.G.size(); ++vi) {
        ScatterGeom& G = *B.G[vi];
        const int n = G.T.n_det;
        for (int b0 = 0; b0 < n; b0 += B.block) {
            const int cnt = std::min(B.block, n - b0);
            for (int term = 0; term < 2; ++term) {  // 0 collimator (writes), 1 crystal (adds)
                ScParams p = G.base;
                p.mode = term == 0 ? 1 : 0;
                p.max_sites = term == 0 ? std::max(1, G.n_holes0) : n;
                p.site_chunk = kPrepChunk;
                p.det_begin = b0;
                p.det_count = cnt;
                p.row_stride = p.V;
                set_tiles(p, true);
                p.n_rgroups = (cnt + 3) / 4;
                p.n_blocks = (long long)p.n_tiles * p.n_rgroups;
                p.accumulate = term;
                if (term == 0 && G.n_holes0 <= 0) {
                    // no layer-0 holes: a zero collimator term; the crystal term then writes
                    continue;
                }
                if (term == 1 && G.n_holes0 <= 0) p.accumulate = 0;
                HIP_TRY(launch_scatter_prep(p, G.rows.ptr, G.cry.ptr, G.hol.ptr, nullptr, B.pairs.ptr, B.counts.ptr, s));
                HIP_TRY(launch_scatter(p, G.rows.ptr, B.pairs.ptr, B.counts.ptr, pe[vi], dst[vi] + (long long)b0 * p.V, s));
            }
        }
    }
    return SM_OK;
}

}  // namespace

extern "C" {

int sm_scatter_batch_create(int n_variants, const float* const* cols, const float* det, const float* img,
                            const float* phy, int rotation, int device, SmScatterBatch** out) {
    clear_error();
    if (n_variants < 1 || !cols || !det || !img || !phy || !out) return set_error(SM_E_ARG, "bad argument");
    int rc = use_device(device);
    if (rc < 0) return rc;
    auto* B = new SmScatterBatch();
    B->device = device;
    auto fail = [&](int code) {
        delete B;
        return code;
    };
    long long max_sites = 1;
    for (int v = 0; v < n_variants; ++v) {
        if (!cols[v]) return fail(set_error(SM_E_ARG, "null collimator parameters"));
        B->G.emplace_back(new ScatterGeom());
        if ((rc = build_geom(cols[v], det, img, phy, device, rotation, *B->G.back())) < 0) return fail(rc);
        if (B->G.back()->base.V != B->G[0]->base.V || B->G.back()->T.n_det != B->G[0]->T.n_det)
            return fail(set_error(SM_E_ARG, "variants must share the detector and the image"));
        max_sites = std::max<long long>(max_sites, std::max(B->G.back()->T.n_det, B->G.back()->n_holes0));
    }
    const long long pair_budget = 2048ll << 20;
    B->block = (int)std::max(4ll, std::min<long long>(B->G[0]->T.n_det, pair_budget / (max_sites * (long long)sizeof(ScPair))));
    B->block = (B->block + 3) / 4 * 4;
    if ((rc = B->pairs.alloc((size_t)B->block * max_sites)) < 0) return fail(rc);
    if ((rc = B->counts.alloc((size_t)B->block * (size_t)((max_sites + kPrepChunk - 1) / kPrepChunk))) < 0)
        return fail(rc);
    *out = B;
    return SM_OK;
}

// dst_v = collimator scatter + crystal scatter of variant v (N x V floats each, device).
// use_graph: the first call captures the launch sequence into a hipGraph (on
// `stream`, or an internal stream when NULL) and every call replays it; a call
// with other buffers re-captures.  Synchronous.
int sm_scatter_batch_run(SmScatterBatch* B, const float* const* pe_devs, float* const* dst_devs, void* stream,
                         int use_graph) {
    clear_error();
    if (!B || !pe_devs || !dst_devs) return set_error(SM_E_ARG, "null argument");
    int rc = use_device(B->device);
    if (rc < 0) return rc;
    hipStream_t s = (hipStream_t)stream;
    if (!use_graph) {
        if ((rc = batch_enqueue(*B, pe_devs, dst_devs, s)) < 0) return rc;
        HIP_TRY(hipStreamSynchronize(s));
        return SM_OK;
    }
    if (!s) {
        if (!B->own_stream) HIP_TRY(hipStreamCreateWithFlags(&B->own_stream, hipStreamNonBlocking));
        s = B->own_stream;
    }
    const size_t nv = B->G.size();
    bool same = B->exec && B->cap_stream == s && B->cap_pe.size() == nv;
    for (size_t v = 0; same && v < nv; ++v) same = B->cap_pe[v] == pe_devs[v] && B->cap_dst[v] == dst_devs[v];
    if (!same) {
        if (B->exec) {
            (void)hipGraphExecDestroy(B->exec);
            B->exec = nullptr;
        }
        if (B->graph) {
            (void)hipGraphDestroy(B->graph);
            B->graph = nullptr;
        }
        HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        rc = batch_enqueue(*B, pe_devs, dst_devs, s);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(s, &g);
        if (rc < 0) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return set_error(SM_E_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        B->graph = g;
        HIP_TRY(hipGraphInstantiate(&B->exec, B->graph, nullptr, nullptr, 0));
        B->cap_pe.assign(pe_devs, pe_devs + nv);
        B->cap_dst.assign(dst_devs, dst_devs + nv);
        B->cap_stream = s;
    }
    HIP_TRY(hipGraphLaunch(B->exec, s));
    HIP_TRY(hipStreamSynchronize(s));
    return SM_OK;
}

int sm_scatter_batch_graph_nodes(const SmScatterBatch* B) {
    if (!B || !B->graph) return 0;
    size_t n = 0;
    if (hipGraphGetNodes(B->graph, nullptr, &n) != hipSuccess) return set_error(SM_E_DEVICE, "hipGraphGetNodes");
    return (int)n;
}

int sm_scatter_batch_destroy(SmScatterBatch* B) {
    if (B) {
        use_device(B->device);
        delete B;
    }
    return SM_OK;
}

}  // extern "C"
