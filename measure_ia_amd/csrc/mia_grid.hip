// Grid autotuner: columns per side, slabs per side and slabs per shape block.
//
// The count's time (pair kernel + planning) is, to within a few percent over
// the measured range,
//     t = t0 + a * candidates + b * items
// (t0: the in-domain pairs, fixed by the data; candidates = (shape, position)
// pairs the work items evaluate; items = packed work items).  For a grid the
// last two are estimated on the host for uniform densities -- the expected
// number of positions in the cells each shape block reaches, and the items the
// planner packs from each group of stencil columns -- and the grid with the
// lowest a*cand + b*items is taken.  The candidate estimate is ~10 % high on
// clustered samples (the same for every grid); the item estimate is scaled by
// its measured ratio (0.7).  a and b were fitted on MI355X sweeps of a
// 1M-position w_gg/w_g+ count and a 2M-position multipole count: a = 0.58 ps
// (r_p, pi) and 0.85 ps (r, mu) per candidate, b = 1.0 ns (r_p, pi) and
// 1.3 ns (r, mu) per item.  Choices are cached per (box, binning reach,
// jackknife, sizes).
#include "mia_common.hpp"

#include <algorithm>
#include <cmath>
#include <mutex>
#include <tuple>

namespace mia {

namespace {

constexpr double kCandCostW = 0.58e-12;  // s per candidate, (r_p, pi)
constexpr double kCandCostMu = 0.85e-12; // s per candidate, (r, mu)
constexpr double kItemCostW = 1.0e-9;    // s per work item, (r_p, pi)
constexpr double kItemCostMu = 1.3e-9;   // s per work item, (r, mu)
constexpr double kItemScale = 0.7;       // measured items / estimated items
constexpr int kBlockShapes = 64;
constexpr int kItemPositions = 1024;
// Dense columns split their position runs into many 1024-position items;
// 2048 halves those splits.  Measured: large20M (runs of thousands of
// positions) 1.5 % faster at 2048, the 1M/2M boxes and the batch within
// 0.5 % either way.  Taken when it cuts the estimated items by a tenth.
constexpr int kItemPositionsDense = 2048;

struct Choice {
    int ncj, nzj, G, pmax;
};

// Expected candidates, work items and position ranges of one grid (uniform
// densities).  Per shape block: every reached stencil column holds
// `pos` positions in (1 + win/jspan + win/L) line-of-sight runs (jackknife
// slab and periodic-wrap splits); the planner packs the runs of each group of
// kGroupColumns columns into items of at most pmax positions and kItemRanges
// runs, one item per (region, shift) class at least.
void estimate(double L, int jk, int ncj, int nzj, int G, double R, double Z, bool rmu, double Ns,
              double Np, double &cand, double &items, double *ranges = nullptr, int pmax = kItemPositions) {
    const int nc = ncj * jk, nz = nzj * jk;
    const double cw = L / nc, hz = L / nz;
    int kx = (int)std::ceil(R / cw) + 1;
    int lo = -kx, hi = kx;
    if (2 * kx + 1 >= nc) {  // the stencil covers the plane: each column once
        lo = -(nc / 2);
        hi = lo + nc - 1;
    }
    const int nxw = hi - lo + 1;
    const double cell_p = Np / ((double)nc * nc * nz);
    const double jspan = L / jk;
    double zsum = 0.0, per_block = 0.0, rec = 0.0;
    double g_pos = 0.0, g_runs = 0.0, g_cls = 0.0;
    int g_at = -1;
    auto close_group = [&]() {
        if (g_cls > 0.0) {
            const double n = std::max({g_cls, std::ceil(g_pos / pmax) + g_cls - 1.0, std::ceil(g_runs / kItemRanges)});
            per_block += n;
            rec += g_runs;
        }
        g_pos = g_runs = g_cls = 0.0;
    };
    for (int i = lo; i <= hi; ++i) {
        int ai = std::abs(i);
        if (2 * kx + 1 >= nc) ai = std::min(ai, nc - ai);
        double u = std::max(ai - 1, 0) * cw;
        for (int j = lo; j <= hi; ++j) {
            const int grp = ((i - lo) * nxw + (j - lo)) / kGroupColumns;
            if (grp != g_at) {
                close_group();
                g_at = grp;
            }
            int aj = std::abs(j);
            if (2 * kx + 1 >= nc) aj = std::min(aj, nc - aj);
            double v = std::max(aj - 1, 0) * cw;
            double d2 = u * u + v * v;
            if (d2 >= R * R) continue;
            double reach = rmu ? std::sqrt(R * R - d2) : Z;
            int nsl = std::min(G + 2 * (int)std::ceil(reach / hz), nz);
            zsum += nsl;
            double win = nsl * hz;
            double pos = cell_p * nsl;
            const double cls = 1.0 + win / jspan + win / L;
            g_pos += pos;
            g_runs += cls * std::max(1.0, std::ceil(pos / pmax));
            g_cls = std::max(g_cls, cls);
        }
    }
    close_group();
    cand = Ns * cell_p * zsum;
    const double groups = (double)nc * nc * (nz / G);
    const double lam = Ns / groups;
    const double nblocks = groups * (1.0 - std::exp(-lam)) * std::max(1.0, lam / kBlockShapes);
    items = kItemScale * nblocks * per_block;
    if (ranges) *ranges = nblocks * rec;
}

std::mutex g_mu;
std::map<std::tuple<double, int, double, double, int, int64_t, int64_t>, Choice> g_cache;

Choice search(double L, int jk, double R, double Z, bool rmu, int64_t ns, int64_t np) {
    auto key = std::make_tuple(L, jk, R, Z, (int)rmu, ns, np);
    {
        std::lock_guard<std::mutex> lk(g_mu);
        auto it = g_cache.find(key);
        if (it != g_cache.end()) return it->second;
    }
    const double Ns = (double)std::max<int64_t>(ns, 1), Np = (double)std::max<int64_t>(np, 1);
    const double span = L / jk;
    const double zr = rmu ? R : Z;
    const double a = rmu ? kCandCostMu : kCandCostW;
    const double b = rmu ? kItemCostMu : kItemCostW;
    // column widths in [R/12, R/2], slab heights in [zr/12, max(zr, R)/2], at most 512 per side
    int c0 = std::max(1, (int)std::floor(span / (R / 2.0)));
    int c1 = std::max(c0, std::min(512 / jk, (int)std::ceil(span / (R / 12.0))));
    int z0 = std::max(1, (int)std::floor(span / (std::max(zr, R) / 2.0)));
    int z1 = std::max(z0, std::min(512 / jk, (int)std::ceil(span / (zr / 12.0))));
    Choice best{c0, z0, 1, 0};
    double best_t = -1.0;
    for (int ncj = c0; ncj <= c1; ++ncj)
        for (int nzj = z0; nzj <= z1; ++nzj)
            for (int G = 1; G <= 4 && G <= nzj; ++G) {
                if (nzj % G) continue;
                double cand, items;
                estimate(L, jk, ncj, nzj, G, R, Z, rmu, Ns, Np, cand, items);
                double t = a * cand + b * items;
                if (best_t < 0 || t < best_t) {
                    best_t = t;
                    best = Choice{ncj, nzj, G, 0};
                }
            }
    {
        double cand, items, items2;
        estimate(L, jk, best.ncj, best.nzj, best.G, R, Z, rmu, Ns, Np, cand, items);
        estimate(L, jk, best.ncj, best.nzj, best.G, R, Z, rmu, Ns, Np, cand, items2, nullptr, kItemPositionsDense);
        if (items2 < 0.9 * items) best.pmax = kItemPositionsDense;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    g_cache[key] = best;
    return best;
}

}  // namespace

double estimate_items(const Grid &g, int G, double reach_r, double reach_z, bool rmu, int64_t ns, int64_t np,
                      int pmax, double *ranges) {
    const int jk = g.jk > 0 ? g.jk : 1;
    double cand = 0, items = 0, rec = 0;
    estimate(g.L, jk, g.ncj, g.nzj, std::max(1, G), reach_r, reach_z, rmu, (double)std::max<int64_t>(ns, 1),
             (double)std::max<int64_t>(np, 1), cand, items, &rec, pmax);
    const bool any = ns > 0 && np > 0;
    if (ranges) *ranges = any ? rec : 0.0;
    return any ? items : 0.0;
}

Grid make_grid(const mia_sample &s, const mia_bins &bins, const mia_tuning *t) {
    Grid g;
    g.L = s.boxsize;
    g.half = s.boxsize / 2.0;
    g.los = s.los;
    int k = 0;
    int na[2];
    for (int c = 0; c < 3; ++c)
        if (c != s.los) na[k++] = c;
    g.ax_a = na[0];
    g.ax_b = na[1];
    g.jk = s.jk_per_side > 0 ? s.jk_per_side : 0;
    g.jk_sub = g.jk > 0 ? s.jk_sub : s.boxsize;
    const bool rmu = bins.mode == MIA_RMU;
    double reach_r = std::sqrt(bins.r2_edges[bins.n_r]);
    double reach_z = rmu ? reach_r
                         : std::fmax(std::fabs(bins.x2_edges[0]), std::fabs(bins.x2_edges[bins.n_2]));
    int jk_eff = g.jk > 0 ? g.jk : 1;
    double span = g.L / jk_eff;
    int ncj, nzj, G = 0;
    if (t && t->cell_shapes > 0) {
        // density rule: cubes holding ~cell_shapes shapes, within [reach/10, reach/2]
        double nobj = (double)(s.n_shape > 0 ? s.n_shape : std::max<int64_t>(s.n_pos, 1));
        double a = std::cbrt((double)t->cell_shapes * g.L * g.L * g.L / nobj);
        double cw_t = std::min(std::max(a, reach_r / 10.0), reach_r / 2.0);
        double hz_t = std::min(std::max(a, reach_z / 10.0), std::max(reach_z / 2.0, reach_r / 2.0));
        ncj = std::max(1, (int)std::lround(span / cw_t));
        nzj = std::max(1, (int)std::lround(span / hz_t));
    } else {
        Choice c = search(g.L, jk_eff, reach_r, reach_z, rmu, s.n_shape, s.n_pos);
        ncj = c.ncj;
        nzj = c.nzj;
        G = c.G;
        g.pmax = c.pmax;
    }
    if (t && t->columns_per_side > 0) ncj = std::max(1, t->columns_per_side / jk_eff);
    if (t && t->slabs_per_side > 0) nzj = std::max(1, t->slabs_per_side / jk_eff);
    while (ncj * jk_eff > 512 && ncj > 1) --ncj;
    while (nzj * jk_eff > 512 && nzj > 1) --nzj;
    g.ncj = ncj;
    g.nzj = nzj;
    g.nc = ncj * jk_eff;
    g.nz = nzj * jk_eff;
    g.cw = g.L / g.nc;
    g.hz = g.L / g.nz;
    g.margin = 1e-9 * g.L + 1e-12;
    g.G = (G > 0 && nzj % G == 0) ? G : 0;
    return g;
}

}  // namespace mia
