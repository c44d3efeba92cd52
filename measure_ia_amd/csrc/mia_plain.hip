// Plain FP64 pair counter: one thread per shape, global-atomic histograms.
//
// The straightforward GPU form of the reference's per-shape tree loop
// (src/measureia/measure_w_box_jk.py:398-461 and
// src/measureia/measure_m_box_jk.py:415-487): each thread walks the
// position cells within reach of its shape, evaluates every pair in FP64 in
// the reference's op order and adds DD, S+D, SxD and the jackknife tallies
// straight into global memory with atomics.  No LDS, no batching.
#include "mia_common.hpp"

#include <cmath>

namespace mia {

namespace {

struct PlainArgs {
    Grid g;
    int mode, n_r, n_2, periodic;
    int weighted, var, K, nbins;
    double reach_r, reach_z;
    const double *T;  // n_r+1 squared-separation thresholds
    const double *X;  // pi or mu thresholds
    int kx;           // column half-width of the search window (in columns)
    uint64_t *npairs;
    double *dd, *splus, *scross, *variance;
};

__device__ __forceinline__ double r2_column_order(double a2, double b2, double l2, int los) {
    // np.sum(separation**2, axis=1) adds columns 0, 1, 2 in order.
    if (los == 2) return (a2 + b2) + l2;
    if (los == 1) return (a2 + l2) + b2;
    return (l2 + a2) + b2;
}

// min |minimum-image separation| between point x and the interval [lo, hi].
__device__ __forceinline__ double dist_point_interval(double x, double lo, double hi, double L,
                                                      bool periodic) {
    double d;
    if (x < lo)
        d = lo - x;
    else if (x > hi)
        d = x - hi;
    else
        return 0.0;
    if (periodic) {
        // images of the interval at +-L
        double d2 = (x < lo) ? (x + L) - hi : lo - (x - L);
        if (d2 < 0) d2 = 0;
        d = d < d2 ? d : d2;
    }
    return d;
}

__device__ __forceinline__ void pair_plain(const PlainArgs &A, double sa, double sb, double sl,
                                           double sax, double say, double se, double sw, int Rs,
                                           double pa, double pb, double pl, double pw, int Rp) {
    const Grid &g = A.g;
    double da = sa - pa, db = sb - pb, dl = sl - pl;
    if (A.periodic) {
        da = wrap_sep(da, g.L, g.half);
        db = wrap_sep(db, g.L, g.half);
        dl = wrap_sep(dl, g.L, g.half);
    }
    double a2 = da * da, b2 = db * db;
    double s = a2 + b2;
    double rv;  // value binned along r
    int i2;
    if (A.mode == MIA_RPPI) {
        if (!(s >= A.T[0] && s < A.T[A.n_r] && dl >= A.X[0] && dl < A.X[A.n_2])) return;
        rv = s;
        i2 = 0;
        for (int k = 1; k < A.n_2; ++k) i2 += (A.X[k] <= dl);
    } else {
        double r2 = r2_column_order(a2, b2, dl * dl, g.los);
        if (!(s > 0.0 && r2 >= A.T[0] && r2 < A.T[A.n_r])) return;
        rv = r2;
        double mu = dl / sqrt(r2);
        if (mu < A.X[0]) {
            i2 = A.n_2 - 1;  // numpy's index -1
        } else {
            i2 = 0;
            for (int k = 1; k < A.n_2; ++k) i2 += (A.X[k] <= mu);
        }
    }
    int ir = 0;
    for (int k = 1; k < A.n_r; ++k) ir += (A.T[k] <= rv);
    // e+ = e cos(2 arccos d), ex = e sin(2 arccos d), d = (a/r_p) ax + (b/r_p) ay,
    // zero when |d| > 1 (arccos NaN) -- measure_w_box_jk.py:166-175.
    double rp = sqrt(s);
    double d = 0.0 + (da / rp) * sax;
    d = d + (db / rp) * say;
    double ep = 0.0, ex = 0.0;
    if (fabs(d) <= 1.0) {
        ep = se * (2.0 * d * d - 1.0);
        ex = se * (2.0 * d * sqrt(1.0 - d * d));
    }
    double wp = pw * sw;
    double vp = wp * ep;
    int b = ir * A.n_2 + i2;
    atomicAdd((unsigned long long *)&A.npairs[b], 1ull);
    if (A.weighted) atomicAdd(&A.dd[b], wp);
    atomicAdd(&A.splus[b], vp);
    atomicAdd(&A.scross[b], wp * ex);
    if (A.var) atomicAdd(&A.variance[b], vp * vp);
    if (A.K > 0) {
        int o = (1 + Rs) * A.nbins + b;
        atomicAdd((unsigned long long *)&A.npairs[o], 1ull);
        if (A.weighted) atomicAdd(&A.dd[o], wp);
        atomicAdd(&A.splus[o], vp);
        if (Rp != Rs) {
            o = (1 + Rp) * A.nbins + b;
            atomicAdd((unsigned long long *)&A.npairs[o], 1ull);
            if (A.weighted) atomicAdd(&A.dd[o], wp);
            atomicAdd(&A.splus[o], vp);
        }
    }
}

__global__ __launch_bounds__(256) void k_plain(PlainArgs A, Sorted pos, Sorted shp) {
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= shp.n) return;
    const Grid &g = A.g;
    double sa = shp.a[s], sb = shp.b[s], sl = shp.l[s];
    double sax = shp.ax[s], say = shp.ay[s], se = shp.e[s], sw = shp.w[s];
    int Rs = shp.reg[s];
    int key = shp.key[s];
    int col = key / (g.nz + 1);
    int ix = col / g.nc, iy = col % g.nc;
    bool full_x = 2 * A.kx + 1 >= g.nc;
    int nx = full_x ? g.nc : 2 * A.kx + 1;
    double m = g.margin;
    double rr = A.reach_r + m;
    // slab window from the shape's own LOS coordinate
    int klo = (int)floor((sl - A.reach_z - m) / g.hz) - 1;
    int khi = (int)floor((sl + A.reach_z + m) / g.hz) + 1;
    bool full_z = (khi - klo + 1) >= g.nz || !A.periodic;
    for (int ox = 0; ox < nx; ++ox) {
        int jx = full_x ? ox : ix - A.kx + ox;
        if (!full_x) {
            if (A.periodic)
                jx = ((jx % g.nc) + g.nc) % g.nc;
            else if (jx < 0 || jx >= g.nc)
                continue;
        }
        double dxm = dist_point_interval(sa, jx * g.cw - m, (jx + 1) * g.cw + m, g.L, A.periodic);
        if (dxm >= rr) continue;
        for (int oy = 0; oy < nx; ++oy) {
            int jy = full_x ? oy : iy - A.kx + oy;
            if (!full_x) {
                if (A.periodic)
                    jy = ((jy % g.nc) + g.nc) % g.nc;
                else if (jy < 0 || jy >= g.nc)
                    continue;
            }
            double dym = dist_point_interval(sb, jy * g.cw - m, (jy + 1) * g.cw + m, g.L,
                                             A.periodic);
            if (dxm * dxm + dym * dym >= rr * rr) continue;
            int64_t base = (int64_t)(jx * g.nc + jy) * (g.nz + 1);
            // nominal part: one or two slab ranges; then the tail
            int r0[3], r1[3], nr = 0;
            if (full_z) {
                r0[nr] = pos.start[base];
                r1[nr++] = pos.start[base + g.nz];
            } else {
                int k0 = klo, k1 = khi;
                if (k0 < 0) {
                    r0[nr] = pos.start[base + (k0 + g.nz)];
                    r1[nr++] = pos.start[base + g.nz];
                    k0 = 0;
                }
                if (k1 >= g.nz) {
                    r0[nr] = pos.start[base];
                    r1[nr++] = pos.start[base + (k1 - g.nz) + 1];
                    k1 = g.nz - 1;
                }
                r0[nr] = pos.start[base + k0];
                r1[nr++] = pos.start[base + k1 + 1];
            }
            for (int q = 0; q < nr; ++q)
                for (int p = r0[q]; p < r1[q]; ++p)
                    pair_plain(A, sa, sb, sl, sax, say, se, sw, Rs, pos.a[p], pos.b[p], pos.l[p],
                               pos.w[p], pos.reg[p]);
            for (int p = pos.start[base + g.nz]; p < pos.start[base + g.nz + 1]; ++p)
                pair_plain(A, sa, sb, sl, sax, say, se, sw, Rs, pos.a[p], pos.b[p], pos.l[p],
                           pos.w[p], pos.reg[p]);
        }
    }
}

}  // namespace

void count_plain(Ctx &ctx, const mia_sample &s, const mia_bins &bins, const Grid &g,
                 const Sorted &pos, const Sorted &shp, const mia_counts &out, bool weighted,
                 hipStream_t st, mia_stats *stats) {
    PlainArgs A{};
    A.g = g;
    A.mode = bins.mode;
    A.n_r = bins.n_r;
    A.n_2 = bins.n_2;
    A.periodic = s.periodic;
    A.weighted = weighted ? 1 : 0;
    A.var = (bins.accumulate_variance && out.variance) ? 1 : 0;
    A.K = g.jk > 0 ? g.jk * g.jk * g.jk : 0;
    A.nbins = bins.n_r * bins.n_2;
    A.reach_r = std::sqrt(bins.r2_edges[bins.n_r]);
    A.reach_z = bins.mode == MIA_RPPI
                    ? std::fmax(std::fabs(bins.x2_edges[0]), std::fabs(bins.x2_edges[bins.n_2]))
                    : A.reach_r;
    int nx2 = bins.mode == MIA_RPPI ? bins.n_2 + 1 : bins.n_2;
    double *T = ctx.arena.get<double>("plain/T", bins.n_r + 1);
    double *X = ctx.arena.get<double>("plain/X", nx2);
    MIA_HIP(hipMemcpyAsync(T, bins.r2_edges, sizeof(double) * (bins.n_r + 1), hipMemcpyHostToDevice, st));
    MIA_HIP(hipMemcpyAsync(X, bins.x2_edges, sizeof(double) * nx2, hipMemcpyHostToDevice, st));
    A.T = T;
    A.X = X;
    A.kx = (int)std::ceil((A.reach_r + g.margin) / g.cw) + 1;
    A.npairs = out.npairs;
    A.dd = out.dd;
    A.splus = out.splus;
    A.scross = out.scross;
    A.variance = out.variance;
    int64_t nb = (int64_t)(1 + A.K) * A.nbins;
    fill<uint64_t>(out.npairs, 0, nb, st);
    if (out.dd) fill<double>(out.dd, 0.0, nb, st);
    fill<double>(out.splus, 0.0, nb, st);
    fill<double>(out.scross, 0.0, A.nbins, st);
    if (out.variance) fill<double>(out.variance, 0.0, A.nbins, st);
    Events ev(2);
    MIA_HIP(hipEventRecord(ev[0], st));
    if (shp.n > 0 && pos.n > 0) {
        unsigned blocks = (unsigned)((shp.n + 255) / 256);
        hipLaunchKernelGGL(k_plain, dim3(blocks), dim3(256), 0, st, A, pos, shp);
        MIA_HIP(hipGetLastError());
    }
    MIA_HIP(hipEventRecord(ev[1], st));
    MIA_HIP(hipEventSynchronize(ev[1]));
    if (stats) stats->ms_pairs = ev.ms(0, 1);
}

}  // namespace mia
