// Fast pair counter: cell-pair work items, per-wavefront LDS histograms.
//
// Replaces the per-shape KDTree loop and np.add.at scatter of the reference
// (src/measureia/measure_w_box_jk.py:386-461, measure_m_box_jk.py:403-487)
// with three steps on the device:
//   1. plan   : shapes of one column and a few line-of-sight slabs form a
//               block (<= 64 shapes, one jackknife region R_s); every block
//               is paired with the position cell ranges within reach, split
//               at jackknife sub-box faces so that every range has one
//               position region R_p and one periodic shift per axis.  The
//               ranges of up to 8 neighbouring stencil columns with the same
//               region and shift are packed into one work item.  Items are
//               counting-sorted by (sample, R_s, R_p).
//   2. pairs  : a persistent grid of wavefronts takes contiguous runs of the
//               sorted list, balanced by estimated cost, then claims tail
//               chunks from a counter.  Each wave reads its
//               shape block through the scalar cache (one 64-byte record per
//               shape, wave-uniform) and streams positions through registers,
//               64 at a time.  Separations are the reference's own FP64
//               differences, so r_p^2 (or r^2) and pi are compared exactly
//               with the host-bisected thresholds (an FP32 log2 only guesses
//               the bin); e+ and ex are FP64 (CDNA4 runs FP64 FMA at the rate
//               of non-packed FP32) and go into the wave's private LDS
//               histogram with LDS atomics.  mu = pi/r is screened in FP32 and recomputed in
//               FP64 when it lies within a guard band of a bin edge; the
//               arccos NaN rule is likewise checked in FP64 for nearly
//               parallel pairs.
//   3. flush  : a wave adds its histogram to the workspace's FP64
//               accumulators of the total and of tally[R_s] / tally[R_p]
//               only when the (R_s, R_p) key changes, i.e. a handful of times
//               per wave; k_finalize then writes the caller's outputs.  In
//               the deterministic mode the flushes go to per-wave slots that
//               k_finalize_det sums in a fixed order instead.
// Pair counts are integers (exact, order independent); the e+ sums are FP64
// throughout, so every bin carries only the rounding of FP64 additions.
#include "mia_common.hpp"

#include <hipcub/hipcub.hpp>

#include <cmath>
#include <memory>
#include <mutex>

namespace mia {

namespace {

constexpr int kMaxWavesPerBlock = 4;
constexpr int kThreads = kWave * kMaxWavesPerBlock;
constexpr size_t kLdsBytes = 160 * 1024;         // LDS per CU (gfx950)
constexpr int64_t kMaxKeysPerSample = 1 << 24;   // K^2 planner keys per sample (K regions)
constexpr double kSinExact = 1e-5;  // FP64 arccos-domain check when |sin phi| < 1e-5 |cos phi|

struct Block {
    int32_t s0, ns;
    int32_t col, grp;   // column index, slab group (== ngroups: tail)
    int32_t sample, rs;
};

// A work item pairs one shape block with up to kItemRanges position ranges of
// one key and one shift class (the ranges of neighbouring stencil columns are
// packed together, so short columns still fill the 64 lanes).  Item
// coordinates k = 0 .. np-1 run over the ranges in order.
struct Item {
    int32_t blk;
    int32_t p0;         // nr == 1: first position; nr > 1: first of nr Range records
    int32_t np;
    int16_t rp;
    uint8_t shifts;     // 2 bits per axis (a, b, l): 0 -> 0, 1 -> +L, 2 -> -L ; bit 6: per-pair wrap
    uint8_t nr;         // position ranges
};
static_assert(sizeof(Item) == 16, "Item layout");
// One range of a multi-range item: item coordinates k in [end of the previous
// range, end) are positions base + k.
struct Range {
    int32_t base, end;
};
static_assert(sizeof(Range) == 8, "Range layout");
static_assert(sizeof(Block) == 24, "Block layout");

struct SampleDev {
    // sorted sample
    const double *prec;          // 4 doubles per position {a, b, l, w}, sorted order
    const double *srec;          // 8 doubles per shape, sorted order
    const int32_t *pstart;
    double L, half;
    int los, K, nbins, pad;
    // accumulators (workspace)
    unsigned long long *npairs;   // (1+K)*B
    double *dd, *sp;              // (1+K)*B (dd: weighted samples only)
    double *sx, *var;             // B
};

// Planner view of one sample (per-call parameter block).  Its (column, slab
// group) cells occupy [cell0, cell0 + ncells) of the global cell arrays, its
// range threads (one per block and stencil column) [t0r, t0r + ntr) and its
// pack threads (one per block and group of stencil columns) [t0, t0 + nt),
// both starting at multiples of kPlanThreads, so a workgroup never straddles
// two samples.
struct PlanSample {
    Grid g;
    int mode, periodic;
    int ng, G;           // slab groups per column, slabs per group
    int kx, nxw;         // stencil half width / width in columns
    int pmax;            // positions per work item at most
    int ngc;             // plan threads per block: groups of kGroupColumns stencil columns
    double reach_r, reach_z;
    const int32_t *sstart, *pstart;
    const int32_t *pfine;  // starts of the positions' fine keys (cell << sbits) + line-of-sight part
    int sbits, pad2;
    double x_lo, x_hi;     // (r_p, pi): the pi domain, for cutting slab runs to a block's window
    const double *srec;  // shape records (block bounding boxes)
    int64_t cell0, ncells;
    int64_t t0, nt;      // pack threads: worst-case blocks x ngc
    int64_t t0r, ntr;    // range threads: worst-case blocks x nxw^2
};

// Finalisation view of one sample: accumulators to caller outputs.
struct FinSample {
    const unsigned long long *np;
    const double *sp, *dd, *sx, *var;
    const unsigned long long *rcnt;   // [2K] objects per region (positions, shapes)
    const double *rsum;               // [2K] sum w_s, sum w_s (1 - e^2/2) per region
    mia_counts out;
    int K, B, weighted, var_on;
};

// Largest s in [0, n) with tab[s] <= v (tab ascending, tab[0] <= v).
__device__ __forceinline__ int find_sample(const int64_t *tab, int n, int64_t v) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tab[mid] <= v)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ int region_of(const Grid &g, int ia, int ib, int il) {
    if (g.jk <= 0) return 0;
    int j[3];
    j[g.ax_a] = ia / g.ncj;
    j[g.ax_b] = ib / g.ncj;
    j[g.los] = il / g.nzj;
    return j[0] * g.jk * g.jk + j[1] * g.jk + j[2];
}

// Separation range [lo, hi] (unwrapped) -> minimum |minimum-image value| and
// the shift class: 0 none, 1 +L for all, 2 -L for all, 3 mixed.
__device__ __forceinline__ void axis_class(double lo, double hi, double L, double half,
                                           bool periodic, double &dmin, int &cls) {
    if (!periodic) {
        dmin = lo > 0 ? lo : (hi < 0 ? -hi : 0.0);
        cls = 0;
        return;
    }
    dmin = 1e300;
    // piece below -half: value + L
    if (lo < -half) {
        double a = lo + L, b = (hi < -half ? hi : -half) + L;
        double d = a > 0 ? a : (b < 0 ? -b : 0.0);
        dmin = d < dmin ? d : dmin;
    }
    if (hi >= -half && lo <= half) {
        double a = lo > -half ? lo : -half, b = hi < half ? hi : half;
        double d = a > 0 ? a : (b < 0 ? -b : 0.0);
        dmin = d < dmin ? d : dmin;
    }
    if (hi > half) {
        double a = (lo > half ? lo : half) - L, b = hi - L;
        double d = a > 0 ? a : (b < 0 ? -b : 0.0);
        dmin = d < dmin ? d : dmin;
    }
    if (lo > half)
        cls = 2;
    else if (hi < -half)
        cls = 1;
    else if (lo >= -half && hi <= half)
        cls = 0;
    else
        cls = 3;
}

__device__ __forceinline__ double shift_val(int cls, double L) {
    return cls == 1 ? L : (cls == 2 ? -L : 0.0);
}

// The slabs within line-of-sight reach of a block (its shapes' extent
// [zlo, zhi]) for a column at squared transverse distance dxy2, merged into
// runs [k0, k) of equal periodic shift class and jackknife slab:
// close(k0, k, class) for each run in slab order.  For (r_p, pi) the runs do
// not depend on the column; for (r, mu) a column's runs are those of
// dxy2 = 0 with slabs out of its reach taken off their ends.
// Whether slab k is within reach of a block's shapes (extent [zlo, zhi]) for
// a column at squared transverse distance dxy2, and its shift class cz.
__device__ __forceinline__ bool slab_in(const PlanSample &P, double zlo, double zhi, int k, double dxy2, double rr,
                                        int &cz) {
    const Grid &g = P.g;
    const double m = g.margin;
    double dzmin;
    axis_class(zlo - (k + 1) * g.hz - 2 * m, zhi - k * g.hz + 2 * m, g.L, g.half, P.periodic, dzmin, cz);
    bool inc = dzmin < P.reach_z + 4 * m;
    if (inc && P.mode == MIA_RMU) inc = dxy2 + dzmin * dzmin < rr * rr;
    return inc;
}

template <class F>
__device__ __forceinline__ void slab_runs(const PlanSample &P, double zlo, double zhi, double dxy2, double rr,
                                          F &&close) {
    const Grid &g = P.g;
    int run_k0 = -1, run_cls = 0, run_jz = 0;
    for (int k = 0; k <= g.nz; ++k) {
        bool inc = false;
        int cz = 0;
        if (k < g.nz) inc = slab_in(P, zlo, zhi, k, dxy2, rr, cz);
        const int jz = k / g.nzj;
        if (run_k0 >= 0 && (!inc || cz != run_cls || jz != run_jz)) {
            close(run_k0, k, run_cls);
            run_k0 = -1;
        }
        if (inc && run_k0 < 0) {
            run_k0 = k;
            run_cls = cz;
            run_jz = jz;
        }
    }
}

// Enumerate the position ranges of one (block, stencil column) pair: slab
// runs within reach of the block, merged while the periodic shift class and
// the jackknife slab stay the same, plus the column's anomalous tail.  For
// each range calls emit(local key, p_begin, p_end, rp, shift byte) with
// local key = R_s * K1 + R_p.  Shift byte: 2 bits per axis (a, b, l) for the
// constant shift class, bit 6 the per-pair wrap.  Reach is tested from the
// bounding box bb of the block's shapes, not from its cells, so a column or
// slab that no shape of the block can reach is left out.  runs: the block's
// slab runs when they were stored (count >= 0), else they are walked here.
template <class F>
__device__ __forceinline__ void plan_ranges(const PlanSample &P, const Block &B, const double *bb, int so, int K1,
                                            const int32_t *runs, F &&emit) {
    const Grid &g = P.g;
    int ox = so / P.nxw, oy = so % P.nxw;
    int ia = B.col / g.nc, ib = B.col % g.nc;
    bool full = P.nxw >= g.nc;
    int jx = full ? ox : ia - P.kx + ox;
    int jy = full ? oy : ib - P.kx + oy;
    if (!full) {
        if (P.periodic) {
            jx = ((jx % g.nc) + g.nc) % g.nc;
            jy = ((jy % g.nc) + g.nc) % g.nc;
        } else if (jx < 0 || jx >= g.nc || jy < 0 || jy >= g.nc) {
            return;
        }
    }
    double m = g.margin;
    double da_min, db_min;
    int ca, cb;
    // separations between the block's shapes (their bounding box bb) and the
    // target column's positions
    axis_class(bb[0] - (jx + 1) * g.cw - 2 * m, bb[1] - jx * g.cw + 2 * m, g.L, g.half, P.periodic, da_min, ca);
    axis_class(bb[2] - (jy + 1) * g.cw - 2 * m, bb[3] - jy * g.cw + 2 * m, g.L, g.half, P.periodic, db_min, cb);
    double rr = P.reach_r + 4 * m;
    double dxy2 = da_min * da_min + db_min * db_min;
    if (dxy2 >= rr * rr) return;
    // line-of-sight extent of the block's shapes (a tail block's shapes may lie anywhere)
    const double zlo = bb[4], zhi = bb[5];
    int col = jx * g.nc + jy;
    int64_t base = (int64_t)col * (g.nz + 1);
    const int kr = B.rs * K1;
    auto range = [&](int p_begin, int p_end, int rp, int cz) {
        if (p_end <= p_begin) return;
        // mixed shift classes take the per-pair minimum-image wrap; a
        // non-periodic box never wraps (classes are always 0 there, and the
        // anomalous tail's "any LOS value" class means no shift)
        bool wrap = P.periodic && ((ca == 3) || (cb == 3) || (cz == 3));
        uint8_t sh = (uint8_t)((ca == 3 ? 0 : ca) | ((cb == 3 ? 0 : cb) << 2) |
                               ((cz == 3 ? 0 : cz) << 4) | (wrap ? 64 : 0));
        emit(kr + rp, p_begin, p_end, rp, sh);
    };
    // the positions of a run of slabs [k0, k) of one shift class
    auto run = [&](int k0, int k, int cls) {
        int pb = P.pstart[base + k0], pe = P.pstart[base + k];
        if (cls != 3 && P.sbits > 0) {
            // the run's positions are in line-of-sight order (to 2^-sbits
            // of a slab): keep those a shape of the block can pair with,
            // l in (zlo + sh - x_hi, zhi + sh - x_lo] for the run's shift
            // sh, one fine bucket of slack on each side
            const double sh = shift_val(cls, g.L);
            double wlo, whi;
            if (P.mode == MIA_RPPI) {
                wlo = zlo + sh - P.x_hi - 4 * m;
                whi = zhi + sh - P.x_lo + 4 * m;
            } else {
                const double rz2 = sqrt(fmax(rr * rr - dxy2, 0.0));
                wlo = zlo + sh - rz2 - 4 * m;
                whi = zhi + sh + rz2 + 4 * m;
            }
            const double hq = g.hz / (double)(1 << P.sbits);
            const int64_t f0 = (int64_t)k0 << P.sbits, f1 = ((int64_t)k << P.sbits) - 1;
            const double flo = floor(wlo / hq) - 1.0, fhi = floor(whi / hq) + 1.0;
            const int64_t fa = flo < (double)f0 ? f0 : (flo > (double)f1 ? f1 + 1 : (int64_t)flo);
            const int64_t fb = fhi > (double)f1 ? f1 : (fhi < (double)f0 ? f0 - 1 : (int64_t)fhi);
            const int64_t fbase = base << P.sbits;
            if (fb < fa) {
                pe = pb;
            } else {
                pb = P.pfine[fbase + fa];
                pe = P.pfine[fbase + fb + 1];
            }
        }
        range(pb, pe, region_of(g, jx, jy, k0), cls);
    };
    if (runs && runs[0] >= 0) {
        // the block's slab runs: (r_p, pi) the same for every column; (r, mu)
        // those of the block's own column, cut at both ends to this column's
        // reach (the slabs a column reaches within one run are contiguous,
        // the distance growing away from the block on either side)
        for (int r = 0; r < runs[0]; ++r) {
            const int v = runs[1 + r];
            int k0 = v & 0x7ff, k1 = (v >> 11) & 0x7ff, cz;
            if (P.mode == MIA_RMU) {
                while (k0 < k1 && !slab_in(P, zlo, zhi, k0, dxy2, rr, cz)) ++k0;
                while (k1 > k0 && !slab_in(P, zlo, zhi, k1 - 1, dxy2, rr, cz)) --k1;
                if (k0 >= k1) continue;
            }
            run(k0, k1, (v >> 22) & 3);
        }
    } else {
        slab_runs(P, zlo, zhi, dxy2, rr, run);
    }
    // anomalous tail of the target column: region 0, any line-of-sight value
    range(P.pstart[base + g.nz], P.pstart[base + g.nz + 1], 0, 3);
}

constexpr int kPlanThreads = 256;
constexpr int kBlockRuns = 15;          // slab runs of a block kept for its columns
constexpr int kPlanGridPerCu = 16;      // workgroups per CU of the planner's persistent grid
constexpr int64_t kItemOverhead = 256;  // per-item cost, in pair evaluations
constexpr int kOpenItems = 4;           // items a plan thread fills at once (one per key and shift)

// One position range of a (block, stencil column) pair as plan_ranges
// enumerates it.
struct RawRange {
    int32_t pb, pe;   // positions [pb, pe)
    int32_t key;      // local key R_s * K1 + R_p
    int32_t w;        // R_p | shift byte << 16
};
static_assert(sizeof(RawRange) == 16, "RawRange layout");

struct RangeLaunch {
    const PlanSample *ps;
    const int64_t *t0r;       // [S + 1] first range thread of each sample (t0r[S]: total)
    int S, K1;
    const int32_t *boff;
    const Block *blocks;
    const double *bbox;
    const int32_t *bruns;     // slab runs of every block (k_blocks)
    RawRange *raw;            // the ranges, one chunk per workgroup
    unsigned long long *raw_cursor;  // ranges taken (may pass raw_cap)
    int64_t raw_cap;
    int32_t *roff, *rcnt;     // [range threads] first range and number of ranges of the thread
    unsigned long long *overflow;
};

// The position ranges of every (block, stencil column) pair: one thread per
// pair counts its ranges, the workgroup reserves one chunk of the range list,
// and each thread stores its ranges there with their first index and count.
__global__ __launch_bounds__(kPlanThreads) void k_plan_ranges(RangeLaunch A) {
    __shared__ int cnt_l;
    __shared__ int64_t base_l;
    const int64_t nvwg = (A.t0r[A.S] + kPlanThreads - 1) / kPlanThreads;
    for (int64_t vb = blockIdx.x; vb < nvwg; vb += gridDim.x) {
        const int s = find_sample(A.t0r, A.S, vb * kPlanThreads);
        const PlanSample &P = A.ps[s];
        const int64_t tl0 = vb * kPlanThreads - P.t0r;
        const int nx2 = P.nxw * P.nxw;
        const int b0 = A.boff[P.cell0], b1 = A.boff[P.cell0 + P.ncells];
        // the whole virtual workgroup lies past the sample's blocks
        if (tl0 >= P.ntr || b0 + tl0 / nx2 >= b1) continue;
        const int64_t tl = tl0 + threadIdx.x;
        const int blk = b0 + (int)(tl / nx2);
        const bool active = tl < P.ntr && blk < b1;
        const int so = (int)(tl % nx2);
        const double *bb = A.bbox + 6 * (int64_t)blk;
        Block B{};
        if (active) B = A.blocks[blk];
        __syncthreads();  // the previous virtual workgroup is done with cnt_l, base_l
        if (threadIdx.x == 0) cnt_l = 0;
        __syncthreads();
        int n = 0;
        const int32_t *br = A.bruns + (kBlockRuns + 1) * (int64_t)blk;
        if (active) plan_ranges(P, B, bb, so, A.K1, br, [&](int, int, int, int, uint8_t) { ++n; });
        const int off = n ? atomicAdd(&cnt_l, n) : 0;
        __syncthreads();
        if (threadIdx.x == 0) base_l = cnt_l ? (int64_t)atomicAdd(A.raw_cursor, (unsigned long long)cnt_l) : 0;
        __syncthreads();
        int64_t o = base_l + off;
        if (o + n > A.raw_cap) {
            if (n) atomicOr(A.overflow, 1ull);
            n = 0;
            o = 0;
        }
        if (active) {
            A.roff[P.t0r + tl] = (int32_t)o;
            A.rcnt[P.t0r + tl] = n;
        }
        if (active && n)
            plan_ranges(P, B, bb, so, A.K1, br, [&](int k, int pb, int pe, int rp, uint8_t sh) {
                A.raw[o++] = RawRange{pb, pe, k, (rp & 0xffff) | ((int)sh << 16)};
            });
    }
}

// A finished item of a plan thread.
struct OpenItem {
    int key, rp, np, nr;
    int sh;
    int32_t p[kItemRanges];  // first position of every range
    int32_t e[kItemRanges];  // end of every range in item coordinates
};

// Pack the ranges of the stencil columns [so0, so1) of one block (range
// threads tr0 + so) into work items: ranges of one (key, shift) go into the
// same item, in column order, until it holds pmax positions or kItemRanges
// ranges (a range that does not fit is split).  Up to kOpenItems items are
// filled at once; a range of a further (key, shift) finishes the fullest.
// close(o) is called once per finished item, in an order fixed by the data
// alone.  The open items live in registers: every array index is a
// compile-time constant after unrolling.
template <class F>
__device__ __forceinline__ void pack_items(int pmax, const RawRange *raw, const int32_t *roff, const int32_t *rcnt,
                                           int64_t tr0, int so0, int so1, F &&close) {
    int okey[kOpenItems], osh[kOpenItems], orp[kOpenItems], onp[kOpenItems], onr[kOpenItems];
    int32_t op[kOpenItems][kItemRanges], oe[kOpenItems][kItemRanges];
#pragma unroll
    for (int i = 0; i < kOpenItems; ++i) {
        okey[i] = -1;
        osh[i] = orp[i] = onp[i] = onr[i] = 0;
#pragma unroll
        for (int j = 0; j < kItemRanges; ++j) op[i][j] = oe[i][j] = 0;
    }
    auto finish = [&](int s) {
        OpenItem o;
#pragma unroll
        for (int i = 0; i < kOpenItems; ++i)
            if (i == s) {
                o.key = okey[i];
                o.rp = orp[i];
                o.sh = osh[i];
                o.np = onp[i];
                o.nr = onr[i];
#pragma unroll
                for (int j = 0; j < kItemRanges; ++j) {
                    o.p[j] = op[i][j];
                    o.e[j] = oe[i][j];
                }
                onp[i] = 0;
                onr[i] = 0;
            }
        close(o);
    };
    for (int so = so0; so < so1; ++so) {
        const int n = rcnt[tr0 + so];
        const int64_t o0 = roff[tr0 + so];
        for (int q = 0; q < n; ++q) {
            const RawRange r = raw[o0 + q];
            const int k = r.key, sh = (r.w >> 16) & 0xff, rp = (int16_t)(r.w & 0xffff);
            int pb = r.pb;
            const int pe = r.pe;
            int s = -1;
#pragma unroll
            for (int i = 0; i < kOpenItems; ++i)
                if (okey[i] == k && osh[i] == sh) s = i;
            int evict = -1;
            if (s < 0) {
                // a free slot, else the fullest item is finished early
#pragma unroll
                for (int i = kOpenItems - 1; i >= 0; --i)
                    if (okey[i] < 0) s = i;
                if (s < 0) {
                    s = 0;
                    int most = onp[0];
#pragma unroll
                    for (int i = 1; i < kOpenItems; ++i)
                        if (onp[i] > most) {
                            most = onp[i];
                            s = i;
                        }
                    evict = s;
                }
            }
            while (evict >= 0 || pb < pe) {
                if (evict >= 0) {
                    finish(evict);
                    evict = -1;
                }
                bool full = false;
                int take = 0;
#pragma unroll
                for (int i = 0; i < kOpenItems; ++i)
                    if (i == s) {
                        okey[i] = k;
                        osh[i] = sh;
                        orp[i] = rp;
                        take = pe - pb < pmax - onp[i] ? pe - pb : pmax - onp[i];
#pragma unroll
                        for (int j = 0; j < kItemRanges; ++j)
                            if (j == onr[i]) {
                                op[i][j] = pb;
                                oe[i][j] = onp[i] + take;
                            }
                        onp[i] += take;
                        onr[i] += 1;
                        full = onp[i] == pmax || onr[i] == kItemRanges;
                    }
                pb += take;
                if (full) evict = s;
                if (pb >= pe && evict >= 0) {
                    finish(evict);
                    evict = -1;
                }
            }
        }
    }
    for (int i = 0; i < kOpenItems; ++i) {
        bool open = false;
#pragma unroll
        for (int q = 0; q < kOpenItems; ++q)
            if (q == i) open = okey[q] >= 0 && onr[q] > 0;
        if (open) finish(i);
    }
}

struct PlanLaunch {
    const PlanSample *ps;
    const int64_t *t0;        // [S + 1] first pack thread of each sample (t0[S]: total)
    int S;
    int K1;                   // regions + 1 of the largest sample's key space (keys per sample K1^2)
    const int32_t *boff;      // block offset of every cell (exclusive scan of the block counts)
    const Block *blocks;
    const RawRange *raw;      // the ranges of every (block, column) pair
    const int32_t *roff, *rcnt;
    unsigned long long *key_count;   // count pass
    unsigned long long *key_cursor;  // emit pass
    const int64_t *key_off;
    Item *items;
    int64_t *cost;            // emit pass: balancing cost of every item
    int64_t cap;              // item capacity
    Range *ranges;            // emit pass: range records of the multi-range items
    unsigned long long *range_cursor;  // emit pass: range records taken (may pass range_cap)
    int64_t range_cap;
    unsigned long long *overflow;  // set when the items or ranges exceed their capacity
    unsigned long long *candidates;  // emit pass: sum of ns * np
    int emit;
    int local;                // 1: aggregate per workgroup in LDS (K1^2 keys fit)
    // deterministic order: with it, the emit pass also writes every item's
    // (key << 32 | pack thread); a stable sort by it fixes the order the
    // cursor atomics gave the items of one key
    unsigned long long *sort_key;
};

// Count (emit == 0) or emit the work items of every (block, group of
// kGroupColumns stencil columns) pair of every sample, counting-sorted by key
// (sample, R_s, R_p).  A persistent grid walks the virtual workgroups of the
// samples' worst-case thread ranges and skips those past a sample's actual
// block count (known only on the device).  With `local` the key counts are
// aggregated per workgroup in LDS, so each key sees one global atomic per
// workgroup instead of one per item.  The emit pass packs the items twice:
// first to reserve the workgroup's slots (keys, range records), then to place
// them.  Items or ranges beyond their capacity are not written and raise the
// overflow flag (the host then grows the capacity and counts again).
__global__ __launch_bounds__(kPlanThreads) void k_plan_items(PlanLaunch A) {
    // local: reserved bases [nk] (64-bit), then counts [nk]
    extern __shared__ int64_t lmem[];
    const int nk = A.K1 * A.K1;
    int64_t *lbase = lmem;
    int32_t *lcnt = (int32_t *)(lmem + nk);
    __shared__ unsigned long long cand_l;
    __shared__ int rcount_l;
    __shared__ int64_t rbase_l;
    const int64_t nvwg = (A.t0[A.S] + kPlanThreads - 1) / kPlanThreads;
    const bool local = A.local != 0;
    unsigned long long cand = 0;
    for (int64_t vb = blockIdx.x; vb < nvwg; vb += gridDim.x) {
        const int s = find_sample(A.t0, A.S, vb * kPlanThreads);
        const PlanSample &P = A.ps[s];
        const int64_t tl0 = vb * kPlanThreads - P.t0;
        const int nx2 = P.nxw * P.nxw;
        const int ngc = P.ngc;
        const int b0 = A.boff[P.cell0], b1 = A.boff[P.cell0 + P.ncells];
        // the whole virtual workgroup lies past the sample's blocks
        if (tl0 >= P.nt || b0 + tl0 / ngc >= b1) continue;
        const int64_t tl = tl0 + threadIdx.x;
        const int blk = b0 + (int)(tl / ngc);
        const bool active = tl < P.nt && blk < b1;
        const int so0 = (int)(tl % ngc) * kGroupColumns;
        const int so1 = so0 + kGroupColumns < nx2 ? so0 + kGroupColumns : nx2;
        const int64_t tr0 = P.t0r + (int64_t)(blk - b0) * nx2;  // the block's first range thread
        const int kb = s * nk;
        const int ns = active ? A.blocks[blk].ns : 0;
        __syncthreads();  // the previous virtual workgroup is done with the LDS tables
        if (local)
            for (int j = threadIdx.x; j < nk; j += blockDim.x) {
                lcnt[j] = 0;
                lbase[j] = 0;
            }
        if (threadIdx.x == 0) rcount_l = 0;
        __syncthreads();
        // pass 1: key counts (count pass; emit pass with local) and the
        // thread's range records (emit pass)
        int nrng = 0;
        if (active)
            pack_items(P.pmax, A.raw, A.roff, A.rcnt, tr0, so0, so1, [&](const OpenItem &o) {
                if (!A.emit || local) {
                    if (local)
                        atomicAdd(&lcnt[o.key], 1);
                    else
                        atomicAdd(&A.key_count[kb + o.key], 1ull);
                }
                if (o.nr > 1) nrng += o.nr;
            });
        if (!A.emit) {
            if (local) {
                __syncthreads();
                for (int j = threadIdx.x; j < nk; j += blockDim.x)
                    if (lcnt[j]) atomicAdd(&A.key_count[kb + j], (unsigned long long)lcnt[j]);
            }
            continue;
        }
        const int roff = nrng ? atomicAdd(&rcount_l, nrng) : 0;
        __syncthreads();
        if (local)
            for (int j = threadIdx.x; j < nk; j += blockDim.x) {
                const int c = lcnt[j];
                lbase[j] = c ? (int64_t)atomicAdd(&A.key_cursor[kb + j], (unsigned long long)c) : 0;
                lcnt[j] = 0;
            }
        if (threadIdx.x == 0)
            rbase_l = rcount_l ? (int64_t)atomicAdd(A.range_cursor, (unsigned long long)rcount_l) : 0;
        __syncthreads();
        int64_t rnext = rbase_l + roff;
        if (rnext + nrng > A.range_cap) {
            atomicOr(A.overflow, 1ull);
            continue;
        }
        // pass 2: place
        if (active)
            pack_items(P.pmax, A.raw, A.roff, A.rcnt, tr0, so0, so1, [&](const OpenItem &o) {
                const int k = o.key;
                const int64_t slot = A.key_off[kb + k] + (local ? lbase[k] + (int64_t)atomicAdd(&lcnt[k], 1)
                                                                : (int64_t)atomicAdd(&A.key_cursor[kb + k], 1ull));
                if (slot >= A.cap) {
                    atomicOr(A.overflow, 1ull);
                    return;
                }
                Item it;
                it.blk = blk;
                it.np = o.np;
                it.rp = (int16_t)o.rp;
                it.shifts = (uint8_t)o.sh;
                it.nr = (uint8_t)o.nr;
                if (o.nr == 1) {
                    it.p0 = o.p[0];
                } else {
                    it.p0 = (int32_t)rnext;
#pragma unroll
                    for (int i = 0; i < kItemRanges; ++i)
                        if (i < o.nr) {
                            Range r;
                            r.base = o.p[i] - (i ? o.e[i - 1] : 0);
                            r.end = o.e[i];
                            A.ranges[rnext + i] = r;
                        }
                    rnext += o.nr;
                }
                A.items[slot] = it;
                if (A.sort_key)
                    A.sort_key[slot] = ((unsigned long long)(kb + k) << 32) | (unsigned long long)(P.t0 + tl);
                // a wave pays for whole 64-lane position batches plus a fixed
                // per-item overhead, whatever the item's true pair count
                A.cost[slot] = (int64_t)ns * (((it.np + kWave - 1) / kWave) * kWave) + kItemOverhead;
                cand += (unsigned long long)ns * it.np;
            });
    }
    if (!A.emit) return;
    __syncthreads();
    if (threadIdx.x == 0) cand_l = 0;
    __syncthreads();
    if (cand) atomicAdd(&cand_l, cand);
    __syncthreads();
    if (threadIdx.x == 0 && cand_l) atomicAdd(A.candidates, cand_l);
}

// Shape blocks: per (column, slab group) cell of every sample, the number of
// blocks of at most bsz shapes (count), then the blocks themselves at the
// scanned offsets (emit) with their bounding boxes and, for (r_p, pi), their
// slab runs (kBlockRuns + 1 ints per block: the count, -1 if they do not fit,
// then k0 | k1 << 11 | class << 22 per run).  cell0[S + 1] locates each
// sample's cells.
__global__ void k_blocks(const PlanSample *ps, const int64_t *cell0, int S, int bsz,
                         int32_t *nb, const int32_t *boff, Block *blocks, double *bbox, int32_t *bruns) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cell0[S]) return;
    const int s = find_sample(cell0, S, c);
    const PlanSample &P = ps[s];
    const Grid &g = P.g;
    const int64_t local = c - P.cell0;
    const int col = (int)(local / (P.ng + 1)), grp = (int)(local % (P.ng + 1));
    const int64_t base = (int64_t)col * (g.nz + 1);
    int s0, n, rs;
    if (grp < P.ng) {
        s0 = P.sstart[base + grp * P.G];
        n = P.sstart[base + (grp + 1) * P.G] - s0;
        rs = region_of(g, col / g.nc, col % g.nc, grp * P.G);
    } else {
        s0 = P.sstart[base + g.nz];
        n = P.sstart[base + g.nz + 1] - s0;
        rs = 0;
    }
    if (!boff) {
        nb[c] = (n + bsz - 1) / bsz;
        return;
    }
    const int o = boff[c];
    for (int b = 0; b * bsz < n; ++b) {
        Block B;
        B.s0 = s0 + b * bsz;
        B.ns = n - b * bsz < bsz ? n - b * bsz : bsz;
        B.col = col;
        B.grp = grp;
        B.sample = s;
        B.rs = rs;
        blocks[o + b] = B;
        // bounding box {a_lo, a_hi, b_lo, b_hi, l_lo, l_hi} of the block's shapes
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (int j = 0; j < B.ns; ++j) {
            const double *r = P.srec + 8 * (int64_t)(B.s0 + j);
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                lo[k] = fmin(lo[k], r[k]);
                hi[k] = fmax(hi[k], r[k]);
            }
        }
        double *bb = bbox + 6 * (int64_t)(o + b);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            bb[2 * k] = lo[k];
            bb[2 * k + 1] = hi[k];
        }
        if (bruns) {
            int32_t *br = bruns + (kBlockRuns + 1) * (int64_t)(o + b);
            // (r, mu): the runs of a column at zero transverse distance, the
            // widest reach of any column
            int nr = 0;
            slab_runs(P, lo[2], hi[2], 0.0, P.reach_r + 4 * g.margin, [&](int k0, int k1, int cls) {
                if (nr < kBlockRuns) br[1 + nr] = k0 | (k1 << 11) | (cls << 22);
                ++nr;
            });
            br[0] = nr <= kBlockRuns ? nr : -1;
        }
    }
}

// n_items = total of the key counts (0 past the capacity); raw total kept for the host.
__global__ void k_items_total(const int64_t *key_off, int64_t nkeys, int64_t cap, int64_t *n_items,
                              unsigned long long *raw) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        const int64_t t = key_off[nkeys];
        // beyond the capacity the host counts again; nothing is scheduled
        *n_items = t <= cap ? t : 0;
        *raw = (unsigned long long)t;
    }
}

// Deterministic order of the items: the permutation of a stable sort by
// (key, plan thread) applied to the items and their costs.
__global__ void k_iota(int32_t *a, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = (int32_t)i;
}

__global__ void k_gather_items(const int32_t *order, const int64_t *n_dev, const Item *items, const int64_t *cost,
                               Item *items_out, int64_t *cost_out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *n_dev) return;
    const int32_t o = order[i];
    items_out[i] = items[o];
    cost_out[i] = cost[o];
}

// Caller outputs from the accumulators: counts, DD (the pair count itself for
// unit weights), S+D, their tallies, SxD, the variance and the region stats.
__global__ void k_finalize(const FinSample *fs, const int64_t *fin0, int S) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= fin0[S]) return;
    const int s = find_sample(fin0, S, i);
    const FinSample &F = fs[s];
    const int64_t j = i - fin0[s];
    const int64_t nb = (int64_t)(1 + F.K) * F.B;
    if (j < nb) {
        const unsigned long long c = F.np[j];
        F.out.npairs[j] = c;
        if (F.out.dd) F.out.dd[j] = F.weighted ? F.dd[j] : (double)c;
        F.out.splus[j] = F.sp[j];
        if (j < F.B) {
            F.out.scross[j] = F.sx[j];
            if (F.out.variance) F.out.variance[j] = F.var_on ? F.var[j] : 0.0;
        }
        return;
    }
    const int64_t k = j - nb;  // region k < K
    if (F.out.region_stats) {
        double *r = F.out.region_stats + 4 * k;
        r[0] = (double)F.rcnt[k];
        r[1] = (double)F.rcnt[F.K + k];
        r[2] = F.rsum[k];
        r[3] = F.rsum[F.K + k];
    }
}

// Deterministic reduction, step 1: the histogram flushes of every wave of an
// all-static schedule (bounds from schedule_bounds(all_static)), which are the
// distinct planner keys among its items: the non-empty keys from key(i0) to
// key(i1 - 1) over its range [i0, i1) of the key-sorted items (key_off:
// nkeys + 1 item offsets), 0 for an empty range.  Their exclusive scan
// numbers the slots.
__device__ __forceinline__ int64_t key_of_item(const int64_t *key_off, int64_t nkeys, int64_t i) {
    int64_t lo = 0, hi = nkeys - 1;  // largest k with key_off[k] <= i
    while (lo < hi) {
        const int64_t mid = (lo + hi + 1) >> 1;
        if (key_off[mid] <= i)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// ne[k] = 1 when key k holds items; its exclusive scan counts the non-empty
// keys below k (empty keys between a wave's first and last key get no flush).
__global__ void k_key_nonempty(const int64_t *key_off, int64_t nkeys, int64_t *ne) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nkeys) ne[k] = key_off[k + 1] > key_off[k] ? 1 : 0;
}

__global__ void k_slot_runs(const int64_t *bounds, int W, const int64_t *key_off, int64_t nkeys,
                            const int64_t *ne_pre, int64_t *runs) {
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= W) return;
    const int64_t i0 = bounds[(int64_t)w * kScheduleUnits], i1 = bounds[(int64_t)(w + 1) * kScheduleUnits];
    runs[w] = i1 > i0 ? ne_pre[key_of_item(key_off, nkeys, i1 - 1) + 1] - ne_pre[key_of_item(key_off, nkeys, i0)]
                      : 0;
}

// Neumaier's compensated addition: s + c carries the sum with the rounding
// error of every addition recovered in c.
__device__ __forceinline__ void neumaier(double &s, double &c, double v) {
    const double t = s + v;
    c += fabs(s) >= fabs(v) ? (s - t) + v : (v - t) + s;
    s = t;
}

// Deterministic reduction, step 2: caller outputs from the slots.  Output
// element (tally t, bin b) of a sample sums, in slot order, the slots of that
// sample whose key (R_s, R_p) includes region t - 1 (every slot for t = 0),
// with compensated FP64 sums; region stats as k_finalize.
__global__ void k_finalize_det(const FinSample *fs, const int64_t *fin0, int S, const int64_t *slot_total,
                               const int64_t *slot_key, const unsigned long long *slot_np,
                               const double *slot_sp, const double *slot_sx, const double *slot_dd,
                               const double *slot_var, int BO) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= fin0[S]) return;
    const int s = find_sample(fin0, S, i);
    const FinSample &F = fs[s];
    const int64_t j = i - fin0[s];
    const int64_t nb = (int64_t)(1 + F.K) * F.B;
    if (j >= nb) {
        const int64_t k = j - nb;  // region k < K
        if (F.out.region_stats) {
            double *r = F.out.region_stats + 4 * k;
            r[0] = (double)F.rcnt[k];
            r[1] = (double)F.rcnt[F.K + k];
            r[2] = F.rsum[k];
            r[3] = F.rsum[F.K + k];
        }
        return;
    }
    const int t = (int)(j / F.B), b = (int)(j % F.B);
    // the sample's slots: keys [s << 32, (s + 1) << 32) of the ascending key list
    const int64_t n = *slot_total;
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (slot_key[mid] < ((int64_t)s << 32))
            lo = mid + 1;
        else
            hi = mid;
    }
    unsigned long long c = 0;
    double sp = 0.0, spc = 0.0, dd = 0.0, ddc = 0.0, sx = 0.0, sxc = 0.0, vr = 0.0, vrc = 0.0;
    for (int64_t q = lo; q < n; ++q) {
        const int64_t key = slot_key[q];
        if ((key >> 32) != s) break;
        const int rs = (int)((key >> 16) & 0xffff), rp = (int)(key & 0xffff);
        if (t > 0 && rs != t - 1 && rp != t - 1) continue;
        const int64_t o = q * BO + b;
        c += slot_np[o];
        neumaier(sp, spc, slot_sp[o]);
        if (F.weighted) neumaier(dd, ddc, slot_dd[o]);
        if (t == 0) {
            neumaier(sx, sxc, slot_sx[o]);
            if (F.var_on) neumaier(vr, vrc, slot_var[o]);
        }
    }
    F.out.npairs[j] = c;
    if (F.out.dd) F.out.dd[j] = F.weighted ? dd + ddc : (double)c;
    F.out.splus[j] = sp + spc;
    if (t == 0) {
        F.out.scross[b] = sx + sxc;
        if (F.out.variance) F.out.variance[b] = F.var_on ? vr + vrc : 0.0;
    }
}

// ---------------------------------------------------------------------------
// pair kernel

struct PairArgs {
    const Item *items;
    const Range *ranges;         // range records of the multi-range items
    const unsigned long long *overflow;  // planner's capacity flag
    const Block *blocks;
    const int64_t *n_items;      // device: work items in the list
    int n_waves;
    const int64_t *bounds;       // schedule unit -> first item, n_waves * kScheduleUnits + 3
    int *queue;                  // tail-chunk counter, zero at launch
    int mode, n_r, n_2, weighted, var, nbins;
    const double *T, *X;         // exact thresholds (global copy)
    double T0, Tn, X0, Xn;       // domain bounds (rppi: s and pi; rmu: r^2)
    float lgs, lgc;              // r bin guess: log2(v) * lgs + lgc
    float xs;                    // mu: (mu + 1) * xs
    double xs64, xc64;           // pi: pi * xs64 + xc64
    float dr;                    // guard bands: exact path when |guess - nearest integer| < d
    float d2;
    int scan;                    // 1: the exact path scans every threshold (guard band of a whole bin)
    const SampleDev *samples;
    int hist_stride;             // bytes per wave histogram
    int bin_off, nbins_out;      // this pass's first output bin (r0 * n_2) and the output's B
    int copies_log2;             // each bin is kept in 2^copies_log2 interleaved copies
    int off_sp, off_sx, off_dd, off_var; // byte offsets inside a wave histogram
    unsigned long long *rechecks;
    // wall-clock span of the launch's busy waves: [0] max of ~start, [1] max of end
    unsigned long long *clock;
    unsigned long long *iterations;  // wave-iterations (a 64-lane position batch against one shape)
    // deterministic reduction: the j-th histogram flush of wave w goes to
    // slot slot0[w] + j (all output bins of the pass, plain stores) with its
    // planner key; k_finalize_det sums the slots in slot order
    int det, K1;
    int64_t slot_cap;            // slots allocated (waves + planner keys bound every wave's flushes)
    const int64_t *slot0;
    int64_t *slot_key;
    unsigned long long *slot_np;
    double *slot_sp, *slot_sx, *slot_dd, *slot_var;
};

// Wave histogram layout (C = 2^copies_log2 interleaved copies of B bins;
// slot h = bin * C + copy).  Separate arrays keep the FP64 slots 8 bytes
// apart, so the lanes of one ds_add_f64 spread over all 16 bank pairs.
//   u32     cnt[B*C]   at 0
//   double  sp[B*C]    at off_sp    (S+D)
//   double  sx[B*C]    at off_sx    (SxD)
//   double  dd[B*C]    at off_dd    (weighted samples)
//   double  var[B*C]   at off_var   (variance requested)

// Shape record {a, b, l, ax, ay, e*w, w, 2*e*w} read through the scalar cache:
// the inner loop is uniform over shapes, so every field lands in SGPRs.
typedef const __attribute__((address_space(4))) double *ShapePtr;

__device__ void flush_hist(const PairArgs &P, char *hist, const SampleDev &S, int rs, int rp,
                           int lane) {
    uint32_t *cnt = (uint32_t *)hist;
    double *sp = (double *)(hist + P.off_sp);
    double *sx = (double *)(hist + P.off_sx);
    double *dd = P.weighted ? (double *)(hist + P.off_dd) : nullptr;
    double *vr = P.var ? (double *)(hist + P.off_var) : nullptr;
    const int B = P.nbins;
    const int BO = P.nbins_out;
    const int C = 1 << P.copies_log2;
    for (int b = lane; b < B; b += kWave) {
        uint32_t c = 0;
        double vsp = 0.0, vsx = 0.0, vdd = 0.0, vvr = 0.0;
        for (int k = 0; k < C; ++k) {
            int h = b * C + k;
            c += cnt[h];
            vsp += sp[h];
            vsx += sx[h];
            cnt[h] = 0;
            sp[h] = 0.0;
            sx[h] = 0.0;
            if (dd) {
                vdd += dd[h];
                dd[h] = 0.0;
            }
            if (vr) {
                vvr += vr[h];
                vr[h] = 0.0;
            }
        }
        if (c == 0) continue;
        const int gb = b + P.bin_off;  // the pass's local bin in the output grid
        atomicAdd(&S.npairs[gb], (unsigned long long)c);
        atomicAdd(&S.sp[gb], vsp);
        atomicAdd(&S.sx[gb], vsx);
        if (dd && S.dd) atomicAdd(&S.dd[gb], vdd);
        if (vr && S.var) atomicAdd(&S.var[gb], vvr);
        if (S.K > 0) {
            int64_t o = (int64_t)(1 + rs) * BO + gb;
            atomicAdd(&S.npairs[o], (unsigned long long)c);
            atomicAdd(&S.sp[o], vsp);
            if (dd && S.dd) atomicAdd(&S.dd[o], vdd);
            if (rp != rs) {
                o = (int64_t)(1 + rp) * BO + gb;
                atomicAdd(&S.npairs[o], (unsigned long long)c);
                atomicAdd(&S.sp[o], vsp);
                if (dd && S.dd) atomicAdd(&S.dd[o], vdd);
            }
        }
    }
}

// A slot's key: sample << 32 | R_s << 16 | R_p (regions < 2^16), ascending in
// the planner's key order.
__device__ __forceinline__ int64_t slot_key_of(int sample, int rs, int rp) {
    return ((int64_t)sample << 32) | ((int64_t)rs << 16) | (int64_t)rp;
}

// The deterministic flush: the wave's histogram (every bin of the pass, the
// empty ones too) is stored into its slot instead of added by atomics.
__device__ void flush_slot(const PairArgs &P, char *hist, int64_t slot, int64_t key, int lane) {
    if (slot >= P.slot_cap) return;  // cannot happen: the slot plan counts every flush
    uint32_t *cnt = (uint32_t *)hist;
    double *sp = (double *)(hist + P.off_sp);
    double *sx = (double *)(hist + P.off_sx);
    double *dd = P.weighted ? (double *)(hist + P.off_dd) : nullptr;
    double *vr = P.var ? (double *)(hist + P.off_var) : nullptr;
    const int C = 1 << P.copies_log2;
    for (int b = lane; b < P.nbins; b += kWave) {
        uint32_t c = 0;
        double vsp = 0.0, vsx = 0.0, vdd = 0.0, vvr = 0.0;
        for (int k = 0; k < C; ++k) {
            const int h = b * C + k;
            c += cnt[h];
            vsp += sp[h];
            vsx += sx[h];
            cnt[h] = 0;
            sp[h] = 0.0;
            sx[h] = 0.0;
            if (dd) {
                vdd += dd[h];
                dd[h] = 0.0;
            }
            if (vr) {
                vvr += vr[h];
                vr[h] = 0.0;
            }
        }
        const int64_t o = slot * P.nbins_out + P.bin_off + b;
        P.slot_np[o] = c;
        P.slot_sp[o] = vsp;
        P.slot_sx[o] = vsx;
        if (dd) P.slot_dd[o] = vdd;
        if (vr) P.slot_var[o] = vvr;
    }
    if (lane == 0) P.slot_key[slot] = key;
}


// r^2 summed in the reference's column order x, y, z: (a^2 + b^2) + l^2 with
// the line of sight along z (LZ), else (a^2 + l^2) + b^2 (line of sight x:
// (l^2 + a^2) + b^2, the same value since FP64 addition commutes).
template <bool LZ>
__device__ __forceinline__ double r2_column_order(double s, double a2, double b2, double l2) {
    return LZ ? s + l2 : (a2 + l2) + b2;
}

// #{k in 1..n-1 : E[k] <= v}, the exact bin of v in [E[0], E[n]).
__device__ __forceinline__ int count_edges(const double *E, int n, double v) {
    int k = 0;
    for (int i = 1; i < n; ++i) k += (E[i] <= v);
    return k;
}

// The same count from a guess g of v's position in bin units that is off by
// less than half a bin: with e the integer nearest g, every edge below E[e]
// lies below v and every edge above it above v, so one comparison decides.
// v lies in [E[0], E[n]) (the domain test), so e = 0 and e = n need none.
__device__ __forceinline__ int edge_near(const double *E, int n, double v, float g) {
    const int e = (int)rintf(g);
    if (e <= 0) return 0;
    if (e >= n) return n - 1;
    return e - 1 + (E[e] <= v ? 1 : 0);
}
__device__ __forceinline__ int edge_near(const double *E, int n, double v, double g) {
    const int e = (int)rint(g);
    if (e <= 0) return 0;
    if (e >= n) return n - 1;
    return e - 1 + (E[e] <= v ? 1 : 0);
}

#define MIA_CONST __attribute__((address_space(4)))

// Per-sample values the pair loop needs, read once per item through the
// scalar cache (they are wave-uniform and end up in SGPRs).
struct SampleView {
    const double *prec;
    const double *srec;
    double L, half;
    int los;
};

__device__ __forceinline__ SampleView load_view(const SampleDev *sd) {
    const MIA_CONST SampleDev *c = (const MIA_CONST SampleDev *)sd;
    SampleView v;
    v.prec = c->prec;
    v.srec = c->srec;
    v.L = c->L;
    v.half = c->half;
    v.los = c->los;
    return v;
}

// Periodic handling of an item: no shift on any axis, a constant shift per
// axis, or the reference's per-pair minimum-image wrap.
enum { SH_NONE = 0, SH_CONST = 1, SH_WRAP = 2 };

// A lane's pending histogram slot: consecutive pairs of one lane (one
// position against the shapes of a block) that land in the same slot are
// summed in registers and reach the LDS histogram as one atomic per run,
// which cuts the same-address collisions of the hot (outer r) bins.
struct Pend {
    int b;
    uint32_t c;
    double sp, sx, dd, vr;
};

template <bool WEIGHTED, bool VAR>
__device__ __forceinline__ void hist_add(const PairArgs &P, char *hist, int b, uint32_t c, double sp,
                                         double sx, double dd, double vr) {
    uint32_t *cnt = (uint32_t *)hist;
    atomicAdd(&cnt[b], c);
    double *hsp = (double *)(hist + P.off_sp);
    double *hsx = (double *)(hist + P.off_sx);
    atomicAdd(&hsp[b], sp);
    atomicAdd(&hsx[b], sx);
    if (WEIGHTED) atomicAdd(&((double *)(hist + P.off_dd))[b], dd);
    if (VAR) atomicAdd(&((double *)(hist + P.off_var))[b], vr);
}

template <int MODE, int SH, bool WEIGHTED, bool VAR, bool LZ>
__device__ __forceinline__ void one_pair(const PairArgs &P, const SampleView &S, bool valid,
                                         double pa, double pb, double pl, double pw, double qa,
                                         double qb, double ql, double qax, double qay, double qew,
                                         double qw, double q2ew, double sha, double shb, double shl,
                                         char *hist, const double *T, const double *X,
                                         uint32_t &nre, Pend &pd, float vlgc, double vxc) {
    const int n_r = P.n_r, n_2 = P.n_2;
    // the reference's own FP64 separation (measure_w_box_jk.py:159-162)
    double da = qa - pa, db = qb - pb, dl = ql - pl;
    if (SH == SH_WRAP) {
        da = wrap_sep(da, S.L, S.half);
        db = wrap_sep(db, S.L, S.half);
        dl = wrap_sep(dl, S.L, S.half);
    } else if (SH == SH_CONST) {
        da = da + sha;
        db = db + shb;
        dl = dl + shl;
    }
    double a2 = da * da, b2 = db * db;
    double s = a2 + b2;
    double rv;
    bool in;
    if (MODE == MIA_RPPI) {
        rv = s;
        in = valid & (s >= P.T0) & (s < P.Tn) & (dl >= P.X0) & (dl < P.Xn);
    } else {
        rv = r2_column_order<LZ>(s, a2, b2, dl * dl);
        in = valid & (s > 0.0) & (rv >= P.T0) & (rv < P.Tn);
    }
    if (in) {
        // bins: FP32 guess; exact FP64 threshold count when the guess lies
        // within the guard band of a bin edge (every edge, the domain edges
        // included, so no clamping is needed)
        float rv32 = (float)rv;
        float gr = __fmaf_rn(__log2f(rv32), P.lgs, vlgc);
        // pi: an FP64 guess (one FMA; its error is ~1e-14 bins, so a 1e-9 guard
        // band); mu = pi / r: an FP32 guess from rsqrt
        double g2;
        bool near2;
        if (MODE == MIA_RPPI) {
            g2 = fma(dl, P.xs64, vxc);
            near2 = fabs(g2 - rint(g2)) < P.d2;
        } else {
            float g2f = __fmaf_rn((float)dl, __builtin_amdgcn_rsqf(rv32), 1.0f) * P.xs;
            near2 = fabsf(g2f - rintf(g2f)) < P.d2;
            g2 = g2f;
        }
        // truncation is floor for the non-negative guesses; a guess within
        // the guard band of an integer (an edge, or a slightly negative
        // guess at the lower domain edge) takes the exact path
        int ir = (int)gr, i2 = (int)g2;
        bool near = ((int)(fabsf(gr - rintf(gr)) < P.dr) | (int)near2) != 0;
        if (near) {
            if (P.scan) {
                ir = count_edges(T, n_r, rv);
                if (MODE == MIA_RPPI) {
                    i2 = count_edges(X, n_2, dl);
                } else {
                    double mu64 = dl / sqrt(rv);
                    i2 = mu64 < X[0] ? n_2 - 1 : count_edges(X, n_2, mu64);
                }
            } else {
                ir = edge_near(T, n_r, rv, gr);
                if (MODE == MIA_RPPI) {
                    i2 = edge_near(X, n_2, dl, g2);
                } else {
                    double mu64 = dl / sqrt(rv);
                    i2 = mu64 < X[0] ? n_2 - 1 : edge_near(X, n_2, mu64, g2);
                }
            }
            ++nre;
        }
        // e+ = e cos 2phi = e (2 cos^2 phi - 1), ex = e sin 2phi = 2 e cos phi |sin phi|
        // (|.| because phi = arccos(d) lies in [0, pi]); cos phi = t/r_p,
        // |sin phi| = |c|/r_p with t, c the dot and cross products.
        double t = fma(da, qax, db * qay);
        double c = fma(da, qay, -(db * qax));
        double r0 = __builtin_amdgcn_rcp(s);
        double inv = fma(r0, fma(-s, r0, 1.0), r0);
        // 2 e w t / r_p^2: e+ = that * t - e w, ex = that * |c|
        double tw = (q2ew * t) * inv;
        double ep = fma(tw, t, -qew);
        double ex = tw * fabs(c);
        if (fabs(c) < fabs(t) * kSinExact) {
            // nearly parallel: the reference's d may round outside [-1, 1],
            // where arccos gives NaN and the pair adds no e+ (measure_w_box_jk.py:170-175)
            double rp = sqrt(s);
            double d = 0.0 + (da / rp) * qax;
            d = d + (db / rp) * qay;
            if (!(fabs(d) <= 1.0)) {
                ep = 0.0;
                ex = 0.0;
            }
        }
        if (WEIGHTED) {
            ep = pw * ep;
            ex = pw * ex;
        }
        int b = __mul24(ir, n_2) + i2;
        double dd = WEIGHTED ? pw * qw : 0.0;
        double vr = VAR ? ep * ep : 0.0;
        // a new slot flushes the pending run; the run then restarts from this
        // pair: keep = 0 drops the old sums (fma(0, sum, x) = x, fma(1, sum, x)
        // = sum + x, both exactly as the plain assignment and addition)
        const bool nw = b != pd.b;
        if (nw & (pd.b >= 0)) hist_add<WEIGHTED, VAR>(P, hist, pd.b, pd.c, pd.sp, pd.sx, pd.dd, pd.vr);
        const double keep = nw ? 0.0 : 1.0;
        pd.b = b;
        pd.c = (nw ? 0u : pd.c) + 1u;
        pd.sp = fma_into(keep, pd.sp, ep);
        pd.sx = fma_into(keep, pd.sx, ex);
        if (WEIGHTED) pd.dd = fma_into(keep, pd.dd, dd);
        if (VAR) pd.vr = fma_into(keep, pd.vr, vr);
    }
}

template <int MODE, int SH, bool WEIGHTED, bool VAR, bool LZ>
__device__ __forceinline__ void run_item(const PairArgs &P, const SampleView &S, int p0, int np, int nr,
                                         int shifts, ShapePtr shp, int ns, char *hist,
                                         const double *T, const double *X, int lane,
                                         uint32_t &nre, uint64_t &iterations) {
    const double L = S.L;
    const double sha = SH == SH_CONST ? shift_val(shifts & 3, L) : 0.0;
    const double shb = SH == SH_CONST ? shift_val((shifts >> 2) & 3, L) : 0.0;
    const double shl = SH == SH_CONST ? shift_val((shifts >> 4) & 3, L) : 0.0;
    // the additive constants of the FP32 bin guesses, held in VGPRs for the
    // whole item (a VOP3 FMA reads one scalar operand; the other would be
    // copied into a VGPR on every pair)
    float vlgc;
    double vxc;
    asm volatile("v_mov_b32 %0, %1" : "=v"(vlgc) : "s"(P.lgc));
    asm volatile("v_mov_b64 %0, %1" : "=v"(vxc) : "s"(P.xc64));
    // Strided lane -> position map: lane l of batch b takes position b + l*nb,
    // so one batch spans the whole line-of-sight range of the item and its
    // pairs spread over more (r, pi) bins (fewer same-address LDS atomics).
    const int nb = (np + kWave - 1) / kWave;
    iterations += (uint64_t)nb * ns;
    // The lane's current range: item coordinate k < ce is position cb + k.
    // Lane l walks k = l*nb .. l*nb + nb - 1, so it starts in the range
    // holding l*nb (found from the item's range records, scalar loads live
    // only here) and steps to the next record when k reaches ce.
    int cb = p0, ce = np, ci = 0;
    const int2 *rg = (const int2 *)(P.ranges + (nr > 1 ? p0 : 0));
    if (nr > 1) {
        const MIA_CONST int32_t *rr = (const MIA_CONST int32_t *)rg;
        const int kf = lane * nb < np ? lane * nb : np - 1;
        cb = rr[0];
        ce = rr[1];
#pragma unroll
        for (int i = 1; i < kItemRanges; ++i)
            if (i < nr && kf >= rr[2 * i - 1]) {
                ci = i;
                cb = rr[2 * i];
                ce = rr[2 * i + 1];
            }
    }
    // the lane's position of batch bt (a lane past the item's end reads a
    // valid position, its range's last)
    auto position = [&](int bt, bool &v) {
        const int k = bt + lane * nb;
        v = k < np;
        while (v && k >= ce) {
            const int2 r = rg[++ci];
            cb = r.x;
            ce = r.y;
        }
        return (const double2 *)(S.prec + 4 * (int64_t)(cb + (v ? k : ce - 1)));
    };
    Pend pd{-1, 0, 0, 0, 0, 0};
    // positions are loaded one batch ahead
    bool vn;
    const double2 *prn = position(0, vn);
    double2 abn = prn[0], lwn = prn[1];
    for (int bt = 0; bt < nb; ++bt) {
        const bool v0 = vn;
        const double2 ab = abn, lw = lwn;
        if (bt + 1 < nb) {
            prn = position(bt + 1, vn);
            abn = prn[0];
            lwn = prn[1];
        }
        double pa0 = ab.x, pb0 = ab.y, pl0 = lw.x;
        double pw0 = WEIGHTED ? lw.y : 1.0;
        // two shapes per iteration, the record after them loaded ahead (scalar
        // loads while the current pairs are evaluated)
        double qa = shp[0], qb = shp[1], ql = shp[2], qax = shp[3], qay = shp[4], qew = shp[5],
               qw = shp[6], q2ew = shp[7];
        int j = 0;
        // byte offsets of the records: 32-bit scalars added to the block's
        // base address by the scalar loads themselves
        const uint32_t olast = (uint32_t)(ns - 1) * 64u;
        uint32_t o = 0;
        for (; j + 1 < ns; j += 2, o += 128u) {
            ShapePtr r1 = (ShapePtr)((const MIA_CONST char *)shp + (o + 64u));
            const double b0 = r1[0], b1 = r1[1], b2 = r1[2], b3 = r1[3], b4 = r1[4], b5 = r1[5],
                         b6 = r1[6], b7 = r1[7];
            one_pair<MODE, SH, WEIGHTED, VAR, LZ>(P, S, v0, pa0, pb0, pl0, pw0, qa, qb, ql, qax, qay, qew, qw,
                                         q2ew, sha, shb, shl, hist, T, X, nre, pd, vlgc, vxc);
            ShapePtr r2 = (ShapePtr)((const MIA_CONST char *)shp + (o + 128u < olast ? o + 128u : olast));
            const double c0 = r2[0], c1 = r2[1], c2 = r2[2], c3 = r2[3], c4 = r2[4], c5 = r2[5],
                         c6 = r2[6], c7 = r2[7];
            one_pair<MODE, SH, WEIGHTED, VAR, LZ>(P, S, v0, pa0, pb0, pl0, pw0, b0, b1, b2, b3, b4, b5, b6,
                                         b7, sha, shb, shl, hist, T, X, nre, pd, vlgc, vxc);
            qa = c0;
            qb = c1;
            ql = c2;
            qax = c3;
            qay = c4;
            qew = c5;
            qw = c6;
            q2ew = c7;
        }
        if (j < ns)
            one_pair<MODE, SH, WEIGHTED, VAR, LZ>(P, S, v0, pa0, pb0, pl0, pw0, qa, qb, ql, qax, qay, qew, qw,
                                         q2ew, sha, shb, shl, hist, T, X, nre, pd, vlgc, vxc);
    }
    if (pd.b >= 0) hist_add<WEIGHTED, VAR>(P, hist, pd.b, pd.c, pd.sp, pd.sx, pd.dd, pd.vr);
}

// An item's pair loop for its periodic-shift variant.
template <int MODE, bool WEIGHTED, bool VAR, bool LZ>
__device__ __forceinline__ void run_shifted(const PairArgs &P, const SampleView &S, int p0, int np, int nr,
                                            int shifts, ShapePtr shp, int ns, char *hist, const double *T,
                                            const double *X, int lane, uint32_t &nre, uint64_t &iterations) {
    if (shifts & 64)
        run_item<MODE, SH_WRAP, WEIGHTED, VAR, LZ>(P, S, p0, np, nr, shifts, shp, ns, hist, T, X, lane, nre,
                                                   iterations);
    else if (shifts & 63)
        run_item<MODE, SH_CONST, WEIGHTED, VAR, LZ>(P, S, p0, np, nr, shifts, shp, ns, hist, T, X, lane, nre,
                                                    iterations);
    else
        run_item<MODE, SH_NONE, WEIGHTED, VAR, LZ>(P, S, p0, np, nr, shifts, shp, ns, hist, T, X, lane, nre,
                                                   iterations);
}

template <int MODE, bool WEIGHTED, bool VAR>
__global__ __launch_bounds__(kThreads) void k_pairs(PairArgs P) {
    extern __shared__ __align__(16) char lds[];
    const int lane = threadIdx.x & (kWave - 1);
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    // exact thresholds for the rare FP64 bin path (shared by the block's waves)
    double *T = (double *)lds;
    double *X = T + (P.n_r + 1);
    int nx = P.n_2 + 1;
    for (int i = threadIdx.x; i < P.n_r + 1; i += blockDim.x) T[i] = P.T[i];
    for (int i = threadIdx.x; i < nx; i += blockDim.x) X[i] = P.X[i];
    size_t thr_bytes = ((size_t)(P.n_r + 1 + nx) * 8 + 15) & ~(size_t)15;
    char *hist = lds + thr_bytes + (size_t)wv * P.hist_stride;
    for (int i = lane * 4; i < P.hist_stride; i += kWave * 4) *(uint32_t *)(hist + i) = 0u;
    __syncthreads();

    // runs of neighbouring items of the key-sorted list (which read the same
    // position cells) stay on one XCD
    const int w = xcd_run_block() * (blockDim.x / kWave) + wv;
    const int64_t n_items = *(const MIA_CONST int64_t *)P.n_items;
    // a plan that overflowed its capacities is counted again by the host
    if (w >= P.n_waves || n_items == 0 || *(const MIA_CONST unsigned long long *)P.overflow) return;
    if (lane == 0) atomicMax(&P.clock[0], ~(unsigned long long)wall_clock64());
    // a contiguous share of the cost-sorted items, then tail chunks from the counter
    const WaveSchedule sch = make_schedule(P.bounds, P.n_waves);
    int64_t it0, it1;
    sch.first(w, it0, it1);
    const MIA_CONST int32_t *items = (const MIA_CONST int32_t *)P.items;
    const MIA_CONST int32_t *blocks = (const MIA_CONST int32_t *)P.blocks;
    int cur_key_s = -1, cur_rs = -1, cur_rp = -1;
    const SampleDev *Sp = nullptr;
    SampleView S{};
    int64_t slot = P.det ? P.slot0[w] : 0;
    uint32_t nre = 0;
    uint64_t iterations = 0;  // wave-uniform
    do {
        for (int64_t i = it0; i < it1; ++i) {
            const MIA_CONST int32_t *ip = items + 4 * i;
            const int blk = ip[0], p0 = ip[1], np = ip[2], w3 = ip[3];
            const int rp = (int16_t)(w3 & 0xffff);
            const int shifts = (w3 >> 16) & 0xff;
            const int nr = (w3 >> 24) & 0xff;
            const MIA_CONST int32_t *bp = blocks + 6 * (int64_t)blk;
            const int s0 = bp[0], ns = bp[1], sample = bp[4], rs = bp[5];
            if (sample != cur_key_s || rs != cur_rs || rp != cur_rp) {
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                if (Sp) {
                    if (P.det)
                        flush_slot(P, hist, slot++, slot_key_of(cur_key_s, cur_rs, cur_rp), lane);
                    else
                        flush_hist(P, hist, *Sp, cur_rs, cur_rp, lane);
                }
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                if (sample != cur_key_s) {
                    Sp = &P.samples[sample];
                    S = load_view(Sp);
                }
                cur_key_s = sample;
                cur_rs = rs;
                cur_rp = rp;
            }
            ShapePtr shp = (ShapePtr)(S.srec + 8 * (int64_t)s0);
            // (r, mu) sums r^2 in an order fixed by the line of sight: one
            // loop per order, so the pair loop holds no branch on it
            if (MODE == MIA_RMU && S.los != 2)
                run_shifted<MODE, WEIGHTED, VAR, false>(P, S, p0, np, nr, shifts, shp, ns, hist, T, X, lane, nre,
                                                        iterations);
            else
                run_shifted<MODE, WEIGHTED, VAR, true>(P, S, p0, np, nr, shifts, shp, ns, hist, T, X, lane, nre,
                                                       iterations);
        }
    } while (sch.next(P.queue, lane, it0, it1));
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    if (Sp) {
        if (P.det)
            flush_slot(P, hist, slot, slot_key_of(cur_key_s, cur_rs, cur_rp), lane);
        else
            flush_hist(P, hist, *Sp, cur_rs, cur_rp, lane);
    }
    if (nre) atomicAdd(P.rechecks, (unsigned long long)nre);
    if (lane == 0) {
        if (iterations) atomicAdd(P.iterations, (unsigned long long)iterations);
        atomicMax(&P.clock[1], (unsigned long long)wall_clock64());
    }
}

template <int MODE, bool WEIGHTED, bool VAR>
void launch_pairs(const PairArgs &A, unsigned grid, int wpb, size_t lds, hipStream_t st) {
    hipLaunchKernelGGL((k_pairs<MODE, WEIGHTED, VAR>), dim3(grid), dim3(kWave * wpb), lds, st, A);
    MIA_HIP(hipGetLastError());
}

template <int MODE>
void launch_pairs_mode(const PairArgs &A, unsigned grid, int wpb, size_t lds, hipStream_t st) {
    if (A.weighted) {
        if (A.var)
            launch_pairs<MODE, true, true>(A, grid, wpb, lds, st);
        else
            launch_pairs<MODE, true, false>(A, grid, wpb, lds, st);
    } else {
        if (A.var)
            launch_pairs<MODE, false, true>(A, grid, wpb, lds, st);
        else
            launch_pairs<MODE, false, false>(A, grid, wpb, lds, st);
    }
}

// LDS bytes of one wave's histogram (one copy) and of the block's thresholds.
size_t wave_hist_bytes(const mia_bins &bins, bool weighted) {
    const size_t B = (size_t)bins.n_r * bins.n_2;
    size_t off = (4 * B + 15) & ~(size_t)15;
    off += 16 * B + (weighted ? 8 * B : 0) + (bins.accumulate_variance ? 8 * B : 0);
    return (off + 15) & ~(size_t)15;
}

size_t threshold_bytes(const mia_bins &bins) {
    return ((size_t)(bins.n_r + 1 + bins.n_2 + 1) * 8 + 15) & ~(size_t)15;
}

inline unsigned nblk(int64_t n, int t) { return (unsigned)((n + t - 1) / t < 1 ? 1 : (n + t - 1) / t); }

// Pair-kernel configuration of a binning: FP32 guess constants and guard
// bands, passes over r-bin ranges, histogram copies and offsets, waves.
struct PairConfig {
    PairArgs A{};        // pointers unset
    int nrp = 0, npass = 0, wpb = 0, waves = 0;
    unsigned grid = 0;
    size_t lds = 0;
    double l0 = 0, lgs = 0;
};

PairConfig pair_config(const mia_bins &bins, bool any_w, bool any_var, const mia_tuning *tuning, int cu_count) {
    PairConfig C;
    PairArgs &A = C.A;
    const int n_r = bins.n_r, n_2 = bins.n_2, B = n_r * n_2;
    A.mode = bins.mode;
    A.n_r = n_r;
    A.n_2 = n_2;
    A.weighted = any_w ? 1 : 0;
    A.var = any_var ? 1 : 0;
    A.nbins = B;
    // FP32 bin guesses and their guard bands (bin units).  The guess error
    // is bounded by a few FP32 ulps of the log2 / linear transform (the
    // bounds below, with a 20x margin); the exact FP64 count runs only for
    // guesses within the band of an edge.
    A.T0 = bins.r2_edges[0];
    A.Tn = bins.r2_edges[n_r];
    double l0 = std::log2(bins.r2_edges[0]), ln = std::log2(bins.r2_edges[n_r]);
    double lgs = n_r / (ln - l0);
    C.l0 = l0;
    C.lgs = lgs;
    A.lgs = (float)lgs;
    A.lgc = (float)(-l0 * lgs);
    const double u = std::ldexp(1.0, -23);  // FP32 ulp of 1
    // (+ 2u n_r for a pass's shifted constant -l0 lgs - r0)
    double err_r = (4 * u * std::fmax(std::fabs(l0), std::fabs(ln)) + 4 * u) * lgs + 4 * u * n_r;
    double err_2;
    double d2;
    if (bins.mode == MIA_RPPI) {
        A.X0 = bins.x2_edges[0];
        A.Xn = bins.x2_edges[n_2];
        // FP64 guess fma(pi, xs, xc): a few FP64 ulps of |pi| xs + n_2 bins,
        // far inside the 1e-9-bin guard band
        const double xs = n_2 / (bins.x2_edges[n_2] - bins.x2_edges[0]);
        A.xs64 = xs;
        A.xc64 = -bins.x2_edges[0] * xs;
        const double mx = std::fmax(std::fabs(bins.x2_edges[0]), std::fabs(bins.x2_edges[n_2]));
        const double e64 = std::ldexp(1.0, -52);
        err_2 = 8 * e64 * (mx * xs + n_2);
        d2 = std::fmax(1e-9, 20 * err_2);
    } else {
        A.X0 = 0;
        A.Xn = 0;
        A.xs = (float)(n_2 / 2.0);
        err_2 = 6 * u * n_2;
        d2 = std::fmax(1e-4, 20 * err_2);
    }
    double dr = std::fmax(1e-4, 20 * err_r);
    // pathological binnings (guard band > 5% of a bin): every pair takes the exact path
    bool exact_only = dr > 0.05 || d2 > 0.05;
    A.dr = exact_only ? 1.0f : (float)dr;
    A.d2 = exact_only ? 1.0f : (float)d2;
    A.scan = exact_only ? 1 : 0;
    // Passes over r-bin ranges.  Pass [r0, r0 + nrp) takes the pairs with
    // T[r0] <= r^2 < T[r0 + nrp] (exact FP64 comparisons, so every pair lands
    // in one pass), bins them locally and flushes to output bins
    // r0 * n_2 + local.  A pass's time goes as 1 / min(1, w / 16), w the
    // waves per CU its histogram leaves room for (latency-bound below 16
    // waves, VALU-bound above; measured on 60 x 60 and 100 x 100 bins), so nrp
    // minimises passes / min(1, w / 16) (ties: fewer passes), with at least
    // 4 waves per CU where any split allows it (at 2 a pass runs slower than
    // that model says).  Ordinary binnings run in one pass.
    const size_t thr_bytes = threshold_bytes(bins);
    int nrp = n_r;
    {
        mia_bins part = bins;
        part.accumulate_variance = any_var ? 1 : 0;
        auto waves_per_cu = [&](int nr_) {
            part.n_r = nr_;
            const size_t hs = wave_hist_bytes(part, any_w);
            int w = kMaxWavesPerBlock;  // as the launch below: 4, 2 or 1 waves per workgroup
            while (w > 1 && thr_bytes + hs * w > kLdsBytes) w /= 2;
            const size_t l = thr_bytes + hs * w;
            if (l > kLdsBytes) return 0;
            return (int)std::min<size_t>(32 / w, kLdsBytes / l) * w;
        };
        if (tuning && tuning->r_bins_per_pass > 0) {
            // as asked, but no more than one wave's histogram can hold
            nrp = std::min(n_r, tuning->r_bins_per_pass);
            while (nrp > 1 && waves_per_cu(nrp) == 0) --nrp;
        } else {
            double best = 1e300;
            int best_p = 0;
            for (int pass_min_waves : {4, 1}) {
                for (int c = n_r; c >= 1; --c) {
                    const int w = waves_per_cu(c);
                    if (w < pass_min_waves) continue;
                    const int passes = (n_r + c - 1) / c;
                    const double cost = (double)passes / std::min(1.0, w / 16.0);
                    if (cost < best * (1 - 1e-9) || (cost <= best * (1 + 1e-9) && passes < best_p)) {
                        best = cost;
                        best_p = passes;
                        nrp = c;
                    }
                }
                if (best_p) break;
            }
        }
    }
    C.nrp = nrp;
    C.npass = (n_r + nrp - 1) / nrp;
    const int Bp = nrp * n_2;     // bins of the largest pass
    A.nbins_out = B;
    A.bin_off = 0;
    // one histogram copy per wave: the pending-slot registers already absorb
    // most same-address collisions, and one copy leaves LDS for 8 waves per
    // SIMD (2 copies measured 20 % slower); mia_tuning.hist_copies is ignored
    const int copies_log2 = 0;
    A.copies_log2 = copies_log2;
    const int BC = Bp << copies_log2;
    int off = 4 * BC;
    off = (off + 15) & ~15;
    A.off_sp = off;
    off += 8 * BC;
    A.off_sx = off;
    off += 8 * BC;
    A.off_dd = off;
    if (A.weighted) off += 8 * BC;
    A.off_var = off;
    if (A.var) off += 8 * BC;
    A.hist_stride = (off + 15) & ~15;
    // waves per block: 4, or fewer when the histograms of 4 waves do not fit
    int wpb = kMaxWavesPerBlock;
    while (wpb > 1 && thr_bytes + (size_t)A.hist_stride * wpb > kLdsBytes) wpb /= 2;
    C.lds = thr_bytes + (size_t)A.hist_stride * wpb;
    if (C.lds > kLdsBytes)
        throw Error{MIA_E_UNSUPPORTED, "histogram too large for LDS (n_r*n_2 too big for MIA_FAST)"};
    C.wpb = wpb;
    // at most 8 waves per SIMD (32 per CU)
    int blocks_per_cu = (int)std::max<size_t>(1, std::min<size_t>(32 / wpb, kLdsBytes / C.lds));
    int waves = tuning && tuning->waves > 0 ? std::min(tuning->waves, kMaxWaves) : cu_count * blocks_per_cu * wpb;
    C.waves = std::max(1, waves);
    A.n_waves = C.waves;
    C.grid = (unsigned)((C.waves + wpb - 1) / wpb);
    return C;
}

}  // namespace

bool fast_path_fits(const mia_bins &bins, bool weighted, int jk) {
    const int64_t K1 = jk > 0 ? (int64_t)jk * jk * jk : 1;
    if (K1 * K1 > kMaxKeysPerSample) return false;
    // larger binnings run in passes over r-bin ranges: one r row must fit
    mia_bins row = bins;
    row.n_r = 1;
    return threshold_bytes(bins) + wave_hist_bytes(row, weighted) <= kLdsBytes;
}

// Sets the dynamic-LDS limit of every pair-kernel variant (once per process).
void pair_kernels_init() {
    static std::mutex mu;
    static uint64_t done = 0;  // one bit per device
    int dev = 0;
    MIA_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(mu);
    if (dev < 64 && (done >> dev & 1)) return;
    const void *fns[] = {
        (const void *)k_pairs<MIA_RPPI, false, false>, (const void *)k_pairs<MIA_RPPI, false, true>,
        (const void *)k_pairs<MIA_RPPI, true, false>,  (const void *)k_pairs<MIA_RPPI, true, true>,
        (const void *)k_pairs<MIA_RMU, false, false>,  (const void *)k_pairs<MIA_RMU, false, true>,
        (const void *)k_pairs<MIA_RMU, true, false>,   (const void *)k_pairs<MIA_RMU, true, true>};
    for (const void *f : fns)
        MIA_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLdsBytes));
    if (dev < 64) done |= 1ull << dev;
}

// ---------------------------------------------------------------------------
// staged count: host plan, workspace binding, launch sequence

struct CountPlan {
    int S = 0;
    int K1 = 1, plan_local = 0, bsz = kMaxBlockShapes;
    int64_t ncells = 0, nthreads = 0, nthreads_r = 0, nkeys = 0, fin_total = 0, item_cap = 0, range_cap = 0;
    int64_t acc_total = 0;           // accumulator words (u64 / f64) of all samples
    double margin = 0;               // largest geometric slack of the samples' grids
    PairConfig pc;
    bool any_w = false, any_var = false;
    bool det = false;                // deterministic reduction (slots instead of atomics)
    int64_t slot_cap = 0;            // slots: at most waves + planner keys
    std::vector<double> T, X;        // thresholds (X padded to n_2 + 1)
    std::vector<PlanSample> ps;
    std::vector<SampleDev> sd;
    std::vector<FinSample> fs;
    std::vector<int64_t> cell0, t0, t0r, fin0;
    std::vector<int64_t> acc0;       // first accumulator word of each sample
    std::vector<int> Bs;             // per sample: K
    // bound workspace
    char *param_dev = nullptr;
    std::vector<char> param;         // host image of the parameter block
    size_t off_T = 0, off_X = 0, off_cell0 = 0, off_t0 = 0, off_t0r = 0, off_fin0 = 0, off_ps = 0, off_sd = 0,
           off_fs = 0;
    size_t param_bytes = 0;
    unsigned long long *zero_begin = nullptr;
    size_t zero_words = 0;
    int32_t *nb = nullptr, *boff = nullptr;
    Block *blocks = nullptr;
    unsigned long long *kc = nullptr, *kcur = nullptr;
    int64_t *ko = nullptr;
    Item *items = nullptr;
    Range *ranges = nullptr;
    RawRange *raw = nullptr;
    int32_t *roff = nullptr, *rcnt = nullptr;
    int64_t *cost = nullptr, *cpre = nullptr, *n_items = nullptr, *bounds = nullptr;
    double *bbox = nullptr;
    int32_t *bruns = nullptr;
    unsigned long long *words = nullptr;   // [kCountReadback] read back after the count
    int *queue = nullptr;
    unsigned long long *acc = nullptr;
    int64_t *runs = nullptr, *slot0 = nullptr, *slot_key = nullptr, *ne = nullptr, *ne_pre = nullptr;
    unsigned long long *skey = nullptr, *skey2 = nullptr;   // deterministic item order
    int32_t *sidx = nullptr, *sidx2 = nullptr;
    Item *items2 = nullptr;
    int64_t *cost2 = nullptr;
    void *stmp = nullptr;
    size_t stmp_bytes = 0;
    unsigned long long *slot_np = nullptr;
    double *slot_sp = nullptr, *slot_sx = nullptr, *slot_dd = nullptr, *slot_var = nullptr;
    ScanBufs<int32_t> scan_nb;
    ScanBufs<int64_t> scan_keys, scan_cost, scan_runs, scan_ne;
    std::vector<int> G;
    std::string key;
};

int count_waves(const CountPlan &p) { return p.pc.waves; }
int count_bsz(const CountPlan &p) { return p.bsz; }
bool count_det(const CountPlan &p) { return p.det; }
int count_G(const CountPlan &p, int s) { return p.G[s]; }
int64_t count_item_cap(const CountPlan &p) { return p.item_cap; }
int64_t count_range_cap(const CountPlan &p) { return p.range_cap; }
std::string count_key(const CountPlan &p) { return p.key; }
size_t count_param_bytes(const CountPlan &p) { return p.param_bytes; }
const void *count_param_image(const CountPlan &p) { return p.param.data(); }
void *count_param_dev(const CountPlan &p) { return p.param_dev; }
const unsigned long long *count_readback(const CountPlan &p) { return p.words; }
void free_plan(CountPlan *p) { delete p; }

// Expected work items and multi-range records of a sample (mia_grid.hip's
// uniform-density model).
double estimate_items(const Grid &g, int G, double reach_r, double reach_z, bool rmu, int64_t ns, int64_t np,
                      int pmax, double *ranges);

CountPlan *plan_count(Ctx &ctx, const std::vector<SortedSample> &ss, const mia_bins &bins,
                      const mia_tuning *tuning, int64_t item_cap, int64_t range_cap, bool cap_exact) {
    pair_kernels_init();
    std::unique_ptr<CountPlan> P(new CountPlan());
    CountPlan &p = *P;
    const int S = (int)ss.size();
    MIA_REQUIRE(S >= 1, "no samples");
    p.S = S;
    const int n_r = bins.n_r, n_2 = bins.n_2, B = n_r * n_2;
    MIA_REQUIRE(n_r >= 1 && n_r <= kMaxBinsR && n_2 >= 1 && n_2 <= kMaxBins2, "bin counts out of range");
    const double reach_r = std::sqrt(bins.r2_edges[n_r]);
    const double reach_z = bins.mode == MIA_RPPI
                               ? std::fmax(std::fabs(bins.x2_edges[0]), std::fabs(bins.x2_edges[n_2]))
                               : reach_r;
    const int nx2 = bins.mode == MIA_RPPI ? n_2 + 1 : n_2;
    p.T.assign(bins.r2_edges, bins.r2_edges + n_r + 1);
    p.X.assign(n_2 + 1, 3.0);
    for (int i = 0; i < nx2; ++i) p.X[i] = bins.x2_edges[i];
    p.bsz = tuning && tuning->block_shapes > 0 ? std::min(tuning->block_shapes, kMaxBlockShapes) : kMaxBlockShapes;
    const int pmax_set = tuning && tuning->item_positions > 0 ? tuning->item_positions : 0;
    p.any_var = bins.accumulate_variance != 0;
    for (const auto &x : ss) p.any_w = p.any_w || x.weighted;
    p.pc = pair_config(bins, p.any_w, p.any_var, tuning, ctx.cu_count);
    p.det = tuning && tuning->deterministic != 0;

    p.ps.resize(S);
    p.sd.resize(S);
    p.fs.resize(S);
    p.G.resize(S);
    p.cell0.assign(S + 1, 0);
    p.t0.assign(S + 1, 0);
    p.t0r.assign(S + 1, 0);
    p.fin0.assign(S + 1, 0);
    p.acc0.assign(S + 1, 0);
    double est_items = 0, est_ranges = 0;
    for (int si = 0; si < S; ++si) {
        const SortedSample &x = ss[si];
        const Grid &g = x.g;
        p.K1 = std::max(p.K1, x.K > 0 ? x.K : 1);
        // slabs per block: explicit, else the grid's choice, else ~16 shapes
        // per block; always a divisor of the per-jk slab count
        int G = 1;
        if (tuning && tuning->slabs_per_block > 0) {
            G = tuning->slabs_per_block;
        } else if (g.G > 0) {
            G = g.G;
        } else {
            double per_slab = (double)x.n_shape / ((double)g.nc * g.nc * g.nz);
            double want = per_slab > 0 ? 16.0 / per_slab : g.nzj;
            G = (int)std::floor(want);
        }
        G = std::max(1, std::min(G, g.nzj));
        while (g.nzj % G) --G;
        p.G[si] = G;
        PlanSample &Q = p.ps[si];
        std::memset(&Q, 0, sizeof(Q));
        p.margin = std::max(p.margin, g.margin);
        Q.g = g;
        Q.mode = bins.mode;
        Q.periodic = x.periodic;
        Q.ng = g.nz / G;
        Q.G = G;
        Q.kx = (int)std::ceil((reach_r + 4 * g.margin) / g.cw) + 1;
        Q.nxw = std::min(g.nc, 2 * Q.kx + 1);
        Q.pmax = pmax_set > 0 ? pmax_set : (g.pmax > 0 ? g.pmax : 1024);
        Q.reach_r = reach_r;
        Q.reach_z = reach_z;
        Q.x_lo = bins.mode == MIA_RPPI ? bins.x2_edges[0] : -reach_r;
        Q.x_hi = bins.mode == MIA_RPPI ? bins.x2_edges[n_2] : reach_r;
        Q.sstart = x.sstart;
        Q.pstart = x.pstart;
        Q.ncells = (int64_t)g.nc * g.nc * (Q.ng + 1);
        Q.cell0 = p.cell0[si];
        p.cell0[si + 1] = Q.cell0 + Q.ncells;
        const int64_t blk_cap = (x.n_shape + p.bsz - 1) / p.bsz + std::min<int64_t>(Q.ncells, x.n_shape);
        Q.ngc = (Q.nxw * Q.nxw + kGroupColumns - 1) / kGroupColumns;
        Q.nt = blk_cap * Q.ngc;
        Q.t0 = p.t0[si];
        p.t0[si + 1] = Q.t0 + (Q.nt + kPlanThreads - 1) / kPlanThreads * kPlanThreads;
        Q.ntr = blk_cap * Q.nxw * Q.nxw;
        Q.t0r = p.t0r[si];
        p.t0r[si + 1] = Q.t0r + (Q.ntr + kPlanThreads - 1) / kPlanThreads * kPlanThreads;
        const int K = x.K;
        p.fin0[si + 1] = p.fin0[si] + (int64_t)(1 + K) * B + K;
        // accumulators: npairs, sp, dd ((1+K) B each), sx, var (B each)
        p.acc0[si + 1] = p.acc0[si] + 3 * (int64_t)(1 + K) * B + 2 * (int64_t)B;
        double er = 0;
        est_items += estimate_items(g, G, reach_r, reach_z, bins.mode == MIA_RMU, x.n_shape, x.n_pos, Q.pmax, &er);
        est_ranges += er;
    }
    p.ncells = p.cell0[S];
    p.nthreads = p.t0[S];
    p.nthreads_r = p.t0r[S];
    p.fin_total = p.fin0[S];
    p.acc_total = p.acc0[S];
    MIA_REQUIRE((double)S * p.K1 * p.K1 < 2e9, "too many planner keys (samples x regions^2)");
    p.nkeys = (int64_t)S * p.K1 * p.K1;
    p.slot_cap = p.det ? (int64_t)p.pc.waves + p.nkeys : 0;
    MIA_REQUIRE(!p.det || p.nthreads < ((int64_t)1 << 32), "plan too large for the deterministic item order");
    const size_t plan_lds = (size_t)p.K1 * p.K1 * (sizeof(int64_t) + sizeof(int32_t));
    p.plan_local = plan_lds <= 64 * 1024 ? 1 : 0;
    // item capacity: the caller's (grown after an overflow), at least 3x the
    // uniform-density estimate plus a margin; range capacity (the planner's
    // ranges and the items' range records): 1.25x the estimate, which runs
    // 1.3-2x above the counts of the clustered benchmark samples
    p.item_cap = cap_exact ? item_cap : std::max<int64_t>(item_cap, (int64_t)(3.0 * est_items) + 65536);
    p.item_cap = std::min<int64_t>(p.item_cap, (int64_t)1 << 31);
    p.range_cap = cap_exact ? range_cap : std::max<int64_t>(range_cap, (int64_t)(1.25 * est_ranges) + 65536);
    p.range_cap = std::min<int64_t>(p.range_cap, ((int64_t)1 << 31) - kItemRanges);

    // the launch-sequence key: everything the launches depend on except the
    // caller's pointers (which live in the parameter block)
    std::string &k = p.key;
    auto put = [&](const void *v, size_t n) { k.append((const char *)v, n); };
    put(&S, sizeof(S));
    put(&bins.mode, sizeof(int) * 4);
    put(p.T.data(), sizeof(double) * p.T.size());
    put(p.X.data(), sizeof(double) * p.X.size());
    if (tuning) put(tuning, sizeof(*tuning));
    put(&p.item_cap, sizeof(p.item_cap));
    put(&p.range_cap, sizeof(p.range_cap));
    for (int si = 0; si < S; ++si) {
        PlanSample q = p.ps[si];
        q.sstart = q.pstart = q.pfine = nullptr;
        q.srec = nullptr;
        put(&q, sizeof(q));
        put(&ss[si].n_pos, sizeof(int64_t));
        put(&ss[si].n_shape, sizeof(int64_t));
        put(&ss[si].K, sizeof(int));
        put(&ss[si].weighted, sizeof(int));
    }
    return P.release();
}

// bits of the deterministic item sort key: 32 of plan thread + the key index
int det_key_bits(const CountPlan &p) {
    int b = 1;
    while (b < 31 && ((int64_t)1 << b) < p.nkeys + 1) ++b;
    return 32 + b;
}

size_t count_bytes(const CountPlan &p) {
    Carve c;
    CountPlan q = p;
    count_bind(q, c, {}, nullptr);
    return c.off;
}

void count_bind(CountPlan &p, Carve &c, const std::vector<SortedSample> &ss, const mia_counts *outs) {
    const int S = p.S;
    const int n_r = p.pc.A.n_r, n_2 = p.pc.A.n_2, B = n_r * n_2;
    const bool dry = c.dry;
    // parameter block
    Carve pc;
    pc.dry = true;
    p.off_T = pc.off; pc.take<double>(n_r + 1);
    p.off_X = pc.off; pc.take<double>(n_2 + 1);
    p.off_cell0 = pc.off; pc.take<int64_t>(S + 1);
    p.off_t0 = pc.off; pc.take<int64_t>(S + 1);
    p.off_t0r = pc.off; pc.take<int64_t>(S + 1);
    p.off_fin0 = pc.off; pc.take<int64_t>(S + 1);
    p.off_ps = pc.off; pc.take<PlanSample>(S);
    p.off_sd = pc.off; pc.take<SampleDev>(S);
    p.off_fs = pc.off; pc.take<FinSample>(S);
    p.param_bytes = pc.off;
    p.param_dev = c.take<char>(p.param_bytes);
    // zeroed at the start of every count: accumulators, key counters,
    // read-back words, tail-chunk counters (one contiguous range)
    char *z0 = c.dry ? nullptr : c.base + c.off;
    p.acc = c.take<unsigned long long>(p.acc_total);
    p.kc = c.take<unsigned long long>(p.nkeys + 1);
    p.kcur = c.take<unsigned long long>(p.nkeys + 1);
    p.words = c.take<unsigned long long>(kCountReadback);
    p.queue = c.take<int>(kMaxBinsR);
    p.zero_begin = (unsigned long long *)z0;
    p.zero_words = dry ? 0 : (size_t)((c.base + c.off) - z0) / 8;
    // plan
    p.nb = c.take<int32_t>(p.ncells + 1);
    p.boff = c.take<int32_t>(p.ncells + 1);
    scan_carve<int32_t>(c, p.ncells, p.scan_nb);
    int64_t blk_total = 0;
    for (int si = 0; si < S; ++si) blk_total += p.ps[si].nt / p.ps[si].ngc;
    p.blocks = c.take<Block>(std::max<int64_t>(blk_total, 1));
    p.bbox = c.take<double>(6 * std::max<int64_t>(blk_total, 1));
    p.bruns = c.take<int32_t>((kBlockRuns + 1) * std::max<int64_t>(blk_total, 1));
    p.ko = c.take<int64_t>(p.nkeys + 1);
    scan_carve<int64_t>(c, p.nkeys, p.scan_keys);
    p.items = c.take<Item>(p.item_cap);
    // kItemRanges records of slack: the pair kernel loads a full set per item
    p.ranges = c.take<Range>(p.range_cap + kItemRanges);
    p.raw = c.take<RawRange>(p.range_cap);
    p.roff = c.take<int32_t>(std::max<int64_t>(p.nthreads_r, 1));
    p.rcnt = c.take<int32_t>(std::max<int64_t>(p.nthreads_r, 1));
    p.cost = c.take<int64_t>(p.item_cap + 1);
    p.cpre = c.take<int64_t>(p.item_cap + 1);
    scan_carve<int64_t>(c, p.item_cap, p.scan_cost);
    p.n_items = c.take<int64_t>(1);
    p.bounds = c.take<int64_t>((size_t)p.pc.waves * kScheduleUnits + 3);
    if (p.det) {
        const int64_t W = p.pc.waves, sw = p.slot_cap * B;
        p.runs = c.take<int64_t>(W + 1);
        p.slot0 = c.take<int64_t>(W + 1);
        scan_carve<int64_t>(c, W, p.scan_runs);
        p.ne = c.take<int64_t>(p.nkeys + 1);
        p.ne_pre = c.take<int64_t>(p.nkeys + 1);
        scan_carve<int64_t>(c, p.nkeys, p.scan_ne);
        p.slot_key = c.take<int64_t>(p.slot_cap);
        p.slot_np = c.take<unsigned long long>(sw);
        p.slot_sp = c.take<double>(sw);
        p.slot_sx = c.take<double>(sw);
        p.slot_dd = p.any_w ? c.take<double>(sw) : nullptr;
        p.slot_var = p.any_var ? c.take<double>(sw) : nullptr;
        p.skey = c.take<unsigned long long>(p.item_cap);
        p.skey2 = c.take<unsigned long long>(p.item_cap);
        p.sidx = c.take<int32_t>(p.item_cap);
        p.sidx2 = c.take<int32_t>(p.item_cap);
        p.items2 = c.take<Item>(p.item_cap);
        p.cost2 = c.take<int64_t>(p.item_cap + 1);
        size_t bytes = 0;
        MIA_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (const unsigned long long *)nullptr,
                                                   (unsigned long long *)nullptr, (const int32_t *)nullptr,
                                                   (int32_t *)nullptr, (int)p.item_cap, 0, det_key_bits(p),
                                                   (hipStream_t)0));
        p.stmp_bytes = std::max<size_t>(bytes, 256);
        p.stmp = c.take<char>(p.stmp_bytes);
    }
    if (dry) return;

    // parameter block image
    p.param.assign(p.param_bytes, 0);
    char *img = p.param.data();
    std::memcpy(img + p.off_T, p.T.data(), sizeof(double) * (n_r + 1));
    std::memcpy(img + p.off_X, p.X.data(), sizeof(double) * (n_2 + 1));
    std::memcpy(img + p.off_cell0, p.cell0.data(), sizeof(int64_t) * (S + 1));
    std::memcpy(img + p.off_t0, p.t0.data(), sizeof(int64_t) * (S + 1));
    std::memcpy(img + p.off_t0r, p.t0r.data(), sizeof(int64_t) * (S + 1));
    std::memcpy(img + p.off_fin0, p.fin0.data(), sizeof(int64_t) * (S + 1));
    unsigned long long *acc = p.acc;
    double *accd = (double *)p.acc;
    for (int si = 0; si < S; ++si) {
        const SortedSample &x = ss[si];
        const int K = x.K;
        const int64_t nb = (int64_t)(1 + K) * B;
        const int64_t a0 = p.acc0[si];
        PlanSample &Q = p.ps[si];
        Q.sstart = x.sstart;
        Q.pstart = x.pstart;
        Q.pfine = x.pfine;
        Q.sbits = x.sbits;
        Q.srec = x.srec;
        SampleDev &D = p.sd[si];
        std::memset(&D, 0, sizeof(D));
        D.prec = x.prec;
        D.srec = x.srec;
        D.pstart = x.pstart;
        D.L = x.g.L;
        D.half = x.g.half;
        D.los = x.g.los;
        D.K = K;
        D.nbins = B;
        D.npairs = acc + a0;
        D.sp = accd + a0 + nb;
        D.dd = x.weighted ? accd + a0 + 2 * nb : nullptr;
        D.sx = accd + a0 + 3 * nb;
        D.var = p.any_var ? accd + a0 + 3 * nb + B : nullptr;
        FinSample &F = p.fs[si];
        std::memset(&F, 0, sizeof(F));
        F.np = acc + a0;
        F.sp = accd + a0 + nb;
        F.dd = accd + a0 + 2 * nb;
        F.sx = accd + a0 + 3 * nb;
        F.var = accd + a0 + 3 * nb + B;
        F.rcnt = x.rcnt;
        F.rsum = x.rsum;
        F.out = outs[si];
        F.K = K;
        F.B = B;
        F.weighted = x.weighted;
        F.var_on = p.any_var ? 1 : 0;
    }
    std::memcpy(img + p.off_ps, p.ps.data(), sizeof(PlanSample) * S);
    std::memcpy(img + p.off_sd, p.sd.data(), sizeof(SampleDev) * S);
    std::memcpy(img + p.off_fs, p.fs.data(), sizeof(FinSample) * S);
}

void count_enqueue(const CountPlan &p, hipStream_t st, const hipEvent_t *ev) {
    const int S = p.S;
    const char *pd = p.param_dev;
    const PlanSample *ps = (const PlanSample *)(pd + p.off_ps);
    const int64_t *cell0 = (const int64_t *)(pd + p.off_cell0);
    const int64_t *t0 = (const int64_t *)(pd + p.off_t0);
    const int64_t *t0r = (const int64_t *)(pd + p.off_t0r);
    const int64_t *fin0 = (const int64_t *)(pd + p.off_fin0);
    if (ev) MIA_HIP(hipEventRecord(ev[0], st));
    fill<unsigned long long>(p.zero_begin, 0ull, (int64_t)p.zero_words, st);
    // shape blocks of every (column, slab group) cell
    hipLaunchKernelGGL(k_blocks, dim3(nblk(p.ncells, 256)), dim3(256), 0, st, ps, cell0, S, p.bsz, p.nb,
                       (const int32_t *)nullptr, (Block *)nullptr, (double *)nullptr, (int32_t *)nullptr);
    MIA_HIP(hipGetLastError());
    scan_ws<int32_t>(p.nb, p.boff, p.ncells, nullptr, p.scan_nb, st);
    hipLaunchKernelGGL(k_blocks, dim3(nblk(p.ncells, 256)), dim3(256), 0, st, ps, cell0, S, p.bsz, p.nb,
                       (const int32_t *)p.boff, p.blocks, p.bbox, p.bruns);
    MIA_HIP(hipGetLastError());
    // the position ranges of every (block, stencil column) pair
    RangeLaunch R{};
    R.ps = ps;
    R.t0r = t0r;
    R.S = S;
    R.K1 = p.K1;
    R.boff = p.boff;
    R.blocks = p.blocks;
    R.bbox = p.bbox;
    R.bruns = p.bruns;
    R.raw = p.raw;
    R.raw_cursor = p.words + 8;
    R.raw_cap = p.range_cap;
    R.roff = p.roff;
    R.rcnt = p.rcnt;
    R.overflow = p.words + 2;
    const int64_t nvwg_r = (p.nthreads_r + kPlanThreads - 1) / kPlanThreads;
    const unsigned rgrid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(nvwg_r, 256 * kPlanGridPerCu));
    hipLaunchKernelGGL(k_plan_ranges, dim3(rgrid), dim3(kPlanThreads), 0, st, R);
    MIA_HIP(hipGetLastError());
    // work items packed from the ranges, counting-sorted by (sample, R_s, R_p)
    PlanLaunch L{};
    L.ps = ps;
    L.t0 = t0;
    L.S = S;
    L.K1 = p.K1;
    L.boff = p.boff;
    L.blocks = p.blocks;
    L.raw = p.raw;
    L.roff = p.roff;
    L.rcnt = p.rcnt;
    L.key_count = p.kc;
    L.key_cursor = p.kcur;
    L.key_off = p.ko;
    L.items = p.items;
    L.cost = p.cost;
    L.cap = p.item_cap;
    L.ranges = p.ranges;
    L.range_cursor = p.words + 7;
    L.range_cap = p.range_cap;
    L.overflow = p.words + 2;
    L.candidates = p.words + 1;
    L.local = p.plan_local;
    const size_t plan_lds = p.plan_local ? (size_t)p.K1 * p.K1 * 12 : 0;
    const int64_t nvwg = (p.nthreads + kPlanThreads - 1) / kPlanThreads;
    const unsigned pgrid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(nvwg, 256 * kPlanGridPerCu));
    L.emit = 0;
    hipLaunchKernelGGL(k_plan_items, dim3(pgrid), dim3(kPlanThreads), plan_lds, st, L);
    MIA_HIP(hipGetLastError());
    scan_ws<int64_t>((const int64_t *)p.kc, p.ko, p.nkeys, nullptr, p.scan_keys, st);
    hipLaunchKernelGGL(k_items_total, dim3(1), dim3(64), 0, st, (const int64_t *)p.ko, p.nkeys, p.item_cap,
                       p.n_items, p.words + 3);
    MIA_HIP(hipGetLastError());
    L.emit = 1;
    const Item *items = p.items;
    const int64_t *cost = p.cost;
    if (p.det) {
        fill<unsigned long long>(p.skey, ~0ull, p.item_cap, st);
        hipLaunchKernelGGL(k_iota, dim3(nblk(p.item_cap, 256)), dim3(256), 0, st, p.sidx, p.item_cap);
        MIA_HIP(hipGetLastError());
        L.sort_key = p.skey;
    }
    hipLaunchKernelGGL(k_plan_items, dim3(pgrid), dim3(kPlanThreads), plan_lds, st, L);
    MIA_HIP(hipGetLastError());
    if (p.det) {
        // items of one key in (plan thread, emission) order: a stable sort of
        // the whole capacity (unused slots keep the largest key and sort last)
        size_t bytes = p.stmp_bytes;
        MIA_HIP(hipcub::DeviceRadixSort::SortPairs(p.stmp, bytes, (const unsigned long long *)p.skey, p.skey2,
                                                   (const int32_t *)p.sidx, p.sidx2, (int)p.item_cap, 0,
                                                   det_key_bits(p), st));
        hipLaunchKernelGGL(k_gather_items, dim3(nblk(p.item_cap, 256)), dim3(256), 0, st, (const int32_t *)p.sidx2,
                           (const int64_t *)p.n_items, (const Item *)p.items, (const int64_t *)p.cost, p.items2,
                           p.cost2);
        MIA_HIP(hipGetLastError());
        items = p.items2;
        cost = p.cost2;
    }
    scan_ws<int64_t>(cost, p.cpre, p.item_cap, p.n_items, p.scan_cost, st);
    schedule_bounds(p.cpre, p.n_items, p.pc.waves, p.bounds, st, p.det);
    if (p.det) {
        // slot numbers of every wave's flushes (slot0[W] = slots in all)
        hipLaunchKernelGGL(k_key_nonempty, dim3(nblk(p.nkeys, 256)), dim3(256), 0, st, (const int64_t *)p.ko,
                           p.nkeys, p.ne);
        MIA_HIP(hipGetLastError());
        scan_ws<int64_t>(p.ne, p.ne_pre, p.nkeys, nullptr, p.scan_ne, st);
        hipLaunchKernelGGL(k_slot_runs, dim3(nblk(p.pc.waves, 256)), dim3(256), 0, st, (const int64_t *)p.bounds,
                           p.pc.waves, (const int64_t *)p.ko, p.nkeys, (const int64_t *)p.ne_pre, p.runs);
        MIA_HIP(hipGetLastError());
        scan_ws<int64_t>(p.runs, p.slot0, p.pc.waves, nullptr, p.scan_runs, st);
    }
    if (ev) MIA_HIP(hipEventRecord(ev[1], st));
    // pair kernel, one launch per r-bin pass
    PairArgs A = p.pc.A;
    A.items = items;
    A.ranges = p.ranges;
    A.overflow = p.words + 2;
    A.blocks = p.blocks;
    A.n_items = p.n_items;
    A.bounds = p.bounds;
    A.T = (const double *)(pd + p.off_T);
    A.X = (const double *)(pd + p.off_X);
    A.samples = (const SampleDev *)(pd + p.off_sd);
    A.rechecks = p.words;
    A.clock = p.words + 4;
    A.iterations = p.words + 6;
    A.det = p.det ? 1 : 0;
    A.K1 = p.K1;
    A.slot_cap = p.slot_cap;
    A.slot0 = p.slot0;
    A.slot_key = p.slot_key;
    A.slot_np = p.slot_np;
    A.slot_sp = p.slot_sp;
    A.slot_sx = p.slot_sx;
    A.slot_dd = p.slot_dd;
    A.slot_var = p.slot_var;
    const int n_r = p.pc.A.n_r, n_2 = p.pc.A.n_2;
    for (int pass = 0; pass < p.pc.npass; ++pass) {
        const int r0 = pass * p.pc.nrp, r1 = std::min(n_r, r0 + p.pc.nrp);
        A.n_r = r1 - r0;
        A.nbins = (r1 - r0) * n_2;
        A.bin_off = r0 * n_2;
        A.T = (const double *)(pd + p.off_T) + r0;
        A.T0 = p.T[r0];
        A.Tn = p.T[r1];
        A.lgc = (float)(-p.pc.l0 * p.pc.lgs - r0);
        A.queue = p.queue + pass;
        if (A.mode == MIA_RPPI)
            launch_pairs_mode<MIA_RPPI>(A, p.pc.grid, p.pc.wpb, p.pc.lds, st);
        else
            launch_pairs_mode<MIA_RMU>(A, p.pc.grid, p.pc.wpb, p.pc.lds, st);
    }
    if (ev) MIA_HIP(hipEventRecord(ev[2], st));
    if (p.det)
        hipLaunchKernelGGL(k_finalize_det, dim3(nblk(p.fin_total, 256)), dim3(256), 0, st,
                           (const FinSample *)(pd + p.off_fs), fin0, S, (const int64_t *)(p.slot0 + p.pc.waves),
                           (const int64_t *)p.slot_key, (const unsigned long long *)p.slot_np,
                           (const double *)p.slot_sp, (const double *)p.slot_sx, (const double *)p.slot_dd,
                           (const double *)p.slot_var, n_r * n_2);
    else
        hipLaunchKernelGGL(k_finalize, dim3(nblk(p.fin_total, 256)), dim3(256), 0, st,
                           (const FinSample *)(pd + p.off_fs), fin0, S);
    MIA_HIP(hipGetLastError());
    if (ev) MIA_HIP(hipEventRecord(ev[3], st));
}

}  // namespace mia
