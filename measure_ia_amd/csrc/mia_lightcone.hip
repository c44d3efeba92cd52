// Lightcone pair counter: angular cells, chi-sorted segments, multi-range
// work items, per-wavefront LDS histograms.
//
// Replaces the per-position loops of the reference's lightcone counters
//   _measure_xi_rp_pi_lightcone_brute     (src/measureia/measure_w_lightcone.py:45-214)
//   _count_pairs_xi_rp_pi_lightcone_brute (measure_w_lightcone.py:216-343)
//   _measure_xi_r_mur_lightcone_brute     (src/measureia/measure_m_lightcone.py:45-219)
//   _count_pairs_xi_r_mur_lightcone_brute (measure_m_lightcone.py:221-360)
// and the leave-one-patch-out loops that call them once per jackknife patch
// (src/measureia/measure_jackknife.py:59-171, 265-483): the per-patch tallies
// of one traversal give every realisation (realisation i = total - tally[i]).
//
// The reference, for every position n and all shapes s at once
// (measure_w_lightcone.py:141-192), computes in the frame of the position
//   pi = chi_s - chi_n
//   dx = ((ra_s - ra_n) / 180 * pi_c) * chi_n * cos(dec_n / 180 * pi_c)
//   dy = ((dec_s - dec_n) / 180 * pi_c) * chi_n
//   (dx, dy) *= h with over_h;   r_p = sqrt(fl(dx^2 + dy^2))
// (multipoles: r = sqrt(fl(fl(dx^2 + dy^2) + pi^2)) from the unscaled dx, dy,
// mu = pi / r, and the scaled r_p only enters the rp_cut test).
//
// Device steps:
//   1. bounds : min/max of ra, dec, chi (per-workgroup partials, one
//               reducing workgroup, one readback) -> the grid.
//   2. sort   : positions by (cell, patch, chi slab), shapes by (cell, patch),
//               chi ascending inside (two stable radix sorts), gathered into
//               position records and shape arrays.
//   3. plan   : positions of one (cell, patch, slab) segment form blocks of at
//               most 64; a block meets every shape cell within its angular
//               reach (set by its smallest chi), and in each the chi window
//               [chi_lo + w_lo, chi_hi + w_hi] of every patch segment is one
//               range.  Consecutive ranges of one shape patch make a work item
//               (<= item_shapes shapes); items are sorted by the patch pair.
//   4. pairs  : persistent waves over the cost-balanced items.  A batch of 64
//               lanes is filled from the item's ranges (one shape per lane);
//               the block's positions are read through the scalar cache, one
//               at a time, and every lane evaluates its pair.  A batch of at
//               most 32 shapes is replicated over the lanes instead, each
//               replica taking every R-th position of the block (records read
//               per lane).  Separations are
//               computed with the position's precomputed scale
//               (pi_c/180 * chi_n * cos dec_n), bins are guessed in FP32 and,
//               within a guard band of any edge, the pair is recomputed in the
//               reference's own FP64 operation order and binned exactly.
//   5. flush  : a wave adds its histogram to the total and to the tallies of
//               both patches when the patch pair changes.
#include "mia_common.hpp"

#include <hipcub/hipcub.hpp>

#include <cmath>
#include <limits>

namespace mia {

namespace {

constexpr int kLcWavesPerBlock = 4;
constexpr int kLcThreads = kWave * kLcWavesPerBlock;
constexpr size_t kLcLds = 160 * 1024;
constexpr double kPiC = 3.141592653589793;  // numpy's np.pi
constexpr double kDeg = kPiC / 180.0;
constexpr int kLcMaxPatches = 4096;           // patch pairs stay below 2^24 keys
constexpr double kDomainSlack = 1e-9;         // relative widening of the approximate domain test

struct LcBlock {
    int32_t p0, np, cell, patch;
    double chi_lo, chi_hi;
};
struct LcRange {
    int32_t s0, n, patch, blk;
};
struct LcItem {
    int32_t blk, r0, nr, ns;
};
struct LcRect {
    int32_t r0, c0, nr, nc;
};
static_assert(sizeof(LcBlock) == 32, "LcBlock layout");
static_assert(sizeof(LcRange) == 16, "LcRange layout");
static_assert(sizeof(LcItem) == 16, "LcItem layout");

struct LcGrid {
    double ra0, dec0, hr, hd;  // grid origin and cell size, degrees
    int nrows, ncols, P1, nslab;
    double chi0, dchi;         // position slabs along chi
    double R;                  // bound on |dx|, |dy| of an in-domain pair (with slack)
    double wlo, whi;           // shape chi window of a position: [chi + wlo, chi + whi]
    double eps_deg;            // slack on cell edges, degrees
};

// Monotone map of doubles onto unsigned 64-bit keys (radix-sortable).
__device__ __forceinline__ uint64_t dkey(double x) {
    uint64_t b = (uint64_t)__double_as_longlong(x);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
inline double dkey_inv(uint64_t k) {
    uint64_t b = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
    double x;
    std::memcpy(&x, &b, 8);
    return x;
}

// Line-of-sight order inside a segment: chi quantised to qmax + 1 steps over
// the sample's range, a non-decreasing function of chi.  One radix sort by
// (segment << bits | quantised chi) orders a sample; searches compare the
// same quantised values, so the order they rely on is exact.  The sort takes
// 8 key bits per pass, so chi gets at least 12 bits and then what rounds the
// key up to a whole pass (at most 24): a coarser quantum only lets a range
// admit shapes up to one quantum outside its window, which the exact pi test
// rejects.
struct ChiQ {
    double lo, scale;
    uint64_t qmax;
    int bits;
};

__host__ __device__ __forceinline__ uint64_t chi_q(const ChiQ &q, double x) {
    double t = (x - q.lo) * q.scale;
    if (!(t > 0.0)) return 0;
    return t >= (double)q.qmax ? q.qmax : (uint64_t)t;
}

inline ChiQ make_chiq(double lo, double hi, int seg_bits) {
    ChiQ q;
    const int whole = (seg_bits + 12 + 7) / 8 * 8;
    q.bits = std::max(1, std::min(std::min(24, whole - seg_bits), 64 - seg_bits));
    q.qmax = ((uint64_t)1 << q.bits) - 1;
    q.lo = lo;
    q.scale = (double)q.qmax / std::fmax(hi - lo, 1e-300);
    return q;
}

// ---------------------------------------------------------------------------
// 1. bounds

struct LcBounds {
    unsigned long long ra_lo, ra_hi, dec_lo, dec_hi, chi_lo, chi_hi;
    double chi_sum;
    int bad;
    int pad;
};

__global__ __launch_bounds__(256) void k_lc_bounds(const double *__restrict__ ra,
                                                   const double *__restrict__ dec,
                                                   const double *__restrict__ chi,
                                                   const int32_t *__restrict__ patch, int64_t n,
                                                   int P, int sum_chi, LcBounds *part) {
    __shared__ unsigned long long s_lo[3][256], s_hi[3][256];
    __shared__ double s_sum[256];
    __shared__ int s_bad;
    if (threadIdx.x == 0) s_bad = 0;
    unsigned long long lo[3] = {~0ull, ~0ull, ~0ull}, hi[3] = {0ull, 0ull, 0ull};
    double sum = 0.0;
    int bad = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        double v[3] = {ra[i], dec[i], chi[i]};
        for (int k = 0; k < 3; ++k) {
            if (!isfinite(v[k])) bad |= 1;
            unsigned long long kk = dkey(v[k]);
            lo[k] = kk < lo[k] ? kk : lo[k];
            hi[k] = kk > hi[k] ? kk : hi[k];
        }
        sum += v[2];
        if (patch && P > 0 && (patch[i] < 0 || patch[i] >= P)) bad |= 2;
    }
    for (int k = 0; k < 3; ++k) {
        s_lo[k][threadIdx.x] = lo[k];
        s_hi[k][threadIdx.x] = hi[k];
    }
    s_sum[threadIdx.x] = sum;
    __syncthreads();
    if (bad) atomicOr(&s_bad, bad);
    for (int st = blockDim.x / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) {
            for (int k = 0; k < 3; ++k) {
                unsigned long long a = s_lo[k][threadIdx.x + st], b = s_hi[k][threadIdx.x + st];
                if (a < s_lo[k][threadIdx.x]) s_lo[k][threadIdx.x] = a;
                if (b > s_hi[k][threadIdx.x]) s_hi[k][threadIdx.x] = b;
            }
            s_sum[threadIdx.x] += s_sum[threadIdx.x + st];
        }
        __syncthreads();
    }
    // one partial per workgroup (reduced by k_lc_bounds_reduce: no contended
    // global atomics, and the chi sum has a fixed order)
    if (threadIdx.x == 0) {
        LcBounds r;
        r.ra_lo = s_lo[0][0];
        r.ra_hi = s_hi[0][0];
        r.dec_lo = s_lo[1][0];
        r.dec_hi = s_hi[1][0];
        r.chi_lo = s_lo[2][0];
        r.chi_hi = s_hi[2][0];
        r.chi_sum = sum_chi ? s_sum[0] : 0.0;
        r.bad = s_bad;
        r.pad = 0;
        part[blockIdx.x] = r;
    }
}

// The partials of k_lc_bounds into one LcBounds (one workgroup).
__global__ __launch_bounds__(256) void k_lc_bounds_reduce(const LcBounds *__restrict__ part, int n,
                                                          LcBounds *__restrict__ out) {
    __shared__ unsigned long long s_lo[3][256], s_hi[3][256];
    __shared__ double s_sum[256];
    __shared__ int s_bad[256];
    unsigned long long lo[3] = {~0ull, ~0ull, ~0ull}, hi[3] = {0ull, 0ull, 0ull};
    double sum = 0.0;
    int bad = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const LcBounds p = part[i];
        const unsigned long long plo[3] = {p.ra_lo, p.dec_lo, p.chi_lo};
        const unsigned long long phi[3] = {p.ra_hi, p.dec_hi, p.chi_hi};
        for (int k = 0; k < 3; ++k) {
            lo[k] = plo[k] < lo[k] ? plo[k] : lo[k];
            hi[k] = phi[k] > hi[k] ? phi[k] : hi[k];
        }
        sum += p.chi_sum;
        bad |= p.bad;
    }
    for (int k = 0; k < 3; ++k) {
        s_lo[k][threadIdx.x] = lo[k];
        s_hi[k][threadIdx.x] = hi[k];
    }
    s_sum[threadIdx.x] = sum;
    s_bad[threadIdx.x] = bad;
    __syncthreads();
    for (int st = blockDim.x / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) {
            for (int k = 0; k < 3; ++k) {
                unsigned long long a = s_lo[k][threadIdx.x + st], b = s_hi[k][threadIdx.x + st];
                if (a < s_lo[k][threadIdx.x]) s_lo[k][threadIdx.x] = a;
                if (b > s_hi[k][threadIdx.x]) s_hi[k][threadIdx.x] = b;
            }
            s_sum[threadIdx.x] += s_sum[threadIdx.x + st];
            s_bad[threadIdx.x] |= s_bad[threadIdx.x + st];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        LcBounds r;
        r.ra_lo = s_lo[0][0];
        r.ra_hi = s_hi[0][0];
        r.dec_lo = s_lo[1][0];
        r.dec_hi = s_hi[1][0];
        r.chi_lo = s_lo[2][0];
        r.chi_hi = s_hi[2][0];
        r.chi_sum = s_sum[0];
        r.bad = s_bad[0];
        r.pad = 0;
        *out = r;
    }
}

// ---------------------------------------------------------------------------
// 2. keys, sort, gather, segments

__device__ __forceinline__ int lc_cell(const LcGrid &g, double ra, double dec) {
    double fr = floor((dec - g.dec0) / g.hd);
    double fc = floor((ra - g.ra0) / g.hr);
    int r = fr < 0 ? 0 : (fr > g.nrows - 1 ? g.nrows - 1 : (int)fr);
    int c = fc < 0 ? 0 : (fc > g.ncols - 1 ? g.ncols - 1 : (int)fc);
    return r * g.ncols + c;
}

__global__ void k_lc_keys(const double *__restrict__ ra, const double *__restrict__ dec,
                          const double *__restrict__ chi, const int32_t *__restrict__ patch,
                          int64_t n, LcGrid g, int with_slab, ChiQ cq, uint64_t *__restrict__ key,
                          int32_t *__restrict__ idx) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int cell = lc_cell(g, ra[i], dec[i]);
    int p = patch ? patch[i] : 0;
    uint64_t k = (uint64_t)cell * g.P1 + (uint64_t)p;
    if (with_slab) {
        double fs = floor((chi[i] - g.chi0) / g.dchi);
        int s = fs < 0 ? 0 : (fs > g.nslab - 1 ? g.nslab - 1 : (int)fs);
        k = k * (uint64_t)g.nslab + (uint64_t)s;
    }
    key[i] = (k << cq.bits) | chi_q(cq, chi[i]);
    idx[i] = (int32_t)i;
}

// Segment keys of the sorted combined keys, and (q != nullptr) the quantised
// chi of every sorted object.
__global__ void k_lc_seg_of(uint64_t *__restrict__ key, int64_t n, int bits, uint32_t *__restrict__ q) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t k = key[i];
    if (q) q[i] = (uint32_t)(k & (((uint64_t)1 << bits) - 1));
    key[i] = k >> bits;
}

// Position record, 8 doubles: {ra, dec, chi, ax, ay, w, cos_dec, 0} with
// ax = pi_c/180 * chi * cos_dec * hs, ay = pi_c/180 * chi * hs (hs = h for
// (r_p, pi), 1 for (r, mu)).
__global__ void k_lc_gather_pos(const int32_t *__restrict__ idx, int64_t n,
                                const double *__restrict__ ra, const double *__restrict__ dec,
                                const double *__restrict__ chi, const double *__restrict__ cosd,
                                const double *__restrict__ w, double hs, double *__restrict__ rec) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int64_t j = idx[i];
    double c = chi[j], cd = cosd[j];
    double2 *r = (double2 *)(rec + 8 * i);
    r[0] = make_double2(ra[j], dec[j]);
    r[1] = make_double2(c, ((kDeg * c) * cd) * hs);
    r[2] = make_double2((kDeg * c) * hs, w ? w[j] : 1.0);
    r[3] = make_double2(cd, 0.0);
}

// Shape arrays in sorted order; c2w = c2*w, s2w = s2*w (zero without shapes).
__global__ void k_lc_gather_shp(const int32_t *__restrict__ idx, int64_t n,
                                const double *__restrict__ ra, const double *__restrict__ dec,
                                const double *__restrict__ chi, const double *__restrict__ c2,
                                const double *__restrict__ s2, const double *__restrict__ w,
                                double *__restrict__ sra, double *__restrict__ sdec,
                                double *__restrict__ schi, double *__restrict__ sc2w,
                                double *__restrict__ ss2w, double *__restrict__ sw) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int64_t j = idx[i];
    double wj = w ? w[j] : 1.0;
    sra[i] = ra[j];
    sdec[i] = dec[j];
    schi[i] = chi[j];
    sw[i] = wj;
    if (c2) {
        double a = c2[j], b = s2[j];
        // a NaN ellipticity adds no e+ / ex (the reference zeroes NaN e+, measure_w_lightcone.py:169-170)
        if (!(a == a) || !(b == b)) {
            a = 0.0;
            b = 0.0;
        }
        sc2w[i] = a * wj;
        ss2w[i] = b * wj;
    }
}

__global__ void k_lc_seg_flags(const uint64_t *__restrict__ key, int64_t n, int32_t *__restrict__ f) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) f[i] = (i == 0 || key[i] != key[i - 1]) ? 1 : 0;
}

// seg_start[id] / seg_key[id] of every segment (flags scanned into ids).
__global__ void k_lc_seg_emit(const uint64_t *__restrict__ key, const int32_t *__restrict__ f,
                              const int32_t *__restrict__ id, int64_t n, int32_t *__restrict__ start,
                              uint64_t *__restrict__ skey) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && f[i]) {
        start[id[i]] = (int32_t)i;
        skey[id[i]] = key[i];
    }
}

// First shape segment of every cell (cell c covers keys [c*P1, (c+1)*P1)).
__global__ void k_lc_cell_first(const uint64_t *__restrict__ skey, int nseg, int ncells, int P1,
                                int32_t *__restrict__ cfs) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c > ncells) return;
    uint64_t t = (uint64_t)c * P1;
    int lo = 0, hi = nseg;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (skey[mid] < t)
            lo = mid + 1;
        else
            hi = mid;
    }
    cfs[c] = lo;
}

// ---------------------------------------------------------------------------
// 3. plan

__global__ void k_lc_blocks_count(const int32_t *__restrict__ seg_start, int nseg, int bsz,
                                  int32_t *__restrict__ nb) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nseg) return;
    int n = seg_start[q + 1] - seg_start[q];
    nb[q] = (n + bsz - 1) / bsz;
}

__global__ void k_lc_blocks_emit(const int32_t *__restrict__ seg_start,
                                 const uint64_t *__restrict__ seg_key, int nseg, int bsz,
                                 const int32_t *__restrict__ off, const double *__restrict__ prec,
                                 int P1, int nslab, LcBlock *__restrict__ blocks) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nseg) return;
    int s0 = seg_start[q], n = seg_start[q + 1] - s0;
    uint64_t k = seg_key[q] / (uint64_t)nslab;
    int cell = (int)(k / (uint64_t)P1), patch = (int)(k % (uint64_t)P1);
    int o = off[q];
    for (int b = 0; b * bsz < n; ++b) {
        LcBlock B;
        B.p0 = s0 + b * bsz;
        B.np = n - b * bsz < bsz ? n - b * bsz : bsz;
        B.cell = cell;
        B.patch = patch;
        // the positions run along chi only to the quantisation of the sort:
        // the block's bounds are its true extremes
        double lo = prec[8 * (int64_t)B.p0 + 2], hi = lo;
        for (int j = 1; j < B.np; ++j) {
            const double c = prec[8 * (int64_t)(B.p0 + j) + 2];
            lo = c < lo ? c : lo;
            hi = c > hi ? c : hi;
        }
        B.chi_lo = lo;
        B.chi_hi = hi;
        blocks[o + b] = B;
    }
}

// Angular bounds of an in-domain partner of any position of the block:
// |dy| = |ddec| k chi_n >= |ddec| k chi_lo and |dx| >= |dra| k chi_lo cos_min,
// both below R.  Infinite reach (chi_lo <= 0, or a pole in reach) covers the grid.
struct Reach {
    double dl, dh, cl, ch;   // block cell bounds (deg)
    double kx, ky;           // lower bounds of the dx / dy scale (0: unbounded reach)
};

__device__ __forceinline__ Reach block_reach(const LcGrid &g, const LcBlock &B) {
    Reach r;
    int row = B.cell / g.ncols, col = B.cell % g.ncols;
    r.dl = g.dec0 + row * g.hd - g.eps_deg;
    r.dh = g.dec0 + (row + 1) * g.hd + g.eps_deg;
    r.cl = g.ra0 + col * g.hr - g.eps_deg;
    r.ch = g.ra0 + (col + 1) * g.hr + g.eps_deg;
    double cmin = 0.0;
    if (r.dl > -90.0 && r.dh < 90.0) {
        cmin = fmin(cos(r.dl * kDeg), cos(r.dh * kDeg)) * (1.0 - 1e-6) - 1e-12;
        if (cmin < 0) cmin = 0.0;
    }
    double chi = B.chi_lo > 0 ? B.chi_lo * (1.0 - 1e-9) : 0.0;
    r.ky = kDeg * chi * (1.0 - 1e-9);
    r.kx = r.ky * cmin;
    return r;
}

__device__ __forceinline__ void axis_span(double lo, double hi, double reach_deg, double o, double h,
                                          int n, int &i0, int &i1) {
    double a = floor((lo - reach_deg - o) / h), b = floor((hi + reach_deg - o) / h);
    a = a < 0 ? 0 : (a > n - 1 ? n - 1 : a);
    b = b < 0 ? 0 : (b > n - 1 ? n - 1 : b);
    i0 = (int)a;
    i1 = (int)b;
}

__global__ void k_lc_stencil(const LcBlock *__restrict__ blocks, int nblocks, LcGrid g,
                             LcRect *__restrict__ rect, int64_t *__restrict__ size) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nblocks) return;
    Reach r = block_reach(g, blocks[b]);
    const double inf = 1e300;
    double rd = r.ky > 0 ? g.R / r.ky : inf;
    double rr = r.kx > 0 ? g.R / r.kx : inf;
    int r0, r1, c0, c1;
    axis_span(r.dl, r.dh, rd, g.dec0, g.hd, g.nrows, r0, r1);
    axis_span(r.cl, r.ch, rr, g.ra0, g.hr, g.ncols, c0, c1);
    LcRect q;
    q.r0 = r0;
    q.c0 = c0;
    q.nr = r1 - r0 + 1;
    q.nc = c1 - c0 + 1;
    rect[b] = q;
    size[b] = (int64_t)q.nr * q.nc;
}

struct LcPlan {
    LcGrid g;
    const LcBlock *blocks;
    const LcRect *rect;
    const int64_t *soff;      // stencil offsets, nblocks + 1
    int nblocks;
    int64_t total;            // stencil cells over all blocks
    const int32_t *cfs;       // first shape segment of every cell, ncells + 1
    const int32_t *sseg;      // shape segment starts, nseg + 1
    const uint64_t *sseg_key;
    const uint32_t *sq;       // quantised chi of every sorted shape (their order inside a segment)
    ChiQ scq;
    int item_shapes;          // ranges are split at this many shapes
    int32_t *rcount;          // count pass: ranges per stencil cell
    const int64_t *roff;      // emit pass: range offsets, total + 1
    LcRange *ranges;
    int emit;
};

// First index in [lo, hi) whose quantised chi is >= q (lower) or > q (upper):
// the segment is sorted by quantised chi, so both are exact, and every shape
// with chi in [v_lo, v_hi] lies in [lower(q(v_lo)), upper(q(v_hi))).
__device__ __forceinline__ int lower_chi(const uint32_t *x, int lo, int hi, uint64_t q) {
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (x[mid] < q)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

__device__ __forceinline__ int upper_chi(const uint32_t *x, int lo, int hi, uint64_t q) {
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (x[mid] <= q)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// One thread per (block, stencil cell): the block's chi window in every shape
// segment (one per patch) of the cell, split into ranges of <= item_shapes.
__global__ void k_lc_ranges(LcPlan P) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= P.total) return;
    int lo = 0, hi = P.nblocks;  // last b with soff[b] <= t
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (P.soff[mid] <= t)
            lo = mid;
        else
            hi = mid;
    }
    const int b = lo;
    const LcBlock B = P.blocks[b];
    const LcRect q = P.rect[b];
    const LcGrid &g = P.g;
    int64_t loc = t - P.soff[b];
    int tr = q.r0 + (int)(loc / q.nc), tc = q.c0 + (int)(loc % q.nc);
    int nout = 0;
    Reach r = block_reach(g, B);
    double tdl = g.dec0 + tr * g.hd - g.eps_deg, tdh = g.dec0 + (tr + 1) * g.hd + g.eps_deg;
    double tcl = g.ra0 + tc * g.hr - g.eps_deg, tch = g.ra0 + (tc + 1) * g.hr + g.eps_deg;
    double ddec = fmax(0.0, fmax(tdl - r.dh, r.dl - tdh));
    double dra = fmax(0.0, fmax(tcl - r.ch, r.cl - tch));
    double y = ddec * r.ky, x = dra * r.kx;
    bool reach = x * x + y * y < g.R * g.R;
    if (reach) {
        int cell = tr * g.ncols + tc;
        double clo = B.chi_lo + g.wlo, chi_hi = B.chi_hi + g.whi;
        int64_t o = P.emit ? P.roff[t] : 0;
        for (int s = P.cfs[cell]; s < P.cfs[cell + 1]; ++s) {
            int a = lower_chi(P.sq, P.sseg[s], P.sseg[s + 1], chi_q(P.scq, clo));
            int e = upper_chi(P.sq, a, P.sseg[s + 1], chi_q(P.scq, chi_hi));
            int patch = (int)(P.sseg_key[s] % (uint64_t)g.P1);
            for (int k = a; k < e; k += P.item_shapes) {
                if (P.emit) {
                    LcRange R;
                    R.s0 = k;
                    R.n = e - k < P.item_shapes ? e - k : P.item_shapes;
                    R.patch = patch;
                    R.blk = b;
                    P.ranges[o + nout] = R;
                }
                ++nout;
            }
        }
    }
    if (!P.emit) P.rcount[t] = nout;
}

// Keys (block, patch) of the ranges, and the gather of the sorted order.
__global__ void k_lc_range_keys(const LcRange *__restrict__ r, int64_t n, int pstride, uint64_t *__restrict__ key,
                                int32_t *__restrict__ idx) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    key[i] = (uint64_t)r[i].blk * (uint64_t)pstride + (uint64_t)r[i].patch;
    idx[i] = (int32_t)i;
}

__global__ void k_lc_gather_ranges(const int32_t *__restrict__ idx, int64_t n, const LcRange *__restrict__ in,
                                   LcRange *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = in[idx[i]];
}

// One thread per block: consecutive ranges of one shape patch form an item of
// at most item_shapes shapes.
__global__ void k_lc_items(const LcRange *__restrict__ ranges, const int64_t *__restrict__ roff,
                           const int64_t *__restrict__ soff, int nblocks, int item_shapes,
                           int32_t *__restrict__ nitem, const int32_t *__restrict__ ioff,
                           LcItem *__restrict__ items, const LcBlock *__restrict__ blocks,
                           int P1, int32_t *__restrict__ ikey, int64_t *__restrict__ cost,
                           unsigned long long *__restrict__ cand) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nblocks) return;
    int64_t r_begin = roff[soff[b]], r_end = roff[soff[b + 1]];
    const bool emit = items != nullptr;
    int n = 0;
    int o = emit ? ioff[b] : 0;
    const int np = emit ? blocks[b].np : 0;
    const int pn = emit ? blocks[b].patch : 0;
    unsigned long long c = 0;
    int64_t r0 = r_begin;
    int acc = 0, patch = -1;
    auto close = [&](int64_t rend) {
        if (acc > 0) {
            if (emit) {
                LcItem it;
                it.blk = b;
                it.r0 = (int32_t)r0;
                it.nr = (int32_t)(rend - r0);
                it.ns = acc;
                items[o + n] = it;
                ikey[o + n] = pn * P1 + patch;
                // one 64-lane batch per 64 shapes, each a pass over the block's positions
                int64_t nbt = (acc + kWave - 1) / kWave;
                cost[o + n] = nbt * ((int64_t)kWave * np + 32) + 128;
                c += (unsigned long long)acc * np;
            }
            ++n;
        }
    };
    for (int64_t r = r_begin; r < r_end; ++r) {
        LcRange R = ranges[r];
        if (R.patch != patch || acc + R.n > item_shapes) {
            close(r);
            r0 = r;
            acc = 0;
            patch = R.patch;
        }
        acc += R.n;
    }
    close(r_end);
    if (!emit) nitem[b] = n;
    if (emit && c) atomicAdd(cand, c);
}

__global__ void k_lc_widen(const int32_t *__restrict__ a, int64_t *__restrict__ b, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

__global__ void k_lc_iota(int32_t *__restrict__ a, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = (int32_t)i;
}

__global__ void k_lc_gather_items(const int32_t *__restrict__ order, int64_t n,
                                  const LcItem *__restrict__ in, const int64_t *__restrict__ cin,
                                  LcItem *__restrict__ out, int64_t *__restrict__ cout) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = in[order[i]];
    cout[i] = cin[order[i]];
}

// ---------------------------------------------------------------------------
// 4. pair kernel

struct LcPairArgs {
    const LcItem *items;
    const LcBlock *blocks;
    const LcRange *ranges;
    const int64_t *cost_prefix;  // exclusive, n_items + 1
    const int64_t *total_cost;
    int64_t n_items;
    int n_waves;
    const int64_t *bounds;       // schedule unit -> first item, n_waves * kScheduleUnits + 3
    int *queue;                  // tail-chunk counter per pass, zero at launch
    const double *prec;          // position records
    const double *sra, *sdec, *schi, *sc2w, *ss2w, *sw;
    int n_r, n_2, nbins, P;      // P: patches (0: no tallies)
    const double *T, *X;
    double T0lo, Tnhi;           // approximate domain, widened by kDomainSlack
    double X0, Xn;               // (r_p, pi): exact pi domain
    double h;                    // projected-separation scale (h with over_h)
    double cut, cut_lo, cut_hi;  // (r, mu): r_p^2 threshold of rp_cut and its guard band
    float lgs, lgc, xs, xc, dr, d2;
    int scan;                    // 1: the exact path scans every threshold (guard band of a whole bin)
    int bin_off, nbins_out;      // this pass's first output bin; bins per output row (B)
    int hist_stride, off_dd, off_sp, off_sx;
    unsigned long long *npairs;  // (1+P)*B
    double *dd, *sp, *sx;        // (1+P)*B (sp, sx NULL for pair counts)
    unsigned long long *rechecks;
};

typedef const __attribute__((address_space(4))) double *CPtrD;
typedef const __attribute__((address_space(4))) int32_t *CPtrI;

struct LcPend {
    int b;
    uint32_t c;
    double dd, sp, sx;
};

template <bool SHAPES>
__device__ __forceinline__ void lc_hist_add(const LcPairArgs &P, char *hist, const LcPend &p) {
    atomicAdd(&((uint32_t *)hist)[p.b], p.c);
    atomicAdd(&((double *)(hist + P.off_dd))[p.b], p.dd);
    if (SHAPES) {
        atomicAdd(&((double *)(hist + P.off_sp))[p.b], p.sp);
        atomicAdd(&((double *)(hist + P.off_sx))[p.b], p.sx);
    }
}

template <bool SHAPES>
__device__ void lc_flush(const LcPairArgs &P, char *hist, int pn, int ps, int lane) {
    uint32_t *cnt = (uint32_t *)hist;
    double *hdd = (double *)(hist + P.off_dd);
    double *hsp = (double *)(hist + P.off_sp);
    double *hsx = (double *)(hist + P.off_sx);
    const int B = P.nbins, BO = P.nbins_out;
    for (int b = lane; b < B; b += kWave) {
        uint32_t c = cnt[b];
        if (c == 0) continue;
        double vdd = hdd[b], vsp = SHAPES ? hsp[b] : 0.0, vsx = SHAPES ? hsx[b] : 0.0;
        cnt[b] = 0;
        hdd[b] = 0.0;
        if (SHAPES) {
            hsp[b] = 0.0;
            hsx[b] = 0.0;
        }
        int tgt[3] = {0, -1, -1};
        if (P.P > 0) {
            tgt[1] = 1 + pn;
            if (ps != pn) tgt[2] = 1 + ps;
        }
        for (int k = 0; k < 3; ++k) {
            if (tgt[k] < 0) continue;
            const int64_t o = (int64_t)tgt[k] * BO + P.bin_off + b;
            atomicAdd(&P.npairs[o], (unsigned long long)c);
            atomicAdd(&P.dd[o], vdd);
            if (SHAPES) {
                atomicAdd(&P.sp[o], vsp);
                atomicAdd(&P.sx[o], vsx);
            }
        }
    }
}

__device__ __forceinline__ int lc_count_edges(const double *E, int n, double v) {
    int k = 0;
    for (int i = 1; i < n; ++i) k += (E[i] <= v);
    return k;
}

// The same count from a guess g of v's position in bin units that is off by
// less than half a bin (the FP32 guess of the approximate separation: the
// approximation differs from the exact value by a few ulps): with e the
// integer nearest g, one comparison with E[e] decides.  v lies in
// [E[0], E[n]), so e = 0 and e = n need none.
__device__ __forceinline__ int lc_edge_near(const double *E, int n, double v, float g) {
    const int e = (int)rintf(g);
    if (e <= 0) return 0;
    if (e >= n) return n - 1;
    return e - 1 + (E[e] <= v ? 1 : 0);
}

template <int MODE, bool SHAPES>
__device__ __forceinline__ void lc_pair(const LcPairArgs &P, bool valid, double qra, double qdec,
                                        double qchi, double qc2, double qs2, double qw, double pra,
                                        double pdec, double pchi, double pax, double pay, double pw,
                                        double pcos, char *hist, const double *T, const double *X,
                                        uint32_t &nre, LcPend &pd, float vlgc, float vxc) {
    const int n_r = P.n_r, n_2 = P.n_2;
    const double dra = qra - pra, ddec = qdec - pdec;
    const double pi = qchi - pchi;  // the reference's LOS, exactly
    const double dx = dra * pax, dy = ddec * pay;
    const double s = fma(dx, dx, dy * dy);  // approximate: guarded below
    double rv;
    bool in;
    if (MODE == MIA_RPPI) {
        rv = s;
        in = valid & (s >= P.T0lo) & (s < P.Tnhi) & (pi >= P.X0) & (pi < P.Xn);
    } else {
        rv = fma(pi, pi, s);
        const double sh = s * (P.h * P.h);
        in = valid & (rv >= P.T0lo) & (rv < P.Tnhi) & (sh > P.cut_lo);
    }
    if (!in) return;
    float rv32 = (float)rv, pif = (float)pi;
    float gr = __fmaf_rn(__log2f(rv32), P.lgs, vlgc);
    float g2 = MODE == MIA_RPPI ? __fmaf_rn(pif, P.xs, vxc)
                                : __fmaf_rn(pif, __builtin_amdgcn_rsqf(rv32), 1.0f) * P.xs;
    int ir = (int)gr, i2 = (int)g2;
    bool near = ((int)(fabsf(gr - rintf(gr)) < P.dr) | (int)(fabsf(g2 - rintf(g2)) < P.d2)) != 0;
    if (MODE == MIA_RMU) near |= s * (P.h * P.h) <= P.cut_hi;
    if (near) {
        // the reference's operation order (measure_w_lightcone.py:143-150)
        const double edra = ((qra - pra) / 180.0) * kPiC;
        const double edx = (edra * pchi) * pcos;
        const double eddec = ((qdec - pdec) / 180.0) * kPiC;
        const double edy = eddec * pchi;
        const double hx = edx * P.h, hy = edy * P.h;
        const double es = hx * hx + hy * hy;
        if (MODE == MIA_RPPI) {
            if (!(es >= T[0] && es < T[n_r])) return;
            ir = P.scan ? lc_count_edges(T, n_r, es) : lc_edge_near(T, n_r, es, gr);
            i2 = P.scan ? lc_count_edges(X, n_2, pi) : lc_edge_near(X, n_2, pi, g2);
        } else {
            // (r, mu): r from the unscaled separation (measure_m_lightcone.py:136-141)
            const double r2 = (edx * edx + edy * edy) + pi * pi;
            if (!(es > P.cut && r2 >= T[0] && r2 < T[n_r])) return;
            ir = P.scan ? lc_count_edges(T, n_r, r2) : lc_edge_near(T, n_r, r2, gr);
            const double mu = pi / sqrt(r2);
            i2 = mu < X[0] ? n_2 - 1 : (P.scan ? lc_count_edges(X, n_2, mu) : lc_edge_near(X, n_2, mu, g2));
        }
        ++nre;
    }
    double ep = 0.0, ex = 0.0;
    if (SHAPES) {
        // e+ = -e cos 2(phi_axis - phi_sep), ex = -e sin 2(phi_axis - phi_sep)
        // (measure_w_lightcone.py:163-168) with cos 2phi_sep = (dx^2 - dy^2)/s,
        // sin 2phi_sep = 2 dx dy / s; qc2 = -e cos 2phi_axis * w_s, qs2 = -e sin 2phi_axis * w_s
        const double A = fma(dx, dx, -(dy * dy));
        const double B2 = 2.0 * dx * dy;
        const double r0 = __builtin_amdgcn_rcp(s);
        const double inv = fma(r0, fma(-s, r0, 1.0), r0);
        ep = fma(qc2, A, qs2 * B2) * inv * pw;
        ex = fma(qs2, A, -(qc2 * B2)) * inv * pw;
        if (MODE == MIA_RMU && !(s > 0.0)) {
            // zero projected separation (only reachable with rp_cut < 0): NaN e+ -> 0
            ep = 0.0;
            ex = 0.0;
        }
    }
    const int b = __mul24(ir, n_2) + i2;
    // a new slot flushes the pending run, which then restarts from this pair:
    // keep = 0 drops the old sums (fma(0, sum, x) = x and fma(1, sum, x) =
    // sum + x, exactly as the plain assignment and addition)
    const bool nw = b != pd.b;
    if (nw & (pd.b >= 0)) lc_hist_add<SHAPES>(P, hist, pd);
    const double keep = nw ? 0.0 : 1.0;
    pd.b = b;
    pd.c = (nw ? 0u : pd.c) + 1u;
    pd.dd = fma_into(keep, pd.dd, pw * qw);
    if (SHAPES) {
        pd.sp = fma_into(keep, pd.sp, ep);
        pd.sx = fma_into(keep, pd.sx, ex);
    }
}

template <int MODE, bool SHAPES>
__global__ __launch_bounds__(kLcThreads) void k_lc_pairs(LcPairArgs P) {
    extern __shared__ __align__(16) char lds[];
    const int lane = threadIdx.x & (kWave - 1);
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    double *T = (double *)lds;
    double *X = T + (P.n_r + 1);
    const int nx = P.n_2 + 1;
    for (int i = threadIdx.x; i < P.n_r + 1; i += blockDim.x) T[i] = P.T[i];
    for (int i = threadIdx.x; i < nx; i += blockDim.x) X[i] = P.X[i];
    const size_t thr_bytes = ((size_t)(P.n_r + 1 + nx) * 8 + 15) & ~(size_t)15;
    char *hist = lds + thr_bytes + (size_t)wv * P.hist_stride;
    for (int i = lane * 4; i < P.hist_stride; i += kWave * 4) *(uint32_t *)(hist + i) = 0u;
    __syncthreads();

    // runs of neighbouring items of the patch-sorted list (which read the
    // same position blocks) stay on one XCD
    const int w = xcd_run_block() * (blockDim.x / kWave) + wv;
    if (w >= P.n_waves || P.n_items == 0) return;
    // a contiguous share of the patch-sorted items, then tail chunks from the counter
    const WaveSchedule sch = make_schedule(P.bounds, P.n_waves);
    int64_t it0, it1;
    sch.first(w, it0, it1);
    CPtrI items = (CPtrI)P.items;
    CPtrI blocks = (CPtrI)P.blocks;
    CPtrI ranges = (CPtrI)P.ranges;
    int cur_pn = -1, cur_ps = -1;
    uint32_t nre = 0;
    // the additive constants of the FP32 bin guesses, held in VGPRs (a VOP3
    // FMA reads one scalar operand; the other would be copied on every pair)
    float vlgc, vxc;
    asm volatile("v_mov_b32 %0, %1" : "=v"(vlgc) : "s"(P.lgc));
    asm volatile("v_mov_b32 %0, %1" : "=v"(vxc) : "s"(P.xc));
    do {
        for (int64_t i = it0; i < it1; ++i) {
            CPtrI ip = items + 4 * i;
            const int blk = ip[0], r0 = ip[1], nr = ip[2], ns = ip[3];
            CPtrI bp = blocks + 8 * (int64_t)blk;
            const int p0 = bp[0], np = bp[1], pn = bp[3];
            const int ps = ranges[4 * (int64_t)r0 + 2];
            if (pn != cur_pn || ps != cur_ps) {
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                if (cur_pn >= 0) lc_flush<SHAPES>(P, hist, cur_pn, cur_ps, lane);
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                cur_pn = pn;
                cur_ps = ps;
            }
            LcPend pd{-1, 0, 0.0, 0.0, 0.0};
            int j = r0, o = 0;
            const int jend = r0 + nr;
            for (int done = 0; done < ns; done += kWave) {
                // fill the 64 lanes from consecutive ranges (wave-uniform walk)
                int my = -1, filled = 0;
                while (filled < kWave && j < jend) {
                    const int rs0 = ranges[4 * (int64_t)j], rn = ranges[4 * (int64_t)j + 1];
                    const int take = rn - o < kWave - filled ? rn - o : kWave - filled;
                    if (lane >= filled && lane < filled + take) my = rs0 + o + (lane - filled);
                    filled += take;
                    o += take;
                    if (o >= rn) {
                        ++j;
                        o = 0;
                    }
                }
                if (filled <= kWave / 2 && np > 1) {
                    // Few shapes: R replicas of the chunk's shapes fill the
                    // lanes, replica r taking positions r, r + R, ... of the
                    // block, whose records are then read per lane.
                    const int R = kWave / filled < np ? kWave / filled : np;
                    const int rep = lane / filled;
                    const int msh = __shfl(my, lane - rep * filled);
                    const bool sv = (rep < R) & (msh >= 0);
                    const int q = sv ? msh : 0;
                    const double qra = P.sra[q], qdec = P.sdec[q], qchi = P.schi[q], qw = P.sw[q];
                    const double qc2 = SHAPES ? P.sc2w[q] : 0.0, qs2 = SHAPES ? P.ss2w[q] : 0.0;
                    int k = rep;
                    bool v = sv & (k < np);
                    const double *rec = P.prec + 8 * ((int64_t)p0 + (v ? k : 0));
                    double a0 = rec[0], a1 = rec[1], a2 = rec[2], a3 = rec[3], a4 = rec[4], a5 = rec[5],
                           a6 = rec[6];
                    const int nt = (np + R - 1) / R;
                    for (int t = 0; t < nt; ++t) {
                        const int kn = k + R;
                        const bool vn = sv & (kn < np);
                        const double *rn = P.prec + 8 * ((int64_t)p0 + (vn ? kn : 0));
                        const double b0 = rn[0], b1 = rn[1], b2 = rn[2], b3 = rn[3], b4 = rn[4], b5 = rn[5],
                                     b6 = rn[6];
                        lc_pair<MODE, SHAPES>(P, v, qra, qdec, qchi, qc2, qs2, qw, a0, a1, a2, a3, a4, a5, a6,
                                              hist, T, X, nre, pd, vlgc, vxc);
                        a0 = b0;
                        a1 = b1;
                        a2 = b2;
                        a3 = b3;
                        a4 = b4;
                        a5 = b5;
                        a6 = b6;
                        v = vn;
                        k = kn;
                    }
                    continue;
                }
                const bool valid = my >= 0;
                const int q = valid ? my : 0;
                const double qra = P.sra[q], qdec = P.sdec[q], qchi = P.schi[q], qw = P.sw[q];
                const double qc2 = SHAPES ? P.sc2w[q] : 0.0, qs2 = SHAPES ? P.ss2w[q] : 0.0;
                // two positions per iteration, the record after them loaded ahead
                // (scalar loads): 10 % faster than one position with a one-ahead
                // prefetch, whose record copies cost scalar instructions each pass
                CPtrD pr = (CPtrD)(P.prec + 8 * (int64_t)p0);
                double a0 = pr[0], a1 = pr[1], a2 = pr[2], a3 = pr[3], a4 = pr[4], a5 = pr[5], a6 = pr[6];
                int k = 0;
                // byte offsets of the records: 32-bit scalars that the scalar
                // loads add to the block's base address themselves
                typedef const __attribute__((address_space(4))) char *CPtrC;
                const uint32_t olast = (uint32_t)(np - 1) * 64u;
                uint32_t ob = 0;
                for (; k + 1 < np; k += 2, ob += 128u) {
                    CPtrD r1 = (CPtrD)((CPtrC)pr + (ob + 64u));
                    const double b0 = r1[0], b1 = r1[1], b2 = r1[2], b3 = r1[3], b4 = r1[4], b5 = r1[5],
                                 b6 = r1[6];
                    lc_pair<MODE, SHAPES>(P, valid, qra, qdec, qchi, qc2, qs2, qw, a0, a1, a2, a3, a4, a5,
                                          a6, hist, T, X, nre, pd, vlgc, vxc);
                    CPtrD r2 = (CPtrD)((CPtrC)pr + (ob + 128u < olast ? ob + 128u : olast));
                    const double c0 = r2[0], c1 = r2[1], c2 = r2[2], c3 = r2[3], c4 = r2[4], c5 = r2[5],
                                 c6 = r2[6];
                    lc_pair<MODE, SHAPES>(P, valid, qra, qdec, qchi, qc2, qs2, qw, b0, b1, b2, b3, b4, b5,
                                          b6, hist, T, X, nre, pd, vlgc, vxc);
                    a0 = c0;
                    a1 = c1;
                    a2 = c2;
                    a3 = c3;
                    a4 = c4;
                    a5 = c5;
                    a6 = c6;
                }
                if (k < np)
                    lc_pair<MODE, SHAPES>(P, valid, qra, qdec, qchi, qc2, qs2, qw, a0, a1, a2, a3, a4, a5,
                                          a6, hist, T, X, nre, pd, vlgc, vxc);
            }
            if (pd.b >= 0) lc_hist_add<SHAPES>(P, hist, pd);
        }
    } while (sch.next(P.queue, lane, it0, it1));
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    if (cur_pn >= 0) lc_flush<SHAPES>(P, hist, cur_pn, cur_ps, lane);
    if (nre) atomicAdd(P.rechecks, (unsigned long long)nre);
}

template <int MODE, bool SHAPES>
void lc_launch(const LcPairArgs &A, unsigned grid, int wpb, size_t lds, hipStream_t st) {
    MIA_HIP(hipFuncSetAttribute((const void *)k_lc_pairs<MODE, SHAPES>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL((k_lc_pairs<MODE, SHAPES>), dim3(grid), dim3(kWave * wpb), lds, st, A);
    MIA_HIP(hipGetLastError());
}

inline unsigned nblk_(int64_t n, int t) {
    int64_t b = (n + t - 1) / t;
    return (unsigned)(b < 1 ? 1 : b);
}

// Stable radix sort of (key, value) pairs with arena scratch.
template <class K>
void radix_sort(Ctx &ctx, const K *kin, K *kout, const int32_t *vin, int32_t *vout, int64_t n,
                int end_bit, hipStream_t st, const std::string &tag) {
    if (n <= 0) return;
    size_t bytes = 0;
    MIA_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, kin, kout, vin, vout, (int)n, 0,
                                               end_bit, st));
    void *tmp = ctx.arena.get(tag + "/tmp", bytes);
    MIA_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, bytes, kin, kout, vin, vout, (int)n, 0, end_bit,
                                               st));
}

int bits_for(uint64_t v) {
    int b = 1;
    while (b < 64 && (v >> b) != 0) ++b;
    return b;
}

struct SortedLc {
    int64_t n = 0;
    int nseg = 0;
    int32_t *seg = nullptr;     // nseg + 1 starts
    uint64_t *seg_key = nullptr;
};

// Sort one sample by (segment key, chi) and return the gather order.
int32_t *lc_sort(Ctx &ctx, const double *ra, const double *dec, const double *chi,
                 const int32_t *patch, int64_t n, const LcGrid &g, bool slab, uint64_t max_key, const ChiQ &cq,
                 const std::string &tag, hipStream_t st, uint64_t **sorted_key, uint32_t *sorted_q) {
    uint64_t *ck = ctx.arena.get<uint64_t>(tag + "/ck", n);
    uint64_t *k2 = ctx.arena.get<uint64_t>(tag + "/k2", n);
    int32_t *i0 = ctx.arena.get<int32_t>(tag + "/i0", n);
    int32_t *i1 = ctx.arena.get<int32_t>(tag + "/i1", n);
    hipLaunchKernelGGL(k_lc_keys, dim3(nblk_(n, 256)), dim3(256), 0, st, ra, dec, chi, patch, n, g,
                       slab ? 1 : 0, cq, ck, i0);
    MIA_HIP(hipGetLastError());
    // one stable sort by (segment key, quantised chi)
    radix_sort<uint64_t>(ctx, ck, k2, i0, i1, n, bits_for(max_key) + cq.bits, st, tag + "/s");
    hipLaunchKernelGGL(k_lc_seg_of, dim3(nblk_(n, 256)), dim3(256), 0, st, k2, n, cq.bits, sorted_q);
    MIA_HIP(hipGetLastError());
    *sorted_key = k2;
    return i1;
}

void lc_segments(Ctx &ctx, const uint64_t *key, int64_t n, const std::string &tag, hipStream_t st,
                 int32_t **flags_scan, int32_t **flags) {
    int32_t *f = ctx.arena.get<int32_t>(tag + "/f", n + 1);
    int32_t *id = ctx.arena.get<int32_t>(tag + "/id", n + 1);
    hipLaunchKernelGGL(k_lc_seg_flags, dim3(nblk_(n, 256)), dim3(256), 0, st, key, n, f);
    MIA_HIP(hipGetLastError());
    exclusive_scan<int32_t>(ctx, f, id, n, st, tag + "/fs", 0);
    *flags = f;
    *flags_scan = id;
}

size_t lc_hist_bytes(int B, bool shapes) {
    size_t off = ((size_t)4 * B + 15) & ~(size_t)15;
    off += (size_t)8 * B * (shapes ? 3 : 1);
    return (off + 15) & ~(size_t)15;
}

}  // namespace

// Larger binnings run in passes over r-bin ranges, so one r row of the
// histogram (with every threshold) has to fit.
bool lightcone_fits(const mia_bins &bins) {
    size_t thr = ((size_t)(bins.n_r + 1 + bins.n_2 + 1) * 8 + 15) & ~(size_t)15;
    return thr + lc_hist_bytes(bins.n_2, true) <= kLcLds;
}

void count_lightcone(Ctx &ctx, const mia_lc_sample &s, const mia_bins &bins, const mia_counts &out,
                     const mia_tuning *tuning, hipStream_t st, mia_stats *stats) {
    const bool shapes = s.c2 != nullptr;
    const int n_r = bins.n_r, n_2 = bins.n_2, B = n_r * n_2;
    const int P = s.n_patches;
    const int P1 = P > 0 ? P : 1;
    const int64_t nb_out = (int64_t)(1 + P) * B;
    const int64_t np_ = s.n_pos, ns_ = s.n_shape;
    MIA_REQUIRE(lightcone_fits(bins), "binning too large for the lightcone kernel's LDS histogram");

    // outputs are zeroed first, so empty samples return zeros
    fill<unsigned long long>((unsigned long long *)out.npairs, 0ull, nb_out, st);
    fill<double>(out.dd, 0.0, nb_out, st);
    if (out.splus) fill<double>(out.splus, 0.0, nb_out, st);
    if (out.scross) fill<double>(out.scross, 0.0, nb_out, st);
    if (stats) std::memset(stats, 0, sizeof(*stats));
    if (np_ == 0 || ns_ == 0) {
        MIA_HIP(hipStreamSynchronize(st));
        return;
    }

    Events ev(5);
    MIA_HIP(hipEventRecord(ev[0], st));

    // ---- 1. bounds (one readback)
    LcBounds *bd = ctx.arena.get<LcBounds>("lc/bounds", 2);
    {
        const unsigned gp = std::min<unsigned>(nblk_(np_, 256), 1024);
        const unsigned gs = std::min<unsigned>(nblk_(ns_, 256), 1024);
        LcBounds *part = ctx.arena.get<LcBounds>("lc/bounds_part", (size_t)gp + gs);
        hipLaunchKernelGGL(k_lc_bounds, dim3(gp), dim3(256), 0, st, s.ra, s.dec, s.chi, s.patch, np_, P, 1,
                           part);
        MIA_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_lc_bounds, dim3(gs), dim3(256), 0, st, s.ra_s, s.dec_s, s.chi_s, s.patch_s, ns_,
                           P, 0, part + gp);
        MIA_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_lc_bounds_reduce, dim3(1), dim3(256), 0, st, part, (int)gp, bd);
        MIA_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_lc_bounds_reduce, dim3(1), dim3(256), 0, st, part + gp, (int)gs, bd + 1);
        MIA_HIP(hipGetLastError());
    }
    LcBounds hb[2];
    MIA_HIP(hipMemcpyAsync(hb, bd, sizeof(hb), hipMemcpyDeviceToHost, st));
    MIA_HIP(hipStreamSynchronize(st));
    if (hb[0].bad & 1 || hb[1].bad & 1)
        throw Error{MIA_E_RANGE, "non-finite ra, dec or comoving distance"};
    if (hb[0].bad & 2 || hb[1].bad & 2) throw Error{MIA_E_ARG, "patch index outside [0, n_patches)"};
    const double ra_lo = std::fmin(dkey_inv(hb[0].ra_lo), dkey_inv(hb[1].ra_lo));
    const double ra_hi = std::fmax(dkey_inv(hb[0].ra_hi), dkey_inv(hb[1].ra_hi));
    const double dec_lo = std::fmin(dkey_inv(hb[0].dec_lo), dkey_inv(hb[1].dec_lo));
    const double dec_hi = std::fmax(dkey_inv(hb[0].dec_hi), dkey_inv(hb[1].dec_hi));
    const double pchi_lo = dkey_inv(hb[0].chi_lo), pchi_hi = dkey_inv(hb[0].chi_hi);
    const double chi_max = std::fmax(std::fabs(pchi_hi),
                                     std::fmax(std::fabs(dkey_inv(hb[1].chi_hi)),
                                               std::fabs(dkey_inv(hb[1].chi_lo))));

    // ---- grid
    LcGrid g{};
    const double hs = bins.mode == MIA_RPPI ? s.proj_scale : 1.0;
    const double Tn = bins.r2_edges[n_r];
    g.R = std::sqrt(Tn) * (1.0 + 1e-7) / hs + 1e-300;
    const double m = 1e-9 * (chi_max + 1.0);
    if (bins.mode == MIA_RPPI) {
        g.wlo = bins.x2_edges[0] - m - 1e-9 * std::fabs(bins.x2_edges[0]);
        g.whi = bins.x2_edges[n_2] + m + 1e-9 * std::fabs(bins.x2_edges[n_2]);
    } else {
        double r3 = std::sqrt(Tn) * (1.0 + 1e-7);
        g.wlo = -r3 - m;
        g.whi = r3 + m;
    }
    double chi_typ = hb[0].chi_sum / (double)np_;
    if (!(chi_typ > 0)) chi_typ = chi_max > 0 ? chi_max : 1.0;
    const double theta = g.R / (kDeg * chi_typ);  // angular reach at the typical distance, degrees
    const double dec_span = std::fmax(dec_hi - dec_lo, 1e-9);
    const double ra_span = std::fmax(ra_hi - ra_lo, 1e-9);
    const double dec_mid = 0.5 * (dec_lo + dec_hi);
    const double cmid = std::fmax(std::cos(dec_mid * kDeg), 0.05);
    // cell: the angular reach at the typical distance, but at least ~cell_shapes
    // shapes per cell.  Finer cells cut candidates but multiply work items and
    // ranges; on MI355X one typical reach per cell gave the shortest counts
    // (pair kernel ~1.4 ps per candidate plus ~1.2 ns per item).
    double cell = theta;
    const double want = tuning && tuning->cell_shapes > 0 ? tuning->cell_shapes : 16.0;
    const double area = dec_span * ra_span * cmid;
    const double cell_occ = std::sqrt(area * want / (double)ns_);
    cell = std::fmax(cell, cell_occ);
    int nrows = (int)std::ceil(dec_span / cell);
    int ncols = (int)std::ceil(ra_span * cmid / cell);
    if (tuning && tuning->slabs_per_side > 0) nrows = tuning->slabs_per_side;
    if (tuning && tuning->columns_per_side > 0) ncols = tuning->columns_per_side;
    nrows = std::max(1, std::min(nrows, 4096));
    ncols = std::max(1, std::min(ncols, 4096));
    while ((int64_t)nrows * ncols > (1 << 22)) {
        nrows = std::max(1, nrows / 2);
        ncols = std::max(1, ncols / 2);
    }
    g.nrows = nrows;
    g.ncols = ncols;
    g.dec0 = dec_lo;
    g.ra0 = ra_lo;
    g.hd = dec_span / nrows;
    g.hr = ra_span / ncols;
    g.eps_deg = 1e-9 * (std::fabs(dec_lo) + std::fabs(dec_hi) + std::fabs(ra_lo) + std::fabs(ra_hi) + 1.0);
    g.P1 = P1;
    // position slabs along chi, in eighths of the chi window: deeper slabs put
    // more positions in a block (fewer work items) at the price of a wider
    // window; measured best near 16 positions per (cell, slab) segment
    int eighths = 4;
    {
        const double ncell = (double)std::max(1, std::min(nrows, 4096)) * std::max(1, std::min(ncols, 4096));
        const double nsl4 = std::floor((pchi_hi - pchi_lo) / std::fmax((g.whi - g.wlo) * 0.5, 1e-9)) + 1;
        const double per_seg = (double)np_ / (ncell * nsl4);
        // ~16 positions per segment, deeper where shapes are sparse next to
        // positions: a block's per-item cost is then spread over more pairs
        // (measured on 1M positions / 300k shapes: 8 eighths 4 % faster than 5;
        // equal samples keep the plain rule)
        const double sparse = std::sqrt(std::fmax((double)np_ / (double)ns_, 1.0));
        eighths = (int)std::lround(4.0 * std::sqrt(16.0 / std::fmax(per_seg, 1e-3)) * sparse);
        eighths = std::max(2, std::min(eighths, 8));
    }
    if (tuning && tuning->slabs_per_block > 0) eighths = tuning->slabs_per_block;
    g.dchi = std::fmax((g.whi - g.wlo) * eighths / 8.0, 1e-9);
    g.chi0 = pchi_lo;
    double nsl = std::floor((pchi_hi - pchi_lo) / g.dchi) + 1;
    g.nslab = (int)std::max(1.0, std::min(nsl, 1.0e6));
    const int ncells = nrows * ncols;
    const int bsz = tuning && tuning->block_shapes > 0 ? std::min(tuning->block_shapes, kWave) : kWave;
    const int item_shapes = tuning && tuning->item_positions > 0 ? tuning->item_positions : 1024;

    // ---- 2. sort both samples
    uint64_t *pkey = nullptr, *skey = nullptr;
    const uint64_t pmax_key = ((uint64_t)ncells * P1) * (uint64_t)g.nslab;
    const uint64_t smax_key = (uint64_t)ncells * P1;
    const ChiQ pcq = make_chiq(pchi_lo, pchi_hi, bits_for(pmax_key));
    const ChiQ scq = make_chiq(dkey_inv(hb[1].chi_lo), dkey_inv(hb[1].chi_hi), bits_for(smax_key));
    uint32_t *sq = ctx.arena.get<uint32_t>("lc/sq", ns_);
    int32_t *pord = lc_sort(ctx, s.ra, s.dec, s.chi, s.patch, np_, g, true, pmax_key, pcq, "lc/p", st, &pkey,
                            nullptr);
    int32_t *sord = lc_sort(ctx, s.ra_s, s.dec_s, s.chi_s, s.patch_s, ns_, g, false, smax_key, scq, "lc/s", st,
                            &skey, sq);
    double *prec = ctx.arena.get<double>("lc/prec", 8 * np_);
    hipLaunchKernelGGL(k_lc_gather_pos, dim3(nblk_(np_, 256)), dim3(256), 0, st, pord, np_, s.ra, s.dec,
                       s.chi, s.cos_dec, s.w, hs, prec);
    MIA_HIP(hipGetLastError());
    double *sra = ctx.arena.get<double>("lc/sra", ns_);
    double *sdec = ctx.arena.get<double>("lc/sdec", ns_);
    double *schi = ctx.arena.get<double>("lc/schi", ns_);
    double *sw = ctx.arena.get<double>("lc/sw", ns_);
    double *sc2w = shapes ? ctx.arena.get<double>("lc/sc2w", ns_) : nullptr;
    double *ss2w = shapes ? ctx.arena.get<double>("lc/ss2w", ns_) : nullptr;
    hipLaunchKernelGGL(k_lc_gather_shp, dim3(nblk_(ns_, 256)), dim3(256), 0, st, sord, ns_, s.ra_s,
                       s.dec_s, s.chi_s, s.c2, s.s2, s.w_s, sra, sdec, schi, sc2w, ss2w, sw);
    MIA_HIP(hipGetLastError());
    int32_t *pf, *pid, *sf, *sid;
    lc_segments(ctx, pkey, np_, "lc/pseg", st, &pid, &pf);
    lc_segments(ctx, skey, ns_, "lc/sseg", st, &sid, &sf);
    int32_t cnt2[2];
    {
        // segment totals (the scans' totals)
        MIA_HIP(hipMemcpyAsync(&cnt2[0], pid + np_, 4, hipMemcpyDeviceToHost, st));
        MIA_HIP(hipMemcpyAsync(&cnt2[1], sid + ns_, 4, hipMemcpyDeviceToHost, st));
        MIA_HIP(hipStreamSynchronize(st));
    }
    const int npseg = cnt2[0], nsseg = cnt2[1];
    int32_t *pseg = ctx.arena.get<int32_t>("lc/pseg/start", npseg + 1);
    uint64_t *pseg_key = ctx.arena.get<uint64_t>("lc/pseg/key", npseg + 1);
    int32_t *sseg = ctx.arena.get<int32_t>("lc/sseg/start", nsseg + 1);
    uint64_t *sseg_key = ctx.arena.get<uint64_t>("lc/sseg/key", nsseg + 1);
    hipLaunchKernelGGL(k_lc_seg_emit, dim3(nblk_(np_, 256)), dim3(256), 0, st, pkey, pf, pid, np_, pseg,
                       pseg_key);
    MIA_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_lc_seg_emit, dim3(nblk_(ns_, 256)), dim3(256), 0, st, skey, sf, sid, ns_, sseg,
                       sseg_key);
    MIA_HIP(hipGetLastError());
    // the end of the last segment of each sample (a fill kernel: no host buffer, no sync)
    fill<int32_t>(pseg + npseg, (int32_t)np_, 1, st);
    fill<int32_t>(sseg + nsseg, (int32_t)ns_, 1, st);
    int32_t *cfs = ctx.arena.get<int32_t>("lc/cfs", ncells + 1);
    hipLaunchKernelGGL(k_lc_cell_first, dim3(nblk_(ncells + 1, 256)), dim3(256), 0, st, sseg_key, nsseg,
                       ncells, P1, cfs);
    MIA_HIP(hipGetLastError());
    MIA_HIP(hipEventRecord(ev[1], st));

    // ---- 3. plan: position blocks
    int32_t *nbk = ctx.arena.get<int32_t>("lc/nbk", npseg + 1);
    int32_t *nbo = ctx.arena.get<int32_t>("lc/nbo", npseg + 1);
    hipLaunchKernelGGL(k_lc_blocks_count, dim3(nblk_(npseg, 256)), dim3(256), 0, st, pseg, npseg, bsz, nbk);
    MIA_HIP(hipGetLastError());
    exclusive_scan<int32_t>(ctx, nbk, nbo, npseg, st, "lc/nbs", 0);
    int32_t nblocks = 0;
    MIA_HIP(hipMemcpyAsync(&nblocks, nbo + npseg, 4, hipMemcpyDeviceToHost, st));
    MIA_HIP(hipStreamSynchronize(st));
    LcBlock *blocks = ctx.arena.get<LcBlock>("lc/blocks", std::max(nblocks, 1));
    hipLaunchKernelGGL(k_lc_blocks_emit, dim3(nblk_(npseg, 256)), dim3(256), 0, st, pseg, pseg_key, npseg,
                       bsz, nbo, prec, P1, g.nslab, blocks);
    MIA_HIP(hipGetLastError());
    LcRect *rect = ctx.arena.get<LcRect>("lc/rect", std::max(nblocks, 1));
    int64_t *ssz = ctx.arena.get<int64_t>("lc/ssz", nblocks + 1);
    int64_t *soff = ctx.arena.get<int64_t>("lc/soff", nblocks + 1);
    hipLaunchKernelGGL(k_lc_stencil, dim3(nblk_(nblocks, 256)), dim3(256), 0, st, blocks, nblocks, g, rect,
                       ssz);
    MIA_HIP(hipGetLastError());
    exclusive_scan<int64_t>(ctx, ssz, soff, nblocks, st, "lc/sts", 0);
    int64_t total = 0;
    MIA_HIP(hipMemcpyAsync(&total, soff + nblocks, 8, hipMemcpyDeviceToHost, st));
    MIA_HIP(hipStreamSynchronize(st));
    // positions at comoving distances near zero reach the whole sky; too many
    // of them would need more (block, cell) pairs than one planning pass holds
    if (total > (int64_t(1) << 31) - 2)
        throw Error{MIA_E_UNSUPPORTED,
                    "angular reach of the positions covers too many cells (comoving distances near zero?)"};

    // ---- plan: ranges (count, scan, emit)
    LcPlan L{};
    L.g = g;
    L.blocks = blocks;
    L.rect = rect;
    L.soff = soff;
    L.nblocks = nblocks;
    L.total = total;
    L.cfs = cfs;
    L.sseg = sseg;
    L.sseg_key = sseg_key;
    L.scq = scq;
    L.sq = sq;
    L.item_shapes = item_shapes;
    int32_t *rc = ctx.arena.get<int32_t>("lc/rc", total + 1);
    int64_t *rc64 = ctx.arena.get<int64_t>("lc/rc64", total + 1);
    int64_t *roff = ctx.arena.get<int64_t>("lc/roff", total + 1);
    L.rcount = rc;
    L.emit = 0;
    if (total > 0) {
        hipLaunchKernelGGL(k_lc_ranges, dim3(nblk_(total, 256)), dim3(256), 0, st, L);
        MIA_HIP(hipGetLastError());
    }
    // widen the counts to 64 bits for the scan
    hipLaunchKernelGGL(k_lc_widen, dim3(nblk_(total, 256)), dim3(256), 0, st, rc, rc64, total);
    MIA_HIP(hipGetLastError());
    exclusive_scan<int64_t>(ctx, rc64, roff, total, st, "lc/rs", 0);
    int64_t nranges = 0;
    MIA_HIP(hipMemcpyAsync(&nranges, roff + total, 8, hipMemcpyDeviceToHost, st));
    MIA_HIP(hipStreamSynchronize(st));
    MIA_REQUIRE(nranges < (int64_t(1) << 31), "too many shape ranges (coarser grid needed)");
    LcRange *ranges = ctx.arena.get<LcRange>("lc/ranges", std::max<int64_t>(nranges, 1));
    L.emit = 1;
    L.roff = roff;
    L.ranges = ranges;
    if (total > 0 && nranges > 0) {
        hipLaunchKernelGGL(k_lc_ranges, dim3(nblk_(total, 256)), dim3(256), 0, st, L);
        MIA_HIP(hipGetLastError());
    }
    if (P > 0 && nranges > 1) {
        // A block's ranges come in stencil-cell order, so a stencil over a
        // patch border alternates patches row by row and the items (one
        // patch pair each) break at every change.  A stable sort by (block,
        // patch) keeps each block's ranges in place and groups its patches:
        // lightcone1M went from 6.2 items per block to about one per patch.
        uint64_t *rk = ctx.arena.get<uint64_t>("lc/rk", nranges);
        uint64_t *rk2 = ctx.arena.get<uint64_t>("lc/rk2", nranges);
        int32_t *ri = ctx.arena.get<int32_t>("lc/ri", nranges);
        int32_t *ri2 = ctx.arena.get<int32_t>("lc/ri2", nranges);
        LcRange *ranges2 = ctx.arena.get<LcRange>("lc/ranges2", nranges);
        hipLaunchKernelGGL(k_lc_range_keys, dim3(nblk_(nranges, 256)), dim3(256), 0, st,
                           (const LcRange *)ranges, nranges, P1 + 1, rk, ri);
        MIA_HIP(hipGetLastError());
        radix_sort<uint64_t>(ctx, rk, rk2, ri, ri2, nranges, bits_for((uint64_t)nblocks * (P1 + 1)), st,
                             "lc/rsort");
        hipLaunchKernelGGL(k_lc_gather_ranges, dim3(nblk_(nranges, 256)), dim3(256), 0, st,
                           (const int32_t *)ri2, nranges, (const LcRange *)ranges, ranges2);
        MIA_HIP(hipGetLastError());
        ranges = ranges2;
    }

    // ---- plan: items (count, scan, emit), sorted by patch pair
    int32_t *nit = ctx.arena.get<int32_t>("lc/nit", nblocks + 1);
    int32_t *ioff = ctx.arena.get<int32_t>("lc/ioff", nblocks + 1);
    hipLaunchKernelGGL(k_lc_items, dim3(nblk_(nblocks, 256)), dim3(256), 0, st, ranges, roff, soff,
                       nblocks, item_shapes, nit, (const int32_t *)nullptr, (LcItem *)nullptr,
                       (const LcBlock *)nullptr, P1, (int32_t *)nullptr, (int64_t *)nullptr,
                       (unsigned long long *)nullptr);
    MIA_HIP(hipGetLastError());
    exclusive_scan<int32_t>(ctx, nit, ioff, nblocks, st, "lc/its", 0);
    int32_t nitems = 0;
    MIA_HIP(hipMemcpyAsync(&nitems, ioff + nblocks, 4, hipMemcpyDeviceToHost, st));
    MIA_HIP(hipStreamSynchronize(st));
    LcItem *items0 = ctx.arena.get<LcItem>("lc/items0", std::max(nitems, 1));
    LcItem *items = ctx.arena.get<LcItem>("lc/items", std::max(nitems, 1));
    int32_t *ikey = ctx.arena.get<int32_t>("lc/ikey", std::max(nitems, 1));
    int32_t *ikey2 = ctx.arena.get<int32_t>("lc/ikey2", std::max(nitems, 1));
    int32_t *iidx = ctx.arena.get<int32_t>("lc/iidx", std::max(nitems, 1));
    int32_t *iord = ctx.arena.get<int32_t>("lc/iord", std::max(nitems, 1));
    int64_t *cost0 = ctx.arena.get<int64_t>("lc/cost0", nitems + 1);
    int64_t *cost = ctx.arena.get<int64_t>("lc/cost", nitems + 1);
    int64_t *cpre = ctx.arena.get<int64_t>("lc/cpre", nitems + 1);
    unsigned long long *counters = ctx.arena.get<unsigned long long>("lc/counters", 2);
    fill<unsigned long long>(counters, 0ull, 2, st);
    // one tail-chunk counter per r-bin pass (n_r <= kMaxBinsR)
    int32_t *queue = ctx.arena.get<int32_t>("lc/queue", kMaxBinsR);
    fill<int32_t>(queue, 0, kMaxBinsR, st);
    hipLaunchKernelGGL(k_lc_items, dim3(nblk_(nblocks, 256)), dim3(256), 0, st, ranges, roff, soff,
                       nblocks, item_shapes, nit, ioff, items0, blocks, P1, ikey, cost0, counters + 1);
    MIA_HIP(hipGetLastError());
    if (nitems > 0) {
        hipLaunchKernelGGL(k_lc_iota, dim3(nblk_(nitems, 256)), dim3(256), 0, st, iidx, (int64_t)nitems);
        MIA_HIP(hipGetLastError());
        radix_sort<int32_t>(ctx, ikey, ikey2, iidx, iord, nitems,
                            bits_for((uint64_t)P1 * (uint64_t)P1), st, "lc/isort");
        hipLaunchKernelGGL(k_lc_gather_items, dim3(nblk_(nitems, 256)), dim3(256), 0, st, iord,
                           (int64_t)nitems, items0, cost0, items, cost);
        MIA_HIP(hipGetLastError());
    }
    exclusive_scan<int64_t>(ctx, cost, cpre, nitems, st, "lc/cs", 0);
    MIA_HIP(hipEventRecord(ev[2], st));

    // ---- 4. pair kernel
    LcPairArgs A{};
    A.items = items;
    A.blocks = blocks;
    A.ranges = ranges;
    A.cost_prefix = cpre;
    A.total_cost = cpre + nitems;
    A.n_items = nitems;
    A.prec = prec;
    A.sra = sra;
    A.sdec = sdec;
    A.schi = schi;
    A.sc2w = sc2w;
    A.ss2w = ss2w;
    A.sw = sw;
    A.n_2 = n_2;
    A.nbins_out = B;
    A.P = P;
    double *Td = ctx.arena.get<double>("lc/T", n_r + 1);
    double *Xd = ctx.arena.get<double>("lc/X", n_2 + 1);
    const int nx2 = bins.mode == MIA_RPPI ? n_2 + 1 : n_2;
    std::vector<double> xh(n_2 + 1, 3.0);
    for (int i = 0; i < nx2; ++i) xh[i] = bins.x2_edges[i];
    MIA_HIP(hipMemcpyAsync(Td, bins.r2_edges, sizeof(double) * (n_r + 1), hipMemcpyHostToDevice, st));
    MIA_HIP(hipMemcpyAsync(Xd, xh.data(), sizeof(double) * (n_2 + 1), hipMemcpyHostToDevice, st));
    A.X = Xd;
    A.h = s.proj_scale;
    A.cut = s.rp2_cut;
    A.cut_lo = s.rp2_cut > 0 ? s.rp2_cut * (1.0 - kDomainSlack) : s.rp2_cut;
    A.cut_hi = s.rp2_cut > 0 ? s.rp2_cut * (1.0 + kDomainSlack) : s.rp2_cut;
    // FP32 guesses and guard bands, as in the box counter
    const double u = std::ldexp(1.0, -23);
    double l0 = std::log2(bins.r2_edges[0]), ln = std::log2(bins.r2_edges[n_r]);
    double lgs = n_r / (ln - l0);
    A.lgs = (float)lgs;
    // (4u n_r: a pass's constant -l0 lgs - r0 is rounded to FP32 once more)
    double err_r = (4 * u * std::fmax(std::fabs(l0), std::fabs(ln)) + 4 * u) * lgs + 4 * u * n_r;
    double err_2;
    if (bins.mode == MIA_RPPI) {
        A.X0 = bins.x2_edges[0];
        A.Xn = bins.x2_edges[n_2];
        double xs = n_2 / (bins.x2_edges[n_2] - bins.x2_edges[0]);
        A.xs = (float)xs;
        A.xc = (float)(-bins.x2_edges[0] * xs);
        double mx = std::fmax(std::fabs(bins.x2_edges[0]), std::fabs(bins.x2_edges[n_2]));
        err_2 = 4 * u * mx * xs + 2 * u * n_2;
    } else {
        A.X0 = 0;
        A.Xn = 0;
        A.xs = (float)(n_2 / 2.0);
        A.xc = 0.0f;
        err_2 = 6 * u * n_2;
    }
    double dr = std::fmax(1e-4, 20 * err_r), d2 = std::fmax(1e-4, 20 * err_2);
    bool exact_only = dr > 0.05 || d2 > 0.05;
    A.dr = exact_only ? 1.0f : (float)dr;
    A.d2 = exact_only ? 1.0f : (float)d2;
    A.scan = exact_only ? 1 : 0;
    // Passes over r-bin ranges, as in the box counter: pass [r0, r1) bins the
    // pairs with T[r0] <= r^2 < T[r1] (the exact path decides at the pass's
    // ends, so every pair lands in one pass) into output bins r0 * n_2 + local.
    // The pass size minimises passes / min(1, w / 16), w the waves per CU
    // that the pass's histograms leave room for, with at least 4 waves per
    // CU where any split allows it (ties: fewer passes).
    const size_t thr_bytes = ((size_t)(n_r + 1 + n_2 + 1) * 8 + 15) & ~(size_t)15;
    auto wpb_for = [&](int nr_) {
        int w = kLcWavesPerBlock;
        while (w > 1 && thr_bytes + lc_hist_bytes(nr_ * n_2, shapes) * w > kLcLds) w /= 2;
        return w;
    };
    auto waves_per_cu = [&](int nr_) {
        const int w = wpb_for(nr_);
        const size_t l = thr_bytes + lc_hist_bytes(nr_ * n_2, shapes) * w;
        if (l > kLcLds) return 0;
        return (int)std::min<size_t>(32 / w, kLcLds / l) * w;
    };
    int nrp = n_r;
    if (tuning && tuning->r_bins_per_pass > 0) {
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
    const int npass = (n_r + nrp - 1) / nrp;
    const int Bp = nrp * n_2;  // bins of the largest pass
    int off = ((4 * Bp) + 15) & ~15;
    A.off_dd = off;
    off += 8 * Bp;
    A.off_sp = off;
    if (shapes) off += 8 * Bp;
    A.off_sx = off;
    if (shapes) off += 8 * Bp;
    A.hist_stride = (off + 15) & ~15;
    A.npairs = (unsigned long long *)out.npairs;
    A.dd = out.dd;
    A.sp = out.splus;
    A.sx = out.scross;
    A.rechecks = counters;
    const int wpb = wpb_for(nrp);
    const size_t lds = thr_bytes + (size_t)A.hist_stride * wpb;
    MIA_REQUIRE(lds <= kLcLds, "binning too large for the lightcone kernel's LDS histogram");
    int blocks_per_cu = (int)std::max<size_t>(1, std::min<size_t>(32 / wpb, kLcLds / lds));
    int waves = tuning && tuning->waves > 0 ? std::min(tuning->waves, kMaxWaves) : ctx.cu_count * blocks_per_cu * wpb;
    if (nitems < waves) waves = std::max(1, nitems);
    A.n_waves = waves;
    unsigned grid = (unsigned)((waves + wpb - 1) / wpb);
    int64_t *bounds = ctx.arena.get<int64_t>("lc/bounds", (size_t)waves * kScheduleUnits + 3);
    if (nitems > 0) schedule_bounds(cpre, nitems, waves, bounds, st);
    A.bounds = bounds;
    for (int pass = 0; pass < npass && nitems > 0; ++pass) {
        const int r0 = pass * nrp, r1 = std::min(n_r, r0 + nrp);
        A.n_r = r1 - r0;
        A.nbins = (r1 - r0) * n_2;
        A.bin_off = r0 * n_2;
        A.T = Td + r0;
        A.T0lo = bins.r2_edges[r0] * (1.0 - kDomainSlack);
        A.Tnhi = bins.r2_edges[r1] * (1.0 + kDomainSlack);
        A.lgc = (float)(-l0 * lgs - r0);
        A.queue = queue + pass;
        if (bins.mode == MIA_RPPI) {
            if (shapes)
                lc_launch<MIA_RPPI, true>(A, grid, wpb, lds, st);
            else
                lc_launch<MIA_RPPI, false>(A, grid, wpb, lds, st);
        } else {
            if (shapes)
                lc_launch<MIA_RMU, true>(A, grid, wpb, lds, st);
            else
                lc_launch<MIA_RMU, false>(A, grid, wpb, lds, st);
        }
    }
    MIA_HIP(hipEventRecord(ev[3], st));
    MIA_HIP(hipEventRecord(ev[4], st));
    MIA_HIP(hipEventSynchronize(ev[4]));
    unsigned long long hc[2] = {0, 0};
    MIA_HIP(hipMemcpy(hc, counters, sizeof(hc), hipMemcpyDeviceToHost));
    if (stats) {
        stats->ms_prepare = ev.ms(0, 1);
        stats->ms_plan = ev.ms(1, 2);
        stats->ms_pairs = ev.ms(2, 3);
        stats->ms_finalize = ev.ms(3, 4);
        stats->candidates = (int64_t)hc[1];
        stats->items = nitems;
        stats->rechecks = (int64_t)hc[0];
        stats->columns_per_side = ncols;
        stats->slabs_per_side = nrows;
        stats->slabs_per_block = eighths;
        stats->block_shapes = bsz;
        stats->waves = waves;
        stats->method = MIA_FAST;
    }
}

}  // namespace mia
