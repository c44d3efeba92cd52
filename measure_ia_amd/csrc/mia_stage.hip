// Staged counts: the cell sort of every sample of a call in one set of
// launches, the drivers that run it with the planner and the pair kernel
// (plain launches, or a captured hipGraph replayed while the launch sequence
// stays the same), and prepared samples that keep their sort between counts.
//
// The reference rebuilds its KDTree for every counter call
// (src/measureia/measure_w_box_jk.py:386, measure_m_box_jk.py:403) and loops
// counters over samples; here a call sorts all of its samples at once:
//   classify : one launch over the positions and shapes of all samples, each
//              workgroup inside one segment (one kind of one sample): region
//              id by the reference's strict sub-box test
//              (src/measureia/measure_IA_base.py:404-452), cell key, and a rank
//              inside the cell from the key counter's atomic (so the scatter
//              needs no second atomic); per-region object counts and shape
//              weight sums, aggregated per workgroup in LDS.
//   scan     : one exclusive scan per kind over the concatenated key counters
//              of all samples, so every start is a global index.
//   scatter  : one launch: positions to 32-byte records {a, b, l, w}, shapes to 64-byte
//              records {a, b, l, ax, ay, e*w, w, 2*e*w}.
// Nothing is read back until the count has finished: sizes the device decides
// stay on the device, and every buffer is carved from one workspace whose
// size the host bounds beforehand.
#include "mia_common.hpp"

#include <hipcub/hipcub.hpp>

#include <cmath>
#include <memory>
#include <mutex>

namespace mia {

// Workspace of the allocation-free entry point (mia_count_noalloc): every
// buffer of one count is carved from this array, so that entry point never
// allocates device memory.  Each device that loads the library holds its own
// copy until the process ends.  In host code the name is only a stand-in;
// the device address comes from hipGetSymbolAddress (stage_init).
constexpr size_t kStaticBytes = (size_t)256 << 20;
static __device__ __align__(256) unsigned char scratch[kStaticBytes];

namespace {

constexpr int kStageThreads = 256;
constexpr int kStageItems = 4;                        // objects per thread
constexpr int kStageWg = kStageThreads * kStageItems; // objects per workgroup
constexpr int kLdsRegions = 4096;                     // region stats in LDS up to this many regions
constexpr int kGroupRounds = 8;                       // cell-key groups formed per wave and object
// Objects are sorted by (cell, line-of-sight 32nd of the slab).  A shape
// block's shapes then run along the line of sight, so the pairs one lane
// makes with consecutive shapes stay longer in one (r, pi) bin and its
// pending-slot runs flush less often (pair kernels 6-31 % faster on the
// benchmark boxes; eighths gave most of it); a column's positions run along
// the line of sight too, so the planner can cut a slab run to the exact
// window of a block.  Cells start at fine key (cell << sbits); sbits is
// kSubBits unless the fine keys would pass 2^31.
constexpr int kSubBits = 5;

// One segment: the positions (kind 0) or the shapes (kind 1) of one sample.
struct PrepSeg {
    const double *xyz, *w, *axis, *e;
    int64_t n;
    int64_t obj0;             // first object of the segment among the objects of its kind
    int64_t key0;             // first cell key of the sample in the concatenated key arrays
    int64_t wg0;              // first workgroup of the segment
    Grid g;
    int kind, K, sample, sbits;
    int32_t *cnt;             // key counters of the kind
    const int32_t *start;     // scanned key starts of the kind
    int2 *kr;                 // (global key, rank in cell) of every object of the kind
    unsigned long long *rcnt; // [2K] of the sample: positions, shapes per region
    double *rsum;             // [2K] of the sample: sum w_s, sum w_s (1 - e^2/2) per region
    double *prec;
    double *srec;
    int *oor;                 // the sample's out-of-box flag
};

__device__ __forceinline__ int find_seg(const int64_t *tab, int n, int64_t v) {
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

__device__ __forceinline__ int stage_axis_cell(double x, int ji, double jk_sub, int per_jk, double width, int n,
                                               int *sub = nullptr, int sbits = 0) {
    int k, k0;
    double q;
    if (ji >= 0) {
        double rel = x - (double)ji * jk_sub;
        double cs = jk_sub / per_jk;
        q = rel / cs;
        q = q < 0.0 ? 0.0 : (q > (double)per_jk ? (double)per_jk : q);
        k0 = clampi((int)floor(q), 0, per_jk - 1);
        k = k0 + ji * per_jk;
    } else {
        q = x / width;
        q = q < 0.0 ? 0.0 : (q > (double)n ? (double)n : q);
        k0 = clampi((int)floor(q), 0, n - 1);
        k = k0;
    }
    // the 2^-sbits part of the cell the coordinate lies in (an ordering hint only)
    if (sub) *sub = clampi((int)((q - (double)k0) * (1 << sbits)), 0, (1 << sbits) - 1);
    return k;
}

// Region id, cell key and rank of every object of every segment.  LDS holds
// the workgroup's per-region counts (and shape weight sums) when the
// sample's regions fit (lds_regions > 0); they reach the sample's totals with
// one atomic per touched region.
__global__ __launch_bounds__(kStageThreads) void k_stage_classify(const PrepSeg *segs, const int64_t *wg0,
                                                                  int nseg, int lds_regions) {
    extern __shared__ double lstat[];  // [2K] sums, then [K] counts (u32)
    // the segment is the workgroup's: a uniform index, so its fields are read
    // once into scalar registers instead of by vector loads at every use
    const int j = __builtin_amdgcn_readfirstlane(find_seg(wg0, nseg, blockIdx.x));
    const PrepSeg G = segs[j];
    const Grid &g = G.g;
    const int K = G.K;
    const bool lds = K > 0 && K <= lds_regions;
    unsigned int *lcnt = (unsigned int *)(lstat + 2 * (lds ? K : 0));
    if (lds) {
        for (int r = threadIdx.x; r < K; r += blockDim.x) {
            lcnt[r] = 0;
            lstat[r] = 0.0;
            lstat[K + r] = 0.0;
        }
        __syncthreads();
    }
    bool bad = false;
    const int64_t base = (int64_t)(blockIdx.x - G.wg0) * kStageWg + threadIdx.x;
    const int lane = threadIdx.x & (kWave - 1);
    // 1: every item's coordinates -> region and cell key (all loads in flight
    // together; a lane past the segment's end reads object 0 and counts nothing)
    bool valid[kStageItems];
    int key[kStageItems], region[kStageItems];
    double xs[kStageItems][3], wq[kStageItems][2];
    const bool stats = K > 0 && G.kind == 1;
#pragma unroll
    for (int it = 0; it < kStageItems; ++it) {
        const int64_t i0 = base + (int64_t)it * kStageThreads;
        valid[it] = i0 < G.n;
        const int64_t i = valid[it] ? i0 : 0;
#pragma unroll
        for (int c = 0; c < 3; ++c) xs[it][c] = G.xyz[3 * i + c];
        wq[it][0] = stats && G.w ? G.w[i] : 1.0;
        wq[it][1] = stats ? G.e[i] : 0.0;
    }
#pragma unroll
    for (int it = 0; it < kStageItems; ++it) {
        double *x = xs[it];
        bool ok = true;
#pragma unroll
        for (int c = 0; c < 3; ++c) ok = ok && (x[c] >= 0.0) && (x[c] <= g.L);
        if (!ok) {
            bad = bad || valid[it];
#pragma unroll
            for (int c = 0; c < 3; ++c) x[c] = 0.0;
        }
        int ji[3] = {-1, -1, -1};
        bool anom = false;
        region[it] = 0;
        if (g.jk > 0) {
#pragma unroll
            for (int c = 0; c < 3; ++c) ji[c] = strict_axis_index(x[c], g.jk_sub, g.jk);
            anom = ji[0] < 0 || ji[1] < 0 || ji[2] < 0;
            if (!anom) region[it] = ji[0] * g.jk * g.jk + ji[1] * g.jk + ji[2];
        }
        const int ca = stage_axis_cell(x[g.ax_a], anom ? -1 : ji[g.ax_a], g.jk_sub, g.ncj, g.cw, g.nc);
        const int cb = stage_axis_cell(x[g.ax_b], anom ? -1 : ji[g.ax_b], g.jk_sub, g.ncj, g.cw, g.nc);
        int sub = 0;
        const int cl = stage_axis_cell(x[g.los], anom ? -1 : ji[g.los], g.jk_sub, g.nzj, g.hz, g.nz, &sub, G.sbits);
        const int64_t cell = G.key0 + (int64_t)(ca * g.nc + cb) * (g.nz + 1) + (anom ? g.nz : cl);
        // fine key (cell << sbits) + part of the slab (0 in the tail slot)
        key[it] = (int)((cell << G.sbits) + (anom ? 0 : sub));
        // per-region sums of shapes: w and w (1 - e^2/2)
        const double ei = wq[it][1];
        wq[it][1] = wq[it][0] * (1 - ei * ei / 2.0);
    }
    // 2: the key counters' atomics of all items, issued back to back.  Lanes
    // of the wave that share a cell key (neighbouring objects of the input
    // often do) take their ranks from one atomic of a leader lane: up to
    // kGroupRounds groups per item, the rest one atomic each.
    int lead[kStageItems], sub[kStageItems], inc[kStageItems], got[kStageItems];
#pragma unroll
    for (int it = 0; it < kStageItems; ++it) {
        lead[it] = lane;
        sub[it] = 0;
        inc[it] = valid[it] ? 1 : 0;
        uint64_t rem = __ballot(valid[it]);
        for (int round = 0; round < kGroupRounds && rem; ++round) {
            const int l = __builtin_ffsll((long long)rem) - 1;
            const int k0 = __builtin_amdgcn_readlane(key[it], l);
            const uint64_t m = __ballot(key[it] == k0) & rem;
            if ((m >> lane) & 1) {
                lead[it] = l;
                sub[it] = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
                inc[it] = lane == l ? __popcll(m) : 0;
            }
            rem &= ~m;
        }
        got[it] = 0;
        if (inc[it] > 0) got[it] = atomicAdd(&G.cnt[key[it]], inc[it]);
    }
    // 3: ranks, (key, rank) records and the per-region statistics
#pragma unroll
    for (int it = 0; it < kStageItems; ++it) {
        const int rank = __shfl(got[it], lead[it]) + sub[it];
        if (!valid[it]) continue;
        G.kr[G.obj0 + base + (int64_t)it * kStageThreads] = make_int2(key[it], rank);
        if (K > 0) {
            // a group shares its cell, hence its region: its leader counts it
            const int rg = region[it];
            if (inc[it] > 0) {
                if (lds)
                    atomicAdd(&lcnt[rg], (unsigned int)inc[it]);
                else
                    atomicAdd(&G.rcnt[G.kind * K + rg], (unsigned long long)inc[it]);
            }
            if (G.kind == 1) {
                if (lds) {
                    atomicAdd(&lstat[rg], wq[it][0]);
                    atomicAdd(&lstat[K + rg], wq[it][1]);
                } else {
                    atomicAdd(&G.rsum[rg], wq[it][0]);
                    atomicAdd(&G.rsum[K + rg], wq[it][1]);
                }
            }
        }
    }
    if (bad) atomicOr(G.oor, 1);
    if (lds) {
        __syncthreads();
        for (int r = threadIdx.x; r < K; r += blockDim.x) {
            const unsigned int c = lcnt[r];
            if (!c) continue;
            atomicAdd(&G.rcnt[G.kind * K + r], (unsigned long long)c);
            if (G.kind == 1) {
                atomicAdd(&G.rsum[r], lstat[r]);
                atomicAdd(&G.rsum[K + r], lstat[K + r]);
            }
        }
    }
}

// Counting-sort scatter to start[key] + rank: all of a thread's loads are
// issued before its stores (the records may alias the inputs as far as the
// compiler knows, which would otherwise serialise the items).
__global__ __launch_bounds__(kStageThreads) void k_stage_scatter(const PrepSeg *segs, const int64_t *wg0,
                                                                 int nseg) {
    // the segment is the workgroup's: a uniform index, so its fields are read
    // once into scalar registers instead of by vector loads at every use
    const int j = __builtin_amdgcn_readfirstlane(find_seg(wg0, nseg, blockIdx.x));
    const PrepSeg G = segs[j];
    const Grid &g = G.g;
    const int64_t base = (int64_t)(blockIdx.x - G.wg0) * kStageWg + threadIdx.x;
    bool valid[kStageItems];
    int64_t idx[kStageItems], d[kStageItems];
#pragma unroll
    for (int it = 0; it < kStageItems; ++it) {
        const int64_t i0 = base + (int64_t)it * kStageThreads;
        valid[it] = i0 < G.n;
        idx[it] = valid[it] ? i0 : 0;
        const int2 kr = G.kr[G.obj0 + idx[it]];
        d[it] = (int64_t)G.start[kr.x] + kr.y;
    }
    double a[kStageItems], b[kStageItems], l[kStageItems], wi[kStageItems];
#pragma unroll
    for (int it = 0; it < kStageItems; ++it) {
        const int64_t i = idx[it];
        a[it] = G.xyz[3 * i + g.ax_a];
        b[it] = G.xyz[3 * i + g.ax_b];
        l[it] = G.xyz[3 * i + g.los];
        wi[it] = G.w ? G.w[i] : 1.0;
    }
    if (G.kind == 0) {
#pragma unroll
        for (int it = 0; it < kStageItems; ++it) {
            if (!valid[it]) continue;
            double2 *r = (double2 *)(G.prec + 4 * d[it]);
            r[0] = make_double2(a[it], b[it]);
            r[1] = make_double2(l[it], wi[it]);
        }
        return;
    }
    double ax[kStageItems], ay[kStageItems], ei[kStageItems];
#pragma unroll
    for (int it = 0; it < kStageItems; ++it) {
        const int64_t i = idx[it];
        ax[it] = G.axis[2 * i];
        ay[it] = G.axis[2 * i + 1];
        ei[it] = G.e[i];
    }
#pragma unroll
    for (int it = 0; it < kStageItems; ++it) {
        if (!valid[it]) continue;
        double qx = ax[it], qy = ay[it], qe = ei[it];
        // a shape whose axis or e is NaN (e.g. a zero axis vector normalised on
        // the host) adds e+ = ex = 0, as the reference's isnan rule does
        // (measure_w_box_jk.py:174-175): zero its axis and e in the record
        if (!(qx == qx) || !(qy == qy) || !(qe == qe)) {
            qx = 0.0;
            qy = 0.0;
            qe = 0.0;
        }
        const double ew = qe * wi[it];
        double2 *r = (double2 *)(G.srec + 8 * d[it]);
        r[0] = make_double2(a[it], b[it]);
        r[1] = make_double2(l[it], qx);
        r[2] = make_double2(qy, ew);
        r[3] = make_double2(wi[it], 2.0 * ew);
    }
}

// Deterministic ranks (mia_tuning.deterministic): the rank the classify
// atomic gave depends on the order in which lanes reached the counter; a
// stable radix sort of (cell key, object index) orders every cell by input
// index instead, and the rank becomes the object's place in that order.
__global__ void k_stage_keys(const int2 *kr, int64_t n, int32_t *keys, int32_t *idx) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    keys[i] = kr[i].x;
    idx[i] = (int32_t)i;
}

__global__ void k_stage_rank(const int32_t *skeys, const int32_t *sidx, int64_t n, const int32_t *start,
                             int2 *kr) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    kr[sidx[q]].y = (int)(q - start[skeys[q]]);
}

// Deterministic per-region shape sums: classify adds them with FP64 atomics
// in arrival order; here each cell's shapes are summed in their sorted (input)
// order, then each region's cells in a fixed order.  Cells lie inside one
// region (the grid is aligned to the sub-boxes); the tail slot of every
// column holds region-0 objects.
__global__ void k_stage_cellsum(const PrepSeg *segs, int S, int64_t keys, const int32_t *sstart,
                                const int32_t *sidx, double *cellsum) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= keys) return;
    int lo = 0, hi = S - 1;  // sample of global key k: shape segments at 2s + 1
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (segs[2 * mid + 1].key0 <= k)
            lo = mid;
        else
            hi = mid - 1;
    }
    const PrepSeg &G = segs[2 * lo + 1];
    double w = 0.0, q = 0.0;
    if (G.K > 0) {
        for (int j = sstart[k]; j < sstart[k + 1]; ++j) {
            const int64_t i = sidx[j] - G.obj0;
            const double wi = G.w ? G.w[i] : 1.0;
            const double ei = G.e[i];
            w += wi;
            q += wi * (1 - ei * ei / 2.0);
        }
    }
    cellsum[2 * k] = w;
    cellsum[2 * k + 1] = q;
}

__global__ void k_stage_regionsum(const PrepSeg *segs, const double *cellsum) {
    const PrepSeg &G = segs[2 * blockIdx.y + 1];
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= G.K) return;
    const Grid &g = G.g;
    const int jk = g.jk;
    int j[3] = {r / (jk * jk), (r / jk) % jk, r % jk};
    const int ja = j[g.ax_a], jb = j[g.ax_b], jl = j[g.los];
    double w = 0.0, q = 0.0;
    for (int ca = ja * g.ncj; ca < (ja + 1) * g.ncj; ++ca)
        for (int cb = jb * g.ncj; cb < (jb + 1) * g.ncj; ++cb) {
            const int64_t k0 = G.key0 + (int64_t)(ca * g.nc + cb) * (g.nz + 1);
            for (int cl = jl * g.nzj; cl < (jl + 1) * g.nzj; ++cl) {
                w += cellsum[2 * (k0 + cl)];
                q += cellsum[2 * (k0 + cl) + 1];
            }
        }
    if (r == 0) {
        // the objects the strict test leaves at region 0 (faces, 0 and L)
        for (int64_t c = 0; c < (int64_t)g.nc * g.nc; ++c) {
            const int64_t k = G.key0 + c * (g.nz + 1) + g.nz;
            w += cellsum[2 * k];
            q += cellsum[2 * k + 1];
        }
    }
    G.rsum[r] = w;
    G.rsum[G.K + r] = q;
}

// Cell starts of the shapes from the starts of their fine keys.
__global__ void k_coarse_starts(const int32_t *fine, int64_t keys, int sbits, int32_t *coarse) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k <= keys) coarse[k] = fine[k << sbits];
}

inline int key_bits(int64_t keys) {
    int b = 1;
    while (b < 31 && ((int64_t)1 << b) < keys) ++b;
    return b;
}

inline unsigned nwg(int64_t n) { return (unsigned)((n + kStageWg - 1) / kStageWg); }

// Host plan of the staged prepare of S samples.
struct PrepPlan {
    int S = 0;
    std::vector<Grid> grids;
    std::vector<PrepSeg> segs;      // 2S: positions of s at 2s, shapes at 2s + 1
    std::vector<int64_t> wg0;       // 2S + 1
    std::vector<int64_t> key0, rs0; // per sample: first key, first region-stat word
    int64_t keys = 0, npos = 0, nshp = 0, rstat = 0;
    int lds_regions = 0;
    // bound
    size_t off_segs = 0, off_wg0 = 0, param_bytes = 0;
    char *param_dev = nullptr;
    std::vector<char> param;
    int32_t *pcnt = nullptr, *scnt = nullptr, *pstart = nullptr, *sstart = nullptr;
    int32_t *pfine = nullptr;  // positions: starts of the fine keys
    int32_t *sfine = nullptr;  // shapes: starts of the fine keys (cell << sbits) + part of the slab
    int sbits = kSubBits;
    ScanBufs<int32_t> scan_p, scan_s;
    int2 *pkr = nullptr, *skr = nullptr;
    double *prec = nullptr, *srec = nullptr;
    unsigned long long *rcnt = nullptr;
    double *rsum = nullptr;
    int *oor = nullptr;
    unsigned long long *zero_begin = nullptr;
    size_t zero_words = 0;
    // deterministic ranks: keys / indices before and after a stable radix sort
    bool det = false;
    int32_t *dkey = nullptr, *didx = nullptr, *dkey_s = nullptr, *didx_s = nullptr;
    double *dcell = nullptr;  // per cell key: shape weight sums (deterministic region sums)
    int kmax = 0;
    void *dtmp = nullptr;
    size_t dtmp_bytes = 0;
    std::string key;
};

void prep_plan(PrepPlan &p, int S, const mia_sample *samples, const mia_bins &bins, const mia_tuning *tuning) {
    p.S = S;
    p.grids.resize(S);
    p.segs.assign(2 * S, PrepSeg{});
    p.wg0.assign(2 * S + 1, 0);
    p.key0.assign(S + 1, 0);
    p.rs0.assign(S + 1, 0);
    int kmax = 0;
    for (int s = 0; s < S; ++s) {
        const mia_sample &m = samples[s];
        Grid &g = p.grids[s] = make_grid(m, bins, tuning);
        const int K = g.jk > 0 ? g.jk * g.jk * g.jk : 0;
        kmax = std::max(kmax, K);
        p.key0[s + 1] = p.key0[s] + g.keys();
        p.rs0[s + 1] = p.rs0[s] + 2 * (int64_t)K;
        for (int kind = 0; kind < 2; ++kind) {
            PrepSeg &G = p.segs[2 * s + kind];
            std::memset(&G, 0, sizeof(G));
            G.n = kind ? m.n_shape : m.n_pos;
            G.obj0 = kind ? p.nshp : p.npos;
            G.key0 = p.key0[s];
            G.g = g;
            G.kind = kind;
            G.K = K;
            G.sample = s;
            const int j = 2 * s + kind;
            G.wg0 = p.wg0[j];
            p.wg0[j + 1] = p.wg0[j] + nwg(G.n);
        }
        p.npos += m.n_pos;
        p.nshp += m.n_shape;
    }
    MIA_REQUIRE(p.npos < (1ll << 31) && p.nshp < (1ll << 31) && p.key0[S] < (1ll << 31),
                "samples of one call hold more than 2^31 objects or cell keys");
    p.sbits = kSubBits;
    while (p.sbits > 0 && (p.key0[S] << p.sbits) >= (1ll << 31)) --p.sbits;
    for (PrepSeg &G : p.segs) G.sbits = p.sbits;
    p.keys = p.key0[S];
    p.rstat = p.rs0[S];
    p.lds_regions = kmax <= kLdsRegions ? kmax : 0;
    p.kmax = kmax;
    p.det = tuning && tuning->deterministic != 0;
    // the launch sequence depends on the object counts, grids and region counts
    auto put = [&](const void *v, size_t n) { p.key.append((const char *)v, n); };
    put(&S, sizeof(S));
    put(&p.det, sizeof(p.det));
    for (const PrepSeg &G : p.segs) {
        put(&G.n, sizeof(G.n));
        put(&G.g, sizeof(G.g));
        put(&G.K, sizeof(G.K));
    }
}

// Carves the prepare's buffers; with samples (not a dry run), binds the
// segments to them and to the callers' inputs and returns the sorted views.
std::vector<SortedSample> prep_bind(PrepPlan &p, Carve &c, const mia_sample *samples) {
    const int S = p.S;
    Carve pc;
    p.off_segs = pc.off; pc.take<PrepSeg>(2 * S);
    p.off_wg0 = pc.off; pc.take<int64_t>(2 * S + 1);
    p.param_bytes = pc.off;
    p.param_dev = c.take<char>(p.param_bytes);
    char *z0 = c.dry ? nullptr : c.base + c.off;
    p.pcnt = c.take<int32_t>((p.keys << p.sbits) + 1);
    p.scnt = c.take<int32_t>((p.keys << p.sbits) + 1);
    p.rcnt = c.take<unsigned long long>(p.rstat);
    p.rsum = c.take<double>(p.rstat);
    p.oor = c.take<int>(S);
    p.zero_begin = (unsigned long long *)z0;
    p.zero_words = c.dry ? 0 : (size_t)((c.base + c.off) - z0) / 8;
    p.pstart = c.take<int32_t>(p.keys + 1);
    p.pfine = c.take<int32_t>((p.keys << p.sbits) + 1);
    p.sstart = c.take<int32_t>(p.keys + 1);
    p.sfine = c.take<int32_t>((p.keys << p.sbits) + 1);
    scan_carve<int32_t>(c, p.keys << p.sbits, p.scan_p);
    scan_carve<int32_t>(c, p.keys << p.sbits, p.scan_s);
    p.pkr = c.take<int2>(p.npos);
    p.skr = c.take<int2>(p.nshp);
    p.prec = c.take<double>(4 * p.npos);
    p.srec = c.take<double>(8 * p.nshp);
    if (p.det) {
        const int64_t nmax = std::max<int64_t>(std::max(p.npos, p.nshp), 1);
        p.dkey = c.take<int32_t>(nmax);
        p.didx = c.take<int32_t>(nmax);
        p.dkey_s = c.take<int32_t>(nmax);
        p.didx_s = c.take<int32_t>(nmax);
        size_t bytes = 0;
        MIA_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (const int32_t *)nullptr, (int32_t *)nullptr,
                                                   (const int32_t *)nullptr, (int32_t *)nullptr, (int)nmax, 0,
                                                   key_bits((p.keys << p.sbits) + 1), (hipStream_t)0));
        p.dtmp_bytes = std::max<size_t>(bytes, 256);
        p.dtmp = c.take<char>(p.dtmp_bytes);
        p.dcell = c.take<double>(2 * (p.keys + 1));
    }
    std::vector<SortedSample> ss(S);
    for (int s = 0; s < S; ++s) {
        SortedSample &x = ss[s];
        const Grid &g = p.grids[s];
        const mia_sample &m = samples[s];
        x.n_pos = m.n_pos;
        x.n_shape = m.n_shape;
        x.g = g;
        x.K = g.jk > 0 ? g.jk * g.jk * g.jk : 0;
        x.periodic = m.periodic;
        x.weighted = (m.pos_w != nullptr || m.shp_w != nullptr) ? 1 : 0;
        if (c.dry) continue;
        x.prec = p.prec;
        x.srec = p.srec;
        x.pstart = p.pstart + p.key0[s];
        x.pfine = p.pfine + (p.key0[s] << p.sbits);
        x.sbits = p.sbits;
        x.sstart = p.sstart + p.key0[s];
        x.rcnt = p.rcnt + p.rs0[s];
        x.rsum = p.rsum + p.rs0[s];
        for (int kind = 0; kind < 2; ++kind) {
            PrepSeg &G = p.segs[2 * s + kind];
            G.xyz = kind ? m.shp : m.pos;
            G.w = kind ? m.shp_w : m.pos_w;
            G.axis = kind ? m.axis : nullptr;
            G.e = kind ? m.e : nullptr;
            G.cnt = kind ? p.scnt : p.pcnt;
            G.start = kind ? p.sfine : p.pfine;
            G.kr = kind ? p.skr : p.pkr;
            G.rcnt = p.rcnt + p.rs0[s];
            G.rsum = p.rsum + p.rs0[s];
            G.prec = p.prec;
            G.srec = p.srec;
            G.oor = p.oor + s;
        }
    }
    if (!c.dry) {
        p.param.assign(p.param_bytes, 0);
        std::memcpy(p.param.data() + p.off_segs, p.segs.data(), sizeof(PrepSeg) * 2 * S);
        std::memcpy(p.param.data() + p.off_wg0, p.wg0.data(), sizeof(int64_t) * (2 * S + 1));
    }
    return ss;
}

size_t prep_bytes(PrepPlan &p, const mia_sample *samples) {
    Carve c;
    prep_bind(p, c, samples);
    return c.off;
}

void prep_enqueue(const PrepPlan &p, hipStream_t st) {
    const int nseg = 2 * p.S;
    const PrepSeg *segs = (const PrepSeg *)(p.param_dev + p.off_segs);
    const int64_t *wg0 = (const int64_t *)(p.param_dev + p.off_wg0);
    fill<unsigned long long>(p.zero_begin, 0ull, (int64_t)p.zero_words, st);
    const int64_t wgs = p.wg0[nseg];
    const size_t lds = (size_t)p.lds_regions * (2 * sizeof(double) + sizeof(unsigned int));
    if (wgs > 0) {
        hipLaunchKernelGGL(k_stage_classify, dim3((unsigned)wgs), dim3(kStageThreads), lds, st, segs, wg0, nseg,
                           p.lds_regions);
        MIA_HIP(hipGetLastError());
    }
    scan_ws<int32_t>(p.pcnt, p.pfine, p.keys << p.sbits, nullptr, p.scan_p, st);
    hipLaunchKernelGGL(k_coarse_starts, dim3((unsigned)((p.keys + 1 + 255) / 256)), dim3(256), 0, st,
                       (const int32_t *)p.pfine, p.keys, p.sbits, p.pstart);
    MIA_HIP(hipGetLastError());
    scan_ws<int32_t>(p.scnt, p.sfine, p.keys << p.sbits, nullptr, p.scan_s, st);
    hipLaunchKernelGGL(k_coarse_starts, dim3((unsigned)((p.keys + 1 + 255) / 256)), dim3(256), 0, st,
                       (const int32_t *)p.sfine, p.keys, p.sbits, p.sstart);
    MIA_HIP(hipGetLastError());
    if (p.det) {
        for (int kind = 0; kind < 2; ++kind) {
            const int64_t n = kind ? p.nshp : p.npos;
            if (n <= 0) continue;
            int2 *kr = kind ? p.skr : p.pkr;
            const unsigned g = (unsigned)((n + 255) / 256);
            hipLaunchKernelGGL(k_stage_keys, dim3(g), dim3(256), 0, st, (const int2 *)kr, n, p.dkey, p.didx);
            MIA_HIP(hipGetLastError());
            size_t bytes = p.dtmp_bytes;
            MIA_HIP(hipcub::DeviceRadixSort::SortPairs(p.dtmp, bytes, (const int32_t *)p.dkey, p.dkey_s,
                                                       (const int32_t *)p.didx, p.didx_s, (int)n, 0,
                                                       key_bits((p.keys << p.sbits) + 1), st));
            hipLaunchKernelGGL(k_stage_rank, dim3(g), dim3(256), 0, st, (const int32_t *)p.dkey_s,
                               (const int32_t *)p.didx_s, n, (const int32_t *)(kind ? p.sfine : p.pfine), kr);
            MIA_HIP(hipGetLastError());
            if (kind == 1 && p.kmax > 0) {
                hipLaunchKernelGGL(k_stage_cellsum, dim3((unsigned)((p.keys + 255) / 256)), dim3(256), 0, st, segs,
                                   p.S, p.keys, (const int32_t *)p.sstart, (const int32_t *)p.didx_s, p.dcell);
                MIA_HIP(hipGetLastError());
                hipLaunchKernelGGL(k_stage_regionsum, dim3((unsigned)((p.kmax + 63) / 64), (unsigned)p.S), dim3(64),
                                   0, st, segs, (const double *)p.dcell);
                MIA_HIP(hipGetLastError());
            }
        }
    }
    if (wgs > 0) {
        hipLaunchKernelGGL(k_stage_scatter, dim3((unsigned)wgs), dim3(kStageThreads), 0, st, segs, wg0, nseg);
        MIA_HIP(hipGetLastError());
    }
}

// Pinned host staging: [0, a) parameter images, then the read-back words.
void *pinned(Ctx &ctx, size_t bytes, bool may_grow) {
    if (ctx.pinned_bytes >= bytes) return ctx.pinned;
    if (!may_grow) throw Error{MIA_E_UNSUPPORTED, "parameter block larger than the preallocated staging"};
    if (ctx.pinned) (void)hipHostFree(ctx.pinned);
    ctx.pinned = nullptr;
    ctx.pinned_bytes = 0;
    size_t b = std::max<size_t>(bytes, (size_t)1 << 20);
    if (hipHostMalloc(&ctx.pinned, b, hipHostMallocDefault) != hipSuccess) {
        ctx.pinned = nullptr;
        throw Error{MIA_E_NOMEM, "hipHostMalloc(staging)"};
    }
    ctx.pinned_bytes = b;
    return ctx.pinned;
}

void *workspace(Ctx &ctx, size_t bytes) {
    if (ctx.ws_bytes >= bytes) return ctx.ws;
    // the graph refers to the old workspace
    if (ctx.gexec) (void)hipGraphExecDestroy(ctx.gexec);
    if (ctx.graph) (void)hipGraphDestroy(ctx.graph);
    ctx.gexec = nullptr;
    ctx.graph = nullptr;
    ctx.gkey.clear();
    if (ctx.ws) (void)hipFree(ctx.ws);
    ctx.ws = nullptr;
    ctx.ws_bytes = 0;
    size_t b = bytes + bytes / 8 + ((size_t)1 << 20);
    if (hipMalloc(&ctx.ws, b) != hipSuccess) {
        ctx.ws = nullptr;
        throw Error{MIA_E_NOMEM, "hipMalloc(workspace of " + std::to_string(b) + " bytes)"};
    }
    ctx.ws_bytes = b;
    return ctx.ws;
}

struct PlanDeleter {
    void operator()(CountPlan *p) const { free_plan(p); }
};
using PlanPtr = std::unique_ptr<CountPlan, PlanDeleter>;

// Uploads the parameter images, runs (or replays) the sequence, reads back
// the words and flags.  Returns the read-back words; oor[S] set per sample.
struct RunOut {
    unsigned long long words[kCountReadback];
    std::vector<int> oor;
    float ms_prepare = 0, ms_plan = 0, ms_pairs = 0, ms_finalize = 0, ms_total = 0;
    int graph = 0;
};

void upload_and_run(Ctx &ctx, const PrepPlan *prep, const CountPlan &cp, hipStream_t st, StageMode mode,
                    RunOut &r) {
    const size_t pb = prep ? (prep->param_bytes + 255) & ~(size_t)255 : 0;
    const size_t cb = (count_param_bytes(cp) + 255) & ~(size_t)255;
    const int S = prep ? prep->S : 0;
    const size_t rb = sizeof(unsigned long long) * kCountReadback + sizeof(int) * (S + 1);
    char *h = (char *)pinned(ctx, pb + cb + rb, mode != STAGE_STATIC);
    if (prep) std::memcpy(h, prep->param.data(), prep->param_bytes);
    std::memcpy(h + pb, count_param_image(cp), count_param_bytes(cp));
    if (mode == STAGE_STATIC) {
        // no events, streams or graphs: nothing is created or allocated
        if (prep) MIA_HIP(hipMemcpyAsync(prep->param_dev, h, prep->param_bytes, hipMemcpyHostToDevice, st));
        MIA_HIP(hipMemcpyAsync(count_param_dev(cp), h + pb, count_param_bytes(cp), hipMemcpyHostToDevice, st));
        if (prep) prep_enqueue(*prep, st);
        count_enqueue(cp, st, nullptr);
    }
    Events ev(mode == STAGE_STATIC ? 0 : 6);
    if (mode != STAGE_STATIC) MIA_HIP(hipEventRecord(ev[0], st));
    if (mode != STAGE_STATIC) {
        if (prep) MIA_HIP(hipMemcpyAsync(prep->param_dev, h, prep->param_bytes, hipMemcpyHostToDevice, st));
        MIA_HIP(hipMemcpyAsync(count_param_dev(cp), h + pb, count_param_bytes(cp), hipMemcpyHostToDevice, st));
    }
    if (mode == STAGE_STATIC) {
        // enqueued above
    } else if (mode == STAGE_GRAPH) {
        std::string key = (prep ? prep->key : std::string()) + count_key(cp);
        key.append((const char *)&ctx.ws, sizeof(ctx.ws));
        if (key != ctx.gkey || !ctx.gexec) {
            if (ctx.gexec) (void)hipGraphExecDestroy(ctx.gexec);
            if (ctx.graph) (void)hipGraphDestroy(ctx.graph);
            ctx.gexec = nullptr;
            ctx.graph = nullptr;
            ctx.gkey.clear();
            MIA_HIP(hipStreamBeginCapture(ctx.cap_stream, hipStreamCaptureModeRelaxed));
            try {
                if (prep) prep_enqueue(*prep, ctx.cap_stream);
                count_enqueue(cp, ctx.cap_stream, nullptr);
            } catch (...) {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(ctx.cap_stream, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            MIA_HIP(hipStreamEndCapture(ctx.cap_stream, &ctx.graph));
            MIA_HIP(hipGraphInstantiate(&ctx.gexec, ctx.graph, nullptr, nullptr, 0));
            ctx.gkey = key;
        } else {
            r.graph = 1;
        }
        MIA_HIP(hipEventRecord(ev[1], st));
        MIA_HIP(hipGraphLaunch(ctx.gexec, st));
        MIA_HIP(hipEventRecord(ev[5], st));
    } else {
        MIA_HIP(hipEventRecord(ev[1], st));
        if (prep) prep_enqueue(*prep, st);
        count_enqueue(cp, st, ev.data() + 2);
    }
    char *back = h + pb + cb;
    MIA_HIP(hipMemcpyAsync(back, count_readback(cp), sizeof(unsigned long long) * kCountReadback,
                           hipMemcpyDeviceToHost, st));
    if (prep && S > 0)
        MIA_HIP(hipMemcpyAsync(back + sizeof(unsigned long long) * kCountReadback, prep->oor, sizeof(int) * S,
                               hipMemcpyDeviceToHost, st));
    MIA_HIP(hipStreamSynchronize(st));
    std::memcpy(r.words, back, sizeof(r.words));
    r.oor.assign(S, 0);
    if (S > 0) std::memcpy(r.oor.data(), back + sizeof(unsigned long long) * kCountReadback, sizeof(int) * S);
    const double span_ms = r.words[5] > ~r.words[4] ? (double)(r.words[5] - ~r.words[4]) / ctx.wall_khz : 0.0;
    if (mode == STAGE_STATIC) {
        r.ms_pairs = (float)span_ms;
    } else if (mode == STAGE_GRAPH) {
        r.ms_total = ev.ms(0, 5);
        r.ms_pairs = (float)span_ms;
    } else {
        // ev[1]: before prepare, ev[2]: before plan, ev[3]: before pairs, ev[4]: after pairs, ev[5]: end
        r.ms_prepare = ev.ms(1, 2);
        r.ms_plan = ev.ms(2, 3);
        r.ms_pairs = ev.ms(3, 4);
        r.ms_finalize = ev.ms(4, 5);
        r.ms_total = ev.ms(0, 5);
    }
}

void fill_stats(mia_stats *stats, const RunOut &r, const CountPlan &cp, const Grid &g0) {
    if (!stats) return;
    std::memset(stats, 0, sizeof(*stats));
    stats->ms_prepare = r.ms_prepare;
    stats->ms_plan = r.ms_plan;
    stats->ms_pairs = r.ms_pairs;
    stats->ms_finalize = r.ms_finalize;
    stats->ms_total = r.ms_total;
    stats->rechecks = (int64_t)r.words[0];
    stats->candidates = (int64_t)r.words[1];
    stats->wave_iterations = (int64_t)r.words[6];
    stats->items = (int64_t)r.words[3];
    stats->columns_per_side = g0.nc;
    stats->slabs_per_side = g0.nz;
    stats->slabs_per_block = count_G(cp, 0);
    stats->block_shapes = count_bsz(cp);
    stats->waves = count_waves(cp);
    stats->method = MIA_FAST;
    stats->graph = r.graph;
    stats->deterministic = count_det(cp) ? 1 : 0;
    stats->item_capacity = count_item_cap(cp);
}

}  // namespace

struct PreparedSample {
    int device = 0;
    void *mem = nullptr;
    size_t bytes = 0;
    SortedSample s;
};

namespace {
std::mutex g_scratch_mu;
void *g_scratch[64] = {};     // device address of `scratch` per device
bool g_scratch_done[64] = {};
}  // namespace

void stage_init(Ctx &ctx) {
    // the static workspace's device address, once per device, with that
    // device current (mia_create has made it so)
    MIA_REQUIRE(ctx.device >= 0 && ctx.device < 64, "device index above 63");
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        if (!g_scratch_done[ctx.device]) {
            void *ptr = nullptr;
            MIA_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
            MIA_REQUIRE(ptr != nullptr, "no device address for the static workspace");
            g_scratch[ctx.device] = ptr;
            g_scratch_done[ctx.device] = true;
        }
        ctx.scratch = g_scratch[ctx.device];
    }
    (void)pinned(ctx, (size_t)1 << 20, true);
    MIA_HIP(hipStreamCreateWithFlags(&ctx.cap_stream, hipStreamNonBlocking));
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, ctx.device) == hipSuccess && khz > 0)
        ctx.wall_khz = khz;
}

void stage_release(Ctx &ctx) {
    if (ctx.gexec) (void)hipGraphExecDestroy(ctx.gexec);
    if (ctx.graph) (void)hipGraphDestroy(ctx.graph);
    if (ctx.cap_stream) (void)hipStreamDestroy(ctx.cap_stream);
    if (ctx.ws) (void)hipFree(ctx.ws);
    if (ctx.pinned) (void)hipHostFree(ctx.pinned);
    ctx.gexec = nullptr;
    ctx.graph = nullptr;
    ctx.cap_stream = nullptr;
    ctx.ws = nullptr;
    ctx.pinned = nullptr;
    ctx.ws_bytes = ctx.pinned_bytes = 0;
}

// After a count whose plan overflowed: larger item and range capacities for
// the next attempt (a quarter of headroom over what the count needed).
void grow_caps(Ctx &ctx, const RunOut &r, int attempt) {
    if (attempt >= 3)
        throw Error{MIA_E_UNSUPPORTED, "work items exceed the workspace capacity (" + std::to_string(r.words[3]) +
                                           " items, " + std::to_string(std::max(r.words[7], r.words[8])) +
                                           " range records)"};
    ctx.item_cap = std::max<int64_t>(ctx.item_cap, (int64_t)(r.words[3] + r.words[3] / 4 + 1024));
    const unsigned long long need = std::max(r.words[7], r.words[8]);
    ctx.range_cap = std::max<int64_t>(ctx.range_cap, (int64_t)(need + need / 4 + 1024));
}

void stage_count(Ctx &ctx, int n, const mia_sample *samples, const mia_bins &bins, const mia_counts *outs,
                 const mia_tuning *tuning, mia_stats *stats, hipStream_t st, StageMode mode) {
    // the deterministic sort's radix-sort launches stay out of graphs
    if (mode == STAGE_GRAPH && tuning && tuning->deterministic) mode = STAGE_EAGER;
    for (int attempt = 0;; ++attempt) {
        PrepPlan prep;
        prep_plan(prep, n, samples, bins, tuning);
        Carve dry;
        std::vector<SortedSample> proto = prep_bind(prep, dry, samples);
        const size_t pbytes = dry.off;
        PlanPtr cp(plan_count(ctx, proto, bins, tuning, ctx.item_cap, ctx.range_cap, false));
        size_t total = pbytes + count_bytes(*cp);
        char *base;
        if (mode == STAGE_STATIC) {
            MIA_REQUIRE(ctx.scratch != nullptr, "no static workspace on this device");
            const size_t room = (size_t)stage_static_bytes();
            if (total > room) {
                // fewer work-item slots (32 bytes each with their costs and
                // scan) and ranges (24 bytes: planner range and item record),
                // down to a floor; a count that needs more reports it
                const int64_t over = (int64_t)((total - room) / 56) + 4096;
                const int64_t cap = count_item_cap(*cp) - over, rcap = count_range_cap(*cp) - over;
                if (cap >= 65536 && rcap >= 65536) {
                    cp.reset(plan_count(ctx, proto, bins, tuning, cap, rcap, true));
                    total = pbytes + count_bytes(*cp);
                }
            }
            if (total > room)
                throw Error{MIA_E_UNSUPPORTED, "the count needs " + std::to_string(total) +
                                                   " bytes of workspace; the static workspace holds " +
                                                   std::to_string(room) + " (mia_count allocates what it needs)"};
            base = (char *)ctx.scratch;
        } else {
            base = (char *)workspace(ctx, total);
        }
        Carve c;
        c.dry = false;
        c.base = base;
        std::vector<SortedSample> ss = prep_bind(prep, c, samples);
        count_bind(*cp, c, ss, outs);
        RunOut r;
        upload_and_run(ctx, &prep, *cp, st, mode, r);
        for (int s = 0; s < n; ++s)
            if (r.oor[s])
                throw Error{MIA_E_RANGE, "coordinates outside [0, boxsize] (or NaN) in sample " + std::to_string(s)};
        if (r.words[2]) {
            // more work items or range records than the capacity: grow it and count again
            if (mode == STAGE_STATIC)
                throw Error{MIA_E_UNSUPPORTED, "the count needs " + std::to_string(r.words[3]) +
                                                   " work items and " +
                                                   std::to_string(std::max(r.words[7], r.words[8])) +
                                                   " range records; the static workspace holds " +
                                                   std::to_string(count_item_cap(*cp)) + " and " +
                                                   std::to_string(count_range_cap(*cp)) +
                                                   " (mia_count allocates what it needs)"};
            grow_caps(ctx, r, attempt);
            continue;
        }
        fill_stats(stats, r, *cp, prep.grids[0]);
        return;
    }
}

PreparedSample *stage_prepare(Ctx &ctx, const mia_sample &s, const mia_bins &bins, const mia_tuning *tuning,
                              hipStream_t st) {
    PrepPlan prep;
    prep_plan(prep, 1, &s, bins, tuning);
    const size_t bytes = prep_bytes(prep, &s);
    std::unique_ptr<PreparedSample> h(new PreparedSample());
    h->device = ctx.device;
    if (hipMalloc(&h->mem, bytes) != hipSuccess) {
        h->mem = nullptr;
        throw Error{MIA_E_NOMEM, "hipMalloc(prepared sample of " + std::to_string(bytes) + " bytes)"};
    }
    h->bytes = bytes;
    try {
        Carve c;
        c.dry = false;
        c.base = (char *)h->mem;
        std::vector<SortedSample> ss = prep_bind(prep, c, &s);
        char *stage = (char *)pinned(ctx, prep.param_bytes + 256, true);
        std::memcpy(stage, prep.param.data(), prep.param_bytes);
        MIA_HIP(hipMemcpyAsync(prep.param_dev, stage, prep.param_bytes, hipMemcpyHostToDevice, st));
        prep_enqueue(prep, st);
        int *flag = (int *)(stage + ((prep.param_bytes + 255) & ~(size_t)255));
        MIA_HIP(hipMemcpyAsync(flag, prep.oor, sizeof(int), hipMemcpyDeviceToHost, st));
        MIA_HIP(hipStreamSynchronize(st));
        if (*flag) throw Error{MIA_E_RANGE, "coordinates outside [0, boxsize] (or NaN)"};
        h->s = ss[0];
    } catch (...) {
        (void)hipFree(h->mem);
        throw;
    }
    return h.release();
}

void stage_count_prepared(Ctx &ctx, int n, const PreparedSample *const *hs, const mia_bins &bins,
                          const mia_counts *outs, const mia_tuning *tuning, mia_stats *stats, hipStream_t st) {
    std::vector<SortedSample> ss(n);
    for (int i = 0; i < n; ++i) {
        MIA_REQUIRE(hs[i] != nullptr, "prepared sample is NULL");
        MIA_REQUIRE(hs[i]->device == ctx.device, "prepared sample belongs to another device");
        ss[i] = hs[i]->s;
    }
    for (int attempt = 0;; ++attempt) {
        PlanPtr cp(plan_count(ctx, ss, bins, tuning, ctx.item_cap, ctx.range_cap, false));
        Carve c;
        c.dry = false;
        c.base = (char *)workspace(ctx, count_bytes(*cp));
        count_bind(*cp, c, ss, outs);
        RunOut r;
        upload_and_run(ctx, nullptr, *cp, st, STAGE_EAGER, r);
        if (r.words[2]) {
            grow_caps(ctx, r, attempt);
            continue;
        }
        fill_stats(stats, r, *cp, ss[0].g);
        return;
    }
}

const SortedSample &prepared_view(const PreparedSample *h) { return h->s; }

void stage_free_prepared(PreparedSample *h) {
    if (!h) return;
    if (h->mem) (void)hipFree(h->mem);
    delete h;
}

int64_t stage_static_bytes() { return (int64_t)kStaticBytes; }

}  // namespace mia
