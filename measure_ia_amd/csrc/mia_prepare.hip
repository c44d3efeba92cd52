// Jackknife region ids and the counting-sort cell list (HBM-bound pass).
//
// Every object gets a region id with the reference's strict sub-box test
// (src/measureia/measure_IA_base.py:404-452) and a cell key: projected-plane
// column x line-of-sight slab, with cells aligned to the jackknife sub-boxes
// so that every nominal cell lies inside one region.  Objects the strict test
// leaves unassigned (exactly on a sub-box face, at 0, or at L) keep region 0,
// as in the reference, and go to a per-column tail so that every cell range
// has one region.  A counting sort then scatters the sample into structure-
// of-arrays order in HBM.
#include "mia_common.hpp"

#include <cmath>

namespace mia {

namespace {

constexpr int kScanThreads = 256;
constexpr int kScanPerThread = 8;
constexpr int kScanTile = kScanThreads * kScanPerThread;

template <class T>
__global__ void k_fill(T *p, T v, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

// bounds[u] = first i in [0, n] with cost_prefix[i] * U >= u * C, for u in
// [0, U]; bounds[U + 1], bounds[U + 2] = static units per wave and units per
// tail chunk.
__global__ void k_schedule_bounds(const int64_t *cpre, int64_t n_host, const int64_t *n_dev, int64_t U,
                                  int waves, int all_static, int64_t *bounds) {
    const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u > U + 2) return;
    const int64_t n = n_dev ? *n_dev : n_host;
    const double C = (double)cpre[n], Ud = (double)U, t = (double)u * C;
    if (u > U) {
        const bool fine = !all_static && C >= kTailFineCost * waves;
        const bool tail = !all_static && C >= kTailMinCost * waves;
        if (u == U + 1)
            bounds[u] = fine ? 2 : (tail ? 4 : kScheduleUnits);
        else
            bounds[u] = fine ? 1 : (tail ? 2 : 1);
        return;
    }
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if ((double)cpre[mid] * Ud >= t)
            hi = mid;
        else
            lo = mid + 1;
    }
    bounds[u] = lo;
}

// Exclusive scan of one tile per block; writes the tile total to sums.
template <class T>
__global__ __launch_bounds__(kScanThreads) void k_scan_tile(const T *in, T *out, int64_t n,
                                                            T *sums) {
    __shared__ T part[kScanThreads];
    int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanPerThread;
    T v[kScanPerThread];
    T acc = 0;
#pragma unroll
    for (int k = 0; k < kScanPerThread; ++k) {
        int64_t i = base + k;
        v[k] = i < n ? in[i] : T(0);
        acc += v[k];
    }
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int off = 1; off < kScanThreads; off <<= 1) {
        T t = threadIdx.x >= off ? part[threadIdx.x - off] : T(0);
        __syncthreads();
        part[threadIdx.x] += t;
        __syncthreads();
    }
    T run = part[threadIdx.x] - acc;  // exclusive prefix of this thread
#pragma unroll
    for (int k = 0; k < kScanPerThread; ++k) {
        int64_t i = base + k;
        if (i < n) out[i] = run;
        run += v[k];
    }
    if (threadIdx.x == kScanThreads - 1) sums[blockIdx.x] = part[threadIdx.x];
}

template <class T>
__global__ void k_scan_add(T *out, int64_t n, const T *offs) {
    int64_t i = (int64_t)blockIdx.x * kScanTile + threadIdx.x;
    T o = offs[blockIdx.x];
    for (int k = 0; k < kScanPerThread; ++k, i += kScanThreads)
        if (i < n) out[i] += o;
}

template <class T>
__global__ void k_set_total(T *out, int64_t n, const T *in, const T *excl) {
    // out[n] = excl[n-1] + in[n-1]
    if (threadIdx.x == 0 && blockIdx.x == 0) out[n] = n > 0 ? excl[n - 1] + in[n - 1] : T(0);
}

// Scans whose length is read on the device: element i takes part iff
// i < n = min(*n_dev, cap) (n = cap without n_dev); tiles past n write a zero sum.
template <class T>
__global__ __launch_bounds__(kScanThreads) void k_scan_tile_n(const T *in, T *out, int64_t cap,
                                                              const int64_t *n_dev, T *sums) {
    __shared__ T part[kScanThreads];
    int64_t n = cap;
    if (n_dev) n = *n_dev < cap ? *n_dev : cap;
    int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanPerThread;
    T v[kScanPerThread];
    T acc = 0;
#pragma unroll
    for (int k = 0; k < kScanPerThread; ++k) {
        int64_t i = base + k;
        v[k] = i < n ? in[i] : T(0);
        acc += v[k];
    }
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int off = 1; off < kScanThreads; off <<= 1) {
        T t = threadIdx.x >= off ? part[threadIdx.x - off] : T(0);
        __syncthreads();
        part[threadIdx.x] += t;
        __syncthreads();
    }
    T run = part[threadIdx.x] - acc;
#pragma unroll
    for (int k = 0; k < kScanPerThread; ++k) {
        int64_t i = base + k;
        if (i < n) out[i] = run;
        run += v[k];
    }
    if (threadIdx.x == kScanThreads - 1) sums[blockIdx.x] = part[threadIdx.x];
}

template <class T>
__global__ void k_scan_add_n(T *out, int64_t cap, const int64_t *n_dev, const T *offs) {
    int64_t n = cap;
    if (n_dev) n = *n_dev < cap ? *n_dev : cap;
    int64_t i = (int64_t)blockIdx.x * kScanTile + threadIdx.x;
    T o = offs[blockIdx.x];
    for (int k = 0; k < kScanPerThread; ++k, i += kScanThreads)
        if (i < n) out[i] += o;
}

// out[n] = total: the last level's single tile sum (every level below adds up
// to it), written at the element count read on the device.
template <class T>
__global__ void k_scan_total_n(T *out, int64_t cap, const int64_t *n_dev, const T *top) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        int64_t n = cap;
        if (n_dev) n = *n_dev < cap ? *n_dev : cap;
        out[n] = top[0];
    }
}

__device__ __forceinline__ int axis_cell(double x, int ji, double jk_sub, int per_jk, double width,
                                         int n) {
    int k;
    if (ji >= 0) {
        double rel = x - (double)ji * jk_sub;
        double cs = jk_sub / per_jk;
        double q = rel / cs;
        q = q < 0.0 ? 0.0 : (q > (double)per_jk ? (double)per_jk : q);
        k = clampi((int)floor(q), 0, per_jk - 1) + ji * per_jk;
    } else {
        double q = x / width;
        q = q < 0.0 ? 0.0 : (q > (double)n ? (double)n : q);
        k = clampi((int)floor(q), 0, n - 1);
    }
    return k;
}

constexpr int kClassifyThreads = 256;

// Region id, cell key and per-region statistics of every object.  With
// `partials` the region statistics are first summed per workgroup in LDS and
// written as one row of K*4 partial sums per workgroup (reduced in a fixed
// order by k_reduce_rows); without it they go straight to global atomics.
__global__ __launch_bounds__(kClassifyThreads) void k_classify(
    const double *__restrict__ xyz, int64_t n, Grid g, const double *__restrict__ w,
    const double *__restrict__ e, int32_t *__restrict__ key, int32_t *__restrict__ reg,
    int32_t *__restrict__ counts, double *__restrict__ rstats, double *__restrict__ partials,
    int which, int *__restrict__ oor) {
    extern __shared__ double rs_l[];
    const int K4 = g.jk > 0 ? 4 * g.jk * g.jk * g.jk : 0;
    const bool stats = rstats != nullptr && g.jk > 0;
    const bool local = stats && partials != nullptr;
    if (local) {
        for (int j = threadIdx.x; j < K4; j += blockDim.x) rs_l[j] = 0.0;
        __syncthreads();
    }
    double *acc = local ? rs_l : rstats;
    bool bad = false;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        double x[3] = {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
        bool ok = true;
#pragma unroll
        for (int c = 0; c < 3; ++c) ok = ok && (x[c] >= 0.0) && (x[c] <= g.L);
        if (!ok) {
            bad = true;
#pragma unroll
            for (int c = 0; c < 3; ++c) x[c] = 0.0;
        }
        int ji[3] = {-1, -1, -1};
        bool anom = false;
        int region = 0;
        if (g.jk > 0) {
#pragma unroll
            for (int c = 0; c < 3; ++c) ji[c] = strict_axis_index(x[c], g.jk_sub, g.jk);
            anom = ji[0] < 0 || ji[1] < 0 || ji[2] < 0;
            if (!anom) region = ji[0] * g.jk * g.jk + ji[1] * g.jk + ji[2];
        }
        int ca = axis_cell(x[g.ax_a], anom ? -1 : ji[g.ax_a], g.jk_sub, g.ncj, g.cw, g.nc);
        int cb = axis_cell(x[g.ax_b], anom ? -1 : ji[g.ax_b], g.jk_sub, g.ncj, g.cw, g.nc);
        int cl = axis_cell(x[g.los], anom ? -1 : ji[g.los], g.jk_sub, g.nzj, g.hz, g.nz);
        int k = (ca * g.nc + cb) * (g.nz + 1) + (anom ? g.nz : cl);
        key[i] = k;
        reg[i] = region;
        atomicAdd(&counts[k], 1);
        if (stats) {
            atomicAdd(&acc[4 * region + which], 1.0);
            if (which == 1) {
                double wi = w ? w[i] : 1.0;
                double ei = e[i];
                atomicAdd(&acc[4 * region + 2], wi);
                atomicAdd(&acc[4 * region + 3], wi * (1 - ei * ei / 2.0));
            }
        }
    }
    if (bad) atomicOr(oor, 1);
    if (local) {
        __syncthreads();
        for (int j = threadIdx.x; j < K4; j += blockDim.x)
            partials[(int64_t)blockIdx.x * K4 + j] = rs_l[j];
    }
}

constexpr int kReduceThreads = 256;

// out[j] += sum over rows r of rows[r*m + j]: one workgroup per column j,
// thread t adds rows t, t+256, ... in order, then a fixed-shape tree in LDS
// (the same partial sums in the same order on every run).
__global__ __launch_bounds__(kReduceThreads) void k_reduce_rows(const double *__restrict__ rows,
                                                                int nrows, int m,
                                                                double *__restrict__ out) {
    __shared__ double part[kReduceThreads];
    const int j = blockIdx.x;
    double s = 0.0;
    for (int r = threadIdx.x; r < nrows; r += kReduceThreads) s += rows[(int64_t)r * m + j];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int h = kReduceThreads / 2; h > 0; h >>= 1) {
        if (threadIdx.x < h) part[threadIdx.x] += part[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[j] += part[0];
}

// Counting-sort scatter.  Positions: a, b, l, w (+ reg, key for the plain
// counter).  Shapes: the 8-double record {a, b, l, ax, ay, e*w, w, 2*e*w}
// (+ the SoA copies the plain counter reads).
__global__ void k_scatter(const double *__restrict__ xyz, int64_t n, Grid g,
                          const double *__restrict__ w, const double *__restrict__ axis,
                          const double *__restrict__ e, const int32_t *__restrict__ key,
                          const int32_t *__restrict__ reg, const int32_t *__restrict__ start,
                          int32_t *__restrict__ cursor, Sorted out, int soa) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int k = key[i];
    int64_t d = (int64_t)start[k] + atomicAdd(&cursor[k], 1);
    double a = xyz[3 * i + g.ax_a], b = xyz[3 * i + g.ax_b], l = xyz[3 * i + g.los];
    double wi = w ? w[i] : 1.0;
    if (!axis || soa) {
        out.a[d] = a;
        out.b[d] = b;
        out.l[d] = l;
        out.w[d] = wi;
    }
    if (soa) {
        out.reg[d] = reg[i];
        out.key[d] = k;
    }
    if (axis) {
        double ax = axis[2 * i], ay = axis[2 * i + 1], ei = e[i];
        if (soa) {
            out.ax[d] = ax;
            out.ay[d] = ay;
            out.e[d] = ei;
        }
        // a shape whose axis or e is NaN (e.g. a zero axis vector normalised on
        // the host) adds e+ = ex = 0, as the reference's isnan rule does
        // (measure_w_box_jk.py:174-175): zero its axis and e in the record
        if (!(ax == ax) || !(ay == ay) || !(ei == ei)) {
            ax = 0.0;
            ay = 0.0;
            ei = 0.0;
        }
        double ew = ei * wi;
        double2 *r = (double2 *)(out.rec + 8 * d);
        r[0] = make_double2(a, b);
        r[1] = make_double2(l, ax);
        r[2] = make_double2(ay, ew);
        r[3] = make_double2(wi, 2.0 * ew);
    }
}

__global__ void k_region_ids(const double *__restrict__ xyz, int64_t n, double jk_sub, int jk,
                             int32_t *__restrict__ region) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int ji[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) ji[c] = strict_axis_index(xyz[3 * i + c], jk_sub, jk);
    bool anom = ji[0] < 0 || ji[1] < 0 || ji[2] < 0;
    region[i] = anom ? 0 : ji[0] * jk * jk + ji[1] * jk + ji[2];
}

inline unsigned grid_for(int64_t n, int threads) {
    int64_t b = (n + threads - 1) / threads;
    return (unsigned)(b < 1 ? 1 : b);
}

}  // namespace

template <class T>
void fill(T *p, T v, int64_t n, hipStream_t st) {
    if (n <= 0) return;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_fill<T>, dim3((unsigned)blocks), dim3(256), 0, st, p, v, n);
    MIA_HIP(hipGetLastError());
}
void schedule_bounds(const int64_t *cost_prefix, int64_t n, int waves, int64_t *bounds, hipStream_t st) {
    const int64_t U = (int64_t)waves * kScheduleUnits;
    hipLaunchKernelGGL(k_schedule_bounds, dim3((unsigned)((U + 3 + 255) / 256)), dim3(256), 0, st, cost_prefix,
                       n, (const int64_t *)nullptr, U, waves, 0, bounds);
    MIA_HIP(hipGetLastError());
}
void schedule_bounds(const int64_t *cost_prefix, const int64_t *n_dev, int waves, int64_t *bounds,
                     hipStream_t st, bool all_static) {
    const int64_t U = (int64_t)waves * kScheduleUnits;
    hipLaunchKernelGGL(k_schedule_bounds, dim3((unsigned)((U + 3 + 255) / 256)), dim3(256), 0, st, cost_prefix,
                       (int64_t)0, n_dev, U, waves, all_static ? 1 : 0, bounds);
    MIA_HIP(hipGetLastError());
}

template void fill<int32_t>(int32_t *, int32_t, int64_t, hipStream_t);
template void fill<int64_t>(int64_t *, int64_t, int64_t, hipStream_t);
template void fill<uint64_t>(uint64_t *, uint64_t, int64_t, hipStream_t);
template void fill<double>(double *, double, int64_t, hipStream_t);
template void fill<unsigned long long>(unsigned long long *, unsigned long long, int64_t,
                                       hipStream_t);

// Exclusive scan of in[0..n) into out[0..n], out[n] = total.  Scratch from the arena.
template <class T>
void exclusive_scan(Ctx &ctx, const T *in, T *out, int64_t n, hipStream_t st,
                    const std::string &tag, int depth) {
    int64_t tiles = (n + kScanTile - 1) / kScanTile;
    if (tiles < 1) tiles = 1;
    std::string nm = tag + "/scan" + std::to_string(depth);
    T *sums = ctx.arena.get<T>(nm + "s", (size_t)tiles + 1);
    T *soff = ctx.arena.get<T>(nm + "o", (size_t)tiles + 1);
    hipLaunchKernelGGL(k_scan_tile<T>, dim3((unsigned)tiles), dim3(kScanThreads), 0, st, in, out,
                       n, sums);
    MIA_HIP(hipGetLastError());
    if (tiles > 1) {
        exclusive_scan<T>(ctx, sums, soff, tiles, st, tag, depth + 1);
        hipLaunchKernelGGL(k_scan_add<T>, dim3((unsigned)tiles), dim3(kScanThreads), 0, st, out, n,
                           soff);
        MIA_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_set_total<T>, dim3(1), dim3(1), 0, st, out, n, in, out);
    MIA_HIP(hipGetLastError());
}
template <class T>
void scan_carve(Carve &c, int64_t cap, ScanBufs<T> &b) {
    b.levels = 0;
    int64_t n = cap < 0 ? 0 : cap;
    // level l scans n[l] elements in tiles (at least one); its tile sums are level l + 1
    while (true) {
        const int64_t tiles = n > 0 ? (n + kScanTile - 1) / kScanTile : 1;
        b.n[b.levels] = n;
        b.sums[b.levels] = c.take<T>((size_t)tiles + 1);
        b.offs[b.levels] = c.take<T>((size_t)tiles + 1);
        ++b.levels;
        if (tiles == 1) break;
        MIA_REQUIRE(b.levels < 4, "scan too long");
        n = tiles;
    }
}

template <class T>
void scan_ws(const T *in, T *out, int64_t cap, const int64_t *n_dev, const ScanBufs<T> &b, hipStream_t st) {
    // level 0: the input; level l > 0 scans the tile sums of level l - 1 into offs
    const T *src = in;
    T *dst = out;
    for (int l = 0; l < b.levels; ++l) {
        const int64_t tiles = b.n[l] > 0 ? (b.n[l] + kScanTile - 1) / kScanTile : 1;
        hipLaunchKernelGGL(k_scan_tile_n<T>, dim3((unsigned)tiles), dim3(kScanThreads), 0, st, src, dst,
                           b.n[l], l == 0 ? n_dev : (const int64_t *)nullptr, b.sums[l]);
        MIA_HIP(hipGetLastError());
        src = b.sums[l];
        dst = b.offs[l];
    }
    // add the scanned tile sums back down, level by level
    for (int l = b.levels - 2; l >= 0; --l) {
        const int64_t tiles = b.n[l] > 0 ? (b.n[l] + kScanTile - 1) / kScanTile : 1;
        T *o = l == 0 ? out : b.offs[l - 1];
        hipLaunchKernelGGL(k_scan_add_n<T>, dim3((unsigned)tiles), dim3(kScanThreads), 0, st, o, b.n[l],
                           l == 0 ? n_dev : (const int64_t *)nullptr, (const T *)b.offs[l]);
        MIA_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_scan_total_n<T>, dim3(1), dim3(64), 0, st, out, b.n[0], n_dev,
                       (const T *)b.sums[b.levels - 1]);
    MIA_HIP(hipGetLastError());
}
template void scan_carve<int32_t>(Carve &, int64_t, ScanBufs<int32_t> &);
template void scan_carve<int64_t>(Carve &, int64_t, ScanBufs<int64_t> &);
template void scan_ws<int32_t>(const int32_t *, int32_t *, int64_t, const int64_t *, const ScanBufs<int32_t> &,
                               hipStream_t);
template void scan_ws<int64_t>(const int64_t *, int64_t *, int64_t, const int64_t *, const ScanBufs<int64_t> &,
                               hipStream_t);

template void exclusive_scan<int32_t>(Ctx &, const int32_t *, int32_t *, int64_t, hipStream_t,
                                      const std::string &, int);
template void exclusive_scan<int64_t>(Ctx &, const int64_t *, int64_t *, int64_t, hipStream_t,
                                      const std::string &, int);

void region_ids(Ctx &ctx, const double *xyz, int64_t n, double jk_sub, int jk, int32_t *region,
                hipStream_t st) {
    (void)ctx;
    if (n <= 0) return;
    hipLaunchKernelGGL(k_region_ids, dim3(grid_for(n, 256)), dim3(256), 0, st, xyz, n, jk_sub, jk,
                       region);
    MIA_HIP(hipGetLastError());
}

static void sort_one(Ctx &ctx, const double *xyz, int64_t n, const double *w, const double *axis,
                     const double *e, const Grid &g, Sorted &out, double *rstats, int which,
                     int *oor, bool soa, hipStream_t st, const std::string &tag) {
    int keys = g.keys();
    out.n = n;
    size_t nn = (size_t)(n > 0 ? n : 1);
    out.a = ctx.arena.get<double>(tag + "/a", nn);
    out.b = ctx.arena.get<double>(tag + "/b", nn);
    out.l = ctx.arena.get<double>(tag + "/l", nn);
    out.w = ctx.arena.get<double>(tag + "/w", nn);
    out.reg = ctx.arena.get<int32_t>(tag + "/reg", nn);
    out.key = ctx.arena.get<int32_t>(tag + "/key", nn);
    if (axis) {
        out.ax = ctx.arena.get<double>(tag + "/ax", nn);
        out.ay = ctx.arena.get<double>(tag + "/ay", nn);
        out.e = ctx.arena.get<double>(tag + "/e", nn);
        out.rec = ctx.arena.get<double>(tag + "/rec", 8 * nn);
    }
    out.start = ctx.arena.get<int32_t>(tag + "/start", (size_t)keys + 1);
    int32_t *counts = ctx.arena.get<int32_t>(tag + "/counts", (size_t)keys + 1);
    int32_t *cursor = ctx.arena.get<int32_t>(tag + "/cursor", (size_t)keys + 1);
    int32_t *key0 = ctx.arena.get<int32_t>(tag + "/key0", nn);
    int32_t *reg0 = ctx.arena.get<int32_t>(tag + "/reg0", nn);
    fill<int32_t>(counts, 0, keys + 1, st);
    fill<int32_t>(cursor, 0, keys + 1, st);
    if (n > 0) {
        const int K4 = g.jk > 0 ? 4 * g.jk * g.jk * g.jk : 0;
        unsigned grid = std::min<unsigned>(grid_for(n, kClassifyThreads), 4u * ctx.cu_count);
        // per-workgroup partial region sums when they fit in LDS
        bool local = rstats && K4 > 0 && K4 * 8 <= 32 * 1024;
        double *part = local ? ctx.arena.get<double>(tag + "/rpart", (size_t)grid * K4) : nullptr;
        size_t lds = local ? (size_t)K4 * 8 : 0;
        hipLaunchKernelGGL(k_classify, dim3(grid), dim3(kClassifyThreads), lds, st, xyz, n, g, w,
                           e, key0, reg0, counts, rstats, part, which, oor);
        MIA_HIP(hipGetLastError());
        if (local) {
            hipLaunchKernelGGL(k_reduce_rows, dim3(K4), dim3(kReduceThreads), 0, st, part,
                               (int)grid, K4, rstats);
            MIA_HIP(hipGetLastError());
        }
    }
    exclusive_scan<int32_t>(ctx, counts, out.start, keys, st, tag, 0);
    if (n > 0) {
        hipLaunchKernelGGL(k_scatter, dim3(grid_for(n, 256)), dim3(256), 0, st, xyz, n, g, w, axis,
                           e, key0, reg0, out.start, cursor, out, soa ? 1 : 0);
        MIA_HIP(hipGetLastError());
    }
}

void prepare_sample(Ctx &ctx, const mia_sample &s, Grid &g, Sorted &pos, Sorted &shp,
                    double *region_stats, int *oor, bool soa, hipStream_t st,
                    const std::string &tag) {
    if (region_stats && g.jk > 0) fill<double>(region_stats, 0.0, (int64_t)g.jk * g.jk * g.jk * 4, st);
    double *rs = g.jk > 0 ? region_stats : nullptr;
    sort_one(ctx, s.pos, s.n_pos, s.pos_w, nullptr, nullptr, g, pos, rs, 0, oor, soa, st,
             tag + "/pos");
    sort_one(ctx, s.shp, s.n_shape, s.shp_w, s.axis, s.e, g, shp, rs, 1, oor, soa, st,
             tag + "/shp");
}

void check_prepared(const int *oor_dev, int n, hipStream_t st) {
    std::vector<int> h(n > 0 ? n : 1, 0);
    MIA_HIP(hipMemcpyAsync(h.data(), oor_dev, sizeof(int) * n, hipMemcpyDeviceToHost, st));
    MIA_HIP(hipStreamSynchronize(st));
    for (int i = 0; i < n; ++i)
        if (h[i])
            throw Error{MIA_E_RANGE, "coordinates outside [0, boxsize] (or NaN) in sample " +
                                         std::to_string(i)};
}

}  // namespace mia
