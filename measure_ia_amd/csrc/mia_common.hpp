// Shared declarations of the MI355X pair-counting engine (internal).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/measure_ia_amd.h"

namespace mia {

constexpr int kWave = 64;
constexpr int kXcds = 8;            // accelerator complex dies, each with its own L2

// Workgroup number that keeps runs of consecutive work on one XCD.
// Workgroups are dispatched round-robin over the XCDs, so hardware workgroups
// b, b + 8, b + 16, ... share an L2.  The logical numbering gives each XCD
// runs of S = 16 consecutive workgroups (64 waves) in turn: neighbouring
// items of a sorted work list, which read the same inputs, meet in one L2,
// while every XCD still sees every part of the list (one contiguous eighth
// per XCD left whole samples of a batched launch to single XCDs and
// unbalanced them).  S halves until it divides the grid; a grid that is not
// a multiple of kXcds keeps its order.
__device__ __forceinline__ int xcd_run_block() {
    const int g = (int)gridDim.x, b = (int)blockIdx.x;
    int S = 16;
    while (S > 1 && g % (kXcds * S) != 0) S >>= 1;
    if (g % (kXcds * S) != 0) return b;
    const int slot = b / kXcds, x = b % kXcds;
    return ((slot / S) * kXcds + x) * S + slot % S;
}
// Persistent-wave schedule over a cost-sorted work list (cost_prefix[i] =
// summed estimated cost of items 0..i-1, cost_prefix[n] = C).  The cost is cut
// into U = W * kScheduleUnits equal units; bounds[u] is the first item whose
// prefix reaches u * C / U (schedule_bounds), so bounds[0] = 0, bounds[U] = n
// (item costs are positive) and each item lies in exactly one unit range.
// Wave w first takes units [w * S, (w + 1) * S) as one contiguous range; then
// the waves claim the rest in chunks of G units, in order, from a counter
// (zeroed before the launch), until it passes the last chunk.  Waves the cost
// estimate under-loaded thus take work that would otherwise wait on the waves
// it over-loaded.  S and G (bounds[U + 1], bounds[U + 2]) follow a wave's
// share of the cost, c = C / W, since every claim queues on the one counter:
//   c >= kTailFineCost: S = 2, G = 1 (1/16 static, 30 claims per wave);
//   c >= kTailMinCost:  S = 4, G = 2 (1/8 static, 14 claims);
//   otherwise:          S = kScheduleUnits (all static, no claims).
// Contiguous static ranges keep the histogram flushes few and neighbouring
// items on one XCD.  Measured on the benchmark workloads: pair kernels
// 20-27 % faster than a static split of the whole cost; with claims, a
// 0.8 ms launch (lightcone1M) was 40 % slower, hence the static case.
constexpr int kScheduleUnits = 32;
constexpr double kTailMinCost = 4.0e5;   // item-cost units per wave (about 5 ms of pair work)
constexpr double kTailFineCost = 8.0e5;  // (about 10 ms)
constexpr int kMaxWaves = 1 << 20;       // persistent waves per launch (tail counter stays in int)

struct WaveSchedule {
    const __attribute__((address_space(4))) int64_t *bounds;
    int W, S, G, tail;

    // the static range of wave w
    __device__ __forceinline__ void first(int w, int64_t &i0, int64_t &i1) const {
        i0 = bounds[(int64_t)w * S];
        i1 = bounds[(int64_t)(w + 1) * S];
    }
    // claims the next tail chunk; false once all are taken
    __device__ __forceinline__ bool next(int *counter, int lane, int64_t &i0, int64_t &i1) const {
        if (tail == 0) return false;
        int k = 0;
        if (lane == 0) k = atomicAdd(counter, 1);
        k = __builtin_amdgcn_readfirstlane(k);
        if (k >= tail) return false;
        const int64_t u = (int64_t)W * S + (int64_t)k * G;
        i0 = bounds[u];
        i1 = bounds[u + G];
        return true;
    }
};

__device__ __forceinline__ WaveSchedule make_schedule(const int64_t *bounds, int W) {
    WaveSchedule s;
    s.bounds = (const __attribute__((address_space(4))) int64_t *)bounds;
    s.W = W;
    const int64_t U = (int64_t)W * kScheduleUnits;
    s.S = (int)s.bounds[U + 1];
    s.G = (int)s.bounds[U + 2];
    s.tail = W * ((kScheduleUnits - s.S) / s.G);
    return s;
}

// bounds[0..waves * kScheduleUnits + 2] of the schedule above, on stream st.
void schedule_bounds(const int64_t *cost_prefix, int64_t n, int waves, int64_t *bounds, hipStream_t st);
// the same with the item count read on the device; all_static: every wave
// takes only its static range (no tail claims), so the item -> wave map is
// fixed by the costs
void schedule_bounds(const int64_t *cost_prefix, const int64_t *n_dev, int waves, int64_t *bounds,
                     hipStream_t st, bool all_static = false);

constexpr int kMaxBins2 = 256;      // max n_2 (pi or mu bins)
constexpr int kMaxBinsR = 256;      // max n_r
constexpr int kMaxBlockShapes = 64;

// Error carrying a MIA_E_* code; caught at the C boundary.
struct Error {
    int code;
    std::string msg;
};

#define MIA_HIP(call)                                                                       \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess)                                                               \
            throw ::mia::Error{MIA_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)}; \
    } while (0)

#define MIA_REQUIRE(cond, msg)                                     \
    do {                                                           \
        if (!(cond)) throw ::mia::Error{MIA_E_ARG, std::string(msg)}; \
    } while (0)

// Growable named device buffers owned by a context.
class Arena {
  public:
    ~Arena() { release(); }
    void *get(const std::string &name, size_t bytes) {
        auto it = bufs_.find(name);
        if (it != bufs_.end() && it->second.bytes >= bytes) return it->second.ptr;
        if (it != bufs_.end()) {
            (void)hipFree(it->second.ptr);
            bufs_.erase(it);
        }
        void *p = nullptr;
        size_t b = bytes < 256 ? 256 : bytes;
        if (hipMalloc(&p, b) != hipSuccess) throw Error{MIA_E_NOMEM, "hipMalloc(" + name + ")"};
        bufs_[name] = {p, b};
        return p;
    }
    template <class T>
    T *get(const std::string &name, size_t count) {
        return static_cast<T *>(get(name, count * sizeof(T)));
    }
    void release() {
        for (auto &kv : bufs_) (void)hipFree(kv.second.ptr);
        bufs_.clear();
    }

  private:
    struct Buf {
        void *ptr;
        size_t bytes;
    };
    std::map<std::string, Buf> bufs_;
};

// Grid of the counting sort: columns in the projected plane (axes a, b =
// the two non-LOS columns in ascending order) and slabs along the LOS axis l.
// With jackknife, cell boundaries are aligned to the sub-box boundaries.
struct Grid {
    double L = 0, half = 0;
    int los = 2, ax_a = 0, ax_b = 1;
    int jk = 0;            // sub-boxes per side (0: none)
    double jk_sub = 0;     // L_sub
    int nc = 1, nz = 1;    // columns / slabs per side
    int ncj = 1, nzj = 1;  // per jackknife sub-box (nc = jk*ncj when jk > 0)
    double cw = 0, hz = 0; // nominal column width and slab height
    double margin = 0;     // geometric slack for rounding at cell edges
    int G = 0;             // slabs per shape block chosen with the grid (0: decide later)
    int pmax = 0;          // positions per work item chosen with the grid (0: the default)
    int keys() const { return nc * nc * (nz + 1); }
};

// Sorted sample in HBM (structure of arrays, sorted by cell key).
struct Sorted {
    int64_t n = 0;
    double *a = nullptr, *b = nullptr, *l = nullptr;  // coordinates in (a, b, los) order
    double *w = nullptr;                              // weights (1 when absent)
    int32_t *reg = nullptr;                           // jackknife region (0 without jk)
    int32_t *key = nullptr;                           // cell key
    double *ax = nullptr, *ay = nullptr, *e = nullptr; // shapes only
    double *rec = nullptr;  // shapes only: 8 doubles per shape {a, b, l, ax, ay, e*w, w, 2*e*w}
    int32_t *start = nullptr;                         // keys+1 offsets
};

struct Ctx;

// HIP events destroyed on scope exit (also when a count throws).
class Events {
  public:
    explicit Events(int n) : ev_(n, nullptr) {
        for (auto &e : ev_) MIA_HIP(hipEventCreate(&e));
    }
    ~Events() {
        for (auto e : ev_)
            if (e) (void)hipEventDestroy(e);
    }
    Events(const Events &) = delete;
    Events &operator=(const Events &) = delete;
    hipEvent_t operator[](int i) const { return ev_[i]; }
    const hipEvent_t *data() const { return ev_.data(); }
    // milliseconds between two recorded events (the later one must be complete)
    float ms(int a, int b) const {
        float t = 0;
        MIA_HIP(hipEventElapsedTime(&t, ev_[a], ev_[b]));
        return t;
    }

  private:
    std::vector<hipEvent_t> ev_;
};

// Host-side helpers implemented in the .hip files.
void region_ids(Ctx &ctx, const double *xyz, int64_t n, double jk_sub, int jk, int32_t *region,
                hipStream_t st);
// Region ids + counting sort of one sample (stream-ordered, no host sync).
// Sets *oor (device) to 1 if a coordinate lies outside [0, L] or is NaN;
// check_prepared() reads the flags of n samples back and throws MIA_E_RANGE.
// soa: also write the structure-of-arrays copies the plain counter reads.
void prepare_sample(Ctx &ctx, const mia_sample &s, Grid &g, Sorted &pos, Sorted &shp,
                    double *region_stats, int *oor, bool soa, hipStream_t st,
                    const std::string &tag);
void check_prepared(const int *oor_dev, int n, hipStream_t st);
void count_plain(Ctx &ctx, const mia_sample &s, const mia_bins &bins, const Grid &g,
                 const Sorted &pos, const Sorted &shp, const mia_counts &out, bool weighted,
                 hipStream_t st, mia_stats *stats);

// Whether the batched kernel can count this binning: one wave's histogram
// must fit in LDS and the (1+K)^2 planner keys must stay bounded.
bool fast_path_fits(const mia_bins &bins, bool weighted, int jk);

// ---------------------------------------------------------------------------
// Staged fast count.  Every size the device decides (shape blocks, work
// items, cost totals) stays on the device, all scratch is carved from one
// workspace whose size the host bounds beforehand, and no stage reads
// anything back.  The whole sequence can therefore be replayed from a
// hipGraph, or run inside a fixed (static) workspace.

// A cell-sorted sample on the device: a prepared handle, or one sample of a
// batch prepared inside the workspace.
struct SortedSample {
    int64_t n_pos = 0, n_shape = 0;
    Grid g;
    int K = 0;                        // jackknife regions (0: none)
    int periodic = 1;
    int weighted = 0;                 // a weight array was given
    const double *prec = nullptr;     // 4 doubles per position {a, b, l, w}, sorted order
    const double *srec = nullptr;     // 8 doubles per shape {a, b, l, ax, ay, e*w, w, 2*e*w}
    const int32_t *pstart = nullptr, *sstart = nullptr;  // keys + 1 start offsets into the arrays above
    const int32_t *pfine = nullptr;   // (keys << sbits) + 1 starts of the positions' fine keys
    int sbits = 0;                    // fine key = (cell << sbits) + line-of-sight part of the slab
    const unsigned long long *rcnt = nullptr;  // [2K] per region: positions, shapes
    const double *rsum = nullptr;     // [2K] per region: sum w_s, sum w_s (1 - e^2/2)
};

// Bytes a workspace needs, and the carving of it (mia_stage.hip).
struct Carve {
    char *base = nullptr;
    size_t off = 0;
    bool dry = true;  // dry run: only sum the sizes
    template <class T>
    T *take(size_t count) {
        size_t b = (count * sizeof(T) + 255) & ~(size_t)255;
        if (b == 0) b = 256;
        T *p = dry ? nullptr : reinterpret_cast<T *>(base + off);
        off += b;
        return p;
    }
};

// Scratch of one exclusive scan of up to `cap` elements (tile sums per level).
template <class T>
struct ScanBufs {
    int levels = 0;
    int64_t n[4] = {0, 0, 0, 0};      // elements per level (level 0 = the input)
    T *sums[4] = {nullptr, nullptr, nullptr, nullptr};
    T *offs[4] = {nullptr, nullptr, nullptr, nullptr};
};
template <class T>
void scan_carve(Carve &c, int64_t cap, ScanBufs<T> &b);
// Exclusive scan of in[0..n) into out[0..n], out[n] = total.  n = min(*n_dev,
// cap) when n_dev is given (read on the device), else cap.
template <class T>
void scan_ws(const T *in, T *out, int64_t cap, const int64_t *n_dev, const ScanBufs<T> &b,
             hipStream_t st);

// Per-call parameters of the count stages, uploaded once per call.
struct CountCall;
// Host plan of a staged count over n sorted samples: worst-case sizes,
// launch shapes and the pair-kernel configuration.
struct CountPlan;
// Work-item packing of the planner: one plan thread walks kGroupColumns
// stencil columns of a shape block and packs their position ranges into items
// of at most kItemRanges ranges (and the item's position limit).
constexpr int kGroupColumns = 8;
constexpr int kItemRanges = 8;
// item_cap, range_cap: at least this many work-item slots and range records
// (exactly this many with cap_exact)
struct CountPlan *plan_count(Ctx &ctx, const std::vector<SortedSample> &ss, const mia_bins &bins,
                             const mia_tuning *tuning, int64_t item_cap, int64_t range_cap, bool cap_exact);
void free_plan(CountPlan *p);
// workspace bytes of the count stages (after plan_count)
size_t count_bytes(const CountPlan &p);
// Carves the count stages' buffers from c and writes the per-call parameter
// block (host image) for these samples and outputs.
void count_bind(CountPlan &p, Carve &c, const std::vector<SortedSample> &ss, const mia_counts *outs);
// host bytes of the parameter block and its image
size_t count_param_bytes(const CountPlan &p);
const void *count_param_image(const CountPlan &p);
void *count_param_dev(const CountPlan &p);
// Enqueues every count stage on st (no host sync, no allocation).  With ev
// (4 events), records them before the plan, before the pairs, after the
// pairs and at the end.
void count_enqueue(const CountPlan &p, hipStream_t st, const hipEvent_t *ev);
// Device words read back after a count: [0] rechecks, [1] candidates of the
// planned items, [2] item or range overflow, [3] items, [4] pair-kernel start
// (~wall clock), [5] end, [6] wave-iterations of the pair kernel (one 64-lane
// position batch against one shape), [7] range records of the multi-range
// items, [8] position ranges of the planner.
const unsigned long long *count_readback(const CountPlan &p);
constexpr int kCountReadback = 9;
int count_waves(const CountPlan &p);
int count_bsz(const CountPlan &p);
bool count_det(const CountPlan &p);
int count_G(const CountPlan &p, int sample);
int64_t count_item_cap(const CountPlan &p);
int64_t count_range_cap(const CountPlan &p);
// a key identifying the launch sequence (everything but caller pointers)
std::string count_key(const CountPlan &p);

// Lightcone counter (mia_lightcone.hip): whether the binning's histogram fits
// one wave's LDS share, and the count itself (outputs zeroed and filled).
bool lightcone_fits(const mia_bins &bins);
struct Ctx;
void count_lightcone(Ctx &ctx, const mia_lc_sample &s, const mia_bins &bins, const mia_counts &out,
                     const mia_tuning *tuning, hipStream_t st, mia_stats *stats);

struct Ctx {
    int device = 0;
    std::string last_error;
    Arena arena;
    int cu_count = 256;
    // staged counts (mia_stage.hip)
    void *ws = nullptr;             // growable workspace
    size_t ws_bytes = 0;
    void *pinned = nullptr;         // host staging of parameter blocks and read-backs
    size_t pinned_bytes = 0;
    hipStream_t cap_stream = nullptr;   // stream the batched count is captured on
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    std::string gkey;               // launch-sequence key of gexec
    int64_t item_cap = 0;           // grown after an item overflow
    int64_t range_cap = 0;          // grown after a range-record overflow
    double wall_khz = 100000.0;     // device wall-clock rate
    void *scratch = nullptr;        // the device's static workspace (allocation-free entry point)
};

// Staged-count resources of a context (mia_stage.hip).
void stage_init(Ctx &ctx);
void stage_release(Ctx &ctx);
struct PreparedSample;
// mia_count (MIA_FAST), mia_count_batch, mia_count_noalloc and the handle calls.
enum StageMode { STAGE_EAGER = 0, STAGE_GRAPH = 1, STAGE_STATIC = 2 };
void stage_count(Ctx &ctx, int n, const mia_sample *samples, const mia_bins &bins, const mia_counts *outs,
                 const mia_tuning *tuning, mia_stats *stats, hipStream_t st, StageMode mode);
PreparedSample *stage_prepare(Ctx &ctx, const mia_sample &s, const mia_bins &bins, const mia_tuning *tuning,
                              hipStream_t st);
void stage_count_prepared(Ctx &ctx, int n, const PreparedSample *const *h, const mia_bins &bins,
                          const mia_counts *outs, const mia_tuning *tuning, mia_stats *stats, hipStream_t st);
void stage_free_prepared(PreparedSample *h);
const SortedSample &prepared_view(const PreparedSample *h);
int64_t stage_static_bytes();

template <class T>
void fill(T *p, T v, int64_t n, hipStream_t st);
template <class T>
void exclusive_scan(Ctx &ctx, const T *in, T *out, int64_t n, hipStream_t st,
                    const std::string &tag, int depth);

// Pick columns/slabs per side and slabs per block for a sample: explicit
// tuning fields win; otherwise a cost-model search (mia_grid.hip).
Grid make_grid(const mia_sample &s, const mia_bins &bins, const mia_tuning *t);

// ---------------------------------------------------------------------------
// device helpers

__device__ __forceinline__ int clampi(int x, int lo, int hi) { return x < lo ? lo : (x > hi ? hi : x); }

// Index i with the reference's strict test i*L_sub < x < (i+1)*L_sub, or -1.
// acc = keep * acc + x (one FP64 FMA, the same rounding as fma()), written
// into acc's own register: for fma(keep, acc, x) the compiler picks the
// two-operand v_fmac form, which overwrites x's register, and then copies the
// result back into acc's register on every pair.
__device__ __forceinline__ double fma_into(double keep, double acc, double x) {
    asm("v_fma_f64 %0, %1, %0, %2" : "+v"(acc) : "v"(keep), "v"(x));
    return acc;
}

__device__ __forceinline__ int strict_axis_index(double x, double Lsub, int jk) {
    if (!(x > 0.0) || !(x < Lsub * jk + Lsub)) return -1;
    double q = x / Lsub;
    int i0 = (int)floor(q);
    for (int i = i0 - 1; i <= i0 + 1; ++i) {
        if (i < 0 || i >= jk) continue;
        double lo = (double)i * Lsub;
        double hi = (double)(i + 1) * Lsub;
        if (x > lo && x < hi) return i;
    }
    return -1;
}

// The reference's minimum-image wrap (measure_w_box_jk.py:160-162).
__device__ __forceinline__ double wrap_sep(double d, double L, double half) {
    if (d > half) d -= L;
    if (d < -half) d += L;
    return d;
}

}  // namespace mia
