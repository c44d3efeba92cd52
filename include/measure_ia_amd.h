/*
 * measure_ia_amd.h -- C ABI of the MI355X periodic-box pair-counting engine.
 *
 * The reference (MeasureIA, pure Python) has no FFI: its boundary is the
 * Python method contract of the private box counters that
 * MeasureIABox.measure_xi_w / measure_xi_multipoles dispatch to
 * (src/measureia/measure_IA.py:120-160, 218-259).  Every counter takes the
 * data dict, prepares the arrays (src/measureia/measure_w_box_jk.py:97-155)
 * and returns or writes DD, S+D, SxD and the jackknife tallies DD_jk,
 * S+D_jk (measure_w_box_jk.py:157-211).  This header replaces the pair
 * loop of those counters -- the KDTree neighbour search, the binning and the
 * jackknife tallies -- and nothing else: input preparation, the analytic
 * randoms, the estimators and the HDF5 output stay in the host code.
 *
 * Conventions
 *   - Plain C types only.  Functions return 0 on success and a negative
 *     MIA_E_* code on error; mia_last_error() gives the message.  No C++
 *     exception crosses this boundary.
 *   - Arrays marked "device" are HIP device pointers owned by the caller;
 *     arrays marked "host" are ordinary host memory.  The context owns all
 *     scratch memory and reuses it between calls.
 *   - Every launch goes on the caller's stream (NULL = default stream).  A
 *     count enqueues its whole launch sequence without reading anything back
 *     (every size the device decides stays on the device), then reads a few
 *     status words and synchronises that stream before returning.
 *   - Output buffers only need to be device-accessible memory: the kernels
 *     accumulate in the context's own workspace and write each output once.
 *   - One context per host thread; a context is not thread-safe.
 */
#ifndef MEASURE_IA_AMD_H
#define MEASURE_IA_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MIA_ABI_VERSION 4

enum mia_status {
    MIA_OK = 0,
    MIA_E_ARG = -1,      /* invalid argument or inconsistent sizes            */
    MIA_E_HIP = -2,      /* a HIP runtime call failed                          */
    MIA_E_NOMEM = -3,    /* device allocation failed                           */
    MIA_E_RANGE = -4,    /* coordinates outside [0, boxsize] (or NaN)            */
    MIA_E_UNSUPPORTED = -5
};

/* Binning of the counter.
 *   MIA_RPPI: (r_p, pi) grid of _measure_xi_rp_pi_box_* (measure_w_box_jk.py:157-211)
 *   MIA_RMU : (r, mu_r) grid of _measure_xi_r_mur_box_* (measure_m_box_jk.py:164-223) */
enum mia_mode { MIA_RPPI = 0, MIA_RMU = 1 };

/* Counting method.
 *   MIA_PLAIN: one thread per shape, FP64 geometry in the reference's op
 *              order, global-atomic histograms (the straightforward GPU form).
 *   MIA_FAST : cell-pair work items sorted by jackknife region pair, a
 *              persistent cost-balanced grid, per-wavefront LDS histograms
 *              (u32 counts, FP64 sums); FP64 separations and e+ in the
 *              reference's rounding, FP32 bin guesses resolved exactly in
 *              FP64 near an edge. */
enum mia_method { MIA_PLAIN = 0, MIA_FAST = 1 };

typedef struct mia_ctx mia_ctx;

/* One galaxy sample: the arrays a box counter prepares on the host
 * (measure_w_box_jk.py:97-139).  All pointers are device pointers. */
typedef struct mia_sample {
    const double *pos;      /* n_pos x 3 positions, row-major, columns x,y,z        */
    const double *pos_w;    /* n_pos position weights, or NULL for unit weights     */
    const double *shp;      /* n_shape x 3 shape-sample positions                   */
    const double *shp_w;    /* n_shape shape weights, or NULL for unit weights      */
    const double *axis;     /* n_shape x 2 normalised projected axis direction      */
    const double *e;        /* n_shape ellipticity |e| (from q, host-side)          */
    int64_t n_pos;
    int64_t n_shape;
    double boxsize;         /* L (periodic box [0, L)^3)                            */
    double jk_sub;          /* L_sub = 2*L_0p5/L_subboxes (measure_IA_base.py:428)  */
    int32_t los;            /* line-of-sight column, 0..2                           */
    int32_t periodic;       /* 1: minimum-image wrap (measure_w_box_jk.py:160-162); 0: none,
                               as MeasureIABox(periodicity=False).  Either way every
                               coordinate must lie in [0, boxsize], as the reference's
                               KDTree over the box requires (measure_w_box_jk.py:386). */
    int32_t jk_per_side;    /* L_subboxes; 0 = no jackknife                         */
    int32_t reserved;
} mia_sample;

/* Bin thresholds, bisected on the host so that the device reproduces the
 * reference's np.log10/np.floor bin index bit-for-bit.  Host pointers.
 *   r2_edges[n_r+1]: squared-separation thresholds [S_lo, T_1..T_{n_r-1}, S_hi]:
 *                    in domain iff S_lo <= s < S_hi, bin = #{k>=1 : T_k <= s}.
 *   MIA_RPPI: x2_edges[n_2+1] = [pi_lo, P_1..P_{n_2-1}, pi_hi]: in domain iff
 *                    pi_lo <= pi < pi_hi, bin = #{k>=1 : P_k <= pi}.
 *   MIA_RMU : x2_edges[n_2]   = [M_0, M_1..M_{n_2-1}] on mu = pi/r:
 *                    bin = n_2-1 if mu < M_0 (numpy's index -1), else
 *                    #{k>=1 : M_k <= mu}.  Domain also needs r_p^2 > 0.   */
typedef struct mia_bins {
    int32_t mode;           /* enum mia_mode                                          */
    int32_t n_r;
    int32_t n_2;            /* pi bins or mu bins                                     */
    int32_t accumulate_variance; /* 1: also sum (w_p w_s e+)^2 (brute-path _sigmasq) */
    const double *r2_edges;
    const double *x2_edges;
} mia_bins;

/* Outputs, device pointers, zeroed and filled by the count call.  With
 * K = jk_per_side^3 (0 without jackknife) and B = n_r*n_2:
 *   npairs [(1+K)*B] exact integer pair counts: [0] all pairs, [1+i] the
 *                    pairs with the shape or the position in region i
 *                    (each pair once per distinct region it touches).
 *   dd     [(1+K)*B] the same tallies of w_p*w_s (DD); may be NULL when both
 *                    weight arrays are NULL (DD == npairs then).
 *   splus  [(1+K)*B] the same tallies of w_p*w_s*e+ (S+D times 2R).
 *   scross [B]       sum of w_p*w_s*ex (SxD times 2R).
 *   variance [B]     sum of (w_p*w_s*e+)^2, or NULL.
 *   region_stats [K*4] per region: #positions, #shapes, sum w_s,
 *                    sum w_s*(1-e^2/2); or NULL.                              */
typedef struct mia_counts {
    uint64_t *npairs;
    double *dd;
    double *splus;
    double *scross;
    double *variance;
    double *region_stats;
} mia_counts;

/* Optional tuning (NULL = automatic).  Zero fields mean "choose": the grid
 * (columns, slabs, slabs per block) then comes from a cost-model search over
 * the sample's sizes, box and binning reach. */
typedef struct mia_tuning {
    int32_t columns_per_side;   /* projected-plane cells per box side                 */
    int32_t slabs_per_side;     /* line-of-sight slabs per box side                   */
    int32_t slabs_per_block;    /* shape-block height in slabs                        */
    int32_t block_shapes;       /* shapes per block, <= 64                            */
    int32_t item_positions;     /* max positions per work item                        */
    int32_t waves;              /* persistent waves of the pair kernel (at most 2^20)  */
    int32_t hist_copies;        /* ignored (one LDS histogram copy per wave)           */
    int32_t cell_shapes;        /* > 0: size cells to hold this many shapes on average
                                   instead of the cost-model search                     */
    int32_t r_bins_per_pass;    /* > 0: r bins per pass of the pair kernel, clamped to
                                   what one wave's LDS histogram holds (large binnings
                                   are counted in passes over r-bin ranges; by default
                                   the library picks the pass size)                    */
    int32_t graph;              /* mia_count_batch: 0 = capture the launch sequence in a
                                   hipGraph and replay it while it stays the same (the
                                   default); -1 = plain launches                        */
    int32_t deterministic;      /* 1 (MIA_FAST): results identical bit for bit from run
                                   to run: the cell sort keeps input order inside a cell
                                   (stable radix sort), every wave takes a fixed share of
                                   the work, and the waves' partial histograms are summed
                                   in a fixed order with compensated FP64 sums instead of
                                   atomics.  Plain launches (no graph); slower.  The sums
                                   of the default path differ from run to run by FP64
                                   rounding only.  Added in ABI version 4.              */
} mia_tuning;

/* Per-call statistics, filled when the pointer is non-NULL. */
typedef struct mia_stats {
    double ms_prepare;      /* region ids + cell sort                                  */
    double ms_plan;         /* shape blocks + work items (MIA_FAST)                    */
    double ms_pairs;        /* pair kernel                                             */
    double ms_finalize;     /* work after the pair kernel (currently none: ~0)          */
    int64_t candidates;     /* (shape, position) pairs evaluated                       */
    int64_t items;          /* work items (MIA_FAST)                                   */
    int64_t rechecks;       /* pairs re-evaluated in FP64 near an edge (MIA_FAST)      */
    int32_t columns_per_side, slabs_per_side, slabs_per_block, block_shapes;
    int32_t waves;
    int32_t method;         /* counting method actually used (MIA_FAST falls back to
                               MIA_PLAIN when its LDS histogram or key space is too big) */
    double ms_total;        /* the whole call on the device (events around it)          */
    int32_t graph;          /* 1: the call replayed a captured hipGraph                 */
    int32_t deterministic;  /* 1: the deterministic reduction ran (ABI version 4)        */
    int64_t item_capacity;  /* work-item slots of the workspace                         */
    int64_t wave_iterations;    /* pair-kernel loop trips: one 64-lane position batch
                                   against one shape                                    */
} mia_stats;

int mia_abi_version(void);

/* Create a context bound to HIP device `device`. */
int mia_create(int32_t device, mia_ctx **out);
void mia_destroy(mia_ctx *ctx);
const char *mia_last_error(const mia_ctx *ctx);

/* Count one sample: replaces the pair loop of
 * _measure_xi_rp_pi_box_jk_{brute,tree,multiprocessing} (measure_w_box_jk.py:69-857),
 * _measure_xi_r_mur_box_jk_* (measure_m_box_jk.py:72-898) and their no-jackknife
 * twins (measure_w_box.py:65-656, measure_m_box.py:255-890). */
int mia_count(mia_ctx *ctx, const mia_sample *sample, const mia_bins *bins,
              const mia_counts *out, int32_t method, const mia_tuning *tuning,
              mia_stats *stats, void *stream);

/* Count n samples of uneven size (MIA_FAST only): all samples are sorted by
 * one set of launches and counted by one persistent pair-kernel launch, and
 * the whole sequence is captured in a hipGraph that later calls with the same
 * sizes, binning and tuning replay (tuning->graph = -1: plain launches).
 * Bins are shared, outputs are per sample.  The reference has no batched
 * form; a caller loops measure_xi_w over samples (measure_IA.py:68-163). */
int mia_count_batch(mia_ctx *ctx, int32_t n, const mia_sample *samples, const mia_bins *bins,
                    const mia_counts *outs, const mia_tuning *tuning, mia_stats *stats,
                    void *stream);

/* mia_count (MIA_FAST) without any allocation.  This entry point has no
 * counterpart in the reference; it is for callers that cannot allocate
 * device memory (code under hipGraph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  Every buffer of the count is carved from a static
 * device workspace of mia_noalloc_workspace_bytes() bytes (256 MiB) that each
 * device holds from mia_create until the process ends; the call makes no
 * hipMalloc, hipFree, event, stream or graph.  A count that needs more
 * workspace (about 1.5 M positions with 0.3 M shapes at the benchmark binning)
 * returns MIA_E_UNSUPPORTED with the size it needs, before anything is
 * launched where the size is known up front.  All calls on one device share
 * its workspace: calls that overlap in time on one device (two host threads,
 * or two streams) race and must be serialised by the caller.  No stats. */
int mia_count_noalloc(mia_ctx *ctx, const mia_sample *sample, const mia_bins *bins,
                      const mia_counts *out, const mia_tuning *tuning, void *stream);
int64_t mia_noalloc_workspace_bytes(void);

/* A sample whose cell sort is kept on the device between counts. */
typedef struct mia_prepared mia_prepared;

/* Sort one sample once (region ids, cell keys, counting sort) into memory
 * the handle owns; the cell grid is chosen for `bins` and `tuning` and serves
 * every later binning.  The reference rebuilds its KDTree on every counter
 * call (measure_w_box_jk.py:386, measure_m_box_jk.py:403), so a
 * measure_xi_w followed by measure_xi_multipoles sorts twice; with a handle
 * the second count skips the sort.  The input arrays may be freed after the
 * call returns.  MIA_E_RANGE as mia_count. */
int mia_prepare_sample(mia_ctx *ctx, const mia_sample *sample, const mia_bins *bins,
                       const mia_tuning *tuning, mia_prepared **out, void *stream);

/* Count n prepared samples (MIA_FAST) in one persistent launch: the same
 * outputs as mia_count / mia_count_batch, without the sort. */
int mia_count_prepared(mia_ctx *ctx, int32_t n, mia_prepared *const *samples, const mia_bins *bins,
                       const mia_counts *outs, const mia_tuning *tuning, mia_stats *stats,
                       void *stream);

/* Free a prepared sample (NULL is ignored). */
void mia_release_prepared(mia_ctx *ctx, mia_prepared *sample);

/* One lightcone measurement: a position sample and a shape sample on the sky
 * (RA, DEC in degrees, comoving distance chi), as the reference's lightcone
 * counters prepare them (src/measureia/measure_w_lightcone.py:82-139,
 * measure_m_lightcone.py:82-139).  All pointers are device pointers.
 * The host supplies chi already multiplied by h when over_h, cos(dec/180*pi)
 * of every position (numpy's value), and per shape the rotated ellipticity
 *   c2 = -e cos(2 phi_axis),  s2 = -e sin(2 phi_axis)
 * (e = |e1 + i e2|, phi_axis the axis angle the reference derives from e1, e2),
 * so that e+ = c2 cos 2phi_sep + s2 sin 2phi_sep, ex = s2 cos 2phi_sep - c2 sin 2phi_sep. */
typedef struct mia_lc_sample {
    const double *ra, *dec, *chi;  /* positions, n_pos each                                  */
    const double *cos_dec;         /* n_pos: cos(dec / 180 * pi)                              */
    const double *w;               /* n_pos weights, or NULL for unit weights                 */
    const int32_t *patch;          /* n_pos jackknife patch in [0, n_patches), or NULL        */
    const double *ra_s, *dec_s, *chi_s; /* shape sample, n_shape each                         */
    const double *w_s;             /* n_shape weights, or NULL                                */
    const double *c2, *s2;         /* n_shape rotated ellipticity; both NULL: pair counts only */
    const int32_t *patch_s;        /* n_shape patch, or NULL                                  */
    int64_t n_pos;
    int64_t n_shape;
    double proj_scale;             /* h with over_h, else 1: multiplies (dx, dy) before r_p    */
    double rp2_cut;                /* MIA_RMU: pairs need fl(dx'^2 + dy'^2) > rp2_cut (the
                                      squared-separation threshold of rp_cut)                 */
    int32_t n_patches;             /* 0: no jackknife tallies; at most 4096                    */
    int32_t reserved;
} mia_lc_sample;

/* Count one lightcone measurement: replaces the per-position loops of
 * _measure_xi_rp_pi_lightcone_brute / _count_pairs_xi_rp_pi_lightcone_brute
 * (measure_w_lightcone.py:45-343), _measure_xi_r_mur_lightcone_brute /
 * _count_pairs_xi_r_mur_lightcone_brute (measure_m_lightcone.py:45-360), and
 * the per-patch reruns of _measure_jackknife_realisations_lightcone[_multiprocessing]
 * (measure_jackknife.py:59-171, 265-483).  Per ordered pair (shape s, position n):
 *   pi = chi_s - chi_n
 *   dx = ((ra_s - ra_n)/180*pi)*chi_n*cos_dec_n,  dy = ((dec_s - dec_n)/180*pi)*chi_n
 *   MIA_RPPI: s = fl((h dx)^2 + (h dy)^2) binned on bins.r2_edges, pi on x2_edges
 *   MIA_RMU : r^2 = fl(fl(dx^2 + dy^2) + pi^2), mu = pi/r, and fl((h dx)^2+(h dy)^2) > rp2_cut
 * Outputs (device, zeroed and filled), K = n_patches, B = n_r*n_2:
 *   npairs [(1+K)*B] integer pair counts, [0] all pairs, [1+i] pairs touching patch i
 *   dd     [(1+K)*B] the same tallies of w_n*w_s (required)
 *   splus  [(1+K)*B] tallies of w_n*w_s*e+   (required with c2/s2, else ignored)
 *   scross [(1+K)*B] tallies of w_n*w_s*ex   (required with c2/s2, else ignored)
 * Realisation i (the sample without patch i) is total - tally[i].  Tuning
 * (optional): slabs_per_side / columns_per_side = DEC rows / RA columns of
 * the angular grid, cell_shapes = target shapes per cell, slabs_per_block =
 * depth of a position block in eighths of the chi window, block_shapes =
 * positions per block (<= 64), item_positions = shapes per work item, waves. */
int mia_count_lightcone(mia_ctx *ctx, const mia_lc_sample *sample, const mia_bins *bins,
                        const mia_counts *out, const mia_tuning *tuning, mia_stats *stats,
                        void *stream);

/* Jackknife region id of each object (measure_IA_base.py:404-452): strict
 * i*L_sub < x < (i+1)*L_sub on every axis, region i*L^2 + j*L + k, else 0.
 * xyz device n x 3, region device n int32; 1 <= jk_per_side <= 1024. */
int mia_region_ids(mia_ctx *ctx, const double *xyz, int64_t n, double jk_sub,
                   int32_t jk_per_side, int32_t *region, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* MEASURE_IA_AMD_H */
