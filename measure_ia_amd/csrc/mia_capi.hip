// The C ABI of include/measure_ia_amd.h: argument checks, dispatch, errors.
#include "mia_common.hpp"

#include <chrono>
#include <cmath>

struct mia_ctx {
    mia::Ctx c;
};
struct mia_prepared;

namespace {

thread_local std::string g_create_error;

__global__ void k_u64_to_f64(const uint64_t *in, double *out, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (double)in[i];
}

void check_sample(const mia_sample *s) {
    MIA_REQUIRE(s != nullptr, "sample is NULL");
    MIA_REQUIRE(s->n_pos >= 0 && s->n_shape >= 0, "negative sample size");
    MIA_REQUIRE(s->n_pos < (1ll << 31) && s->n_shape < (1ll << 31), "sample larger than 2^31 objects");
    MIA_REQUIRE(s->n_pos == 0 || s->pos != nullptr, "pos is NULL");
    MIA_REQUIRE(s->n_shape == 0 || (s->shp && s->axis && s->e), "shape arrays are NULL");
    MIA_REQUIRE(s->boxsize > 0 && std::isfinite(s->boxsize), "boxsize must be positive");
    MIA_REQUIRE(s->los >= 0 && s->los <= 2, "los must be 0, 1 or 2");
    MIA_REQUIRE(s->jk_per_side >= 0 && s->jk_per_side <= 64, "jk_per_side out of range");
    MIA_REQUIRE(s->jk_per_side == 0 || s->jk_sub > 0, "jk_sub must be positive with jackknife");
    MIA_REQUIRE(s->periodic == 0 || s->periodic == 1, "periodic must be 0 or 1");
}

void check_bins(const mia_bins *b) {
    MIA_REQUIRE(b != nullptr, "bins is NULL");
    MIA_REQUIRE(b->mode == MIA_RPPI || b->mode == MIA_RMU, "unknown bin mode");
    MIA_REQUIRE(b->n_r >= 1 && b->n_r <= mia::kMaxBinsR, "n_r out of range");
    MIA_REQUIRE(b->n_2 >= 1 && b->n_2 <= mia::kMaxBins2, "n_2 out of range");
    MIA_REQUIRE(b->r2_edges && b->x2_edges, "bin edges are NULL");
    for (int k = 0; k < b->n_r; ++k)
        MIA_REQUIRE(b->r2_edges[k] <= b->r2_edges[k + 1], "r2_edges not sorted");
    MIA_REQUIRE(b->r2_edges[0] > 0, "r2_edges[0] must be positive");
    int nx = b->mode == MIA_RPPI ? b->n_2 + 1 : b->n_2;
    for (int k = 0; k + 1 < nx; ++k)
        MIA_REQUIRE(b->x2_edges[k] <= b->x2_edges[k + 1], "x2_edges not sorted");
}

void check_out(const mia_counts *o, bool weighted) {
    MIA_REQUIRE(o != nullptr, "counts is NULL");
    MIA_REQUIRE(o->npairs && o->splus && o->scross, "npairs/splus/scross are required");
    MIA_REQUIRE(!weighted || o->dd, "dd is required with weights");
}

template <class F>
int guard(mia_ctx *ctx, F &&f) {
    if (!ctx) return MIA_E_ARG;
    try {
        f();
        ctx->c.last_error.clear();
        return MIA_OK;
    } catch (const mia::Error &e) {
        ctx->c.last_error = e.msg;
        return e.code;
    } catch (const std::exception &e) {
        ctx->c.last_error = e.what();
        return MIA_E_HIP;
    }
}

void unit_dd(const mia_counts &o, bool weighted, int64_t nb, hipStream_t st) {
    if (!weighted && o.dd && nb > 0) {
        hipLaunchKernelGGL(k_u64_to_f64, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, st,
                           (const uint64_t *)o.npairs, o.dd, nb);
        MIA_HIP(hipGetLastError());
    }
}

}  // namespace

extern "C" {

int mia_abi_version(void) { return MIA_ABI_VERSION; }

int mia_create(int32_t device, mia_ctx **out) {
    if (!out) return MIA_E_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) {
        g_create_error = "no such HIP device";
        return MIA_E_ARG;
    }
    if (hipSetDevice(device) != hipSuccess) return MIA_E_HIP;
    mia_ctx *c = new (std::nothrow) mia_ctx();
    if (!c) return MIA_E_NOMEM;
    c->c.device = device;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess &&
        cus > 0)
        c->c.cu_count = cus;
    try {
        mia::stage_init(c->c);
    } catch (const mia::Error &e) {
        g_create_error = e.msg;
        mia::stage_release(c->c);
        delete c;
        return e.code;
    }
    *out = c;
    return MIA_OK;
}

void mia_destroy(mia_ctx *ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->c.device);
    mia::stage_release(ctx->c);
    delete ctx;
}

const char *mia_last_error(const mia_ctx *ctx) {
    if (!ctx) return g_create_error.c_str();
    return ctx->c.last_error.c_str();
}

int mia_count(mia_ctx *ctx, const mia_sample *sample, const mia_bins *bins, const mia_counts *out,
              int32_t method, const mia_tuning *tuning, mia_stats *stats, void *stream) {
    return guard(ctx, [&] {
        check_sample(sample);
        check_bins(bins);
        bool weighted = sample->pos_w != nullptr || sample->shp_w != nullptr;
        check_out(out, weighted);
        MIA_REQUIRE(method == MIA_PLAIN || method == MIA_FAST, "unknown method");
        MIA_HIP(hipSetDevice(ctx->c.device));
        hipStream_t st = (hipStream_t)stream;
        if (stats) std::memset(stats, 0, sizeof(*stats));
        // binnings or jackknife counts the batched kernel cannot hold go to the
        // plain kernel, which has no LDS histogram or planner key space
        if (method == MIA_FAST && !mia::fast_path_fits(*bins, weighted, sample->jk_per_side))
            method = MIA_PLAIN;
        if (method == MIA_FAST) {
            mia::stage_count(ctx->c, 1, sample, *bins, out, tuning, stats, st, mia::STAGE_EAGER);
            return;
        }
        // the plain comparator: global-atomic histograms in context memory,
        // copied to the caller's outputs at the end
        mia::Grid g = mia::make_grid(*sample, *bins, tuning);
        const int K = g.jk > 0 ? g.jk * g.jk * g.jk : 0;
        const int64_t B = (int64_t)bins->n_r * bins->n_2, nb = (1 + K) * B;
        double *acc = ctx->c.arena.get<double>("plain/acc", (size_t)(3 * nb + 2 * B + 4 * K));
        mia_counts a{};
        a.npairs = (uint64_t *)acc;
        a.dd = acc + nb;
        a.splus = acc + 2 * nb;
        a.scross = acc + 3 * nb;
        a.variance = out->variance ? acc + 3 * nb + B : nullptr;
        a.region_stats = out->region_stats ? acc + 3 * nb + 2 * B : nullptr;
        mia::fill<double>(acc, 0.0, 3 * nb + 2 * B + 4 * K, st);
        mia::Sorted pos, shp;
        int *oor = ctx->c.arena.get<int>("oor", 1);
        mia::Events ev(2);
        MIA_HIP(hipEventRecord(ev[0], st));
        mia::fill<int32_t>(oor, 0, 1, st);
        mia::prepare_sample(ctx->c, *sample, g, pos, shp, a.region_stats, oor, true, st, "s0");
        MIA_HIP(hipEventRecord(ev[1], st));
        mia::check_prepared(oor, 1, st);
        mia::count_plain(ctx->c, *sample, *bins, g, pos, shp, a, weighted, st, stats);
        unit_dd(a, weighted, nb, st);
        auto copy = [&](void *dst, const void *src, int64_t words) {
            if (dst && words > 0)
                MIA_HIP(hipMemcpyAsync(dst, src, 8 * words, hipMemcpyDeviceToDevice, st));
        };
        copy(out->npairs, a.npairs, nb);
        copy(out->dd, a.dd, nb);
        copy(out->splus, a.splus, nb);
        copy(out->scross, a.scross, B);
        copy(out->variance, a.variance, B);
        copy(out->region_stats, a.region_stats, 4 * K);
        MIA_HIP(hipStreamSynchronize(st));
        float ms = ev.ms(0, 1);
        if (stats) {
            stats->ms_prepare = ms;
            stats->columns_per_side = g.nc;
            stats->slabs_per_side = g.nz;
            stats->method = MIA_PLAIN;
        }
    });
}

int mia_count_batch(mia_ctx *ctx, int32_t n, const mia_sample *samples, const mia_bins *bins,
                    const mia_counts *outs, const mia_tuning *tuning, mia_stats *stats,
                    void *stream) {
    return guard(ctx, [&] {
        MIA_REQUIRE(n >= 1 && samples && outs, "empty batch");
        MIA_REQUIRE(n <= 4096, "more than 4096 samples in one batch");
        check_bins(bins);
        MIA_HIP(hipSetDevice(ctx->c.device));
        hipStream_t st = (hipStream_t)stream;
        if (stats) std::memset(stats, 0, sizeof(*stats));
        for (int i = 0; i < n; ++i) {
            check_sample(&samples[i]);
            bool weighted = samples[i].pos_w != nullptr || samples[i].shp_w != nullptr;
            check_out(&outs[i], weighted);
            if (!mia::fast_path_fits(*bins, true, samples[i].jk_per_side))
                throw mia::Error{MIA_E_UNSUPPORTED,
                                 "binning or jackknife count too large for the batched kernel; "
                                 "count the samples one by one (mia_count falls back to MIA_PLAIN)"};
        }
        const bool graph = !(tuning && tuning->graph < 0);
        mia::stage_count(ctx->c, n, samples, *bins, outs, tuning, stats, st,
                         graph ? mia::STAGE_GRAPH : mia::STAGE_EAGER);
    });
}

int mia_count_noalloc(mia_ctx *ctx, const mia_sample *sample, const mia_bins *bins, const mia_counts *out,
                      const mia_tuning *tuning, void *stream) {
    return guard(ctx, [&] {
        check_sample(sample);
        check_bins(bins);
        bool weighted = sample->pos_w != nullptr || sample->shp_w != nullptr;
        check_out(out, weighted);
        if (!mia::fast_path_fits(*bins, weighted, sample->jk_per_side))
            throw mia::Error{MIA_E_UNSUPPORTED, "binning or jackknife count too large for MIA_FAST"};
        MIA_HIP(hipSetDevice(ctx->c.device));
        mia::stage_count(ctx->c, 1, sample, *bins, out, tuning, nullptr, (hipStream_t)stream,
                         mia::STAGE_STATIC);
    });
}

int64_t mia_noalloc_workspace_bytes(void) { return mia::stage_static_bytes(); }

int mia_prepare_sample(mia_ctx *ctx, const mia_sample *sample, const mia_bins *bins,
                       const mia_tuning *tuning, mia_prepared **out, void *stream) {
    return guard(ctx, [&] {
        MIA_REQUIRE(out != nullptr, "out is NULL");
        *out = nullptr;
        check_sample(sample);
        check_bins(bins);
        MIA_HIP(hipSetDevice(ctx->c.device));
        *out = (mia_prepared *)mia::stage_prepare(ctx->c, *sample, *bins, tuning, (hipStream_t)stream);
    });
}

int mia_count_prepared(mia_ctx *ctx, int32_t n, mia_prepared *const *samples, const mia_bins *bins,
                       const mia_counts *outs, const mia_tuning *tuning, mia_stats *stats,
                       void *stream) {
    return guard(ctx, [&] {
        MIA_REQUIRE(n >= 1 && samples && outs, "empty batch");
        MIA_REQUIRE(n <= 4096, "more than 4096 samples in one batch");
        check_bins(bins);
        for (int i = 0; i < n; ++i) {
            MIA_REQUIRE(samples[i] != nullptr, "prepared sample is NULL");
            const mia::SortedSample &x = mia::prepared_view((const mia::PreparedSample *)samples[i]);
            check_out(&outs[i], x.weighted != 0);
            if (!mia::fast_path_fits(*bins, true, x.g.jk))
                throw mia::Error{MIA_E_UNSUPPORTED, "binning or jackknife count too large for MIA_FAST"};
        }
        MIA_HIP(hipSetDevice(ctx->c.device));
        if (stats) std::memset(stats, 0, sizeof(*stats));
        mia::stage_count_prepared(ctx->c, n, (const mia::PreparedSample *const *)samples, *bins, outs, tuning,
                                  stats, (hipStream_t)stream);
    });
}

void mia_release_prepared(mia_ctx *ctx, mia_prepared *sample) {
    if (ctx) (void)hipSetDevice(ctx->c.device);
    mia::stage_free_prepared((mia::PreparedSample *)sample);
}

int mia_count_lightcone(mia_ctx *ctx, const mia_lc_sample *s, const mia_bins *bins,
                        const mia_counts *out, const mia_tuning *tuning, mia_stats *stats,
                        void *stream) {
    return guard(ctx, [&] {
        MIA_REQUIRE(s != nullptr, "sample is NULL");
        check_bins(bins);
        MIA_REQUIRE(s->n_pos >= 0 && s->n_shape >= 0, "negative sample size");
        MIA_REQUIRE(s->n_pos < (1ll << 31) && s->n_shape < (1ll << 31), "sample larger than 2^31 objects");
        MIA_REQUIRE(s->n_pos == 0 || (s->ra && s->dec && s->chi && s->cos_dec), "position arrays are NULL");
        MIA_REQUIRE(s->n_shape == 0 || (s->ra_s && s->dec_s && s->chi_s), "shape arrays are NULL");
        MIA_REQUIRE((s->c2 == nullptr) == (s->s2 == nullptr), "c2 and s2 must both be given or both NULL");
        MIA_REQUIRE(s->n_patches >= 0 && s->n_patches <= 4096, "n_patches out of range");
        MIA_REQUIRE(s->n_patches == 0 || s->n_pos == 0 || s->patch, "patch is NULL");
        MIA_REQUIRE(s->n_patches == 0 || s->n_shape == 0 || s->patch_s, "patch_s is NULL");
        MIA_REQUIRE(s->proj_scale > 0 && std::isfinite(s->proj_scale), "proj_scale must be positive");
        MIA_REQUIRE(!std::isnan(s->rp2_cut), "rp2_cut is NaN");
        MIA_REQUIRE(out && out->npairs && out->dd, "npairs and dd are required");
        MIA_REQUIRE(s->c2 == nullptr || (out->splus && out->scross),
                    "splus and scross are required with shapes");
        if (!mia::lightcone_fits(*bins))
            throw mia::Error{MIA_E_UNSUPPORTED, "binning too large for the lightcone kernel"};
        MIA_HIP(hipSetDevice(ctx->c.device));
        hipStream_t st = (hipStream_t)stream;
        mia::count_lightcone(ctx->c, *s, *bins, *out, tuning, st, stats);
        MIA_HIP(hipStreamSynchronize(st));
    });
}

int mia_region_ids(mia_ctx *ctx, const double *xyz, int64_t n, double jk_sub, int32_t jk_per_side,
                   int32_t *region, void *stream) {
    return guard(ctx, [&] {
        MIA_REQUIRE(n >= 0 && (n == 0 || (xyz && region)), "bad arrays");
        MIA_REQUIRE(jk_per_side >= 1 && jk_sub > 0, "bad jackknife split");
        // region ids i*L^2 + j*L + k must fit int32
        MIA_REQUIRE(jk_per_side <= 1024, "jk_per_side above 1024 (region ids would overflow int32)");
        MIA_HIP(hipSetDevice(ctx->c.device));
        mia::region_ids(ctx->c, xyz, n, jk_sub, jk_per_side, region, (hipStream_t)stream);
        MIA_HIP(hipStreamSynchronize((hipStream_t)stream));
    });
}

}  // extern "C"
