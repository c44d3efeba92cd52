"""Benchmark of the MI355X pair-counting engine (one JSON line on stdout).

Workloads (synthetic, fixed seeds, built by measure_ia_amd/synth.py):

  clustered1M (default)  1,000,000 positions in a periodic box of L = 205 Mpc/h
      drawn as 5,000 Gaussian clusters (sigma = 1, Poisson(150) members) plus a
      uniform background; a random 200,000-galaxy shape subset with ellipticity
      aligned to the cluster centre (A = 0.1, noise N(0, 0.2)); w_gg + w_g+
      binning: 10 log r_p bins over 0.1-20, 20 linear pi bins over |pi| < 60,
      27 jackknife regions.
  uniform20k     20,000 uniform galaxies (shapes = positions), same binning.
  multipoles2M   2,000,000 positions / 500,000 shapes as above with 10,000
      clusters; (r, mu) binning: 15 log r bins over 0.1-60, 20 mu bins, 64
      jackknife regions.
  batch12        12 samples cut from one 3,000,000-galaxy clustered box, sizes
      log-uniform in 10,000-800,000, shapes = 30 % of each, weights
      U(0.5, 1.5); w_gg + w_g+ binning as clustered1M, 27 jackknife regions
      each; all samples counted by one pair-kernel launch (mia_count_batch).
  large20M       L = 500, 20,000,000 positions (20,000 clusters, sigma = 1.5,
      30 % background) / 5,000,000 shapes (A = 0.05, noise N(0, 0.25));
      15 log r_p bins over 0.1-100, 40 pi bins over |pi| < 100, 125 regions.
  lightcone1M    a lightcone wedge RA 0-60, DEC -15..15 deg, comoving distance
      300-2000 Mpc: 1,000,000 positions (half in Gaussian clusters of sigma 1.5)
      and a 300,000-galaxy shape subset with noise ellipticities, weights
      U(0.5, 1.5); the S+D count of measure_xi_w on lightcones (10 log r_p
      bins over 0.1-20, 20 pi bins over |pi| < 60) with 27 RA x DEC jackknife
      patches (mia_count_lightcone).  Its CPU baseline is the oracle's
      restatement of the reference's per-position brute loop
      (measure_w_lightcone.py:141-192) on a seeded sample of positions.
  lightcone_rr4M the same wedge filled with 4,000,000 uniform randoms counted
      against themselves (the RR pair count of the lightcone estimators,
      _count_pairs_xi_rp_pi_lightcone_brute), unit weights, 27 patches.

A step is one full jackknife count of the workload (region ids, cell sort,
work plan, pair kernel, histogram readback) with the inputs already in HBM.
``value`` = in-domain (shape, position) pairs per second over the timed
steps.  The pair kernel's own time comes from HIP events the library records
on the launch stream.  Beside it: the plain FP64 GPU path (one thread per
shape, global atomics) timed on a seeded shape subset (2%, at most about
4e9 pairs), the end-to-end
measure_xi_w call (host prep + count + estimators + HDF5), and the CPU
baseline: the oracle's restatement of the reference's KDTree +
multiprocessing counter timed on a seeded sample of shape chunks.
"""
import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6      # MI355X FP64 vector (= FP32 without packed instructions)
PEAK_FP32_PACKED_TFLOPS = 157.3
PEAK_HBM_GBPS = 8000.0       # MI355X HBM3E
# algorithmic bytes of the counting sort per object: classify reads the
# coordinates (24 B; shapes also e, 8 B) and writes (key, rank) (8 B); scatter
# reads (key, rank), the key's start (4 B) and the coordinates again and writes
# the 32-byte position record or the 64-byte shape record (axis 16 B, e 8 B
# read again); a weight array adds 8 B to each read of the object.
SORT_BYTES = {"pos": 100, "shp": 164, "w_pos": 8, "w_shp": 16}
FLOP_PER_PAIR = {"rppi": 28, "rmu": 32}   # algorithmic FLOP per in-domain pair

W_BINS = dict(mode="rppi", r=(0.1, 20.0), nr=10, n2=20, pi_max=60.0, jk=3)
WORKLOADS = {
    "clustered1M": dict(W_BINS, kind="single", cat="clustered1M",
                        desc="1,000,000 clustered positions / 200,000 aligned shapes, L=205, 27 JK, "
                             "10 log r_p bins 0.1-20, 20 pi bins |pi|<60"),
    "uniform20k": dict(W_BINS, kind="single", cat="uniform20k",
                       desc="20,000 uniform galaxies (shapes = positions), L=205, 27 JK"),
    "multipoles2M": dict(kind="single", cat="multipoles2M", mode="rmu", r=(0.1, 60.0), nr=15, n2=20,
                         pi_max=None, jk=4,
                         desc="2,000,000 clustered positions / 500,000 shapes, L=205, 64 JK, "
                              "15 log r bins 0.1-60, 20 mu bins"),
    "batch12": dict(W_BINS, kind="batch",
                    desc="12 weighted samples (10k-800k, shapes 30%) from a 3,000,000-galaxy clustered "
                         "box, L=205, 27 JK each, one batched launch"),
    "large20M": dict(kind="single", cat="large20M", mode="rppi", r=(0.1, 100.0), nr=15, n2=40,
                     pi_max=100.0, jk=5,
                     desc="20,000,000 clustered positions / 5,000,000 shapes, L=500, 125 JK, "
                          "15 log r_p bins 0.1-100, 40 pi bins |pi|<100"),
    "lightcone1M": dict(kind="lightcone", mode="rppi", r=(0.1, 20.0), nr=10, n2=20, pi_max=60.0, jk=0,
                        patches=(9, 3), n_pos=1_000_000, n_shape=300_000,
                        sky=dict(ra=(0.0, 60.0), dec=(-15.0, 15.0), chi=(300.0, 2000.0)),
                        desc="lightcone wedge RA 0-60, DEC -15..15, chi 300-2000 Mpc: 1,000,000 positions / "
                             "300,000 shapes, weighted, 27 patches, 10 log r_p bins 0.1-20, 20 pi bins |pi|<60"),
    "lightcone_rr4M": dict(kind="lightcone", mode="rppi", r=(0.1, 20.0), nr=10, n2=20, pi_max=60.0, jk=0,
                           patches=(9, 3), n_pos=4_000_000, n_shape=4_000_000, randoms=True,
                           sky=dict(ra=(0.0, 60.0), dec=(-15.0, 15.0), chi=(300.0, 2000.0)),
                           desc="lightcone wedge RA 0-60, DEC -15..15, chi 300-2000 Mpc: 4,000,000 uniform "
                                "randoms x themselves (RR pair count), 27 patches, 10 log r_p bins 0.1-20, "
                                "20 pi bins |pi|<60"),
}
FLOP_PER_PAIR_LC = 30   # lightcone (r_p, pi) pair: see the lightcone section of the design notes
CPU_SHARE = 16           # host cores per GPU on an MI355X node
PRECISION = ("FP64 separations and e+ in the reference's own rounding, FP32 bin guesses with an exact "
             "FP64 recheck near every bin edge, FP64 tile sums")


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def binning_for(w, L):
    from measure_ia_amd import binning
    r_bins = np.logspace(np.log10(w["r"][0]), np.log10(w["r"][1]), w["nr"] + 1)
    if w["mode"] == "rppi":
        pi_max = w["pi_max"] if w["pi_max"] is not None else L / 2.0
        sb = np.linspace(-pi_max, pi_max, w["n2"] + 1)
        th = binning.rppi_thresholds(r_bins, sb, *w["r"])
    else:
        sb = np.linspace(-1, 1, w["n2"] + 1)
        th = binning.rmu_thresholds(r_bins, sb, *w["r"])
    return th, r_bins, sb


def build_catalogues(w):
    from measure_ia_amd import synth
    if w["kind"] == "batch":
        return synth.uneven_samples(12, parent=3_000_000, L=205.0, lo=10_000, hi=800_000,
                                    shape_frac=0.3, seed=3)
    return [synth.config(w["cat"])]


def measured_traffic(workload, kernel="k_pairs"):
    """HBM bytes per pair-kernel launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE passes."""
    fn = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_traffic.json")
    if not os.path.exists(fn):
        return None
    with open(fn) as f:
        return float(json.load(f)["bytes_per_launch"])


def cpu_baseline(cat, w, r_bins, sb, pairs_per_shape):
    """Oracle tree + multiprocessing counter on a seeded sample of shape chunks.

    Chunks hold the reference's default 1000 shapes; where one such chunk would
    take a core much longer than ~20 s (dense multipole searches), chunks are
    made smaller so that one round of chunks over the workers stays near 20 s.
    """
    from oracle import prep as oprep
    from oracle import tree
    # the pool is held to the node's CPU share for one GPU (16 cores); the
    # affinity mask can list more cores than that share
    workers = min(CPU_SHARE, tree.default_workers())
    core_rate = 3e6  # in-domain pairs per second per core, measured for this restatement
    chunk = int(np.clip(20.0 * core_rate / max(pairs_per_shape, 1.0), 10, 1000))
    budget_chunks = 3 * workers if chunk == 1000 else workers
    pr = oprep.prepare(cat.data)
    L = cat.boxsize
    jp, js = oprep.jackknife_region_indices(pr["positions"], pr["shapes"], L, w["jk"])
    n_chunks = (len(pr["shapes"]) + chunk - 1) // chunk
    rng = np.random.default_rng(12345)
    pick = sorted(rng.choice(n_chunks, size=min(budget_chunks, n_chunks), replace=False).tolist())
    out = tree.count_pairs(pr, r_bins, sb, w["r"][0], w["r"][1], L, mode=w["mode"], jk_pos=jp,
                           jk_shape=js, num_box=w["jk"] ** 3, chunk_size=chunk, num_nodes=workers,
                           chunk_subset=pick)
    rate = out["npairs"] / out["seconds"]
    return dict(value=rate, unit="pairs/s", cores=workers, kind="port", cpu_model=cpu_model(),
                cores_in_affinity_mask=tree.default_workers(),
                sample=f"{len(pick)} of {n_chunks} seeded {chunk}-shape chunks ({out['npairs']} pairs, "
                       f"{out['seconds']:.1f} s), KDTree + numpy + Pool restatement of "
                       f"measure_w_box_jk.py:543-857",
                seconds_full_estimate=out["seconds"] * n_chunks / len(pick))


def lightcone_inputs(w):
    """The workload's arrays (host) and their patch labels."""
    from measure_ia_amd import synth
    from measure_ia_amd.engine import LightconeArrays
    from measure_ia_amd.lightcone import rotated_ellipticity
    if w.get("randoms"):
        ra, dec, chi = synth.lightcone_points(np.random.default_rng(13), w["n_pos"], clustered_frac=0.0,
                                              **w["sky"])
        s = {"RA": ra, "DEC": dec, "chi": chi, "RA_shape_sample": ra, "DEC_shape_sample": dec,
             "chi_shape_sample": chi, "weight": np.ones(len(ra)), "weight_shape_sample": np.ones(len(ra))}
    else:
        s = synth.lightcone(w["n_pos"], w["n_shape"], seed=11, weights=True, **w["sky"])
    nra, ndec = w["patches"]

    def patch(ra, dec):
        i = np.clip(((ra - w["sky"]["ra"][0]) / (w["sky"]["ra"][1] - w["sky"]["ra"][0]) * nra).astype(int), 0, nra - 1)
        j = np.clip(((dec - w["sky"]["dec"][0]) / (w["sky"]["dec"][1] - w["sky"]["dec"][0]) * ndec).astype(int),
                    0, ndec - 1)
        return (i * ndec + j).astype(np.int32)

    c2, s2 = rotated_ellipticity(s["e1"], s["e2"]) if "e1" in s else (None, None)
    lc = LightconeArrays(ra=s["RA"], dec=s["DEC"], chi=s["chi"], cos_dec=np.cos(s["DEC"] / 180 * np.pi),
                         ra_s=s["RA_shape_sample"], dec_s=s["DEC_shape_sample"], chi_s=s["chi_shape_sample"],
                         weight=s["weight"], weight_shape=s["weight_shape_sample"], c2=c2, s2=s2,
                         patch=patch(s["RA"], s["DEC"]),
                         patch_s=patch(s["RA_shape_sample"], s["DEC_shape_sample"]), n_patches=nra * ndec)
    return s, lc


def lightcone_cpu_baseline(s, w, r_bins, sb, n_sample):
    """The oracle's per-position brute loop (the reference's lightcone algorithm,
    one core) on a seeded sample of positions against every shape."""
    from oracle import lightcone as olc
    rng = np.random.default_rng(12345)
    pick = np.sort(rng.choice(len(s["RA"]), n_sample, replace=False))
    t = time.perf_counter()
    out = olc.count(w["mode"], s["RA"][pick], s["DEC"][pick], s["chi"][pick], s["weight"][pick],
                    s["RA_shape_sample"], s["DEC_shape_sample"], s["chi_shape_sample"],
                    s["weight_shape_sample"], r_bins, sb, w["r"][0], w["r"][1], e1=s.get("e1"),
                    e2=s.get("e2"))
    sec = time.perf_counter() - t
    n = int(out["NP"][0].sum())
    return dict(value=n / sec, unit="pairs/s", cores=1, kind="port", cpu_model=cpu_model(),
                sample=f"{n_sample} seeded positions x all {len(s['RA_shape_sample'])} shapes ({n} pairs, "
                       f"{sec:.1f} s), numpy restatement of measure_w_lightcone.py:141-192",
                seconds_full_estimate=sec * len(s["RA"]) / n_sample)


def run_lightcone(a, w):
    import torch
    from measure_ia_amd.engine import get_engine
    t0 = time.perf_counter()
    s, lc = lightcone_inputs(w)
    th, r_bins, sb = binning_for(w, None)
    log(f"lightcone inputs built in {time.perf_counter() - t0:.1f} s")
    eng = get_engine(0)
    dev = torch.device("cuda", 0)
    for k in ("ra", "dec", "chi", "cos_dec", "ra_s", "dec_s", "chi_s", "weight", "weight_shape", "c2", "s2"):
        if getattr(lc, k) is not None:
            setattr(lc, k, torch.from_numpy(np.ascontiguousarray(getattr(lc, k))).to(dev))
    for k in ("patch", "patch_s"):
        setattr(lc, k, torch.from_numpy(getattr(lc, k)).to(dev))
    torch.cuda.synchronize()
    for _ in range(a.warmup):
        res = eng.count_lightcone(lc, th)
    torch.cuda.synchronize()
    kst = []
    t1 = time.perf_counter()
    for _ in range(a.steps):
        res = eng.count_lightcone(lc, th)
        kst.append(res.stats)
    torch.cuda.synchronize()
    ms_step = (time.perf_counter() - t1) / a.steps * 1e3
    pairs = int(res.npairs[0].sum())
    value = pairs / (ms_step / 1e3)
    k_ms = float(np.mean([x["ms_pairs"] for x in kst]))
    achieved = pairs * FLOP_PER_PAIR_LC / (k_ms / 1e3) / 1e12
    log(f"lightcone: {ms_step:.2f} ms/step, pair kernel {k_ms:.2f} ms, {pairs} pairs")
    cpu = None
    if not a.no_cpu:
        # about 10 s of one host core of an MI355X node: the loop's cost is one
        # pass over all shapes per position
        cpu = lightcone_cpu_baseline(s, w, r_bins, sb, n_sample=int(np.clip(1.2e10 / w["n_shape"] / 8, 8, 8000)))
        log(f"cpu baseline: {cpu['value']:.3e} pairs/s on 1 core")
    st = res.stats
    line = {
        "metric": "pairs/s (lightcone S+D count, 27 patches)", "value": value, "unit": "pairs/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": ms_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{a.workload}: {w['desc']}", "mode": w["mode"], "patches": lc.n_patches,
                   "positions": w["n_pos"], "shapes": w["n_shape"], "pairs": pairs},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS, "traffic": measured_traffic(a.workload, "k_lc_pairs"),
                     "kernel": "k_lc_pairs",
                     "kernel_ms": k_ms, "candidate_efficiency": pairs / max(1, st["candidates"]),
                     "model": f"{FLOP_PER_PAIR_LC} FLOP per in-domain pair; FP64 vector peak"},
        "cpu_baseline": cpu,
        "speedup_vs_cpu": (value / cpu["value"]) if cpu else None,
        "engine": dict(st),
    }
    print(json.dumps(line))


def weighted_prep(p):
    return not (np.all(p.weight == 1.0) and np.all(p.weight_shape == 1.0))


def sort_model_bytes(p):
    w = weighted_prep(p)
    return ((SORT_BYTES["pos"] + w * SORT_BYTES["w_pos"]) * p.n_pos +
            (SORT_BYTES["shp"] + w * SORT_BYTES["w_shp"]) * p.n_shape)


def sort_only(a, eng, samples, preps, th, batch):
    """The cell sort alone, --steps times after --warmup: a batch through
    mia_count_batch with plain launches (the pair count runs too; its stats
    give the sort's own event time), a single sample through
    mia_prepare_sample (timed on the host around a synchronised call)."""
    import torch
    times = []
    for k in range(a.warmup + a.steps):
        if batch:
            ms = eng.count_batch(samples, th, tuning={"graph": -1})[0].stats["ms_prepare"]
        else:
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            h = eng.prepare(samples[0], th)
            torch.cuda.synchronize()
            ms = (time.perf_counter() - t0) * 1e3
            h.release()
        if k >= a.warmup:
            times.append(ms)
        log(f"sort {k}: {ms:.3f} ms")
    nbytes = sum(sort_model_bytes(p) for p in preps)
    ms = float(np.median(times))
    print(json.dumps({"sort_only": a.workload, "ms_median": ms, "runs_ms": times, "model_bytes": nbytes,
                      "model_GBps": nbytes / (ms / 1e3) / 1e9,
                      "timing": "event time of the sort inside the count" if batch else
                                "host time around mia_prepare_sample + synchronise (includes the launch "
                                "and parameter upload)"}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="clustered1M", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain GPU comparator")
    ap.add_argument("--no-e2e", action="store_true", help="skip the end-to-end measure_xi_w call")
    ap.add_argument("--sort-only", action="store_true",
                    help="only run the cell sort of the workload's samples --steps times (single samples "
                         "through mia_prepare_sample, batches through mia_count_batch with plain launches) "
                         "and print its timings; for profiling the sort kernels")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run three 512 MiB device copies (the reference kernel for the FETCH_SIZE / "
                         "WRITE_SIZE scale factors) and exit")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: the counter runs on one GPU only (asked for --gpus {a.gpus}); it "
              f"would shard by sample or shape-cell slab with one histogram all-reduce, which is not "
              f"built.", file=sys.stderr)
        sys.exit(2)
    import torch
    if a.copy_calibration:
        n = 64 * 1024 * 1024
        x = torch.ones(n, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": n * 8, "copies": 3}))
        return
    from measure_ia_amd import hostprep
    from measure_ia_amd.engine import SampleArrays, get_engine

    w = WORKLOADS[a.workload]
    if w["kind"] == "lightcone":
        run_lightcone(a, w)
        return
    batch = w["kind"] == "batch"
    jk = w["jk"]
    t0 = time.perf_counter()
    cats = build_catalogues(w)
    L = cats[0].boxsize
    th, r_bins, sb = binning_for(w, L)
    preps = [hostprep.prepare(c.data) for c in cats]
    log(f"inputs built in {time.perf_counter() - t0:.1f} s: " +
        ", ".join(f"{p.n_pos}/{p.n_shape}" for p in preps))
    eng = get_engine(0)
    dev = torch.device("cuda", 0)
    to = lambda x: torch.from_numpy(np.ascontiguousarray(x)).to(dev)  # noqa: E731

    def dev_sample(p, sel=None):
        weighted = not (np.all(p.weight == 1.0) and np.all(p.weight_shape == 1.0))
        sh = slice(None) if sel is None else sel
        return SampleArrays(to(p.positions), to(p.shapes[sh]), to(p.axis[sh]), to(p.e[sh]),
                            to(p.weight) if weighted else None,
                            to(p.weight_shape[sh]) if weighted else None,
                            L, p.los, jk, (L / 2.0) * 2.0 / jk)

    # inputs resident in HBM before the timed region
    samples = [dev_sample(p) for p in preps]
    torch.cuda.synchronize()

    if a.sort_only:
        sort_only(a, eng, samples, preps, th, batch)
        return

    def step():
        if batch:
            return eng.count_batch(samples, th)
        return [eng.count(samples[0], th, method="fast")]

    for _ in range(a.warmup):
        res = step()
    torch.cuda.synchronize()
    kstats = []
    t1 = time.perf_counter()
    for _ in range(a.steps):
        res = step()
        kstats.append(res[0].stats)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t1
    ms_step = elapsed / a.steps * 1e3
    pairs = int(sum(r.npairs[0].sum() for r in res))
    value = pairs / (ms_step / 1e3)
    # pair-kernel and sort times from HIP events on the launch stream; a
    # graph-replayed batch has no events between its launches, so these come
    # from the same count with plain launches
    ev_stats = kstats
    graph_pairs_ms = None
    if batch:
        graph_pairs_ms = float(np.mean([s["ms_pairs"] for s in kstats]))  # wall-clock span inside the graph
        ev_stats = []
        for _ in range(max(3, min(a.steps, 5))):
            ev_stats.append(eng.count_batch(samples, th, tuning={"graph": -1})[0].stats)
    k_ms = float(np.mean([s["ms_pairs"] for s in ev_stats]))
    flops = pairs * FLOP_PER_PAIR[w["mode"]]
    achieved = flops / (k_ms / 1e3) / 1e12
    log(f"fast: {ms_step:.2f} ms/step, pair kernel {k_ms:.2f} ms, {pairs} pairs")
    prepared = None
    if not batch:
        # the same count on a prepared sample: the sort is done once beforehand
        h = eng.prepare(samples[0], th)
        eng.count_prepared([h], th)
        torch.cuda.synchronize()
        tp = time.perf_counter()
        for _ in range(a.steps):
            rp_ = eng.count_prepared([h], th)[0]
        torch.cuda.synchronize()
        assert np.array_equal(rp_.npairs, res[0].npairs), "prepared and unprepared counts differ"
        prepared = dict(ms_per_step=(time.perf_counter() - tp) / a.steps * 1e3, ms_plan=rp_.stats["ms_plan"],
                        ms_pairs=rp_.stats["ms_pairs"])
        h.release()
        log(f"prepared sample: {prepared['ms_per_step']:.2f} ms/step")

    loop = None
    if batch:
        # the same samples counted one call at a time
        torch.cuda.synchronize()
        tl = time.perf_counter()
        single = [eng.count(s, th, method="fast") for s in samples]
        torch.cuda.synchronize()
        loop_ms = (time.perf_counter() - tl) * 1e3
        for r1, rb in zip(single, res):
            assert np.array_equal(r1.npairs, rb.npairs), "batched and single counts differ"
        loop = dict(ms=loop_ms, pairs_per_s=pairs / (loop_ms / 1e3))
        log(f"one call per sample: {loop_ms:.2f} ms")

    big = int(np.argmax([p.n_shape for p in preps]))
    plain = None
    if not a.no_plain:
        # 2 % of the shapes, capped near 4e9 pairs (the plain kernel's rate
        # is near 2e8 pairs/s, so the large box stays at about 20 s)
        frac = min(0.02, 4e9 / max(pairs, 1))
        p = preps[big]
        ns = max(1, int(p.n_shape * frac))
        sel = np.sort(np.random.default_rng(7).choice(p.n_shape, ns, replace=False))
        s_pl = dev_sample(p, sel)
        eng.count(s_pl, th, method="plain")
        torch.cuda.synchronize()
        tp = time.perf_counter()
        rp = eng.count(s_pl, th, method="plain")
        torch.cuda.synchronize()
        pl_s = time.perf_counter() - tp
        pl_pairs = int(rp.npairs[0].sum())
        plain_rate = pl_pairs / pl_s
        rf = eng.count(s_pl, th, method="fast")
        assert np.array_equal(rf.npairs, rp.npairs), "plain and fast counts differ"
        plain = dict(pairs_per_s=plain_rate, sample=f"seeded {ns} of {p.n_shape} shapes ({pl_pairs} pairs)",
                     s_per_full_count_est=pairs / plain_rate, kernel_ms=rp.stats["ms_pairs"])
        log(f"plain: {plain_rate:.3e} pairs/s on {frac:.2%} of shapes")

    e2e = e2e_parts = e2e_runs = e2e_repeat = None
    if not a.no_e2e and not batch:
        # the whole front-end call from the host data dict: host prep, H2D copy,
        # device count, estimators and HDF5 writes; median of 5 after a warm-up
        from measure_ia_amd import MeasureIABox
        with tempfile.TemporaryDirectory() as td:
            fn = os.path.join(td, "bench.hdf5")

            def new_box():
                # a fresh object per call: it holds no host prep, device copy
                # or device sort from an earlier call
                return MeasureIABox(cats[0].data, fn, None, None, list(w["r"]), w["nr"], w["n2"],
                                    w["pi_max"] if w["mode"] == "rppi" else None, boxsize=L, verbose=False)

            def caller(box):
                return box.measure_xi_w if w["mode"] == "rppi" else box.measure_xi_multipoles

            caller(new_box())("warm", "both", jk ** 3, temp_file_path=None)
            e2e_runs, parts = [], []
            for k in range(5):
                te = time.perf_counter()
                box = new_box()
                call = caller(box)
                call(f"bench{k}", "both", jk ** 3, temp_file_path=None)
                e2e_runs.append(time.perf_counter() - te)
                parts.append(dict(box.last_timing))
            e2e = float(np.median(e2e_runs))
            mid = parts[int(np.argsort(e2e_runs)[2])]
            e2e_parts = {k: round(v, 5) for k, v in mid.items() if k.endswith("_s")}
            # the same call repeated on the same object and unchanged data: the
            # host prep, device copy and device sort are re-used
            rep = []
            for k in range(5):
                te = time.perf_counter()
                call(f"again{k}", "both", jk ** 3, temp_file_path=None)
                rep.append(time.perf_counter() - te)
            e2e_repeat = float(np.median(rep))
        log(f"end-to-end {call.__name__}: median {e2e:.4f} s of {[round(x, 4) for x in e2e_runs]} {e2e_parts}")

    cpu = None
    if not a.no_cpu:
        tc = time.perf_counter()
        pps = res[big].npairs[0].sum() / max(preps[big].n_shape, 1)
        cpu = cpu_baseline(cats[big], w, r_bins, sb, pps)
        log(f"cpu baseline: {cpu['value']:.3e} pairs/s on {cpu['cores']} cores "
            f"({time.perf_counter() - tc:.1f} s)")

    st = res[0].stats
    # the counting sort (region ids, cell keys, scan, scatter of every sample) against HBM
    sort_ms = float(np.mean([s["ms_prepare"] for s in ev_stats]))
    sort_bytes = sum(sort_model_bytes(p) for p in preps)
    sort_gbps = sort_bytes / (sort_ms / 1e3) / 1e9
    cell_sort = {"bound": "hbm", "achieved": sort_gbps, "peak": PEAK_HBM_GBPS, "unit": "GB/s",
                 "frac": sort_gbps / PEAK_HBM_GBPS, "ms": sort_ms, "bytes": sort_bytes,
                 "model": "100 B per position + 164 B per shape (+8 / +16 B when weighted): classify reads "
                          "coordinates, writes (key, rank); scatter reads them again and writes the records; "
                          "the time covers the zero fill, classify, two scans and the scatter"}
    line = {
        "metric": "pairs/s & s per wgg+wg+ (1M gal, 27 JK): batched-FP32 vs plain GPU",
        "value": value, "unit": "pairs/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{a.workload}: {w['desc']}", "mode": w["mode"], "jackknife": jk ** 3,
                   "samples": len(preps), "positions": int(sum(p.n_pos for p in preps)),
                   "shapes": int(sum(p.n_shape for p in preps)), "pairs": pairs},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS, "traffic": measured_traffic(a.workload),
                     "kernel": "k_pairs", "kernel_ms": k_ms,
                     "frac_of_fp32_packed_peak": achieved / PEAK_FP32_PACKED_TFLOPS,
                     "candidate_efficiency": pairs / max(1, st["candidates"]),
                     "model": f"{FLOP_PER_PAIR[w['mode']]} FLOP per in-domain pair; FP64 vector peak "
                              f"(the kernel is vector-ALU + LDS-atomic bound, no matrix work)"},
        "cell_sort": cell_sort,
        "cpu_baseline": cpu,
        "plain_gpu": plain,
        "one_call_per_sample": loop,
        "graph_replayed": bool(st.get("graph")),
        "graph_pair_kernel_ms_wallclock": graph_pairs_ms,
        "step_over_pair_kernel": ms_step / k_ms,
        "prepared_sample": prepared,
        "speedup_vs_plain": (value / plain["pairs_per_s"]) if plain else None,
        "speedup_vs_cpu": (value / cpu["value"]) if cpu else None,
        "s_per_wgg_wgplus": e2e,
        "s_per_wgg_wgplus_def": ("median of 5 whole MeasureIABox.measure_xi_w('both', num_jk) calls, each on a "
                                 "new MeasureIABox built from the host data dict (measure_xi_multipoles for the "
                                 "(r, mu) workload): host prep, H2D copy, device sort, count, estimators, HDF5 "
                                 "writes; end_to_end_repeat_call_s is the median of 5 more calls on the last "
                                 "object, which re-use its host prep, device copy and sort"),
        "end_to_end_runs_s": e2e_runs,
        "end_to_end_parts_s": e2e_parts,
        "end_to_end_repeat_call_s": e2e_repeat,
        "precision": PRECISION,
        "engine": dict(st),
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
