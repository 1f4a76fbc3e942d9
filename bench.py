#!/usr/bin/env python3
"""Benchmark of the LUT precompute + sky lookups on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Workload (one step): a full precompute (Model::Init, 4 scattering orders) of
the spectral atmosphere - 47-band 5778 K solar spectrum, Rayleigh lambda^-4,
Angstrom Mie, ozone tent, 15 precomputed wavelengths (5 Precompute
iterations), FP32 LUTs at the reference sizes (transmittance 256x64,
scattering 32x128x32x8 packed as 256x128x32, irradiance 64x16) - followed by
one 1920x1080 batch of each sky lookup (GetSkyRadiance, GetSkyRadianceToPoint,
GetSunAndSkyIrradiance) against the resulting LUTs. Inputs are resident on
the GPU before the timed region. The printed JSON line reports precomputes
per second (value), ms per atmosphere, sky rays per second, p50/p99 step
latency, the roofline of the dominant kernel (scattering density) from HIP
events around its launches, and the FP64 CPU oracle timed on a bounded sample
of the same workload.
"""
import argparse
import json
import math
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "LUT precompute ms/atmosphere + sky rays/s, p99 latency, at ≤1e-3 rel err"
PEAK_FP32_TFLOPS = 157.3   # MI355X FP32 vector (FMA = 2 flops)
PEAK_HBM_GBS = 8000.0
# Algorithmic work per integration sample (VALU lane-ops, transcendentals
# weighted 4, loop invariants hoisted); frozen estimates of the reference's
# per-sample work, used for every 'achieved' figure.
LANE_OPS = dict(transmittance=40, single_scattering=178, scattering_density_2=111,
                scattering_density_n=68, indirect_irradiance=166,
                multiple_scattering=260)
SAMPLES = dict(transmittance=501, single_scattering=51, scattering_density=512,
               indirect_irradiance=1024, multiple_scattering=51)
QUERY_BYTES_PER_RAY = 124
QUERY_LANE_OPS_PER_RAY = 1500
WIDTH, HEIGHT = 1920, 1080


def percentile(xs, q):
    xs = sorted(xs)
    if not xs:
        return float("nan")
    k = (len(xs) - 1) * q
    f, c = math.floor(k), math.ceil(k)
    return xs[f] + (xs[c] - xs[f]) * (k - f)


def texel_counts(sizes):
    n_s = (sizes["scattering_r"] * sizes["scattering_mu"] * sizes["scattering_mu_s"] *
           sizes["scattering_nu"])
    n_t = sizes["transmittance_width"] * sizes["transmittance_height"]
    n_e = sizes["irradiance_width"] * sizes["irradiance_height"]
    return n_t, n_s, n_e


def sd_flops(n_s, order):
    ops = LANE_OPS["scattering_density_2" if order == 2 else "scattering_density_n"]
    return 2.0 * n_s * SAMPLES["scattering_density"] * ops


PROFILE_TRAFFIC = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                               "hbm_traffic_spectral_workload.md")


def profiled_traffic(kernel):
    """HBM bytes per dispatch of a kernel of the spectral workload, from the
    FETCH_SIZE and WRITE_SIZE passes summarized in the committed profile
    (calibrated against a copy of known size): (bytes, source) or (None,
    None) if the profile has no row for it."""
    try:
        rows = open(PROFILE_TRAFFIC).read().splitlines()
    except OSError:
        return None, None
    for r in rows:
        cells = [c.strip() for c in r.strip().strip("|").split("|")]
        if len(cells) >= 4 and cells[0] == kernel:
            try:
                return round((float(cells[2]) + float(cells[3])) * 1e6), \
                    "profiles/hbm_traffic_spectral_workload.md"
            except ValueError:
                return None, None
    return None, None


def cpu_baseline(cfg, gpu_luts, budget_s, threads):
    """FP64 oracle (this repository's CPU restatement of the reference) timed
    on a bounded sample of one precompute of the same atmosphere: the whole
    transmittance and irradiance passes, and a random subset of texels of the
    single-scattering, scattering-density (orders 2 and >= 3) and
    multiple-scattering passes, extrapolated by texel count and launch count.
    Returns atmospheres/s plus the sample description."""
    import ctypes as C
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_lib as ol
    from precomputed_atmospheric_scattering_amd import configs

    L = ol.lib()
    L.or_set_threads(threads)
    kw, orders = configs.model_kwargs(cfg)
    sizes = kw.pop("lut_sizes")
    spec = ol.ModelSpec(**kw)
    s = ol.sizes(**sizes)
    iters = L.or_num_iterations(C.byref(spec.c))
    lam = (C.c_double * 3)()
    mat = (C.c_double * 9)()
    L.or_iteration(C.byref(spec.c), 0, lam, mat)
    a = ol.lower(spec, list(lam))
    n_t, n_s, n_e = texel_counts(sizes)
    T = np.zeros(ol.transmittance_shape(s))
    t0 = time.perf_counter()
    L.or_pass_transmittance(C.byref(a), s, ol.ptr(T))
    t_T = time.perf_counter() - t0
    dE = np.zeros(ol.irradiance_shape(s))
    t0 = time.perf_counter()
    L.or_pass_direct_irradiance(C.byref(a), s, ol.ptr(T), ol.ptr(dE))
    t_E0 = time.perf_counter() - t0
    # Scattering inputs: the GPU's own final LUTs stand in for the deltas
    # (the per-texel cost does not depend on the values).
    S = np.ascontiguousarray(gpu_luts["scattering"], dtype=np.float64)
    M = np.ascontiguousarray(gpu_luts["single_mie_scattering"], dtype=np.float64)
    E = np.ascontiguousarray(gpu_luts["irradiance"], dtype=np.float64)
    rng = np.random.Generator(np.random.PCG64(11))
    per_pass = budget_s / 6.0

    def timed_texels(fn, probe=64):
        n = probe
        while True:
            idx = np.ascontiguousarray(rng.integers(0, n_s, n), dtype=np.int32)
            out = np.zeros((n, 4))
            t0 = time.perf_counter()
            fn(idx, out)
            dt = time.perf_counter() - t0
            if dt >= per_pass or n >= n_s:
                return dt / n, n
            n = min(n_s, int(n * max(2.0, 1.2 * per_pass / max(dt, 1e-4))))

    ip = lambda idx: idx.ctypes.data_as(C.POINTER(C.c_int))
    def ss_fn(idx, out):
        mie = np.zeros_like(out)
        L.or_texels_single_scattering(C.byref(a), s, ol.ptr(T), len(idx), ip(idx),
                                      ol.ptr(out), ol.ptr(mie))
    ss, n_ss = timed_texels(ss_fn)
    sd2, n_sd2 = timed_texels(lambda idx, out: L.or_texels_scattering_density(
        C.byref(a), s, ol.ptr(T), ol.ptr(S), ol.ptr(M), ol.ptr(S), ol.ptr(E), 2, len(idx),
        ip(idx), ol.ptr(out)))
    sdn, n_sdn = timed_texels(lambda idx, out: L.or_texels_scattering_density(
        C.byref(a), s, ol.ptr(T), ol.ptr(S), ol.ptr(M), ol.ptr(S), ol.ptr(E), 3, len(idx),
        ip(idx), ol.ptr(out)))
    ms, n_ms = timed_texels(lambda idx, out: L.or_texels_multiple_scattering(
        C.byref(a), s, ol.ptr(T), ol.ptr(S), len(idx), ip(idx), ol.ptr(out)))
    II = np.zeros(ol.irradiance_shape(s))
    t0 = time.perf_counter()
    L.or_pass_indirect_irradiance(C.byref(a), s, ol.ptr(S), ol.ptr(M), ol.ptr(S), 1,
                                  ol.ptr(II))
    t_II1 = time.perf_counter() - t0
    t0 = time.perf_counter()
    L.or_pass_indirect_irradiance(C.byref(a), s, ol.ptr(S), ol.ptr(M), ol.ptr(S), 2,
                                  ol.ptr(II))
    t_IIn = time.perf_counter() - t0
    per_iter = (t_T + t_E0 + ss * n_s + sd2 * n_s + (orders - 2) * sdn * n_s + t_II1 +
                (orders - 2) * t_IIn + (orders - 1) * ms * n_s)
    total = iters * per_iter + (t_T if iters > 1 else 0.0)
    sample = ("FP64 oracle, %d threads: full transmittance and irradiance passes; "
              "%d/%d/%d/%d random texels of the single-scattering / scattering-density "
              "order 2 / order 3 / multiple-scattering passes, extrapolated to %d "
              "texels x %d iterations x %d orders (est. %.1f s per atmosphere)" % (
                  threads, n_ss, n_sd2, n_sdn, n_ms, n_s, iters, orders, total))
    return dict(value=1.0 / total, unit="atmospheres/s", cores=threads, kind="port",
                sample=sample), total


def cpu_query_baseline(cfg, gpu_luts, queries, threads, n=20000):
    import ctypes as C
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_lib as ol
    from precomputed_atmospheric_scattering_amd import configs
    L = ol.lib()
    L.or_set_threads(threads)
    kw, _ = configs.model_kwargs(cfg)
    sizes = kw.pop("lut_sizes")
    spec = ol.ModelSpec(**kw)
    s = ol.sizes(**sizes)
    a = ol.lower(spec, (680.0, 550.0, 440.0))
    f = lambda x: np.ascontiguousarray(x[:n], dtype=np.float64)
    T = np.ascontiguousarray(gpu_luts["transmittance"], dtype=np.float64)
    S = np.ascontiguousarray(gpu_luts["scattering"], dtype=np.float64)
    M = np.ascontiguousarray(gpu_luts["single_mie_scattering"], dtype=np.float64)
    E = np.ascontiguousarray(gpu_luts["irradiance"], dtype=np.float64)
    o1, o2 = np.zeros((n, 3)), np.zeros((n, 3))
    t0 = time.perf_counter()
    L.or_sky_radiance(C.byref(a), s, ol.ptr(T), ol.ptr(S), ol.ptr(M), 0, n,
                      ol.ptr(f(queries["camera"])), ol.ptr(f(queries["view_ray"])),
                      ol.ptr(f(queries["shadow_length"])), ol.ptr(f(queries["sun_direction"])),
                      ol.ptr(o1), ol.ptr(o2))
    L.or_sky_radiance_to_point(C.byref(a), s, ol.ptr(T), ol.ptr(S), ol.ptr(M), 0, n,
                               ol.ptr(f(queries["camera"])), ol.ptr(f(queries["point"])),
                               ol.ptr(f(queries["shadow_length"])),
                               ol.ptr(f(queries["sun_direction"])), ol.ptr(o1), ol.ptr(o2))
    L.or_sun_and_sky_irradiance(C.byref(a), s, ol.ptr(T), ol.ptr(E), n,
                                ol.ptr(f(queries["point"])), ol.ptr(f(queries["normal"])),
                                ol.ptr(f(queries["sun_direction"])), ol.ptr(o1), ol.ptr(o2))
    dt = time.perf_counter() - t0
    return 3.0 * n / dt


def _json_line(result):
    print(json.dumps(result), flush=True)


def run_batched(args):
    """64 seeded random atmospheres (radii, scale heights, coefficients,
    albedo, per-dimension LUT size factors in {0.5, 1, 2}, 2..8 orders, 3
    wavelengths, FP32): the plain one-at-a-time Init against one batched
    submission (every pass level of all models in one launch, replayed from
    a HIP graph)."""
    import numpy as np
    import precomputed_atmospheric_scattering_amd as atmo
    from precomputed_atmospheric_scattering_amd import _capi, configs

    cfgs = configs.batch(args.seed, 64)
    kws = [configs.model_kwargs(c) for c in cfgs]
    orders = [o for _, o in kws]
    n_s_total = sum(texel_counts(kw["lut_sizes"])[1] for kw, _ in kws)
    stream = atmo.Stream()

    models = [atmo.Model(**kw) for kw, _ in kws]
    atmo.reset_tiles()  # the plain baseline runs at the default tile shapes

    # Plain: Model.init one model after the other, as a caller of the
    # reference's API would (temporaries allocated and freed by each Init).
    models[0].init(orders[0], stream)
    t0 = time.perf_counter()
    for m, o in zip(models, orders):
        m.init(o, stream)
    stream.synchronize()
    plain_s = time.perf_counter() - t0
    sample = (0, 21, 42, 63)
    ref = {i: models[i].lut("scattering") for i in sample}
    # Plain with the temporaries kept between Inits (no allocation in Init).
    for m in models:
        _capi.check(atmo.lib().atmo_model_set_option(
            m.handle, _capi.OPTION_RETAIN_INTERMEDIATES, 1), "set_option")
    for m, o in zip(models, orders):
        m.init(o, stream)
    stream.synchronize()
    t0 = time.perf_counter()
    for m, o in zip(models, orders):
        m.init(o, stream)
    stream.synchronize()
    plain_kept_s = time.perf_counter() - t0

    # Tile shapes: the autotuner times every shape per size class of the
    # batch (once; the table is then reused).
    sizes = [kw["lut_sizes"] for kw, _ in kws]
    t0 = time.perf_counter()
    atmo.autotune(sizes, stream=stream)
    tune_s = time.perf_counter() - t0
    shapes = {}
    for z in sizes:
        t = atmo.get_tiles(z)
        key = "sd %dx%d, ms %d%s" % (t["sd_nu_per_thread"], t["sd_block_threads"],
                                    t["ms_texels_per_thread"],
                                    " mu-major" if t["ms_mu_major"] else "")
        shapes[key] = shapes.get(key, 0) + 1

    def timed_batch(use_graph, steps, warmup):
        b = atmo.Batch(models, orders, use_graph=use_graph)
        for _ in range(warmup):
            b.launch(stream)
        stream.synchronize()
        ms = []
        t_all = time.perf_counter()
        for _ in range(steps):
            t = time.perf_counter()
            b.launch(stream)
            stream.synchronize()
            ms.append((time.perf_counter() - t) * 1e3)
        total = time.perf_counter() - t_all
        launches = b.num_launches
        b.close()
        return ms, total, launches

    eager_ms, eager_s, _ = timed_batch(False, max(1, args.steps // 2), 1)
    graph_ms, graph_s, launches = timed_batch(True, args.steps, max(1, args.warmup))
    exact = all(np.array_equal(models[i].lut("scattering"), ref[i]) for i in sample)
    # The same batch with every size class at the default shapes.
    tuned_table = os.path.join(tempfile.mkdtemp(prefix="atmo_tiles_"), "tiles.txt")
    atmo.save_tiles(tuned_table)
    atmo.reset_tiles()
    _, default_s, default_launches = timed_batch(True, args.steps, max(1, args.warmup))
    exact = exact and all(np.array_equal(models[i].lut("scattering"), ref[i])
                          for i in sample)
    atmo.load_tiles(tuned_table)
    value = 64 * args.steps / graph_s
    plain = 64 / plain_s
    _json_line({
        "metric": "batched LUT precompute throughput (64 random atmospheres)",
        "value": round(value, 2), "unit": "atmospheres/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(graph_s * 1e3 / args.steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "64 random atmospheres from seed %d: bottom radius "
                               "3000-7000 km, thickness 40-200 km, Rayleigh scale height "
                               "5-15 km and beta x0.3-3, Mie scale height 0.5-3 km, g "
                               "0.6-0.9, albedo 0-0.5, LUT sizes x{0.5,1,2} per dimension, "
                               "2-8 orders, 3 wavelengths, FP32" % args.seed,
                   "atmospheres": 64, "scattering_texels_total": n_s_total,
                   "orders": orders},
        "batch_p50_ms": round(percentile(graph_ms, 0.5), 3),
        "batch_p99_ms": round(percentile(graph_ms, 0.99), 3),
        "launches_per_batch": launches,
        "autotune_s": round(tune_s, 3),
        "tile_shapes": shapes,
        "default_tiles_atmospheres_per_s": round(64 * args.steps / default_s, 2),
        "default_tiles_launches_per_batch": default_launches,
        "plain_atmospheres_per_s": round(plain, 2),
        "plain_kept_temporaries_atmospheres_per_s": round(64 / plain_kept_s, 2),
        "batched_eager_atmospheres_per_s": round(64 * len(eager_ms) / eager_s, 2),
        "speedup_vs_plain": round(value / plain, 2),
        "speedup_vs_plain_kept_temporaries": round(value * plain_kept_s / 64, 2),
        "bit_exact_vs_plain": {"models_checked": list(sample), "equal": exact},
    })
    for m in models:
        m.close()


def run_serving(args):
    """Serving loop under a Poisson stream of precompute requests (random
    atmospheres of the batched workload) at --rate per second for --duration
    seconds. Each finished model then serves one 1920x1080 query batch
    (GetSkyRadiance, GetSkyRadianceToPoint, GetSunAndSkyIrradiance; inputs
    made on the device from a seed: camera altitude 0-100 km, view direction
    uniform, sun zenith 0-100 deg, hit distance 0-50 km) and is released."""
    import precomputed_atmospheric_scattering_amd as atmo
    from precomputed_atmospheric_scattering_amd import _capi, configs

    rng = configs.SplitMix64(args.seed)
    arrivals, t = [], 0.0
    while True:
        t += -math.log(1.0 - rng.uniform()) / args.rate
        if t > args.duration:
            break
        arrivals.append((t, configs.random_atmosphere(rng)))
    n = WIDTH * HEIGHT
    pool = []
    for _ in range(4):
        q = {k: atmo.DeviceArray((n,) if k == "shadow_length" else (n, 3))
             for k in atmo.QUERY_FIELDS}
        pool.append((q, [atmo.DeviceArray((n, 3)) for _ in range(6)]))
    gen = atmo.Stream()
    # Tile shapes of every size class of the request distribution, tuned
    # before the server starts (the table is shared by the process).
    t0 = time.perf_counter()
    atmo.autotune(configs.batch_size_classes(), stream=gen)
    tune_s = time.perf_counter() - t0
    srv = atmo.Server(max_inflight_batches=3, max_batch_models=16, lookup_streams=2)

    # Warm-up: one request and one query batch, not counted.
    kw, o = configs.model_kwargs(configs.random_atmosphere(configs.SplitMix64(args.seed + 1)))
    wid = srv.submit_precompute(num_scattering_orders=o, **kw)
    srv.wait([wid])
    srv.release(wid)

    info, ready, batches = {}, [], {}
    pre_lat, qry_lat, batch_sizes = [], [], []
    failed = 0
    i = 0
    start = time.perf_counter()
    outstanding = 0
    peak_resident = 0
    while i < len(arrivals) or outstanding > 0 or ready:
        now = time.perf_counter() - start
        while i < len(arrivals) and arrivals[i][0] <= now:
            kw, o = configs.model_kwargs(arrivals[i][1])
            rid = srv.submit_precompute(num_scattering_orders=o, **kw)
            info[rid] = kw
            outstanding += 1
            i += 1
        for c in srv.poll(timeout_ms=1):
            if c["status"] != 0:
                failed += 1
            if c["kind"] == _capi.REQUEST_PRECOMPUTE:
                outstanding -= 1
                pre_lat.append(c["end_ms"] - c["submit_ms"])
                batch_sizes.append(c["batch_models"])
                if c["status"] == 0:
                    ready.append(c["request_id"])
            else:
                b = batches[c["model_id"]]
                b["left"] -= 1
                b["end"] = max(b["end"], c["end_ms"])
                b["submit"] = min(b["submit"], c["submit_ms"])
                if b["left"] == 0:
                    qry_lat.append(b["end"] - b["submit"])
                    srv.release(c["model_id"])
                    pool.append(b["buffers"])
                    del batches[c["model_id"]]
                    outstanding -= 1
        while ready and pool:
            mid = ready.pop(0)
            kw = info[mid]
            q, o = pool.pop()
            rb, rt = kw["bottom_radius"] / 1000.0, kw["top_radius"] / 1000.0
            atmo.generate_sky_queries(mid, n, rb, rt, out=q, stream=gen)
            gen.synchronize()
            ids = [srv.submit_lookup(mid, "sky", n, [q["camera"], q["view_ray"],
                                                      q["shadow_length"], q["sun_direction"]],
                                     o[0:2]),
                   srv.submit_lookup(mid, "sky_to_point", n,
                                     [q["camera"], q["point"], q["shadow_length"],
                                      q["sun_direction"]], o[2:4]),
                   srv.submit_lookup(mid, "sun_and_sky", n,
                                     [q["point"], q["normal"], q["sun_direction"]], o[4:6])]
            assert len(ids) == 3
            batches[mid] = dict(left=3, end=0.0, submit=float("inf"), buffers=(q, o))
            outstanding += 1
        peak_resident = max(peak_resident, srv.stats()["resident_bytes"])
    wall = time.perf_counter() - start
    st = srv.stats()
    srv.close()
    done = len(pre_lat)
    _json_line({
        "metric": "served precompute latency p50/p99 and throughput (Poisson arrivals)",
        "value": round(done / wall, 3), "unit": "atmospheres/s", "n_gpus": 1,
        "steps": done, "warmup": 1, "ms_per_step": round(wall * 1e3 / max(done, 1), 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "Poisson %.1f req/s for %.1f s of random atmospheres "
                               "(batched-workload distribution, seed %d), one 1920x1080 "
                               "query batch of 3 lookups per finished model"
                               % (args.rate, args.duration, args.seed),
                   "requests": len(arrivals), "query_rays_per_batch": n},
        "precompute_latency_p50_ms": round(percentile(pre_lat, 0.5), 3),
        "precompute_latency_p99_ms": round(percentile(pre_lat, 0.99), 3),
        "query_batch_latency_p50_ms": round(percentile(qry_lat, 0.5), 3),
        "query_batch_latency_p99_ms": round(percentile(qry_lat, 0.99), 3),
        "sky_rays_per_s": round(3.0 * n * len(qry_lat) / wall, 1),
        "mean_batch_models": round(sum(batch_sizes) / max(1, len(batch_sizes)), 2),
        "max_batch_models": max(batch_sizes) if batch_sizes else 0,
        "failed": failed, "peak_resident_bytes": peak_resident,
        "budget_bytes": st["budget_bytes"], "autotune_s": round(tune_s, 3),
    })


def run_high_resolution(args):
    """High-resolution stress: scattering LUT 64x256x64x16, transmittance
    512x128, irradiance 128x32, the spectral atmosphere with ozone (15
    wavelengths), 8 scattering orders, FP16 storage with FP32 accumulation."""
    import precomputed_atmospheric_scattering_amd as atmo
    from precomputed_atmospheric_scattering_amd import configs

    cfg = configs.high_resolution(half_precision=True)
    kw, orders = configs.model_kwargs(cfg)
    n_t, n_s, n_e = texel_counts(kw["lut_sizes"])
    stream = atmo.Stream()
    # Default tile shapes (the autotuner keeps them for this size class), and
    # no tuning launches, so that a kernel profile of this run holds the
    # workload's launches only.
    atmo.reset_tiles()
    tiles = atmo.get_tiles(kw["lut_sizes"])
    m = atmo.Model(retain_intermediates=True, pass_timing=True, **kw)
    steps = max(1, min(args.steps, 5))
    for _ in range(max(1, min(args.warmup, 2))):
        m.init(orders, stream)
    stream.synchronize()
    ms, pass_ms, pass_n = [], {}, {}
    for _ in range(steps):
        t0 = time.perf_counter()
        m.init(orders, stream)
        stream.synchronize()
        ms.append((time.perf_counter() - t0) * 1e3)
        for name, order, t in m.pass_timings():
            key = name if name not in ("scattering_density",) else (
                "scattering_density_2" if order == 2 else "scattering_density_n")
            pass_ms[key] = pass_ms.get(key, 0.0) + t
            pass_n[key] = pass_n.get(key, 0) + 1
    # Compulsory bytes per launch (FP32 float4 texels of the delta LUTs),
    # each launch covering the 5 Precompute iterations: SD order 2 reads
    # delta Rayleigh + Mie + irradiance and writes the density; SD order >= 3
    # reads the multiple-scattering delta; MS reads the density and
    # transmittance, writes its delta and updates the iteration's FP64
    # scattering accumulator (32 B read + 32 B write per texel).
    tex = 16.0
    iters = m.params.num_iterations()
    bytes_per = {
        "scattering_density_2": iters * (3 * n_s + n_e) * tex,
        "scattering_density_n": iters * (2 * n_s + n_e) * tex,
        "multiple_scattering": iters * ((2 * n_s + n_t) * tex + 64.0 * n_s),
    }
    gbs = {k: round(bytes_per[k] * pass_n[k] / (pass_ms[k] * 1e-3) / 1e9, 1)
           for k in bytes_per if k in pass_ms}
    _json_line({
        "metric": "high-resolution LUT precompute ms/atmosphere",
        "value": round(sum(ms) / steps, 3), "unit": "ms/atmosphere", "n_gpus": 1,
        "steps": steps, "warmup": args.warmup, "ms_per_step": round(sum(ms) / steps, 3),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "spectral atmosphere with ozone, 15 precomputed "
                               "wavelengths, 8 scattering orders, FP16 LUT storage with "
                               "FP32 accumulation, LUTs 512x128 / 64x256x64x16 / 128x32",
                   "scattering_texels": n_s, "storage": "f16"},
        "p50_ms": round(percentile(ms, 0.5), 3), "p99_ms": round(percentile(ms, 0.99), 3),
        "pass_ms_per_atmosphere": {k: round(v / steps, 3) for k, v in pass_ms.items()},
        "compulsory_gb_per_s": gbs,
        "tiles": {"scattering_density": "%d nu x %d threads" % (
                      tiles["sd_nu_per_thread"], tiles["sd_block_threads"]),
                  "multiple_scattering": "%d texels per thread, %s rows" % (
                      tiles["ms_texels_per_thread"],
                      "mu-major" if tiles["ms_mu_major"] else "r-major")},
        "footprint_bytes": dict(zip(("resident", "scratch"), m.footprint())),
    })
    m.close()


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--workload", default="spectral",
                    choices=("spectral", "batched", "serving", "high-resolution"),
                    help="spectral (default): the spectral atmosphere's line; batched: 64 seeded "
                         "random atmospheres, batched vs one at a time; serving: Poisson "
                         "precompute requests + query batches through the serving loop; "
                         "high-resolution: 64x256x64x16 LUT, FP16, 8 orders")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--duration", type=float, default=10.0,
                    help="serving: seconds of request arrivals")
    ap.add_argument("--rate", type=float, default=20.0,
                    help="serving: precompute requests per second (Poisson)")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-budget", type=float, default=20.0,
                    help="seconds of CPU oracle work for the cpu_baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="first copy MIB MiB device-to-device once (one blit kernel "
                         "reading and writing MIB MiB), as the reference point of "
                         "HBM byte counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the library has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import precomputed_atmospheric_scattering_amd as atmo
    from precomputed_atmospheric_scattering_amd import configs, frames

    if args.copy_probe > 0:
        nbytes = args.copy_probe << 20
        src, dst = atmo.DeviceArray((nbytes,), np.uint8), atmo.DeviceArray((nbytes,), np.uint8)
        atmo._capi.check(atmo.lib().atmo_memcpy_to_device(dst.ptr, src.ptr, nbytes, None),
                         "copy probe")
        atmo.synchronize()
        src.free()
        dst.free()

    if args.workload == "batched":
        return run_batched(args)
    if args.workload == "serving":
        return run_serving(args)
    if args.workload == "high-resolution":
        return run_high_resolution(args)

    cfg = configs.spectral(half_precision=False)
    kw, orders = configs.model_kwargs(cfg)
    sizes = kw["lut_sizes"]
    n_t, n_s, n_e = texel_counts(sizes)
    stream = atmo.Stream()
    # Default tile shapes (no tuning launches in this run, so that a kernel
    # profile of it holds the workload's launches only).
    atmo.reset_tiles()
    tiles = atmo.get_tiles(sizes)
    model = atmo.Model(retain_intermediates=True, pass_timing=True, **kw)
    n_rays = WIDTH * HEIGHT
    q = configs.sky_queries(2024, n_rays, kw["bottom_radius"] / 1000.0,
                            kw["top_radius"] / 1000.0)
    dev = {k: atmo.DeviceArray.from_host(v) for k, v in q.items()}
    outs = [atmo.DeviceArray((n_rays, 3)) for _ in range(6)]
    ev = [atmo.Event() for _ in range(4)]

    # 15 precomputed wavelengths: the LUTs hold luminance, so the lookups are
    # the luminance API (GetSkyLuminance, GetSkyLuminanceToPoint,
    # GetSunAndSkyIlluminance), which differ from the radiance ones only by a
    # per-channel factor.
    def run_queries():
        ev[0].record(stream)
        model.sky_luminance_device(n_rays, dev["camera"], dev["view_ray"],
                                   dev["shadow_length"], dev["sun_direction"],
                                   outs[0], outs[1], stream)
        ev[1].record(stream)
        model.sky_luminance_to_point_device(n_rays, dev["camera"], dev["point"],
                                            dev["shadow_length"], dev["sun_direction"],
                                            outs[2], outs[3], stream)
        ev[2].record(stream)
        model.sun_and_sky_illuminance_device(n_rays, dev["point"], dev["normal"],
                                             dev["sun_direction"], outs[4], outs[5], stream)
        ev[3].record(stream)

    for _ in range(args.warmup):
        model.init(orders, stream)
        run_queries()
    stream.synchronize()

    pre_ms, qry_ms, step_ms, lookup_ms = [], [], [], [[], [], []]
    sd_ms, sd_flops_total, pass_ms = 0.0, 0.0, {}
    sd_launches = 0
    # One launch of a pass covers every Precompute iteration (each of N_S
    # texels).
    iters = model.params.num_iterations()
    atmo.synchronize()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        t0 = time.perf_counter()
        model.init(orders, stream)
        stream.synchronize()
        t1 = time.perf_counter()
        run_queries()
        stream.synchronize()
        t2 = time.perf_counter()
        pre_ms.append((t1 - t0) * 1e3)
        qry_ms.append((t2 - t1) * 1e3)
        step_ms.append((t2 - t0) * 1e3)
        for k in range(3):
            lookup_ms[k].append(ev[k].elapsed_ms(ev[k + 1]))
        for name, order, ms in model.pass_timings():
            pass_ms[name] = pass_ms.get(name, 0.0) + ms
            if name == "scattering_density":
                sd_ms += ms
                sd_flops_total += iters * sd_flops(n_s, order)
                sd_launches += 1
    atmo.synchronize()
    elapsed = time.perf_counter() - t_start

    # The same lookups on coherent batches, outside the timed steps: the
    # frames of the demo's nine preset views at 1920x1080 (the scene of the
    # reference's demo, a sphere on the planet with light shafts; views from
    # 7 km to 12000 km away). Per frame: GetSkyLuminance for every pixel,
    # GetSunAndSkyIlluminance and GetSkyLuminanceToPoint for the pixels
    # that hit the sphere or the ground.
    frame_ms = {}
    frame_rays = {}
    fev = [atmo.Event() for _ in range(2)]
    frame_reps = 5
    for view in sorted(frames.DEMO_VIEWS):
        fr, _ = frames.demo_frame(view, kw["bottom_radius"] / 1000.0,
                                  kw["sun_angular_radius"], WIDTH, HEIGHT)
        fq = fr.queries()
        hit = {k: np.concatenate([fq["sphere"][k], fq["ground"][k]])
               for k in ("camera", "point", "normal", "shadow", "sun")}
        nh = len(hit["point"])
        sky = {k: atmo.DeviceArray.from_host(fq["sky"][k])
               for k in ("camera", "view", "shadow", "sun")}
        hd = {k: atmo.DeviceArray.from_host(v) for k, v in hit.items()} if nh else None
        # A pixel can hit both the sphere and the ground behind it.
        hout = [atmo.DeviceArray((max(nh, 1), 3)) for _ in range(4)]
        fev[0].record(stream)
        for _ in range(frame_reps):
            model.sky_luminance_device(n_rays, sky["camera"], sky["view"], sky["shadow"],
                                       sky["sun"], outs[0], outs[1], stream)
            if nh:
                model.sun_and_sky_illuminance_device(nh, hd["point"], hd["normal"],
                                                     hd["sun"], hout[0], hout[1], stream)
                model.sky_luminance_to_point_device(nh, hd["camera"], hd["point"],
                                                    hd["shadow"], hd["sun"], hout[2],
                                                    hout[3], stream)
        fev[1].record(stream)
        stream.synchronize()
        frame_ms[view] = fev[0].elapsed_ms(fev[1]) / frame_reps
        frame_rays[view] = n_rays + 2 * nh
        for d in list(sky.values()) + (list(hd.values()) if hd else []) + hout:
            d.free()

    # The same atmosphere with FP16 LUT storage (FP32 accumulation, then the
    # per-channel maxima and the FP16 conversion), after the timed steps and
    # outside them: Init time only, with the temporaries kept between calls
    # as for the FP32 model above.
    kw16, _ = configs.model_kwargs(configs.spectral(half_precision=True))
    m16 = atmo.Model(retain_intermediates=True, **kw16)
    for _ in range(args.warmup):
        m16.init(orders, stream)
    stream.synchronize()
    half_ms = []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        m16.init(orders, stream)
        stream.synchronize()
        half_ms.append((time.perf_counter() - t0) * 1e3)
    m16.close()

    steps = args.steps
    sd_avg_ms = sd_ms / sd_launches
    sd_traffic, traffic_src = profiled_traffic("k_scattering_density")
    achieved_tflops = sd_flops_total / (sd_ms * 1e-3) / 1e12
    ms_per_atmo = sum(pre_ms) / steps
    query_s = sum(sum(l) for l in lookup_ms) * 1e-3
    rays_per_s = 3.0 * n_rays * steps / query_s
    q_ideal = max(QUERY_BYTES_PER_RAY * n_rays / (PEAK_HBM_GBS * 1e9),
                  QUERY_LANE_OPS_PER_RAY * n_rays / (PEAK_FP32_TFLOPS / 2 * 1e12))
    q_frac = q_ideal / (query_s / steps)

    result = {
        "metric": METRIC,
        "value": round(steps / (sum(pre_ms) * 1e-3), 3),
        "unit": "atmospheres/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed * 1e3 / steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "spectral atmosphere (47-band 5778 K solar spectrum, ozone), "
                        "15 precomputed wavelengths = 5 iterations, 4 scattering orders, "
                        "FP32 LUTs 256x64 / 32x128x32x8 / 64x16; then 1920x1080 rays of "
                        "each of GetSkyLuminance, GetSkyLuminanceToPoint, "
                        "GetSunAndSkyIlluminance",
            "precomputed_wavelengths": 15, "scattering_orders": orders,
            "scattering_texels": n_s, "query_rays": n_rays,
        },
        "ms_per_atmosphere": round(ms_per_atmo, 4),
        "precompute_p50_ms": round(percentile(pre_ms, 0.5), 4),
        "precompute_p99_ms": round(percentile(pre_ms, 0.99), 4),
        "step_p50_ms": round(percentile(step_ms, 0.5), 4),
        "step_p99_ms": round(percentile(step_ms, 0.99), 4),
        "sky_rays_per_s": round(rays_per_s, 1),
        "lookup_ms": {n: round(sum(l) / steps, 4) for n, l in zip(
            ("sky_luminance", "sky_luminance_to_point", "sun_and_sky_illuminance"),
            lookup_ms)},
        "lookup_roofline_frac": round(q_frac, 4),
        "demo_frames": {
            "frames_per_s": round(len(frame_ms) / (sum(frame_ms.values()) * 1e-3), 1),
            "lookups_per_s": round(sum(frame_rays.values()) /
                                   (sum(frame_ms.values()) * 1e-3), 1),
            "ms_per_frame": {str(v): round(t, 4) for v, t in frame_ms.items()},
            "lookups_per_frame": {str(v): n for v, n in frame_rays.items()},
        },
        "pass_ms_per_atmosphere": {k: round(v / steps, 4) for k, v in pass_ms.items()},
        "fp16_storage": {"ms_per_atmosphere": round(sum(half_ms) / len(half_ms), 4),
                         "precompute_p99_ms": round(percentile(half_ms, 0.99), 4)},
        "tiles": {"scattering_density": "%d nu x %d threads" % (
                      tiles["sd_nu_per_thread"], tiles["sd_block_threads"]),
                  "multiple_scattering": "%d texels per thread, %s rows" % (
                      tiles["ms_texels_per_thread"],
                      "mu-major" if tiles["ms_mu_major"] else "r-major")},
        "roofline": {
            "kernel": "k_scattering_density",
            "bound": "valu",
            "achieved": round(achieved_tflops, 3),
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved_tflops / PEAK_FP32_TFLOPS, 4),
            "traffic": sd_traffic,
            "traffic_source": traffic_src,
            "avg_launch_ms": round(sd_avg_ms, 5),
            "algorithmic_flops_per_launch": round(sd_flops_total / sd_launches),
        },
    }
    if not args.no_cpu_baseline:
        luts = {k: model.lut(k) for k in ("transmittance", "scattering",
                                          "single_mie_scattering", "irradiance")}
        threads = min(16, os.cpu_count() or 1)
        base, cpu_s = cpu_baseline(cfg, luts, args.cpu_budget, threads)
        base["sky_rays_per_s"] = round(cpu_query_baseline(cfg, luts, q, threads), 1)
        base["value"] = round(base["value"], 8)
        result["cpu_baseline"] = base
        result["speedup_vs_cpu"] = round(result["value"] / base["value"], 1)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
