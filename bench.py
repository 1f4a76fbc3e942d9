#!/usr/bin/env python3
"""Benchmark of the LUT precompute + sky lookups on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

Default (--workload headline): one JSON line whose headline steps are the
high-resolution stress workload - the spectral atmosphere (47-band 5778 K
solar spectrum, Rayleigh lambda^-4, Angstrom Mie, ozone tent) with 15
precomputed wavelengths (5 Precompute iterations), 8 scattering orders, FP16
LUT storage with FP32 accumulation, LUTs 512x128 (transmittance),
64x256x64x16 (scattering, packed 1024x256x64) and 128x32 (irradiance). One
step is Model::Init as the reference's API runs it (temporaries allocated and
freed by the call), then one 1920x1080 batch of each sky lookup
(GetSkyLuminance, GetSkyLuminanceToPoint, GetSunAndSkyIlluminance) on the
resulting LUTs. Inputs are resident on the GPU before the timed region. The
same line carries, measured in the same run after the headline steps:

- "spectral": the same atmosphere at the reference LUT sizes, 4 orders, FP32
  and FP16 storage;
- "batched": 64 seeded random atmospheres, one-at-a-time Init against one
  batched submission replayed from a HIP graph;
- "serving": Poisson precompute requests at 20/s plus one 1920x1080 query
  batch per finished model through the serving loop, and the same loop
  under a saturating burst;
- "roofline": the dominant precompute kernel's counter-measured VALU rate
  (instructions from the committed rocprofv3 counter summary, launch times
  from HIP events in this run) against the FP32 VALU peak, and in
  "fp_arith" its FP32/FP64 add, mul and FMA rate against their
  single-issue ceiling; "passes" gives the same per pass, with HBM GB/s;
- "measured_peaks": on-device rate ceilings of this GPU (csrc/peaks.hip);
- "lookup_roofline": the lookups' bytes (inputs, outputs and the distinct
  LUT texels they read) per second against the HBM peak;
- "cpu_baseline": the FP64 oracle on the host cores, on a bounded sample.

The other --workload values run one section alone (for profiling).
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
PEAK_FP32_TFLOPS = 157.3   # MI355X FP32 vector (packed FMA = 2 flops per lane)
# VALU lane-operations per second at the FP32 peak: 256 CUs x 2.4 GHz x 128
# FMA lanes per clock (two VALU instructions of a wave per 4 clocks per SIMD,
# or one packed). Every VALU instruction of a wave counts 64 lane-ops.
PEAK_VALU_TOPS = PEAK_FP32_TFLOPS / 2.0
# FP arithmetic (add, mul, FMA) lane-instructions per second when they issue
# one per SIMD per 4 clocks: 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz. On
# MI355X two FP32 arithmetic instructions do not issue together (csrc/
# peaks.hip measures FMA + add pairs at the single rate, FMA + integer or
# conversion pairs at twice it), and packed FP32 is not used here, so this
# is the ceiling of a kernel made of FP32 (or FP64) arithmetic.
PEAK_FP_ARITH_TOPS = 256 * 4 * 16 * 2.4e9 / 1e12
PEAK_HBM_GBS = 8000.0
# Algorithmic work per integration sample (VALU lane-ops, transcendentals
# weighted 4, loop invariants hoisted): frozen estimates of the reference's
# per-sample work (a texel evaluated on its own). The kernels here share work
# across samples and texels, so these give an "algorithmic equivalent" rate,
# not a roofline fraction.
LANE_OPS = dict(transmittance=40, single_scattering=178, scattering_density_2=111,
                scattering_density_n=68, indirect_irradiance=166,
                multiple_scattering=260)
SAMPLES = dict(transmittance=501, single_scattering=51, scattering_density=512,
               indirect_irradiance=1024, multiple_scattering=51)
# The survey's compulsory traffic of the three lookups: inputs and outputs
# only (LUT texels taken as cache hits).
QUERY_BYTES_PER_RAY = 124
QUERY_LANE_OPS_PER_RAY = 1500
# What the three lookup kernels read and write per ray: GetSkyLuminance
# (camera, view, sun, shadow length; 2 RGB outputs) 40 + 24 B,
# GetSkyLuminanceToPoint (camera, point, sun, shadow length) 40 + 24 B,
# GetSunAndSkyIlluminance (point, normal, sun) 36 + 24 B.
QUERY_IO_BYTES_PER_RAY = 188
# A scattering lookup reads 16 texels (2 r x 2 mu x 2 mu_s x 2 nu) of S and
# of S_Mie.
QUERY_TEXELS_PER_LOOKUP = 16
WIDTH, HEIGHT = 1920, 1080
PROFILES = os.path.join(ROOT, "profiles")
COUNTERS_HIGH_RES = os.path.join(PROFILES, "counters_high_resolution.md")
VALU_MIX_HIGH_RES = os.path.join(PROFILES, "valu_mix_high_resolution.md")
# Pass -> kernel of the counter summary (the transmittance pass's launches
# are two kernels, the optical depths and the LUT, and carry no counters).
PASS_KERNELS = dict(single_scattering="k_single_scattering",
                    scattering_density="k_scattering_density",
                    indirect_irradiance="k_indirect_irradiance",
                    multiple_scattering="k_multiple_scattering")


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


def profile_table(path):
    """Rows of the first markdown table of a committed profile summary, as
    dicts keyed by the header cells ({} if the file is absent)."""
    try:
        lines = open(path).read().splitlines()
    except OSError:
        return {}
    rows, header = {}, None
    for r in lines:
        if not r.startswith("|"):
            if header is not None and rows:
                break
            continue
        cells = [c.strip() for c in r.strip().strip("|").split("|")]
        if header is None:
            header = cells
            continue
        if set(cells[0]) <= set("-: "):
            continue
        rows[cells[0]] = dict(zip(header, cells))
    return rows


def counter_figures(path=COUNTERS_HIGH_RES):
    """Per kernel: VALU wave-instructions, fetched and written HBM bytes per
    dispatch, from the committed counter summary of the high-resolution
    workload (rocprofv3 --pmc runs; HBM bytes calibrated against a copy of
    known size)."""
    out = {}
    for k, row in profile_table(path).items():
        try:
            out[k] = dict(valu=float(row["VALU instr / dispatch"]),
                          fetch=float(row["fetch MB / dispatch"]) * 1e6,
                          write=float(row["write MB / dispatch"]) * 1e6)
        except (KeyError, ValueError):
            continue
    return out


def valu_mix_figures(path=VALU_MIX_HIGH_RES):
    """Per kernel: FP32 and FP64 arithmetic (add, mul, FMA) wave-instructions
    per dispatch, from the committed instruction-mix summary of the
    high-resolution workload."""
    out = {}
    for k, row in profile_table(path).items():
        try:
            out[k] = float(row["FP32 arith / dispatch"]) + float(row["FP64 arith / dispatch"])
        except (KeyError, ValueError):
            continue
    return out


def cpu_threads():
    """The host threads the CPU baseline uses: the CPUs this process may run
    on, at most 16 (a one-GPU share of the GPU box's host)."""
    try:
        allowed = len(os.sched_getaffinity(0))
    except AttributeError:
        allowed = os.cpu_count() or 1
    return max(1, min(16, allowed)), allowed


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


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


def _progress(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


LOOKUP_NAMES = ("sky_luminance", "sky_luminance_to_point", "sun_and_sky_illuminance")


def scattering_lookups(q, top_radius):
    """Scattering lookups that a batch of the three lookups makes on the
    rays q (length units), in units of 16 distinct texels of each scattering
    LUT: GetSkyRadiance returns without one for a camera above the
    atmosphere whose ray misses it (functions.glsl:1715-1729), and
    GetSkyRadianceToPoint makes two per ray. A lookup from a camera above
    the atmosphere is made at (or clamped to) the top radius, whose two r
    taps are the same LUT layer: 8 distinct texels."""
    cam = q["camera"].astype("float64")
    v = q["view_ray"].astype("float64")
    r = (cam * cam).sum(1) ** 0.5
    rmu = (cam * v).sum(1)
    disc = rmu * rmu - r * r + top_radius * top_radius
    hits = (disc >= 0.0) & (-rmu - abs(disc) ** 0.5 > 0.0)
    above = r > top_radius
    sky = (~above).sum() + 0.5 * (above & hits).sum()
    to_point = 2.0 * len(r) - 0.5 * above.sum()
    return float(sky + to_point)


class Lookups:
    """One 1920x1080 batch of each luminance lookup (GetSkyLuminance,
    GetSkyLuminanceToPoint, GetSunAndSkyIlluminance) on a model, inputs
    resident on the device, each timed with HIP events on the lookup
    stream, after a fill that evicts the caches. Rays: camera altitude 0-100 km, view direction uniform on the
    sphere, sun zenith 0-100 deg, hit distance 0-50 km (seeded)."""

    def __init__(self, atmo, configs, kw, seed=2024):
        self.n = WIDTH * HEIGHT
        self.half = bool(kw.get("half_precision", False))
        self.luts = 1 if kw.get("combine_scattering_textures", False) else 2
        self.q = configs.sky_queries(seed, self.n, kw["bottom_radius"] / 1000.0,
                                     kw["top_radius"] / 1000.0)
        self.scattering_lookups = scattering_lookups(self.q, kw["top_radius"] / 1000.0)
        self.dev = {k: atmo.DeviceArray.from_host(v) for k, v in self.q.items()}
        self.outs = [atmo.DeviceArray((self.n, 3)) for _ in range(6)]
        self.ev = [atmo.Event() for _ in range(4)]
        self.ms = [[], [], []]

    def run(self, model, stream):
        d, o, ev, n = self.dev, self.outs, self.ev, self.n
        # Without it the GPU would sit idle after the Init's synchronize
        # until the first launch arrives, and ev[0]..ev[1] would count the
        # host's enqueue time (about 25 us) as GetSkyLuminance's.
        flush_caches(stream)
        ev[0].record(stream)
        model.sky_luminance_device(n, d["camera"], d["view_ray"], d["shadow_length"],
                                   d["sun_direction"], o[0], o[1], stream)
        ev[1].record(stream)
        model.sky_luminance_to_point_device(n, d["camera"], d["point"], d["shadow_length"],
                                            d["sun_direction"], o[2], o[3], stream)
        ev[2].record(stream)
        model.sun_and_sky_illuminance_device(n, d["point"], d["normal"], d["sun_direction"],
                                             o[4], o[5], stream)
        ev[3].record(stream)

    def collect(self):
        for k in range(3):
            self.ms[k].append(self.ev[k].elapsed_ms(self.ev[k + 1]))

    def roofline(self, batch_s):
        """Roofline of one batch of the three lookups. On random rays no two
        rays share a texel, so every texel a lookup reads (the 16 taps of S
        and of S_Mie, in the LUTs' storage format, 8 B per texel FP16, 16 B
        FP32; GetSkyRadiance makes one lookup unless its ray misses the
        atmosphere, GetSkyRadianceToPoint two) is read from beyond the
        caches when the LUTs exceed them (the
        high-resolution LUTs do: 268 MB FP16, 537 MB FP32, against 256 MB of
        Infinity Cache); those bytes plus the rays' inputs and outputs are
        the algorithmic bytes. Transmittance and irradiance LUTs (1 MB, 64
        KB) stay cached. The lookup copy reads those texels as 128-byte
        lines, 4 per scattering lookup FP32 and 2 FP16; random gathers cost
        per line, so their rate is set against the measured line rate of a
        random 128-byte gather over a buffer of the copy's size."""
        texel = 8 if self.half else 16
        texel_bytes = (self.scattering_lookups * QUERY_TEXELS_PER_LOOKUP * self.luts * texel /
                       self.n)
        per_ray = QUERY_IO_BYTES_PER_RAY + texel_bytes
        achieved = per_ray * self.n / batch_s / 1e9
        ideal_io = max(QUERY_BYTES_PER_RAY * self.n / (PEAK_HBM_GBS * 1e9),
                       QUERY_LANE_OPS_PER_RAY * self.n / (PEAK_VALU_TOPS * 1e12))
        out = {"bound": "hbm", "bytes_per_ray": round(per_ray, 1),
               "io_bytes_per_ray": QUERY_IO_BYTES_PER_RAY,
               "texel_bytes_per_ray": round(texel_bytes, 1),
               "scattering_lookups_per_ray": round(self.scattering_lookups / self.n, 4),
               "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
               "frac": round(achieved / PEAK_HBM_GBS, 4),
               "io_only_frac": round(ideal_io / batch_s, 4)}
        if self.luts == 2:
            lines = self.scattering_lookups / self.n * (2 if self.half else 4)
            out["lines_per_ray"] = round(lines, 4)
            out["lines_per_s"] = round(lines * self.n / batch_s, 1)
            rate = MEASURED_PEAKS.get("gather_lines_per_s_2_25gib")
            if rate:
                out["frac_of_measured_gather_line_rate"] = round(
                    lines * self.n / batch_s / rate, 4)
        return out

    def summary(self):
        steps = len(self.ms[0])
        batch_s = sum(sum(l) for l in self.ms) * 1e-3 / steps
        rl = self.roofline(batch_s)
        return {
            "sky_rays_per_s": round(3.0 * self.n / batch_s, 1),
            "lookup_ms": {k: round(sum(l) / steps, 4) for k, l in zip(LOOKUP_NAMES, self.ms)},
            "lookup_p99_ms": {k: round(percentile(l, 0.99), 4)
                              for k, l in zip(LOOKUP_NAMES, self.ms)},
            "lookup_roofline": rl,
            "lookup_roofline_frac": rl["frac"],
        }

    def host_buffers(self, model, reps=2):
        """The same three lookups from and to host (numpy) arrays through the
        host wrappers: device buffers allocated, inputs copied in, outputs
        copied back, per call. Rays per second over the three, best of
        reps."""
        q, n = self.q, self.n
        best = float("inf")
        for _ in range(reps):
            t0 = time.perf_counter()
            model.sky_luminance(q["camera"], q["view_ray"], q["shadow_length"],
                                q["sun_direction"])
            model.sky_luminance_to_point(q["camera"], q["point"], q["shadow_length"],
                                         q["sun_direction"])
            model.sun_and_sky_illuminance(q["point"], q["normal"], q["sun_direction"])
            best = min(best, time.perf_counter() - t0)
        return round(3.0 * n / best, 1)

    def free(self):
        for d in list(self.dev.values()) + self.outs:
            d.free()


# On-device rate ceilings (libatmo_peaks.so, csrc/peaks.hip), measured once
# per run by measure_peaks().
MEASURED_PEAKS = {}
PEAK_KEYS = ("fp32_fma_lane_instr_per_s", "valu_dual_issue_lane_instr_per_s",
             "fp64_fma_lane_instr_per_s", "hbm_read_bytes_per_s", "hbm_copy_bytes_per_s",
             "gather_lines_per_s_1gib", "gather_lines_per_s_2_25gib",
             "fp32_fma_add_lane_instr_per_s", "fp32_fma_int_lane_instr_per_s")


_PEAKS_LIB = []


def peaks_lib():
    """libatmo_peaks.so (csrc/peaks.hip), loaded once."""
    if not _PEAKS_LIB:
        import ctypes as C
        lib = C.CDLL(os.path.join(ROOT, "precomputed_atmospheric_scattering_amd",
                                  "libatmo_peaks.so"))
        lib.atmo_peaks_measure.restype = C.c_int
        lib.atmo_peaks_measure.argtypes = [C.POINTER(C.c_double), C.c_int]
        lib.atmo_peaks_flush.restype = C.c_int
        lib.atmo_peaks_flush.argtypes = [C.c_void_p]
        _PEAKS_LIB.append(lib)
    return _PEAKS_LIB[0]


def flush_caches(stream):
    """A 512 MiB fill on 'stream': what is enqueued after it starts with
    cold caches and does not wait for the host's enqueue time."""
    rc = peaks_lib().atmo_peaks_flush(stream.handle)
    if rc != 0:
        raise RuntimeError("atmo_peaks_flush failed: HIP error %d" % rc)


def measure_peaks():
    """Microbenchmarks of the rate ceilings on this GPU: FP32 FMA issue
    (plain, non-packed instructions), the issue rate of an FMA/add/fract mix
    that the SIMDs issue two at a time, FP64 FMA issue, HBM streaming read
    and copy, random 128-byte line gathers over 1 GiB and 2.25 GiB, and the
    issue rates of FMA + FP32 add and FMA + integer instruction pairs."""
    import ctypes as C
    lib = peaks_lib()
    out = (C.c_double * len(PEAK_KEYS))()
    rc = lib.atmo_peaks_measure(out, len(PEAK_KEYS))
    if rc != 0:
        raise RuntimeError("atmo_peaks_measure failed: HIP error %d" % rc)
    MEASURED_PEAKS.update({k: float(v) for k, v in zip(PEAK_KEYS, out)})
    return {
        "fp32_fma_tops": round(out[0] / 1e12, 2),
        "valu_dual_issue_tops": round(out[1] / 1e12, 2),
        "fp64_fma_tops": round(out[2] / 1e12, 2),
        "hbm_read_gb_per_s": round(out[3] / 1e9, 1),
        "hbm_copy_gb_per_s": round(out[4] / 1e9, 1),
        "gather_lines_per_s_1gib": round(out[5], 1),
        "gather_lines_per_s_2_25gib": round(out[6], 1),
        "fp32_fma_plus_add_tops": round(out[7] / 1e12, 2),
        "fp32_fma_plus_int_tops": round(out[8] / 1e12, 2),
        "unit_note": "TOP/s are VALU lane-instructions per second (an FMA counts 1)",
    }


def pass_key(name, order):
    if name == "scattering_density":
        return "scattering_density_2" if order == 2 else "scattering_density_n"
    return name


def timed_inits(atmo, model, orders, stream, steps, lookups=None):
    """steps x (Init, then the lookups if any); returns (precompute ms,
    step ms, wall s, {pass: [launch ms]})."""
    pre_ms, step_ms, launches = [], [], {}
    atmo.synchronize()
    t_start = time.perf_counter()
    for _ in range(steps):
        t0 = time.perf_counter()
        model.init(orders, stream)
        stream.synchronize()
        t1 = time.perf_counter()
        if lookups is not None:
            lookups.run(model, stream)
            stream.synchronize()
        t2 = time.perf_counter()
        pre_ms.append((t1 - t0) * 1e3)
        step_ms.append((t2 - t0) * 1e3)
        if lookups is not None:
            lookups.collect()
        if model.pass_timing:
            for name, order, ms in model.pass_timings():
                launches.setdefault(pass_key(name, order), []).append(ms)
    atmo.synchronize()
    return pre_ms, step_ms, time.perf_counter() - t_start, launches


def kept_init_ms(atmo, kw, orders, stream, steps):
    """Init of the same atmosphere with its temporaries kept between calls
    (no allocation or free inside the call): mean ms over 'steps' calls."""
    m = atmo.Model(retain_intermediates=True, **kw)
    m.init(orders, stream)
    stream.synchronize()
    ms = []
    for _ in range(steps):
        t0 = time.perf_counter()
        m.init(orders, stream)
        stream.synchronize()
        ms.append((time.perf_counter() - t0) * 1e3)
    m.close()
    return sum(ms) / len(ms)


def section_high_resolution(args, steps, warmup, with_lookups=True, keep_luts=False):
    """The high-resolution stress workload: Init as the reference's API runs
    it (temporaries allocated and freed by the call; launch times from HIP
    events), then the three lookups; afterwards the same Init with the
    temporaries kept between calls."""
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
    m = atmo.Model(pass_timing=True, **kw)
    lk = Lookups(atmo, configs, kw) if with_lookups else None
    for _ in range(warmup):
        m.init(orders, stream)
        if lk is not None:
            lk.run(m, stream)
    stream.synchronize()
    _progress("high resolution: warm-up done, %d timed steps" % steps)
    pre_ms, step_ms, wall, launches = timed_inits(atmo, m, orders, stream, steps, lk)
    host_rays = lk.host_buffers(m) if lk is not None else None
    resident, temporaries = m.footprint()
    luts = None
    if keep_luts:
        luts = {k: m.lut(k) for k in ("transmittance", "scattering",
                                      "single_mie_scattering", "irradiance")}
    iters = m.params.num_iterations()
    m.close()
    kept = kept_init_ms(atmo, kw, orders, stream, max(1, min(steps, 5)))
    stream.close()
    res = {
        "value": round(steps / (sum(pre_ms) * 1e-3), 4),
        "ms_per_step": round(wall * 1e3 / steps, 3),
        "ms_per_atmosphere": round(sum(pre_ms) / steps, 3),
        "precompute_p50_ms": round(percentile(pre_ms, 0.5), 3),
        "precompute_p99_ms": round(percentile(pre_ms, 0.99), 3),
        "step_p50_ms": round(percentile(step_ms, 0.5), 3),
        "step_p99_ms": round(percentile(step_ms, 0.99), 3),
        "init_kept_temporaries_ms": round(kept, 3),
        "tiles": {"scattering_density": "%d nu x %d threads" % (
                      tiles["sd_nu_per_thread"], tiles["sd_block_threads"]),
                  "multiple_scattering": "%d texels x %d threads, %s rows" % (
                      tiles["ms_texels_per_thread"], tiles["ms_block_threads"],
                      tiles["ms_row_order"])},
        "footprint_bytes": {"resident": resident, "init_temporaries": temporaries},
    }
    if lk is not None:
        res.update(lk.summary())
        # Not the headline rate: host arrays in and out (PCIe and
        # allocations included).
        res["sky_rays_per_s_host_buffers"] = host_rays
        lk.free()
    meta = dict(cfg=cfg, steps=steps, n_s=n_s, iters=iters, orders=orders, luts=luts)
    return res, launches, meta


def pass_figures(launches, steps, n_s, iters):
    """Per pass, from this run's launch times (HIP events on the Init
    stream) and the committed per-dispatch counters of the same workload:
    ms per atmosphere, VALU lane-op rate and HBM rate, each as a fraction of
    its MI355X peak."""
    cnt = counter_figures()
    mix = valu_mix_figures()
    merged = {}
    for k, v in launches.items():
        merged.setdefault("scattering_density" if k.startswith("scattering_density") else k,
                          []).extend(v)
    out = {}
    for name, ms_list in merged.items():
        avg_s = sum(ms_list) / len(ms_list) * 1e-3
        e = {"ms_per_atmosphere": round(sum(ms_list) / steps, 3),
             "launches_per_atmosphere": len(ms_list) // steps,
             "avg_launch_ms": round(avg_s * 1e3, 4)}
        c = cnt.get(PASS_KERNELS.get(name, ""))
        if c is not None:
            e["valu_tops"] = round(c["valu"] * 64.0 / avg_s / 1e12, 3)
            e["valu_frac"] = round(c["valu"] * 64.0 / avg_s / 1e12 / PEAK_VALU_TOPS, 4)
            e["hbm_bytes_per_launch"] = round(c["fetch"] + c["write"])
            e["hbm_gb_per_s"] = round((c["fetch"] + c["write"]) / avg_s / 1e9, 1)
            e["hbm_frac"] = round((c["fetch"] + c["write"]) / avg_s / 1e9 / PEAK_HBM_GBS, 4)
        fa = mix.get(PASS_KERNELS.get(name, ""))
        if fa is not None:
            e["fp_arith_tops"] = round(fa * 64.0 / avg_s / 1e12, 3)
            e["fp_arith_frac"] = round(fa * 64.0 / avg_s / 1e12 / PEAK_FP_ARITH_TOPS, 4)
        out[name] = e
    if "scattering_density" in out:
        # The frozen per-sample estimate of the reference's work, for scale.
        flops = sum(iters * sd_flops(n_s, 2 if k.endswith("_2") else 3) * len(v)
                    for k, v in launches.items() if k.startswith("scattering_density"))
        sd_s = sum(sum(v) for k, v in launches.items()
                   if k.startswith("scattering_density")) * 1e-3
        out["scattering_density"]["algorithmic_equivalent_tflops"] = round(
            flops / sd_s / 1e12, 1)
    return out


def roofline(passes):
    """The roofline object of the dominant precompute pass (the largest ms
    per atmosphere among passes with counters): its bound is the larger of
    its VALU and HBM fractions."""
    have = {k: v for k, v in passes.items() if "valu_frac" in v}
    if not have:
        return None
    name = max(have, key=lambda k: have[k]["ms_per_atmosphere"])
    p = have[name]
    valu = p["valu_frac"] >= p["hbm_frac"]
    measured = MEASURED_PEAKS.get("valu_dual_issue_lane_instr_per_s" if valu
                                  else "hbm_read_bytes_per_s")
    extra = {}
    if measured:
        m = measured / 1e12 if valu else measured / 1e9
        a = p["valu_tops"] if valu else p["hbm_gb_per_s"]
        extra = {"peak_measured": round(m, 2), "frac_of_measured_peak": round(a / m, 4)}
    fp_arith = {}
    if "fp_arith_tops" in p:
        fp_arith = {"fp_arith": {
            "achieved": p["fp_arith_tops"], "peak": round(PEAK_FP_ARITH_TOPS, 2),
            "unit": "TOP/s (FP32/FP64 add, mul and FMA lane-instructions; one per SIMD "
                    "per 4 clocks: 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz)",
            "frac": p["fp_arith_frac"],
            "counters": os.path.relpath(VALU_MIX_HIGH_RES, ROOT)}}
    return dict({
        "kernel": PASS_KERNELS[name],
        "bound": "valu" if valu else "hbm",
        "achieved": p["valu_tops"] if valu else p["hbm_gb_per_s"],
        "peak": round(PEAK_VALU_TOPS, 2) if valu else PEAK_HBM_GBS,
        "unit": "TOP/s (VALU lane-ops; FP32 peak 157.3 TFLOP/s = 78.65e12 FMA lanes/s)"
                if valu else "GB/s",
        "frac": p["valu_frac"] if valu else p["hbm_frac"],
        "traffic": p["hbm_bytes_per_launch"],
        "avg_launch_ms": p["avg_launch_ms"],
        "valu_frac": p["valu_frac"],
        "hbm_frac": p["hbm_frac"],
        "counters": os.path.relpath(COUNTERS_HIGH_RES, ROOT),
    }, **extra, **fp_arith)


def section_spectral(args, steps, warmup):
    """The spectral atmosphere at the reference LUT sizes (4 orders): FP32
    and FP16 storage, Init as the API runs it and with kept temporaries, the
    lookups, and the demo's nine preset views as coherent lookup frames."""
    import numpy as np
    import precomputed_atmospheric_scattering_amd as atmo
    from precomputed_atmospheric_scattering_amd import configs, frames

    out = {}
    stream = atmo.Stream()
    atmo.reset_tiles()
    for half in (False, True):
        cfg = configs.spectral(half_precision=half)
        kw, orders = configs.model_kwargs(cfg)
        m = atmo.Model(pass_timing=not half, **kw)
        lk = Lookups(atmo, configs, kw) if not half else None
        for _ in range(warmup):
            m.init(orders, stream)
            if lk is not None:
                lk.run(m, stream)
        stream.synchronize()
        pre_ms, _, _, launches = timed_inits(atmo, m, orders, stream, steps, lk)
        e = {"ms_per_atmosphere": round(sum(pre_ms) / steps, 4),
             "precompute_p99_ms": round(percentile(pre_ms, 0.99), 4),
             "init_kept_temporaries_ms": round(
                 kept_init_ms(atmo, kw, orders, stream, steps), 4)}
        if launches:
            e["pass_ms_per_atmosphere"] = {k: round(sum(v) / steps, 4)
                                           for k, v in launches.items()}
        if lk is not None:
            e.update(lk.summary())
            lk.free()
            e["demo_frames"] = demo_frames(atmo, frames, np, m, kw, stream)
        out["fp16" if half else "fp32"] = e
        m.close()
    stream.close()
    out["atmospheres_per_s"] = round(1e3 / out["fp32"]["ms_per_atmosphere"], 2)
    out["workload"] = ("spectral atmosphere, 15 precomputed wavelengths = 5 iterations, "
                       "4 orders, LUTs 256x64 / 32x128x32x8 / 64x16")
    return out


def demo_frames(atmo, frames, np, model, kw, stream, reps=3):
    """The lookups on coherent batches: the frames of the demo's nine preset
    views at 1920x1080 (a sphere on the planet with light shafts; views
    from 7 km to 12000 km away). Per frame: GetSkyLuminance for every
    pixel, GetSunAndSkyIlluminance and GetSkyLuminanceToPoint for the pixels
    that hit the sphere or the ground."""
    n_rays = WIDTH * HEIGHT
    frame_ms, frame_rays = {}, {}
    fev = [atmo.Event() for _ in range(2)]
    outs = [atmo.DeviceArray((n_rays, 3)) for _ in range(2)]
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
        hout = [atmo.DeviceArray((max(nh, 1), 3)) for _ in range(4)]
        flush_caches(stream)
        fev[0].record(stream)
        for _ in range(reps):
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
        frame_ms[view] = fev[0].elapsed_ms(fev[1]) / reps
        frame_rays[view] = n_rays + 2 * nh
        for d in list(sky.values()) + (list(hd.values()) if hd else []) + hout:
            d.free()
    for d in outs:
        d.free()
    total_s = sum(frame_ms.values()) * 1e-3
    return {"frames_per_s": round(len(frame_ms) / total_s, 1),
            "lookups_per_s": round(sum(frame_rays.values()) / total_s, 1)}


def section_batched(args, steps, warmup):
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
    for m, o in zip(models, orders):
        m.init(o, stream)
    stream.synchronize()
    plain_runs = []
    for _ in range(2):
        t0 = time.perf_counter()
        for m, o in zip(models, orders):
            m.init(o, stream)
        stream.synchronize()
        plain_runs.append(time.perf_counter() - t0)
    plain_s = min(plain_runs)
    sample = (0, 21, 42, 63)
    ref = {i: models[i].lut("scattering") for i in sample}
    # Where the plain path's time goes: the same loop with every launch timed
    # by HIP events. Their sum is the GPU's busy time; the rest of the wall
    # time is idle (allocation, free, launch and host synchronisation), all
    # that batching can remove for models that fill the GPU.
    for m in models:
        _capi.check(atmo.lib().atmo_model_set_option(m.handle, _capi.OPTION_PASS_TIMING, 1),
                    "set_option")
    t0 = time.perf_counter()
    busy_ms = 0.0
    for m, o in zip(models, orders):
        m.init(o, stream)
        busy_ms += sum(ms for _, _, ms in m.pass_timings())
    stream.synchronize()
    timed_wall_ms = (time.perf_counter() - t0) * 1e3
    for m in models:
        _capi.check(atmo.lib().atmo_model_set_option(m.handle, _capi.OPTION_PASS_TIMING, 0),
                    "set_option")
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

    def timed_batch(use_graph, n_steps, n_warmup):
        b = atmo.Batch(models, orders, use_graph=use_graph)
        for _ in range(n_warmup):
            b.launch(stream)
        stream.synchronize()
        ms = []
        t_all = time.perf_counter()
        for _ in range(n_steps):
            t = time.perf_counter()
            b.launch(stream)
            stream.synchronize()
            ms.append((time.perf_counter() - t) * 1e3)
        total = time.perf_counter() - t_all
        launches = b.num_launches
        b.close()
        return ms, total, launches

    eager_ms, eager_s, _ = timed_batch(False, max(1, steps // 2), 1)
    graph_ms, graph_s, launches = timed_batch(True, steps, max(1, warmup))
    exact = all(np.array_equal(models[i].lut("scattering"), ref[i]) for i in sample)
    tuned_table = os.path.join(tempfile.mkdtemp(prefix="atmo_tiles_"), "tiles.txt")
    atmo.save_tiles(tuned_table)
    atmo.reset_tiles()
    _, default_s, default_launches = timed_batch(True, steps, max(1, warmup))
    exact = exact and all(np.array_equal(models[i].lut("scattering"), ref[i])
                          for i in sample)
    atmo.load_tiles(tuned_table)
    for m in models:
        m.close()
    stream.close()
    value = 64 * steps / graph_s
    plain = 64 / plain_s
    return {
        "workload": "64 random atmospheres from seed %d: bottom radius 3000-7000 km, "
                    "thickness 40-200 km, Rayleigh scale height 5-15 km and beta x0.3-3, "
                    "Mie scale height 0.5-3 km, g 0.6-0.9, albedo 0-0.5, LUT sizes "
                    "x{0.5,1,2} per dimension, 2-8 orders, 3 wavelengths, FP32" % args.seed,
        "scattering_texels_total": n_s_total,
        "atmospheres_per_s": round(value, 2),
        "batch_p50_ms": round(percentile(graph_ms, 0.5), 3),
        "batch_p99_ms": round(percentile(graph_ms, 0.99), 3),
        "launches_per_batch": launches,
        "autotune_s": round(tune_s, 3),
        "default_tiles_atmospheres_per_s": round(64 * steps / default_s, 2),
        "default_tiles_launches_per_batch": default_launches,
        "plain_atmospheres_per_s": round(plain, 2),
        "plain_kept_temporaries_atmospheres_per_s": round(64 / plain_kept_s, 2),
        "batched_eager_atmospheres_per_s": round(64 * len(eager_ms) / eager_s, 2),
        "speedup_vs_plain": round(value / plain, 3),
        # The plain loop's launches summed (HIP events) against its wall time:
        # the GPU is idle for the rest; 64 / busy time bounds any batching
        # that keeps the kernels as they are.
        "plain_gpu_busy_ms": round(busy_ms, 2),
        "plain_wall_ms_with_launch_timing": round(timed_wall_ms, 2),
        "plain_gpu_busy_frac": round(busy_ms / timed_wall_ms, 4),
        "plain_kernel_time_bound_atmospheres_per_s": round(64e3 / busy_ms, 2),
        "speedup_vs_plain_kept_temporaries": round(value * plain_kept_s / 64, 3),
        "bit_exact_vs_plain": {"models_checked": list(sample), "equal": exact},
    }


def section_serving(args, rate, duration, seed):
    """Serving loop under a Poisson stream of precompute requests (random
    atmospheres of the batched workload) at 'rate' per second for 'duration'
    seconds. Each finished model then serves one 1920x1080 query batch
    (GetSkyRadiance, GetSkyRadianceToPoint, GetSunAndSkyIrradiance; inputs
    made on the device from a seed: camera altitude 0-100 km, view direction
    uniform, sun zenith 0-100 deg, hit distance 0-50 km) and is released."""
    import precomputed_atmospheric_scattering_amd as atmo
    from precomputed_atmospheric_scattering_amd import _capi, configs

    rng = configs.SplitMix64(seed)
    arrivals, t = [], 0.0
    while True:
        t += -math.log(1.0 - rng.uniform()) / rate
        if t > duration:
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

    # Warm-up: one request, not counted.
    kw, o = configs.model_kwargs(configs.random_atmosphere(configs.SplitMix64(seed + 1)))
    wid = srv.submit_precompute(num_scattering_orders=o, **kw)
    srv.wait([wid])
    srv.release(wid)

    info, ready, batches = {}, [], {}
    pre, qry, batch_sizes = [], [], []
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
                pre.append(c)
                batch_sizes.append(c["batch_models"])
                if c["status"] == 0:
                    ready.append(c["request_id"])
            else:
                b = batches[c["model_id"]]
                b["left"] -= 1
                b["end"] = max(b["end"], c["end_ms"])
                b["start"] = min(b["start"], c["start_ms"])
                b["submit"] = min(b["submit"], c["submit_ms"])
                if b["left"] == 0:
                    qry.append(b)
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
            srv.submit_lookup(mid, "sky", n, [q["camera"], q["view_ray"],
                                              q["shadow_length"], q["sun_direction"]],
                              o[0:2])
            srv.submit_lookup(mid, "sky_to_point", n,
                              [q["camera"], q["point"], q["shadow_length"],
                               q["sun_direction"]], o[2:4])
            srv.submit_lookup(mid, "sun_and_sky", n,
                              [q["point"], q["normal"], q["sun_direction"]], o[4:6])
            batches[mid] = dict(left=3, end=0.0, start=float("inf"), submit=float("inf"),
                                buffers=(q, o))
            outstanding += 1
        peak_resident = max(peak_resident, srv.stats()["resident_bytes"])
    wall = time.perf_counter() - start
    st = srv.stats()
    srv.close()
    gen.close()
    for q, o in pool:
        for d in list(q.values()) + o:
            d.free()
    pre_lat = [c["end_ms"] - c["submit_ms"] for c in pre]
    qry_lat = [b["end"] - b["submit"] for b in qry]
    # Where the query-batch tail comes from: batches whose device work
    # overlapped a precompute batch's (they share the GPU, at the same stream
    # priority) against those that ran alone; and the wait from submission
    # to the start of the first lookup.
    busy = [(c["start_ms"], c["end_ms"]) for c in pre]
    over = [b["end"] - b["submit"] for b in qry
            if any(s < b["end"] and e > b["start"] for s, e in busy)]
    alone = [b["end"] - b["submit"] for b in qry
             if not any(s < b["end"] and e > b["start"] for s, e in busy)]
    return {
        "workload": "Poisson %.1f req/s for %.2f s of random atmospheres (batched-workload "
                    "distribution, seed %d), one 1920x1080 query batch of 3 lookups per "
                    "finished model" % (rate, duration, seed),
        "requests": len(arrivals),
        "atmospheres_per_s": round(len(pre) / wall, 3),
        "precompute_latency_p50_ms": round(percentile(pre_lat, 0.5), 3),
        "precompute_latency_p99_ms": round(percentile(pre_lat, 0.99), 3),
        "query_batch_latency_p50_ms": round(percentile(qry_lat, 0.5), 3),
        "query_batch_latency_p99_ms": round(percentile(qry_lat, 0.99), 3),
        "query_batches_overlapping_precompute": len(over),
        "query_batch_p99_ms_overlapping_precompute": round(percentile(over, 0.99), 3),
        "query_batch_p99_ms_alone": round(percentile(alone, 0.99), 3),
        "query_wait_to_start_p99_ms": round(percentile(
            [b["start"] - b["submit"] for b in qry], 0.99), 3),
        "sky_rays_per_s": round(3.0 * n * len(qry) / wall, 1),
        "mean_batch_models": round(sum(batch_sizes) / max(1, len(batch_sizes)), 2),
        "max_batch_models": max(batch_sizes) if batch_sizes else 0,
        "failed": failed, "peak_resident_bytes": peak_resident,
        "budget_bytes": st["budget_bytes"], "autotune_s": round(tune_s, 3),
    }


def run_headline(args):
    _progress("on-device rate ceilings")
    peaks = measure_peaks()
    res, launches, meta = section_high_resolution(
        args, args.steps, args.warmup, keep_luts=not args.no_cpu_baseline)
    passes = pass_figures(launches, meta["steps"], meta["n_s"], meta["iters"])
    result = {
        "metric": METRIC,
        "value": res.pop("value"),
        "unit": "atmospheres/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res.pop("ms_per_step"),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "high-resolution stress: spectral atmosphere (47-band 5778 K solar "
                        "spectrum, ozone), 15 precomputed wavelengths = 5 iterations, 8 "
                        "scattering orders, FP16 LUT storage with FP32 accumulation, LUTs "
                        "512x128 / 64x256x64x16 / 128x32; Init as the API runs it "
                        "(temporaries allocated and freed by the call), then 1920x1080 rays "
                        "of each of GetSkyLuminance, GetSkyLuminanceToPoint, "
                        "GetSunAndSkyIlluminance",
            "precomputed_wavelengths": 15, "scattering_orders": meta["orders"],
            "scattering_texels": meta["n_s"], "storage": "f16", "query_rays": WIDTH * HEIGHT,
        },
    }
    result.update(res)
    result["passes"] = passes
    result["roofline"] = roofline(passes)
    result["measured_peaks"] = peaks
    _progress("spectral section")
    result["spectral"] = section_spectral(args, max(3, args.steps // 2), max(1, args.warmup))
    _progress("batched section")
    result["batched"] = section_batched(args, max(3, args.steps // 2), max(1, args.warmup))
    _progress("serving section (Poisson 20/s)")
    result["serving"] = section_serving(args, args.rate, args.duration, args.seed)
    _progress("serving section (saturating burst)")
    result["serving_saturated"] = section_serving(args, 2000.0, 0.25, args.seed + 7)
    if not args.no_cpu_baseline:
        _progress("cpu baseline")
        threads, allowed = cpu_threads()
        base, _ = cpu_baseline(meta["cfg"], meta["luts"], args.cpu_budget, threads)
        from precomputed_atmospheric_scattering_amd import configs
        kw, _ = configs.model_kwargs(meta["cfg"])
        q = configs.sky_queries(2024, 20000, kw["bottom_radius"] / 1000.0,
                                kw["top_radius"] / 1000.0)
        base["sky_rays_per_s"] = round(cpu_query_baseline(meta["cfg"], meta["luts"], q,
                                                          threads), 1)
        base["value"] = round(base["value"], 8)
        base["cpu_model"] = cpu_model()
        base["host_cpus"] = os.cpu_count()
        base["allowed_cpus"] = allowed
        result["cpu_baseline"] = base
        result["speedup_vs_cpu"] = round(result["value"] / base["value"], 1)
    _json_line(result)


# ---------------------------------------------------------------------------
# Summaries of rocprofv3 outputs (CSV) as the markdown tables of profiles/:
#   python bench.py --summarize stats DIR
#   python bench.py --summarize counters TRACE_DIR SQ_DIR FETCH_DIR WRITE_DIR
#   python bench.py --summarize utilization TRACE_DIR SQ_DIR
#   python bench.py --summarize valu-mix TRACE_DIR MIX32_DIR MIX64_DIR
#   python bench.py --summarize plain-gaps TRACE_DIR
# ---------------------------------------------------------------------------
def _csv_rows(d, suffix):
    import csv
    import glob
    for f in sorted(glob.glob(os.path.join(d, "**", "*" + suffix), recursive=True)):
        with open(f) as fh:
            yield from csv.DictReader(fh)


def _kernel_base(name):
    """The kernel's own name in a (possibly demangled, qualified) trace name."""
    import re
    m = re.search(r"\b(k_[A-Za-z0-9_]+|__amd_[A-Za-z0-9_]+)", name)
    return m.group(1) if m else name.split("(")[0].split("<")[0].strip()


def _kernel_stats(d):
    return {_kernel_base(r["Name"]): r for r in _csv_rows(d, "kernel_stats.csv")}


def _counters(d):
    """kernel -> counter -> {dispatch id: value summed over dimensions}."""
    out = {}
    for r in _csv_rows(d, "counter_collection.csv"):
        per = out.setdefault(_kernel_base(r["Kernel_Name"]), {}).setdefault(
            r["Counter_Name"], {})
        per[r["Dispatch_Id"]] = per.get(r["Dispatch_Id"], 0.0) + float(r["Counter_Value"])
    return out


def summarize_stats(d):
    lines = ["| kernel | calls | avg ms | min ms | max ms | total ms | % |",
             "|---|---|---|---|---|---|---|"]
    for k, r in sorted(_kernel_stats(d).items(), key=lambda e: -float(e[1]["TotalDurationNs"])):
        lines.append("| %s | %s | %.3f | %.3f | %.3f | %.2f | %.1f |" % (
            k, r["Calls"], float(r["AverageNs"]) / 1e6, float(r["MinNs"]) / 1e6,
            float(r["MaxNs"]) / 1e6, float(r["TotalDurationNs"]) / 1e6, float(r["Percentage"])))
    return "\n".join(lines)


def summarize_counters(trace_dir, sq_dir, fetch_dir, write_dir, probe_mib=256):
    """Per kernel: trace durations, SQ counters, and HBM bytes from FETCH_SIZE
    and WRITE_SIZE (KiB) scaled by the factor that makes the copy probe (the
    largest copyBuffer dispatch, --copy-probe MiB read and written) read and
    write its own size."""
    trace, sq = _kernel_stats(trace_dir), _counters(sq_dir)
    fe, wr = _counters(fetch_dir), _counters(write_dir)
    cal_f = cal_w = None
    for k in fe:
        if "copyBuffer" in k:
            cal_f = probe_mib * 1024.0 / max(fe[k]["FETCH_SIZE"].values())
            cal_w = probe_mib * 1024.0 / max(wr[k]["WRITE_SIZE"].values())

    def avg(c, k, n):
        v = c.get(k, {}).get(n)
        return sum(v.values()) / len(v) if v else None

    fmt = lambda x, p=1: "-" if x is None else "%.*f" % (p, x)
    lines = ["calibration: FETCH_SIZE x %s, WRITE_SIZE x %s" % (fmt(cal_f, 3), fmt(cal_w, 3)),
             "",
             "| kernel | dispatches (trace) | avg ms (trace) | waves / dispatch | VALU instr / "
             "dispatch | VALU instr / wave | LDS instr / dispatch | SQ_ACTIVE_INST_VALU / "
             "dispatch | GRBM_GUI_ACTIVE / dispatch | fetch MB / dispatch | write MB / dispatch "
             "| HBM GB/s | VALU Tlane-op/s |",
             "|---|---|---|---|---|---|---|---|---|---|---|---|---|"]
    keys = sorted((k for k in set(sq) | set(fe) if not k.startswith("__amd")),
                  key=lambda k: -float(trace.get(k, {}).get("TotalDurationNs", 0) or 0))
    for k in keys:
        t = trace.get(k)
        ms = float(t["AverageNs"]) / 1e6 if t else None
        waves, valu = avg(sq, k, "SQ_WAVES"), avg(sq, k, "SQ_INSTS_VALU")
        f, w = avg(fe, k, "FETCH_SIZE"), avg(wr, k, "WRITE_SIZE")
        fmb = f * 1024 * cal_f / 1e6 if f is not None and cal_f else None
        wmb = w * 1024 * cal_w / 1e6 if w is not None and cal_w else None
        gbs = (fmb + wmb) / ms if ms and fmb is not None and wmb is not None else None
        vt = valu * 64 / (ms * 1e-3) / 1e12 if ms and valu else None
        lines.append("| %s | %d | %s | %s | %s | %s | %s | %s | %s | %s | %s | %s | %s |" % (
            k, int(t["Calls"]) if t else 0, fmt(ms, 3), fmt(waves, 0), fmt(valu, 0),
            fmt(valu / waves if valu and waves else None, 0), fmt(avg(sq, k, "SQ_INSTS_LDS"), 0),
            fmt(avg(sq, k, "SQ_ACTIVE_INST_VALU"), 0), fmt(avg(sq, k, "GRBM_GUI_ACTIVE"), 0),
            fmt(fmb, 1), fmt(wmb, 1), fmt(gbs, 0), fmt(vt, 2)))
    return "\n".join(lines)


UTILIZATION_COUNTERS = ("SQ_WAVES SQ_INSTS_VALU SQ_ACTIVE_INST_VALU SQ_ACTIVE_INST_VALU2 "
                        "SQ_LDS_IDX_ACTIVE SQ_LDS_BANK_CONFLICT SQ_WAIT_INST_LDS SQ_WAVE_CYCLES "
                        "GRBM_GUI_ACTIVE")


def summarize_utilization(trace_dir, sq_dir):
    """Issue and LDS utilization per kernel from one rocprofv3 --pmc run of
    UTILIZATION_COUNTERS (summed over the 8 XCDs by rocprofv3): cycles =
    GRBM_GUI_ACTIVE / 8; SIMD quad-cycles = 1024 SIMDs x cycles / 4; VALU
    issue = (SQ_ACTIVE_INST_VALU - SQ_ACTIVE_INST_VALU2) / SIMD quad-cycles,
    the share of SIMD quad-cycles issuing at least one VALU instruction;
    dual = 2 SQ_ACTIVE_INST_VALU2 / SQ_ACTIVE_INST_VALU, the share of VALU
    instructions issued in pairs; LDS busy = SQ_LDS_IDX_ACTIVE / (cycles x
    256 CUs); bank conflicts = SQ_LDS_BANK_CONFLICT / SQ_LDS_IDX_ACTIVE; LDS
    wait = SQ_WAIT_INST_LDS / SQ_WAVE_CYCLES."""
    trace, sq = _kernel_stats(trace_dir), _counters(sq_dir)

    def avg(k, n):
        v = sq.get(k, {}).get(n)
        return sum(v.values()) / len(v) if v else None

    pct = lambda x: "-" if x is None else "%.0f%%" % (100.0 * x)
    lines = ["| kernel | dispatches (trace) | avg ms (trace) | VALU instr / wave | VALU issue | "
             "dual | LDS busy | LDS bank conflicts | LDS wait / wave-cycles |",
             "|---|---|---|---|---|---|---|---|---|"]
    keys = sorted((k for k in sq if not k.startswith("__amd")),
                  key=lambda k: -float(trace.get(k, {}).get("TotalDurationNs", 0) or 0))
    for k in keys:
        t = trace.get(k)
        g, waves, valu = avg(k, "GRBM_GUI_ACTIVE"), avg(k, "SQ_WAVES"), avg(k, "SQ_INSTS_VALU")
        act, act2 = avg(k, "SQ_ACTIVE_INST_VALU"), avg(k, "SQ_ACTIVE_INST_VALU2")
        lds, conf = avg(k, "SQ_LDS_IDX_ACTIVE"), avg(k, "SQ_LDS_BANK_CONFLICT")
        wait, wcyc = avg(k, "SQ_WAIT_INST_LDS"), avg(k, "SQ_WAVE_CYCLES")
        cycles = g / 8.0 if g else None
        quad = 1024.0 * cycles / 4.0 if cycles else None
        issue = (act - act2) / quad if quad and act is not None and act2 is not None else None
        dual = 2.0 * act2 / act if act and act2 is not None else None
        busy = lds / (cycles * 256.0) if cycles and lds is not None else None
        conflicts = conf / lds if lds and conf is not None else None
        lw = wait / wcyc if wcyc and wait is not None else None
        lines.append("| %s | %d | %s | %s | %s | %s | %s | %s | %s |" % (
            k, int(t["Calls"]) if t else 0,
            "%.3f" % (float(t["AverageNs"]) / 1e6) if t else "-",
            "%.0f" % (valu / waves) if valu and waves else "-", pct(issue), pct(dual),
            pct(busy), "-" if conflicts is None else "%.1f%%" % (100.0 * conflicts),
            "-" if lw is None else "%.1f%%" % (100.0 * lw)))
    return "\n".join(lines)


MIX32_COUNTERS = ("SQ_WAVES SQ_INSTS_VALU SQ_INSTS_VALU_ADD_F32 SQ_INSTS_VALU_MUL_F32 "
                  "SQ_INSTS_VALU_FMA_F32 SQ_INSTS_VALU_TRANS_F32 SQ_INSTS_VALU_CVT "
                  "SQ_INSTS_VALU_INT32")
MIX64_COUNTERS = ("SQ_WAVES SQ_INSTS_VALU SQ_INSTS_VALU_ADD_F64 SQ_INSTS_VALU_MUL_F64 "
                  "SQ_INSTS_VALU_FMA_F64 SQ_INSTS_VALU_TRANS_F64 SQ_INSTS_VALU_INT64")


def summarize_valu_mix(trace_dir, mix32_dir, mix64_dir):
    """The VALU instruction mix per kernel from two rocprofv3 --pmc runs
    (MIX32_COUNTERS, MIX64_COUNTERS): FP32 and FP64 arithmetic (add + mul +
    FMA), transcendental, conversion and integer instructions per dispatch,
    as shares of SQ_INSTS_VALU, and the FP arithmetic lane-instruction rate
    over the trace's average duration against PEAK_FP_ARITH_TOPS."""
    trace, m32, m64 = _kernel_stats(trace_dir), _counters(mix32_dir), _counters(mix64_dir)

    def avg(c, k, n):
        v = c.get(k, {}).get(n)
        return sum(v.values()) / len(v) if v else 0.0

    lines = ["| kernel | dispatches (trace) | avg ms (trace) | VALU instr / dispatch | "
             "FP32 arith / dispatch | FP64 arith / dispatch | FP32 arith | FP64 arith | "
             "transcendental | conversions | integer | FP arith Tlane-op/s | of single issue |",
             "|---|---|---|---|---|---|---|---|---|---|---|---|---|"]
    keys = sorted((k for k in m32 if not k.startswith("__amd")),
                  key=lambda k: -float(trace.get(k, {}).get("TotalDurationNs", 0) or 0))
    for k in keys:
        t = trace.get(k)
        ms = float(t["AverageNs"]) / 1e6 if t else None
        valu = avg(m32, k, "SQ_INSTS_VALU")
        f32 = sum(avg(m32, k, "SQ_INSTS_VALU_%s_F32" % o) for o in ("ADD", "MUL", "FMA"))
        f64 = sum(avg(m64, k, "SQ_INSTS_VALU_%s_F64" % o) for o in ("ADD", "MUL", "FMA"))
        trans = avg(m32, k, "SQ_INSTS_VALU_TRANS_F32") + avg(m64, k, "SQ_INSTS_VALU_TRANS_F64")
        cvt = avg(m32, k, "SQ_INSTS_VALU_CVT")
        ints = avg(m32, k, "SQ_INSTS_VALU_INT32") + avg(m64, k, "SQ_INSTS_VALU_INT64")
        share = lambda x: "%.0f%%" % (100.0 * x / valu) if valu else "-"
        rate = (f32 + f64) * 64 / (ms * 1e-3) / 1e12 if ms else None
        lines.append("| %s | %d | %s | %.0f | %.0f | %.0f | %s | %s | %s | %s | %s | %s | %s |" % (
            k, int(t["Calls"]) if t else 0, "%.3f" % ms if ms else "-", valu, f32, f64,
            share(f32), share(f64), share(trans), share(cvt), share(ints),
            "-" if rate is None else "%.2f" % rate,
            "-" if rate is None else "%.2f" % (rate / PEAK_FP_ARITH_TOPS)))
    return "\n".join(lines)


def summarize_plain_gaps(trace_dir, seed=2024):
    """Where the plain one-at-a-time Init loop of the batched workload spends
    its time, from a kernel trace (--kernel-trace --hip-runtime-trace) of
    `bench.py --workload batched`: the section runs the 64 Inits three times
    (warm-up and two timed loops) before anything else; each Init of these
    3-wavelength models starts with its one k_transmittance dispatch. For the two timed loops, per size class
    (scattering texels): the Init's span (its first kernel's start to the next
    Init's), the union of its kernels' intervals, and the HIP runtime calls of
    the Init's host thread that overlap its idle time."""
    from precomputed_atmospheric_scattering_amd import configs
    ks = sorted(((int(r["Start_Timestamp"]), int(r["End_Timestamp"]),
                  _kernel_base(r["Kernel_Name"])) for r in _csv_rows(trace_dir,
                                                                     "kernel_trace.csv")))
    starts = [i for i, k in enumerate(ks) if k[2] == "k_transmittance"]
    if len(starts) < 3 * 64 + 1:
        raise SystemExit("expected at least %d Inits in the trace, found %d" %
                         (3 * 64 + 1, len(starts)))
    api = sorted((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), r["Function"])
                 for r in _csv_rows(trace_dir, "hip_api_trace.csv"))
    cfgs = configs.batch(seed, 64)
    n_s = [texel_counts(c["lut_sizes"])[1] for c in cfgs]
    classes = {}
    for loop in (1, 2):
        for m in range(64):
            i = loop * 64 + m
            a, b = starts[i], starts[i + 1]
            t0, t1 = ks[a][0], ks[b][0]
            if i == 3 * 64 - 1:
                # The last plain Init: what follows it is other work of the
                # section (LUT copies among it), so its span ends with its
                # last pass kernel.
                t1 = max(e for _, e, k in ks[a:b] if k.startswith("k_"))
            busy, cur_s, cur_e = 0, None, None
            for s, e, _ in ks[a:b]:
                if s >= t1:
                    break
                e = min(e, t1)
                if cur_e is None or s > cur_e:
                    if cur_e is not None:
                        busy += cur_e - cur_s
                    cur_s, cur_e = s, e
                else:
                    cur_e = max(cur_e, e)
            if cur_e is not None:
                busy += cur_e - cur_s
            calls = {}
            for s, e, f in api:
                if e < t0 or s > t1:
                    continue
                calls[f] = calls.get(f, 0) + min(e, t1) - max(s, t0)
            c = classes.setdefault(n_s[m], dict(n=0, span=0, busy=0, orders=0, calls={}))
            c["n"] += 1
            c["span"] += t1 - t0
            c["busy"] += busy
            c["orders"] += cfgs[m]["orders"]
            for f, d in calls.items():
                c["calls"][f] = c["calls"].get(f, 0) + d
    lines = ["| scattering texels | Inits | mean orders | Init span ms | kernels busy ms | "
             "idle ms | busy fraction | HIP calls in the span (ms per Init) |",
             "|---|---|---|---|---|---|---|---|"]
    tot = dict(n=0, span=0, busy=0)
    for k in sorted(classes):
        c = classes[k]
        top = sorted(c["calls"].items(), key=lambda e: -e[1])[:4]
        lines.append("| %d | %d | %.1f | %.3f | %.3f | %.3f | %.3f | %s |" % (
            k, c["n"], c["orders"] / c["n"], c["span"] / c["n"] / 1e6,
            c["busy"] / c["n"] / 1e6, (c["span"] - c["busy"]) / c["n"] / 1e6,
            c["busy"] / c["span"],
            ", ".join("%s %.3f" % (f, d / c["n"] / 1e6) for f, d in top)))
        for key in tot:
            tot[key] += c[key]
    lines.append("| all | %d | | %.3f | %.3f | %.3f | %.3f | |" % (
        tot["n"], tot["span"] / tot["n"] / 1e6, tot["busy"] / tot["n"] / 1e6,
        (tot["span"] - tot["busy"]) / tot["n"] / 1e6, tot["busy"] / tot["span"]))
    return "\n".join(lines)


def one_line(metric, value, unit, steps, warmup, ms_per_step, higher, workload, extra):
    d = {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": steps,
         "warmup": warmup, "ms_per_step": ms_per_step, "higher_is_better": higher,
         "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
         "config": {"workload": workload}}
    d.update(extra)
    _json_line(d)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--workload", default="headline",
                    choices=("headline", "high-resolution", "spectral", "batched", "serving"),
                    help="headline (default): the high-resolution workload's line with the "
                         "spectral, batched and serving sections; the others run one section "
                         "alone")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--duration", type=float, default=10.0,
                    help="serving: seconds of request arrivals")
    ap.add_argument("--rate", type=float, default=20.0,
                    help="serving: precompute requests per second (Poisson)")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-lookups", action="store_true",
                    help="high-resolution: Init only (for counter runs)")
    ap.add_argument("--cpu-budget", type=float, default=20.0,
                    help="seconds of CPU oracle work for the cpu_baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="first copy MIB MiB device-to-device once (one blit kernel "
                         "reading and writing MIB MiB), as the reference point of "
                         "HBM byte counters")
    ap.add_argument("--summarize", nargs="+", metavar="ARG",
                    help="print a profiles/ table from rocprofv3 CSV output and exit: "
                         "'stats DIR', 'counters TRACE SQ FETCH WRITE', 'utilization TRACE "
                         "SQ', 'valu-mix TRACE MIX32 MIX64' or 'plain-gaps TRACE'")
    args = ap.parse_args()
    if args.summarize:
        kind, dirs = args.summarize[0], args.summarize[1:]
        fn = {"stats": summarize_stats, "counters": summarize_counters,
              "utilization": summarize_utilization, "valu-mix": summarize_valu_mix,
              "plain-gaps": summarize_plain_gaps}.get(kind)
        if fn is None:
            ap.error("unknown --summarize kind %r" % kind)
        print(fn(*dirs))
        return None
    if args.gpus != 1:
        print("bench.py measures one GPU only: the library has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import precomputed_atmospheric_scattering_amd as atmo

    if args.copy_probe > 0:
        nbytes = args.copy_probe << 20
        src, dst = atmo.DeviceArray((nbytes,), np.uint8), atmo.DeviceArray((nbytes,), np.uint8)
        atmo._capi.check(atmo.lib().atmo_memcpy_to_device(dst.ptr, src.ptr, nbytes, None),
                         "copy probe")
        atmo.synchronize()
        src.free()
        dst.free()

    if args.workload == "headline":
        return run_headline(args)
    if args.workload == "high-resolution":
        res, launches, meta = section_high_resolution(args, args.steps, args.warmup,
                                                      with_lookups=not args.no_lookups)
        extra = dict(res)
        extra.pop("value")
        extra.pop("ms_per_step")
        extra["passes"] = pass_figures(launches, meta["steps"], meta["n_s"], meta["iters"])
        return one_line("high-resolution LUT precompute", res["value"], "atmospheres/s",
                        args.steps, args.warmup, res["ms_per_step"], True,
                        "high-resolution stress (64x256x64x16, 15 wavelengths, 8 orders, "
                        "FP16 storage)", extra)
    if args.workload == "spectral":
        res = section_spectral(args, args.steps, args.warmup)
        return one_line("spectral LUT precompute", res["atmospheres_per_s"], "atmospheres/s",
                        args.steps, args.warmup, res["fp32"]["ms_per_atmosphere"], True,
                        res.pop("workload"), res)
    if args.workload == "batched":
        res = section_batched(args, args.steps, args.warmup)
        return one_line("batched LUT precompute throughput (64 random atmospheres)",
                        res["atmospheres_per_s"], "atmospheres/s", args.steps, args.warmup,
                        res["batch_p50_ms"], True, res.pop("workload"), res)
    res = section_serving(args, args.rate, args.duration, args.seed)
    return one_line("served precompute latency p50/p99 and throughput (Poisson arrivals)",
                    res["atmospheres_per_s"], "atmospheres/s", res["requests"], 1,
                    round(1e3 / max(res["atmospheres_per_s"], 1e-9), 3), True,
                    res.pop("workload"), res)


if __name__ == "__main__":
    main()
