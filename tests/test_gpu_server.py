"""Serving loop (atmo_server_*): continuous batching of precompute requests,
admission by device-memory footprint, lookups chained on in-flight models.

The bar for "same alone and in the mix": every LUT of a model precomputed by
the server, among other requests, in concurrent batches on several streams,
equals bit for bit the LUT of the same model initialised alone with
Model.init, and so do the lookups run on it (the kernels and their per-texel
arithmetic do not depend on the batch)."""
import numpy as np
import pytest

import parity
import precomputed_atmospheric_scattering_amd as atmo
from precomputed_atmospheric_scattering_amd import _capi, configs

pytestmark = pytest.mark.gpu

LUTS = ("transmittance", "scattering", "single_mie_scattering", "irradiance")
TINY = dict(transmittance_width=32, transmittance_height=8, scattering_r=4,
            scattering_mu=16, scattering_mu_s=4, scattering_nu=4,
            irradiance_width=8, irradiance_height=4)


def mix(seed, count):
    """Random atmospheres of the batched workload at reduced LUT sizes (each
    dimension 1/8 .. 1/2 of the default, as drawn), orders 2..8."""
    rng = configs.SplitMix64(seed)
    out = []
    for _ in range(count):
        c = configs.random_atmosphere(rng)
        c["lut_sizes"] = {k: max(4, v // 4) for k, v in c["lut_sizes"].items()}
        c["lut_sizes"]["scattering_mu"] += c["lut_sizes"]["scattering_mu"] % 2
        out.append(c)
    return out


def queries(cfg, n, seed):
    q = configs.sky_queries(seed, n, cfg["bottom_radius"] / 1000.0,
                            cfg["top_radius"] / 1000.0)
    return {k: atmo.DeviceArray.from_host(v) for k, v in q.items()}


def outputs(n):
    return [atmo.DeviceArray((n, 3)) for _ in range(2)]


def alone(cfg, q, n):
    kw, orders = configs.model_kwargs(cfg)
    m = atmo.Model(**kw)
    m.init(orders)
    luts = {k: m.lut(k) for k in LUTS}
    o = outputs(n)
    m.sky_radiance_device(n, q["camera"], q["view_ray"], q["shadow_length"],
                          q["sun_direction"], o[0], o[1])
    p = outputs(n)
    m.sun_and_sky_irradiance_device(n, q["point"], q["normal"], q["sun_direction"],
                                    p[0], p[1])
    atmo.synchronize()
    res = luts, [a.to_host() for a in o + p]
    m.close()
    return res


def test_mix_equals_alone():
    cfgs = mix(2024, 8)
    n = 4096
    qs = [queries(c, n, 100 + i) for i, c in enumerate(cfgs)]
    ref = [alone(c, q, n) for c, q in zip(cfgs, qs)]
    srv = atmo.Server(max_inflight_batches=2, max_batch_models=3, lookup_streams=2)
    ids, outs, lookup_ids = [], [], []
    for i, c in enumerate(cfgs):
        kw, orders = configs.model_kwargs(c)
        mid = srv.submit_precompute(num_scattering_orders=orders, **kw)
        ids.append(mid)
        # Lookups submitted right behind their precompute wait for it on the
        # device.
        o, p = outputs(n), outputs(n)
        q = qs[i]
        lookup_ids.append(srv.submit_lookup(
            mid, "sky", n, [q["camera"], q["view_ray"], q["shadow_length"],
                            q["sun_direction"]], o))
        lookup_ids.append(srv.submit_lookup(
            mid, "sun_and_sky", n, [q["point"], q["normal"], q["sun_direction"]], p))
        outs.append(o + p)
    done = srv.wait(ids + lookup_ids)
    assert all(done[i]["status"] == 0 for i in ids + lookup_ids), done
    # Batches of up to 3 models: requests really were batched together.
    assert max(done[i]["batch_models"] for i in ids) > 1
    for i, mid in enumerate(ids):
        m = srv.model(mid)
        for k in LUTS:
            assert np.array_equal(m.lut(k), ref[i][0][k]), (i, k)
        for a, b in zip(outs[i], ref[i][1]):
            assert np.array_equal(a.to_host(), b), i
    st = srv.stats()
    assert st["completed_precomputes"] == len(ids)
    assert st["completed_lookups"] == len(lookup_ids)
    assert st["failed"] == 0
    srv.close()


def test_matches_oracle():
    cfg = configs.earth(lut_sizes=parity.SMALL_SIZES, orders=3)
    ref = parity.oracle_init(cfg)
    srv = atmo.Server()
    kw, orders = configs.model_kwargs(cfg)
    mid = srv.submit_precompute(num_scattering_orders=orders, **kw)
    assert srv.wait([mid])[mid]["status"] == 0
    m = srv.model(mid)
    for name, r in ref.items():
        ch = 4 if name == "scattering" else 3
        e = parity.rel_err(m.lut(name).reshape(r.shape), r, channels=ch)[0]
        assert e <= parity.TOL_FP32, (name, e)
    srv.close()


def footprint_of(cfg):
    kw, _ = configs.model_kwargs(cfg)
    m = atmo.Model(**kw)
    fp = m.footprint()
    m.close()
    return fp


def test_admission_by_footprint():
    cfg = configs.earth(lut_sizes=parity.MID_SIZES, orders=2)
    resident, scratch = footprint_of(cfg)
    # Room for two resident models and one Init's temporaries, not more.
    budget = 2 * resident + scratch + resident // 2
    srv = atmo.Server(memory_budget=budget, max_inflight_batches=2, max_batch_models=4)
    kw, orders = configs.model_kwargs(cfg)
    ids = [srv.submit_precompute(num_scattering_orders=orders, **kw) for _ in range(5)]
    finished = {}
    peak = 0
    import time
    deadline = time.monotonic() + 60
    while len(finished) < len(ids):
        assert time.monotonic() < deadline, srv.stats()
        st = srv.stats()
        peak = max(peak, st["resident_bytes"] + st["scratch_bytes"])
        assert st["resident_models"] <= 2, st
        for c in srv.poll(timeout_ms=10):
            finished[c["request_id"]] = c
            assert c["status"] == 0, c
            # Free the model so that the next request can be admitted.
            srv.release(c["request_id"])
    assert peak <= budget
    # Everything released: nothing stays resident once the server is idle.
    deadline = time.monotonic() + 10
    while srv.stats()["resident_models"] > 0:
        assert time.monotonic() < deadline
        time.sleep(0.01)
    assert srv.stats()["resident_bytes"] == 0
    srv.close()


def test_request_larger_than_budget_fails():
    cfg = configs.earth(lut_sizes=parity.SMALL_SIZES)
    srv = atmo.Server(memory_budget=1 << 16)
    kw, orders = configs.model_kwargs(cfg)
    mid = srv.submit_precompute(num_scattering_orders=orders, **kw)
    c = srv.wait([mid])[mid]
    assert c["status"] == 2  # ATMO_ERROR_OUT_OF_MEMORY
    with pytest.raises(atmo.AtmoError):
        srv.model(mid)
    srv.close()


def test_lookup_on_unknown_or_released_model():
    srv = atmo.Server()
    n = 16
    q = queries(configs.earth(), n, 5)
    o = outputs(n)
    ins = [q["camera"], q["view_ray"], q["shadow_length"], q["sun_direction"]]
    rid = srv.submit_lookup(12345, "sky", n, ins, o)
    assert srv.wait([rid])[rid]["status"] == 4  # ATMO_ERROR_NOT_INITIALIZED
    kw, orders = configs.model_kwargs(configs.earth(lut_sizes=TINY, orders=2))
    mid = srv.submit_precompute(num_scattering_orders=orders, **kw)
    assert srv.wait([mid])[mid]["status"] == 0
    srv.release(mid)
    rid = srv.submit_lookup(mid, "sky", n, ins, o)
    assert srv.wait([rid])[rid]["status"] == 4
    srv.close()


def test_invalid_request_rejected_at_submit():
    srv = atmo.Server()
    kw, orders = configs.model_kwargs(configs.earth(lut_sizes=TINY))
    kw["top_radius"] = kw["bottom_radius"]
    with pytest.raises(atmo.AtmoError):
        srv.submit_precompute(num_scattering_orders=orders, **kw)
    srv.close()


def test_latencies_are_ordered():
    srv = atmo.Server(max_inflight_batches=3, max_batch_models=2)
    ids = []
    for c in mix(7, 6):
        kw, orders = configs.model_kwargs(c)
        ids.append(srv.submit_precompute(num_scattering_orders=orders, **kw))
    done = srv.wait(ids)
    for c in done.values():
        assert c["status"] == 0
        assert 0.0 <= c["submit_ms"] <= c["start_ms"] <= c["end_ms"]
    srv.close()


def test_full_size_requests_and_query_batches_equal_alone():
    """The serving workload's shapes: atmospheres of the batched workload at
    their drawn (full) LUT sizes, served in batches, with 1920x1080-ray query
    batches of each lookup chained on them; the LUTs and every lookup output
    equal those of the same model initialised alone, bit for bit."""
    cfgs = configs.batch(31, 3)
    n = 1920 * 1080
    qs = [queries(c, n, 500 + i) for i, c in enumerate(cfgs)]
    ref = []
    for c, q in zip(cfgs, qs):
        kw, orders = configs.model_kwargs(c)
        m = atmo.Model(**kw)
        m.init(orders)
        o = [outputs(n) for _ in range(3)]
        m.sky_radiance_device(n, q["camera"], q["view_ray"], q["shadow_length"],
                              q["sun_direction"], *o[0])
        m.sky_radiance_to_point_device(n, q["camera"], q["point"], q["shadow_length"],
                                       q["sun_direction"], *o[1])
        m.sun_and_sky_irradiance_device(n, q["point"], q["normal"], q["sun_direction"],
                                        *o[2])
        atmo.synchronize()
        ref.append(({k: m.lut(k) for k in LUTS}, [a.to_host() for p in o for a in p]))
        for p in o:
            for a in p:
                a.free()
        m.close()
    srv = atmo.Server(max_inflight_batches=2, max_batch_models=3, lookup_streams=2)
    ids, lookup_ids, outs = [], [], []
    for c, q in zip(cfgs, qs):
        kw, orders = configs.model_kwargs(c)
        mid = srv.submit_precompute(num_scattering_orders=orders, **kw)
        ids.append(mid)
        o = [outputs(n) for _ in range(3)]
        lookup_ids.append(srv.submit_lookup(
            mid, "sky", n, [q["camera"], q["view_ray"], q["shadow_length"],
                            q["sun_direction"]], o[0]))
        lookup_ids.append(srv.submit_lookup(
            mid, "sky_to_point", n, [q["camera"], q["point"], q["shadow_length"],
                                     q["sun_direction"]], o[1]))
        lookup_ids.append(srv.submit_lookup(
            mid, "sun_and_sky", n, [q["point"], q["normal"], q["sun_direction"]], o[2]))
        outs.append(o)
    done = srv.wait(ids + lookup_ids)
    assert all(done[i]["status"] == 0 for i in ids + lookup_ids), done
    for i, mid in enumerate(ids):
        m = srv.model(mid)
        for k in LUTS:
            assert np.array_equal(m.lut(k), ref[i][0][k]), (i, k)
        got = [a.to_host() for p in outs[i] for a in p]
        for j, (a, b) in enumerate(zip(got, ref[i][1])):
            assert np.array_equal(a, b), (i, j)
    srv.close()


def test_illuminance_requests_in_mix_equal_alone():
    """Illuminance-mode requests (15 wavelengths: 5 Precompute iterations
    run as parallel jobs with their own accumulators), FP32 and FP16, batched
    by the server with 3-wavelength requests, give the LUTs of the same
    models initialised alone."""
    small = dict(transmittance_width=64, transmittance_height=16, scattering_r=8,
                 scattering_mu=32, scattering_mu_s=8, scattering_nu=4,
                 irradiance_width=16, irradiance_height=4)
    cfgs = [configs.spectral(orders=3, lut_sizes=small),
            configs.spectral(half_precision=True, orders=2, lut_sizes=small)]
    cfgs += mix(77, 2)
    ref = []
    for c in cfgs:
        kw, orders = configs.model_kwargs(c)
        m = atmo.Model(**kw)
        m.init(orders)
        ref.append({k: m.lut(k) for k in LUTS})
        m.close()
    srv = atmo.Server(max_inflight_batches=1, max_batch_models=4)
    ids = []
    for c in cfgs:
        kw, orders = configs.model_kwargs(c)
        ids.append(srv.submit_precompute(num_scattering_orders=orders, **kw))
    done = srv.wait(ids)
    assert all(done[i]["status"] == 0 for i in ids), done
    for i, mid in enumerate(ids):
        m = srv.model(mid)
        for k in LUTS:
            assert np.array_equal(m.lut(k), ref[i][k]), (i, k)
    srv.close()
