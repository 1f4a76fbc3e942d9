"""Writes tests/golden/high_resolution_fp64.npz: the FP64 oracle's full Init
(8 scattering orders, all 5 Precompute iterations) of the high-resolution
workload (the spectral atmosphere with ozone, 15 precomputed wavelengths,
LUTs 512x128 / 64x256x64x16 / 128x32), sampled at seeded texels.

The full chain takes about 3.6 hours of host time on 6 threads, too
long for a test run, so it is computed once here and the GPU tests compare
the product's LUTs (FP32 and FP16 storage) with it:
  - scattering (RGBA) and single_mie_scattering (RGB) at N_TEXELS texels of
    the packed LUT (flat texel indices in 'texels', seeded);
  - the whole irradiance LUT (RGB).
The transmittance LUT is not stored: the tests recompute it in full.

    ORACLE_THREADS=8 python tests/golden/make_high_resolution_golden.py
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

import oracle_lib as ol  # noqa: E402
from precomputed_atmospheric_scattering_amd import configs  # noqa: E402

N_TEXELS = 10000
SEED = 2027
OUT = os.path.join(HERE, "high_resolution_fp64.npz")


def main():
    cfg = configs.high_resolution(half_precision=False)
    kw, orders = configs.model_kwargs(cfg)
    sizes = kw.pop("lut_sizes")
    spec = ol.ModelSpec(**kw)
    s = ol.sizes(**sizes)
    t0 = time.time()
    ref = ol.model_init(spec, s, orders)
    dt = time.time() - t0
    n_s = s.scattering_r * s.scattering_mu * s.scattering_mu_s * s.scattering_nu
    rng = np.random.Generator(np.random.PCG64(SEED))
    texels = np.sort(rng.choice(n_s, N_TEXELS, replace=False)).astype(np.int64)
    S = ref["scattering"].reshape(-1, 4)
    M = ref["single_mie_scattering"].reshape(-1, 4)
    E = ref["irradiance"].reshape(-1, 4)
    np.savez_compressed(
        OUT, texels=texels, scattering=S[texels], single_mie_scattering=M[texels, :3],
        irradiance=E[:, :3],
        # Channel maxima over the whole LUTs, for the error metric's floor.
        scattering_max=np.abs(S).max(axis=0), single_mie_scattering_max=np.abs(M[:, :3]).max(axis=0),
        orders=orders, seconds=dt, threads=ol.lib().or_get_threads())
    print("wrote %s (%.0f s, %d threads)" % (OUT, dt, ol.lib().or_get_threads()))


if __name__ == "__main__":
    main()
