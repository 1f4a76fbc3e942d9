"""The tile-shape table of the autotuner (atmo_get_tiles, atmo_set_tiles,
atmo_tune_save/load/reset): host-only calls, no GPU needed."""
import pytest

import precomputed_atmospheric_scattering_amd as atmo
from precomputed_atmospheric_scattering_amd import AtmoError

SMALL = dict(scattering_r=16, scattering_mu=64, scattering_mu_s=16, scattering_nu=4)
WIDE = dict(scattering_mu_s=64, scattering_nu=16)   # rows of 1024 texels
LONG = dict(scattering_mu_s=128, scattering_nu=16)  # 2048: gather kernel


@pytest.fixture(autouse=True)
def clean_table():
    atmo.reset_tiles()
    yield
    atmo.reset_tiles()


def test_candidates():
    sd, ms = atmo.tile_candidates()
    assert sd[0] == (4, 256) and (4, 128) in sd
    assert len(set(sd)) == len(sd) and all(b % 64 == 0 for _, b in sd)
    assert ms == [(1, 256), (2, 256), (2, 512)]


def test_defaults_follow_row_length():
    t = atmo.get_tiles()
    # 4 nu per thread; 128-thread blocks up to 256 threads per (r, mu_s)
    # plane (MU x ceil(NU / 4) = 128 x 2 here), 256 above.
    assert (t["sd_nu_per_thread"], t["sd_block_threads"]) == (4, 128)
    big = atmo.get_tiles(dict(scattering_mu=256, scattering_nu=16))
    assert (big["sd_nu_per_thread"], big["sd_block_threads"]) == (4, 256)
    def ms(z=None):
        g = atmo.get_tiles(z)
        return g["ms_texels_per_thread"], g["ms_block_threads"]
    assert ms() == (1, 256)  # 8 x 32 = 256 texels per row
    assert not t["sd_tuned"] and not t["ms_tuned"]
    assert ms(dict(scattering_mu_s=64)) == (2, 256)
    assert ms(WIDE) == (2, 512)


def test_default_row_order_is_ray_order():
    # Ray order at every size (at the reference sizes it fetches a quarter of
    # the r-major order's bytes).
    for z in (None, dict(scattering_r=64, scattering_mu_s=48),
              dict(scattering_r=64, scattering_mu=256, scattering_mu_s=64, scattering_nu=16),
              dict(scattering_r=4, scattering_mu=16, scattering_mu_s=4, scattering_nu=4)):
        assert atmo.get_tiles(z)["ms_row_order"] == "ray"


def test_set_get_by_size_class():
    atmo.set_tiles(SMALL, 1, 256, 2, 256)
    t = atmo.get_tiles(SMALL)
    assert (t["sd_nu_per_thread"], t["sd_block_threads"], t["ms_texels_per_thread"],
            t["ms_block_threads"]) == (1, 256, 2, 256)
    assert t["sd_tuned"] and t["ms_tuned"]
    # The density class is (MU, NU) and the multiple-scattering class
    # (R, MU, MU_S, NU): other R (or other MU_S) is another class.
    other = dict(SMALL, scattering_r=32)
    t2 = atmo.get_tiles(other)
    assert (t2["sd_nu_per_thread"], t2["sd_block_threads"]) == (1, 256)
    assert t2["sd_tuned"] and not t2["ms_tuned"] and t2["ms_texels_per_thread"] == 1
    assert atmo.get_tiles()["sd_nu_per_thread"] == 4  # default class untouched
    # The row order is pinned with the rest of the shape.
    atmo.set_tiles(SMALL, 4, 256, 1, 256, ms_row_order="mu-major")
    assert atmo.get_tiles(SMALL)["ms_row_order"] == "mu-major"
    atmo.set_tiles(SMALL, 4, 256, 2, 512)  # keeps the class's order
    assert atmo.get_tiles(SMALL)["ms_row_order"] == "mu-major"
    atmo.set_tiles(SMALL, 4, 256, 2, 256, ms_row_order="ray")
    assert atmo.get_tiles(SMALL)["ms_row_order"] == "ray"
    atmo.set_tiles(SMALL, 4, 256, 2, 256, ms_row_order="r-major")
    assert atmo.get_tiles(SMALL)["ms_row_order"] == "r-major"
    assert atmo.get_tiles(other)["ms_row_order"] == "ray"


@pytest.mark.parametrize("shape", [(3, 256, 1, 256), (4, 96, 1, 256), (4, 256, 3, 256),
                                   (4, 512, 1, 256), (4, 256, 4, 256), (4, 256, 1, 512),
                                   (4, 256, 2, 128)])
def test_unknown_shapes_rejected(shape):
    with pytest.raises(AtmoError):
        atmo.set_tiles(None, *shape)


def test_multiple_scattering_tile_must_hold_the_row():
    with pytest.raises(AtmoError):
        atmo.set_tiles(WIDE, 4, 256, 2, 256)  # 1024 texels > 2 x 256
    atmo.set_tiles(WIDE, 4, 256, 2, 512)
    # Gather-path rows take any shape.
    atmo.set_tiles(LONG, 4, 256, 1, 256)


def test_save_load_round_trip(tmp_path):
    atmo.set_tiles(SMALL, 2, 128, 2, 512)
    atmo.set_tiles(WIDE, 4, 64, 2, 512, ms_row_order="ray")
    path = tmp_path / "tiles.txt"
    atmo.save_tiles(path)
    before = {k: atmo.get_tiles(z) for k, z in (("s", SMALL), ("w", WIDE), ("d", None))}
    atmo.reset_tiles()
    assert not atmo.get_tiles(SMALL)["sd_tuned"]
    atmo.load_tiles(path)
    after = {k: atmo.get_tiles(z) for k, z in (("s", SMALL), ("w", WIDE), ("d", None))}
    assert after == before
    lines = [l for l in path.read_text().splitlines() if not l.startswith("#")]
    assert sorted(lines) == sorted(["sd 64 4 2 128", "sd 128 16 4 64",
                                    "ms 16 64 16 4 2 512 2", "ms 32 128 64 16 2 512 2"])


@pytest.mark.parametrize("text", ["sd 64 4 3 256\n", "ms 32 128 64 16 1 256 0\n",
                                  "ms 32 128 8 4 1 256 3\n", "ms 128 8 4 1 256\n",
                                  "ms 32 128 8 4 4 256 0\n", "ms 32 128 8 4 1 0\n",
                                  "xx 1 2 3 4\n", "sd 64 4\n"])
def test_load_rejects_bad_entries_atomically(tmp_path, text):
    path = tmp_path / "bad.txt"
    path.write_text("sd 128 8 1 256\n" + text)
    with pytest.raises(AtmoError):
        atmo.load_tiles(path)
    assert not atmo.get_tiles()["sd_tuned"]  # nothing merged


def test_null_and_missing_arguments(tmp_path):
    L = atmo.lib()
    assert L.atmo_get_tiles(None, None) != 0
    assert L.atmo_set_tiles(None, None) != 0
    assert L.atmo_tune_save(None) != 0
    assert L.atmo_tune_load(None) != 0
    assert L.atmo_autotune(None, -1, 0, None) != 0
    with pytest.raises(AtmoError):
        atmo.load_tiles(tmp_path / "missing.txt")
