"""Host side of the batched Brazilian discs and of the autotuner's table (no GPU).

* The batch geometry: particle counts log-uniform in [1k, 40k] as drawn (each
  disc holds its drawn count to within 5 %), per-specimen seeds and strength
  means, discs apart along the stacking axis by more than any contact reach.
* The tuning table's text file survives a save / load / save round trip and
  ignores lines whose shapes are not ones the library has.
"""
import numpy as np

from rfd import native, native_cases


def test_batch_geometry_follows_the_drawn_counts():
    c = native_cases.brazil_batch(None, S=256, build=False)
    n = np.diff(c.first)
    assert len(n) == 256 and c.first[0] == 0
    assert np.all(n >= c.n_target) and np.all(n <= 1.05 * c.n_target)
    assert c.n_target.min() >= 1000 and c.n_target.max() <= 40000
    # log-uniform: the log counts spread over the whole range
    lg = np.log(c.n_target)
    assert lg.min() < np.log(1500) and lg.max() > np.log(25000)
    assert np.all((c.mu_t >= 10) & (c.mu_t <= 30))
    assert len(set(c.seeds.tolist())) == 256
    # stacking: the gap between neighbouring discs exceeds the largest contact
    # reach (2 r_max + skin = 1.25 mm) by a margin for the lateral bulge
    zc = np.array([L.cz for L in c.loadings])
    gaps = (zc[1:] - 0.5 * c.T[1:]) - (zc[:-1] + 0.5 * c.T[:-1])
    assert np.all(gaps > 2.5)
    assert all(L.mode == 2 and L.cut == 0 for L in c.loadings)
    # the footprint of each disc's platens covers its diameter
    for L, D in zip(c.loadings, c.D):
        assert np.sqrt(L.RR) >= 0.5 * D - 1.0


def test_tune_table_round_trip(tmp_path):
    good = tmp_path / "tune.txt"
    good.write_text("gfx950/n21/single 64 128 1\n"
                    "gfx950/n17/batch 32 256 0\n"
                    "gfx950/n9/single 48 128 1\n")  # 48 is not a tile shape
    native.tune_cache_load(good)
    out = tmp_path / "out.txt"
    native.tune_cache_save(out)
    lines = set(out.read_text().split("\n")) - {""}
    assert "gfx950/n21/single 64 128 1" in lines
    assert "gfx950/n17/batch 32 256 0" in lines
    assert not any(l.startswith("gfx950/n9/") for l in lines)
    native.tune_cache_load(out)
    out2 = tmp_path / "out2.txt"
    native.tune_cache_save(out2)
    assert set(out2.read_text().split("\n")) - {""} == lines
