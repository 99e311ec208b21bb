"""Launch configuration words on the CPU: the 14 words of ``fdtd_sim_set_config`` round-trip
through ``engine.LaunchConfig``, a 5-entry pair shape reads as runs of cells (no halo tile),
and the autotuner's D shapes carry their halo tile rows."""

from meep_gpu import autotune
from meep_gpu.engine import LaunchConfig


def test_words_round_trip_with_the_halo_tile():
    cfg = LaunchConfig(h=(4, 0, 17, True, 2, 0), d=(1, 8, 3, False, 1, 8), graphs=False,
                       flux_mfma=2)
    w = cfg.words()
    assert len(w) == LaunchConfig.WORDS == 2 * LaunchConfig.SHAPE_WORDS + 2 == 14
    assert list(w[:6]) == [4, 0, 17, 1, 2, 0] and list(w[6:12]) == [1, 8, 3, 0, 1, 8]
    assert list(w[12:]) == [0, 2]
    assert LaunchConfig.from_words(w) == cfg


def test_a_five_entry_shape_is_runs_of_cells():
    cfg = LaunchConfig(h=(1, 8, 0, False, 1), d=(2, 0, 1, True, 2))
    assert list(cfg.words()[5::6][:2]) == [0, 0]
    of = LaunchConfig.of_shapes(((1, 8, 0, False), (1, 0, 17, True)), shell=(1, 2), tile=4)
    assert of.d == (1, 0, 17, True, 2, 4) and of.h == (1, 8, 0, False, 1, 0)
    assert LaunchConfig.of_shapes(((1, 8, 0, False), (1, 0, 17, True, 8))).d[5] == 8


def test_the_tile_candidates():
    assert autotune.TILE_ROWS[0] == 0 and set(autotune.TILE_ROWS[1:]) == {2, 4, 8}
    assert all(256 % rows == 0 and 256 // rows >= 32 for rows in autotune.TILE_ROWS[1:])
