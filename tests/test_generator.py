"""The synthetic workload generator (CPU): reproducible, and free of input
geometries that make the SSA system singular."""
import numpy as np

from oracle import oracle
from yelmo_amd import generator
from yelmo_amd.params import SsaParams


def test_one_cell_tongues_are_pruned():
    """A one-cell-wide ice tongue has two calving-front faces whose
    lateral-stress rows are the same equation; the generator removes such
    cells until none is left (a two-cell tongue then loses both cells)."""
    H = np.zeros((7, 7))
    H[1:6, 1:4] = 500.0       # a block of ice
    H[3, 4:6] = 80.0          # a tongue one cell wide in y, two long in x
    out = generator._prune_tongues(H, "zeros")
    assert np.all(out[1:6, 1:4] == 500.0)
    assert np.all(out[3, 4:6] == 0.0)
    ice = out > 0
    nb = lambda a, di, dj: np.roll(np.roll(a, di, 0), dj, 1)
    assert not np.any(ice & ~nb(ice, 1, 0) & ~nb(ice, -1, 0))
    assert not np.any(ice & ~nb(ice, 0, 1) & ~nb(ice, 0, -1))


def test_small_served_grids_solve():
    """Served requests on the coarsest grid (81^2, dx = 19 km), where the
    noisy margin is ragged: every first solve converges in the oracle."""
    par = SsaParams(precision="fp64")
    for seed in range(1000, 1006):
        p = generator.served_sequence(81, seed, 1, dtype=np.float64)[0]
        _, info = oracle.calc_velocity_ssa(p, par, dp=True)
        assert info["rc"] == 0, seed


def test_served_sequence_thickness_increments():
    """Solves of a served request differ by smooth thickness increments of
    at most 10 m per step, only where there is ice; the bed is fixed."""
    seq = generator.served_sequence(97, 5, 3, dtype=np.float64)
    H = [p.fields["H_ice"] for p in seq]
    for a, b in zip(H, H[1:]):
        d = b - a
        assert np.abs(d).max() <= 10.0 + 1e-9
        assert np.all(d[a == 0] == 0)
    assert np.array_equal(seq[0].fields["z_bed"], seq[2].fields["z_bed"])
    again = generator.served_sequence(97, 5, 3, dtype=np.float64)
    assert np.array_equal(again[2].fields["H_ice"], H[2])
