"""Host-side pieces of the tuner: re-gridding and instance extraction."""
import numpy as np
import pytest

from demso_amd import scene as sc
from demso_amd import tuning


def test_cell_multiplier_grids():
    s = sc.column_collapse()
    d = tuning.max_contact_diameter(s)
    assert d == pytest.approx(2 * s.radius.max())
    t = tuning.with_cell_multiplier(s, 1.0)
    assert t.instances[0].extent[0] >= s.instances[0].box[0]
    assert t.instances[0].extent[0] / t.instances[0].grid[0] == pytest.approx(d)
    t2 = tuning.with_cell_multiplier(s, 2.0)
    assert t2.total_cells < t.total_cells
    with pytest.raises(ValueError):
        tuning.with_cell_multiplier(s, 0.9)


def test_ensemble_regrid_keeps_disjoint_keys():
    s = sc.ensemble(instances=5, counts=[100, 2000, 50, 700, 3000], seed=1)
    t = tuning.with_cell_multiplier(s, 1.5)
    base = 0
    for g in t.instances:
        assert g.cell_base == base
        base += g.cells
    assert t.total_cells == base


def test_instance_scene():
    s = sc.ensemble(instances=3, counts=[10, 20, 30], seed=2)
    sub = tuning.instance_scene(s, 1)
    assert sub.n == 20
    assert np.array_equal(sub.pos, s.pos[10:30])
    assert sub.materials[0].kn == s.materials[1].kn and sub.materials[0].instance == 0
    assert sub.instances[0].cell_base == 0 and sub.instances[0].grid == s.instances[1].grid
    assert np.array_equal(sub.id, np.arange(20))


@pytest.mark.parametrize("mult", [1.0, 1.5])
def test_regrid_keeps_discs_above_an_open_top_inside_the_grid(mult):
    """The ensemble's lattices stand above their boxes' open tops; a re-grid
    must still cover them, or every disc above the box would be clamped
    into the top cell row (a crowded row that multiplies collide's work)."""
    s = sc.ensemble(instances=4, counts=[2000, 9000, 3000, 60000], seed=2)
    t = tuning.with_cell_multiplier(s, mult)
    for k, g in enumerate(t.instances):
        a, b = int(s.segments[k]), int(s.segments[k + 1])
        p = s.pos[a:b].astype(np.float64)
        assert p[:, 1].max() > s.instances[k].box[1]          # the lattice does reach above the box
        assert p[:, 0].max() < g.extent[0] and p[:, 1].max() < g.extent[1]
        cell = g.extent[0] / g.grid[0]
        cx = np.floor(p[:, 0] / cell).astype(np.int64)
        cy = np.floor(p[:, 1] / cell).astype(np.int64)
        per_cell = np.bincount(cy * g.grid[0] + cx)
        assert per_cell.max() <= 2 * mult * mult + 2
