"""Host side of the warm start (no device calls): the grid, the node-sample
handling of the wall spike and the mesh <-> grid round trip, through the
library's host spline."""
import numpy as np
import pytest

from tests.helpers import workloads
from tests.test_capi_symbols import _lib

WL = workloads()


@pytest.fixture(scope="module")
def S():
    _lib()  # builds the library if needed
    import scft_amd
    return scft_amd


def test_grid_shape_and_wall_layer(S):
    from scft_amd import warmstart as WS
    g = WS.grid(np.array([2.0, 7.5]))
    assert g.shape == (2, WS.GRID)
    assert np.all(np.diff(g, axis=1) > 0)
    assert g[0, 0] == 0.0 and g[0, -1] == 2.0 and g[1, -1] == 7.5
    assert np.allclose(g[0, :WS.GRID_WALL], g[1, :WS.GRID_WALL])  # absolute wall layer


@pytest.mark.parametrize("m", [40, 300, 1500, 4000])
def test_round_trip_of_a_smooth_field(S, m):
    from scft_amd import warmstart as WS
    x = WL.uniform_mesh(m, 4.0)
    w = -3.0 * np.sin(2.0 * x) * np.exp(-x)
    Y = WS.targets(S, [x], [w])
    wb = WS.to_mesh(S, [x], Y)[0]
    assert wb[0] == 0.0
    assert np.abs(wb - w)[1:].max() <= 2e-4


def test_spike_nodes_are_carried_exactly_at_the_samples(S):
    from scft_amd import warmstart as WS
    x = WL.uniform_mesh(1500, 4.0)
    w = np.where(x < 0.03, 800.0 * np.exp(-x / 0.003), -2.0 * np.sin(x))
    k = WS.spike_nodes(x)
    assert k == 14  # nodes 1..14 lie below x = 0.04 (h = 2.67e-3)
    wb = WS.to_mesh(S, [x], WS.targets(S, [x], [w]))[0]
    for i in WS.NODE_IDX[WS.NODE_IDX <= k]:
        assert abs(wb[i] - w[i]) <= 1e-9 * abs(w[i]) + 1e-12
    assert np.abs(wb - w)[k + 1:].max() <= 1e-3
    assert WS.spike_nodes(WL.uniform_mesh(40, 4.0)) == 0


def test_prepared_prediction_matches_the_per_instance_path(S):
    from scft_amd import warmstart as WS
    cfg = WL.config3(n_inst=64, seed=3)
    prep = WS.Prepared(S, cfg)
    Y = np.random.default_rng(2).standard_normal((64, WS.DIMS[-1])).astype(np.float32) * 0.3

    class Net:
        def forward(self, F):
            assert F.shape == (64, WS.DIMS[0])
            return Y

    ws = WS.WarmStart.__new__(WS.WarmStart)
    ws.S, ws.cuts, ws.nets = S, (), [Net()]
    flat = ws.predict_unknowns(prep)
    ref = np.concatenate([w[1:] for w in WS.to_mesh(S, cfg["xs"], Y)])
    assert np.array_equal(flat, ref)


def test_each_instance_takes_the_network_of_its_mesh_size_range(S):
    """With cuts (128, 512) every instance is predicted by the network of its
    element-count range ([0,128), [128,512), [512,...)), the same mapping as
    one network applied to the rows so assembled."""
    from scft_amd import warmstart as WS
    cfg = WL.config3(n_inst=64, seed=5)
    prep = WS.Prepared(S, cfg)
    cuts = (128, 512)
    rng = np.random.default_rng(6)
    Ys = [rng.standard_normal((64, WS.DIMS[-1])).astype(np.float32) * 0.3 for _ in range(3)]

    class Net:
        def __init__(self, k):
            self.k, self.rows = k, None

        def forward(self, F):
            self.rows = [int(np.flatnonzero((prep.F == f).all(axis=1))[0]) for f in F]
            return Ys[self.k][self.rows]

    ws = WS.WarmStart.__new__(WS.WarmStart)
    ws.S, ws.cuts, ws.nets = S, cuts, [Net(k) for k in range(3)]
    flat = ws.predict_unknowns(prep)
    g = np.searchsorted(cuts, prep.m, side="right")
    assert set(np.unique(g)) == {0, 1, 2}
    assert np.array_equal(ws.group([127, 128, 511, 512]), [0, 1, 1, 2])
    for k in range(3):
        assert ws.nets[k].rows == list(np.flatnonzero(g == k))
    Ymix = np.stack([Ys[g[i]][i] for i in range(64)])
    ref = np.concatenate([w[1:] for w in WS.to_mesh(S, cfg["xs"], Ymix)])
    assert np.array_equal(flat, ref)
    assert np.array_equal(WS.elements_of(prep.F), prep.m)  # m from the log-m feature
    with pytest.raises(ValueError):
        WS.WarmStart(S, cuts=(128, 128))
    with pytest.raises(ValueError):
        WS.WarmStart(S, cuts=(0,))
    with pytest.raises(ValueError):  # one epoch factor per range
        WS.WarmStart(S, cuts=(128,), range_epochs=(2,))
    with pytest.raises(ValueError):
        WS.WarmStart(S, cuts=(128,), range_epochs=(0, 1))
