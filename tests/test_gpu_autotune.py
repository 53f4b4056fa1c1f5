"""The autotuner on a small graph: two candidates per search form, the
parity gate on every candidate, and a saved configuration that loads back
and gives the same depths."""
import numpy as np
import pytest

from enterprise_amd import DeviceGraph, HostGraph, autotune

pytestmark = pytest.mark.gpu


def test_autotune_single_and_batch(oracle, tmp_path):
    g = HostGraph.rmat(14, 16, seed=14)
    roots = g.pick_roots(8, 3)
    quiet = lambda m: None  # noqa: E731
    with DeviceGraph(g) as dg:
        base = dg.config.as_dict()
        space = {"bu_probe": [8, 32]}               # two candidates (the default is 16)
        best, score, hist = autotune.tune(dg, roots, roots[:2], rounds=1, log=quiet, space=space, reps=1)
        assert len(hist) == 3 and score > 0
        assert {k: v for k, v in best.items() if k != "bu_probe"} == {k: v for k, v in base.items()
                                                                       if k != "bu_probe"}
        path = autotune.save("rmat14_smoke", best, score, {"graph": "rmat14 ef16 seed 14"}, str(tmp_path))
        assert path.startswith(str(tmp_path))
        cfg = autotune.load("rmat14_smoke", str(tmp_path))
        assert cfg == best
        dg.configure(**cfg)
        for r in roots[:3]:
            got, _ = dg.bfs(int(r))
            assert np.array_equal(got, oracle.bfs(g.beg_pos, g.adj, int(r)))
        bspace = {"batch_alpha": [2, 8]}            # two candidates (the default is 4)
        broots = g.pick_roots_mixed(64, 32, 5)
        bbest, bscore, bhist = autotune.tune_batch(dg, broots, rounds=1, log=quiet, space=bspace, reps=1)
        assert len(bhist) == 3 and bscore > 0 and bbest["batch_alpha"] in (2, 4, 8)
        autotune.save("rmat14_smoke_batch", bbest, bscore, {}, str(tmp_path))
        assert autotune.load("rmat14_smoke_batch", str(tmp_path)) == bbest
        assert autotune.load("no_such_graph", str(tmp_path)) is None
