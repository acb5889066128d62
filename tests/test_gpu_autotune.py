"""Autotuner (puzzlesearch_amd/autotune.py): the options it picks reproduce the
reference's layer counts (PuzzlesTests/PuzzleSearchTests.cpp:550, 9-pancake),
every candidate it measured gave the same counts and hashes, and the stored
choice loads back."""
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu


def test_tune_pancake9(ps, golden, tmp_path):
    from puzzlesearch_amd import autotune as AT
    r = golden["pancake_9"]
    p = ps.pancake(9)
    cand = {"block_threads": [512, 256], "tile_bits_log2": [0, -1], "grid_blocks": [0, 64],
            "segment_bits": [0, 12], "cross_marks": [0, 1], "skip_visited": [0, 1], "graph_layers": [4, 1]}
    res = AT.tune(p, 0, passes=1, runs=1, candidates=cand)
    assert res["counts"] == r["counts"]
    path = str(tmp_path / "tuning.json")
    AT.store("dev", p, res, path)
    assert AT.load("dev", p, path)["options"] == res["options"]
    with ps.Bfs(p, **res["options"]) as b:
        counts, hashes = b.run(0)
    assert counts == r["counts"]
    assert (counts, hashes) == tuple(O.bfs3(O.pancake(9), 0))
