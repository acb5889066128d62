"""Multi-search engine (ps_batch_*, csrc/batch.hip), per-job-stream mode:
searches of different puzzles running side by side.  Each job's layer counts and frontier
hashes equal the oracle's (InMemoryThreeBitBFS restated,
Common/InMemoryThreeBitBFS.cpp:4-48) and the same search run alone on a
ps_bfs, whatever else runs beside it; admission respects the HBM budget; the
serving loop completes a seeded Poisson stream with every result checked."""
import math

import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

SMALL_MIX = [("sliding", 2, 4), ("sliding", 3, 3), ("sliding", 2, 5), ("pancake", 8, 0), ("pancake", 9, 0),
             ("hanoi", 10, 0), ("hanoi", 11, 0), ("pancake", 10, 0)]


def _puzzles(ps, kind, a, b):
    if kind == "sliding":
        return ps.sliding(a, b), O.sliding(a, b), math.factorial(a * b)
    if kind == "pancake":
        return ps.pancake(a), O.pancake(a), math.factorial(a)
    return ps.hanoi(a), O.hanoi(a), 4 ** a


_alone = {}


def alone(ps, spec):
    if spec not in _alone:
        p, op, n = _puzzles(ps, *spec)
        with ps.Bfs(p, index_count=n) as b:
            _alone[spec] = b.run(0)
    return _alone[spec]


@pytest.mark.parametrize("policy", [{}, {"shortest_first": 0, "reserved_cus": 32}, {"shortest_first": 1},
                                    {"job_grid_blocks": 64}])
def test_mix_equals_alone_and_oracle(ps, policy):
    # the scheduling options change only when each job's layers run
    with ps.Batch(max_jobs=16, **policy) as b:
        ids = {}
        for spec in SMALL_MIX + SMALL_MIX[:3]:  # repeats run side by side too
            p, op, n = _puzzles(ps, *spec)
            ids[b.submit(p, 0, n)] = spec
        steps = 0
        while True:
            running, finished, queued = b.step()
            steps += 1
            if steps % 2000 == 0:
                print(f"step {steps}: running {running}, queued {queued}", flush=True)
            if running == 0 and queued == 0:
                break
            assert steps < 1000000
        for jid, spec in ids.items():
            r = b.job(jid)
            assert r["state"] == ps.JOB_DONE, r
            counts, hashes = alone(ps, spec)
            assert (r["counts"], r["hashes"]) == (counts, hashes), spec
            p, op, n = _puzzles(ps, *spec)
            oc, oh = O.bfs3(op, 0, index_count=n)
            assert (r["counts"], r["hashes"]) == (oc, oh), spec


def _oracle_fixture(name):
    import json
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "oracle_fixtures.json")
    with open(path) as f:
        return json.load(f)[name]


def test_large_jobs_on_masked_streams(ps):
    # pancake 11 and Hanoi 13 have at least 2^25 states: with reserved CUs they
    # run on CU-masked streams, held back while the small jobs run
    specs = {("pancake", 11, 0): "pancake_11", ("hanoi", 13, 0): "hanoi_13"}
    small = [("sliding", 3, 3), ("pancake", 9, 0), ("hanoi", 10, 0)]
    with ps.Batch(max_jobs=8, reserved_cus=32, shortest_first=4) as b:
        ids = {}
        for spec in list(specs) + small + list(specs):
            p, op, n = _puzzles(ps, *spec)
            ids[b.submit(p, 0, n)] = spec
        steps = 0
        while True:
            running, finished, queued = b.step()
            steps += 1
            if running == 0 and queued == 0:
                break
            assert steps < 1000000
        for jid, spec in ids.items():
            r = b.job(jid)
            assert r["state"] == ps.JOB_DONE, r
            if spec in specs:
                fx = _oracle_fixture(specs[spec])
                assert r["counts"] == fx["counts"], spec
                assert [str(h) for h in r["hashes"]] == fx["hashes"], spec
            else:
                assert (r["counts"], r["hashes"]) == alone(ps, spec), spec


def test_admission_respects_budget(ps):
    specs = [("pancake", 10, 0)] * 6
    p, _, n = _puzzles(ps, *specs[0])
    # the engine runs 10..12 pancakes with cross-tile lists
    fp = ps.bfs_footprint(p, index_count=n, segment_bits=0, cross_marks=ps.CROSS_LISTS)
    with ps.Batch(max_jobs=8, hbm_budget=2 * fp + fp // 2) as b:
        ids = [b.submit(p, 0, n) for _ in specs]
        saw_queue = False
        while True:
            running, finished, queued = b.step()
            st = b.stats()
            assert st["hbm_used"] <= st["hbm_budget"]
            assert running <= 2
            saw_queue |= queued > 0
            if running == 0 and queued == 0:
                break
        assert saw_queue
        ref = alone(ps, specs[0])
        for jid in ids:
            r = b.job(jid)
            assert (r["counts"], r["hashes"]) == ref


def test_fused_mode_is_refused(ps):
    # its first device run stalled inside a step for a reason not yet found
    with pytest.raises(ps.PSError):
        ps.Batch(mode=ps.BATCH_FUSED)


def test_rejections(ps):
    for bad in ({"shortest_first": -1}, {"reserved_cus": -1}, {"reserved_cus": 1 << 20}, {"job_grid_blocks": -5}):
        with pytest.raises(ps.PSError):
            ps.Batch(**bad)
    with ps.Batch(max_jobs=4, hbm_budget=1 << 20) as b:
        with pytest.raises(ps.PSError):
            b.submit(ps.pancake(13), 0)          # not in the multi-search table
        with pytest.raises(ps.PSError):
            b.submit(ps.pancake(12), 0, math.factorial(12))   # footprint above the budget
        with pytest.raises(ps.PSError):
            b.submit(ps.sliding(3, 3), 10 ** 9)  # start outside the index space


def test_serving_poisson_small(ps):
    from puzzlesearch_amd import serve as S
    reqs = [r for r in S.draw(200, 200.0, 11) if r[1][0] != "hanoi" or r[1][1] <= 13][:24]
    res = S.serve(rate=200.0, n_jobs=len(reqs), seed=11, requests=reqs)
    assert res["all_complete_and_summed"]
    assert res["p99_ms"] >= res["p50_ms"] > 0
    for r in res["results"]:
        kind = r["kind"]
        name, args = kind.split("(")
        nums = [int(x) for x in args.rstrip(")").split(",")]
        spec = (name, nums[0], nums[1] if len(nums) > 1 else 0)
        assert (r["counts"], r["hashes"]) == alone(ps, spec), kind
