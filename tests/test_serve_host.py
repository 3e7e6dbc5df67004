"""Host logic of the serving loop (CPU): the request stream and admission."""
import numpy as np


def _server(pkg, budget, max_batch=64):
    srv = pkg.serve.Server.__new__(pkg.serve.Server)
    srv.budget = budget
    srv.max_batch = max_batch
    srv.bytes_per_point = dict(pkg.serve.PRIOR_BYTES_PER_POINT)
    srv._seen = set()
    return srv


def test_request_stream_is_seeded_and_poisson(pkg):
    a = pkg.synth.request_stream(count=40, rate_hz=10.0, hi=30000)
    b = pkg.synth.request_stream(count=40, rate_hz=10.0, hi=30000)
    for (ta, xa, na, da), (tb, xb, nb, db) in zip(a, b):
        assert ta == tb and da == db
        np.testing.assert_array_equal(xa, xb)
    t = np.array([r[0] for r in a])
    assert t[0] == 0.0 and np.all(np.diff(t) > 0)
    # mean gap near 1 / rate (40 exponential gaps: within 40 %)
    assert 0.06 <= np.diff(t).mean() <= 0.14
    for _, x, n, d in a:
        assert 10000 <= len(x) <= 30000 and d == pkg.synth.depth_for_size(len(x))
        np.testing.assert_allclose(np.linalg.norm(n, axis=1), 1.0, atol=1e-5)


def test_depth_by_size(pkg):
    f = pkg.synth.depth_for_size
    assert [f(n) for n in (10000, 39999, 40000, 149999, 150000, 449999, 450000, 1000000)] == [6, 6, 7, 7, 8, 8, 9, 9]


def test_admission_by_depth_budget_and_order(pkg):
    srv = _server(pkg, budget=10 * pkg.serve.PRIOR_BYTES_PER_POINT[7] * 100000)
    w = [{"depth": 7, "points": 100000, "i": 0}, {"depth": 6, "points": 20000, "i": 1},
         {"depth": 7, "points": 100000, "i": 2}] + [{"depth": 7, "points": 100000, "i": 3 + k} for k in range(20)]
    depth, batch = srv._admit(w)
    # the oldest request's depth; requests of other depths are skipped, not blocking
    assert depth == 7 and [r["i"] for r in batch] == [0, 2] + list(range(3, 11))
    assert sum(srv.estimate(r["points"], 7) for r in batch) <= srv.budget
    # a request larger than the whole budget is served alone
    srv2 = _server(pkg, budget=1000)
    d2, b2 = srv2._admit(w)
    assert d2 == 7 and [r["i"] for r in b2] == [0]
    # at most max_batch requests
    srv3 = _server(pkg, budget=1 << 60, max_batch=4)
    assert len(srv3._admit(w)[1]) == 4


def test_learned_footprint_only_grows_after_the_first(pkg):
    srv = _server(pkg, budget=1 << 40)
    srv._learn(8, 100000, 100000 * 400)
    assert srv.bytes_per_point[8] == 500.0  # replaces the prior, with a 1.25 margin
    srv._learn(8, 100000, 100000 * 200)
    assert srv.bytes_per_point[8] == 500.0
    srv._learn(8, 100000, 100000 * 800)
    assert srv.bytes_per_point[8] == 1000.0
