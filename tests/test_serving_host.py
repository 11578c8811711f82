"""Host-side pieces of the serving loop (no GPU)."""
def test_poisson_stream_is_seeded():
    import __graft_entry__

    S = __graft_entry__.load_package().serving
    a = S.poisson_requests(rate=20, duration=5, seed=0)
    b = S.poisson_requests(rate=20, duration=5, seed=0)
    assert [(r.arrival, r.problem, r.n, r.steps) for r in a] == \
        [(r.arrival, r.problem, r.n, r.steps) for r in b]
    assert 60 < len(a) < 140
    assert {r.problem for r in a} == {"vortex", "mhd"}
    assert all(100 <= r.steps <= 5000 and r.n in (256, 512, 1024, 2048) for r in a)
