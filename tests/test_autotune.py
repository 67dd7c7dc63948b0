"""Autotuner bookkeeping on the CPU: candidate shapes, the options each
stands for, and the cache key (timing itself needs a GPU)."""
from bitgen_amd import autotune, compiler


def test_candidates_are_distinct_and_clamped():
    one = autotune.candidates(1)
    assert len(one) == len(set(one))
    assert all(c[0] == 1 for c in one)  # a merge wider than the set is the whole set
    many = autotune.candidates(16)
    assert (4, 5, 64, 1, 256) in many and (1, 0, 64, 0, 256) in many
    assert {c[4] for c in many} == {128, 256, 512}
    assert len(many) == len(autotune.CANDIDATES)


def test_options_of_candidate_keep_the_base():
    o = autotune.options_of((8, 5, 64, 1, 512), {"lookback": "speculate", "min_waves": 2})
    assert o == {"lookback": "speculate", "groups_per_kernel": 8, "min_waves": 5, "bigram_filter": 1,
                 "block_threads": 512, "shared_filter": 0}
    assert autotune.options_of((4, 5, 64, 2, 256))["shared_filter"] == 1
    assert autotune.options_of((4, 5, 64, 2, 256))["bigram_filter"] == 1
    for key in ("groups_per_kernel", "min_waves", "bigram_filter", "block_threads", "shared_filter"):
        assert key in compiler.DEFAULT_OPTIONS


def test_cache_key_depends_on_set_sample_device_and_code():
    g = [["ab", "c.*d"], ["xyz"]]
    k = autotune._key(g, None, 1 << 20, "dev")
    assert k == autotune._key([list(x) for x in g], None, 1 << 20, "dev")
    assert k != autotune._key(g, None, 2 << 20, "dev")
    assert k != autotune._key(g, None, 1 << 20, "other")
    assert k != autotune._key(g[:1], None, 1 << 20, "dev")
    assert k != autotune._key(g, {"bigram_filter": 1}, 1 << 20, "dev")


def test_pick_breaks_near_ties_by_candidate_order():
    rows = [{"shape": c, "ms": ms} for c, ms in zip("abcd", (0.70, 0.526, 0.527, 0.5237))]
    # b, c, d are within 2% of the fastest (d): b is listed first
    assert autotune.pick(rows)["shape"] == "b"
    rows[1]["ms"] = 0.60
    assert autotune.pick(rows)["shape"] == "c"
    rows[2]["ms"] = 0.60
    assert autotune.pick(rows)["shape"] == "d"
    # the usual shape is listed before the same shape at other block sizes
    order = [c for c in autotune.CANDIDATES if c[:4] == autotune.USUAL[:4]]
    assert order[0] == autotune.USUAL
