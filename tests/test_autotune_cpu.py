"""The autotuner's choice among timed candidates (no GPU needed): the fastest candidate, except
that among candidates within 1% of it (2% for the block-target split, whose chunks each write half
a launch's algorithmic bytes of partial sums) the one that writes the fewest partial sums wins."""
from torch_amd import autotune


def test_pick_prefers_fewer_chunks_within_the_tie():
    table = [dict(nsplit=0, ms=90.475), dict(nsplit=8, ms=99.807), dict(nsplit=32, ms=90.325),
             dict(nsplit=64, ms=89.944), dict(nsplit=256, ms=89.741)]
    assert autotune.pick(table, autotune._chunks)["nsplit"] == 32


def test_pick_takes_the_fastest_outside_the_tie():
    table = [dict(nsplit=8, ms=4.335), dict(nsplit=64, ms=2.417), dict(nsplit=256, ms=2.357)]
    assert autotune.pick(table, autotune._chunks)["nsplit"] == 256


def test_heuristic_split_ranks_last_among_ties():
    table = [dict(nsplit=0, ms=10.0), dict(nsplit=12, ms=10.05)]
    assert autotune.pick(table, autotune._chunks)["nsplit"] == 12
    table = [dict(nsplit=0, ms=10.0), dict(nsplit=12, ms=10.2)]
    assert autotune.pick(table, autotune._chunks)["nsplit"] == 0


def test_batch_rows_tie_break_on_work_items():
    table = [dict(item=0.0, items_n=49152, ms=190.5), dict(item=2.5e7, items_n=24576, ms=191.9),
             dict(item=1e8, items_n=8192, ms=200.0)]
    assert autotune.pick(table, lambda r: r["items_n"])["item"] == 2.5e7


def test_block_target_split_takes_the_wider_tie():
    # the c2 mixed forward row of one tuning run: 8 chunks fastest, 2 chunks 2.0% slower
    table = [dict(nsplit=0, ms=2.657), dict(nsplit=1, ms=2.482), dict(nsplit=2, ms=2.405),
             dict(nsplit=3, ms=2.436), dict(nsplit=4, ms=2.384), dict(nsplit=6, ms=2.377),
             dict(nsplit=8, ms=2.358), dict(nsplit=12, ms=2.613), dict(nsplit=16, ms=2.673)]
    assert autotune.pick(table, autotune._chunks)["nsplit"] == 6
    assert autotune.pick(table, autotune._chunks, autotune.TIE_BLOCK_TARGETS)["nsplit"] == 2


def test_list_scales_kept_only_when_as_fast():
    # c5 mixed plan rows (one box): larger column / block-source items cut the partial sums
    key = lambda r: (r["partial_bytes"], r["items_n"])
    rows = [dict(item=1e7, col_scale=1.0, gs_scale=1.0, ms=175.54, items_n=49152, partial_bytes=978e6),
            dict(item=2.5e7, col_scale=1.0, gs_scale=1.0, ms=177.7, items_n=24576, partial_bytes=472e6),
            dict(item=1e7, col_scale=2.5, gs_scale=1.0, ms=175.85, items_n=24576, partial_bytes=576e6),
            dict(item=1e7, col_scale=2.5, gs_scale=1.5, ms=176.1, items_n=24576, partial_bytes=518e6)]
    best = autotune.pick(rows, key)
    assert (best["item"], best["col_scale"], best["gs_scale"]) == (1e7, 2.5, 1.5)
    rows[3]["ms"] = 178.0                 # outside the 1% band of the fastest: not taken
    best = autotune.pick(rows, key)
    assert (best["item"], best["col_scale"], best["gs_scale"]) == (1e7, 2.5, 1.0)
