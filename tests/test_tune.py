"""Host logic of the launch autotuner (no GPU): which timed candidate it installs."""
from celeris_amd import tune


def test_pick_prefers_fewest_work_groups_within_tie():
    table = {
        "4x4/8192": {"ms": 1.000, "groups": 8192},
        "4x4/4096": {"ms": 1.004, "groups": 4096},   # within 0.5 %: fewer partial slots win
        "4x4/2048": {"ms": 1.020, "groups": 2048},   # 2 % slower: not a tie
    }
    assert tune.pick(table)["groups"] == 4096


def test_pick_takes_the_fastest_when_nothing_ties():
    table = {"a": {"ms": 0.90, "groups": 8192}, "b": {"ms": 1.00, "groups": 512}}
    assert tune.pick(table)["groups"] == 8192


def test_pick_breaks_equal_group_counts_by_time():
    # Targets above a call's piece cap launch the same work-groups; the faster timing wins.
    table = {"x": {"ms": 1.003, "groups": 462}, "y": {"ms": 1.001, "groups": 462}}
    assert tune.pick(table)["ms"] == 1.001
