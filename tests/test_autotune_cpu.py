"""Host logic of the solver-option autotuner (no GPU: the timer is a stub)."""
import json

import pytest

from lf_dem_amd import autotune, capi


def test_tune_picks_fastest_and_caches(tmp_path):
    rates = {"a": 10.0, "b": 30.0, "c": 20.0}
    cands = [("a", dict(precond=0)), ("b", dict(precond=1, regroup_interval=8)), ("c", dict(precond=1))]
    seen = []

    def timer(o):
        label = {(0, 4): "a", (1, 8): "b", (1, 4): "c"}[(o.precond, o.regroup_interval)]
        seen.append(label)
        return rates[label]

    path = str(tmp_path / "tune.json")
    base = capi.default_opts(tol=1e-10)
    opts, rep = autotune._tune("sig", cands, timer, base, path, True)
    assert rep["choice"] == "b" and not rep["cached"]
    assert opts.precond == 1 and opts.regroup_interval == 8
    assert opts.tol == 1e-10                      # the base options are kept
    assert seen == ["a", "b", "c"]
    assert json.load(open(path))["sig"]["choice"] == "b"
    seen.clear()
    opts2, rep2 = autotune._tune("sig", cands, timer, base, path, True)
    assert rep2["cached"] and seen == [] and opts2.regroup_interval == 8


def test_tune_skips_candidates_that_fail():
    def timer(o):
        if o.precond == 1:
            raise capi.LfdError(-5, "capacity")
        return 5.0
    opts, rep = autotune._tune("s", autotune.BATCH_CANDIDATES, timer, None, None, False)
    assert opts.precond == 0
    assert any(t["steps_per_s"] is None for t in rep["trials"])
    with pytest.raises(RuntimeError):
        autotune._tune("s", [("x", dict(precond=1))], timer, None, None, False)


def test_signature_groups_similar_workloads():
    a = autotune._signature("batch", [300, 400], [0.5, 0.52], "rate")
    b = autotune._signature("batch", [300, 400], [0.51, 0.51], "rate")
    assert a == b
    assert a != autotune._signature("system", [700], [0.51], "rate")


def test_aggregate_candidates_by_instance_size():
    """Up to 8192 particles every candidate; above, the aggregate ones without
    the waves-per-workgroup choice (one solver form for large instances);
    above the large instances' limit none of them, since they would time
    block-Jacobi under another name."""
    from lf_dem_amd import autotune
    small = autotune.system_candidates(4000)
    mid = autotune.system_candidates(32000)
    large = autotune.system_candidates(autotune.AGGREGATE_MAX_N + 1)
    assert small == autotune.SYSTEM_CANDIDATES
    assert autotune.system_candidates(autotune.AGGREGATE_WHOLE_MAX_N) == small
    assert any(c[1].get("precond") == 1 for c in mid)
    assert all("agg_waves" not in c[1] for c in mid)
    assert {c[0] for c in mid} == {c[0] for c in small if "agg_waves" not in c[1]}
    assert all(c[1].get("precond") == 0 for c in large)
    assert {c[0] for c in large} == {c[0] for c in small if c[1].get("precond") == 0}


def test_near_ties_keep_the_candidate_listed_first():
    """Trials that differ by less than the margin (run-to-run noise) keep the
    earlier candidate; a clear win takes over."""
    cands = [("a", dict(precond=0)), ("b", dict(precond=1))]
    near = {0: 100.0, 1: 100.0 * (1 + autotune.MARGIN * 0.5)}
    opts, rep = autotune._tune("s", cands, lambda o: near[o.precond], None, None, False)
    assert rep["choice"] == "a"
    far = {0: 100.0, 1: 100.0 * (1 + autotune.MARGIN * 2)}
    opts, rep = autotune._tune("s", cands, lambda o: far[o.precond], None, None, False)
    assert rep["choice"] == "b"
