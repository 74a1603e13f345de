"""CPU check of the plain-loop records the design workload saves and merges across calls
(3d_fdtd_amd/design.py: plain_part, merge_plain_parts)."""
import importlib

import pytest

design = importlib.import_module("3d_fdtd_amd.design")


def test_merged_record_sums_counts_and_takes_the_rate_over_the_total_time():
    a = design.plain_part(2, 8, 78, 589.6, 0.7556, [[1.0, 2.0], [3.0, 4.0]])
    b = design.plain_part(6, 11, 96, 592.1, 0.7556, [[5.0, 6.0]] * 6)
    assert a["designs_per_hour"] == round(2 / 589.6 * 3600, 1) and a["fdtd_runs"] == 156
    m = design.merge_plain_parts(a, b)
    assert (m["accepted"], m["starts"], m["iterations"], m["fdtd_runs"]) == (8, 19, 174, 348)
    assert m["seconds"] == 1181.7
    # the pooled rate is accepted / total time, not the mean of the two calls' rates (12.2 and 36.5)
    assert m["designs_per_hour"] == round(8 / 1181.7 * 3600, 1) == 24.4
    assert m["accepted_designs"] == a["accepted_designs"] + b["accepted_designs"]
    assert len(m["accepted_designs"]) == m["accepted"]


def test_records_of_different_specs_are_not_merged():
    a = design.plain_part(1, 3, 20, 100.0, 0.70, [[1.0]])
    b = design.plain_part(1, 3, 20, 100.0, 0.75, [[2.0]])
    with pytest.raises(ValueError, match="different specs"):
        design.merge_plain_parts(a, b)
