"""The packaged run-time tuning table parses as xnyz_tuning_load reads it
and every row lies in the ranges xnyz_tuning_set accepts."""
import os

from xnyzip_amd import tune


def test_packaged_table_rows():
    assert os.path.exists(tune.DEFAULT_TABLE)
    rows = {}
    with open(tune.DEFAULT_TABLE) as f:
        for line in f:
            if line.lstrip().startswith("#") or not line.strip():
                continue
            c, chunk, slice_log, share, blocks, first, rbits, rtile = line.split()
            rows[int(c)] = (int(chunk), int(slice_log), float(share), int(blocks), int(first),
                            int(rbits), int(rtile))
    assert sorted(rows) == list(range(33))
    for c, (chunk, slice_log, share, blocks, first, rbits, rtile) in rows.items():
        assert 16 <= chunk <= 24 and 20 <= slice_log <= 30, c
        assert 0.0 <= share <= 0.9 and 64 <= blocks <= 16384 and first in (0, 1), c
        assert 6 <= rbits <= 8 and rtile in (2048, 4096), c
