"""The kernel-trace summary's launch windows, on a small synthetic trace
database with rocprofv3's `kernels` view columns the summary reads."""

import sqlite3

from gpumap_amd import profiling


def _db(tmp_path):
    p = tmp_path / "run_results.db"
    con = sqlite3.connect(p)
    con.execute("CREATE TABLE kernels (name TEXT, start INTEGER, end INTEGER)")
    t = 0
    rows = []
    # three sessions of (screen, sgd) launches; session s's sgd takes 10 s + 5
    for sess in range(3):
        for _ in range(2):
            rows.append(("k_screen", t, t + 3))
            t += 4
            rows.append(("k_sgd_waves<2>", t, t + 10 * sess + 5))
            t += 10 * sess + 6
    con.executemany("INSERT INTO kernels VALUES (?, ?, ?)", rows)
    con.commit()
    con.close()
    return str(p)


def test_window_selects_one_session(tmp_path):
    db = _db(tmp_path)
    whole = profiling.kernel_table(db)
    assert "| `k_sgd_waves<2>` | 6 |" in whole
    # the second session: after the 2nd sgd launch, until the 4th
    mid = profiling.kernel_table(db, after="k_sgd_waves:2", until="k_sgd_waves:4")
    assert "| `k_sgd_waves<2>` | 2 | 0.00 | 0.01 | 0.01 | 0.01 |" in mid
    assert "| `k_screen` | 2 |" in mid
    last = profiling.kernel_table(db, after="k_sgd_waves:4")
    assert "| `k_sgd_waves<2>` | 2 |" in last


def test_window_past_the_launches_is_an_error(tmp_path):
    db = _db(tmp_path)
    try:
        profiling.kernel_table(db, after="k_sgd_waves:7")
    except ValueError:
        return
    raise AssertionError("expected ValueError")


def test_pmc_table_reports_mean_and_max(tmp_path):
    # a calibration copy among small copies: its reading is the maximum
    p = tmp_path / "pmc.db"
    con = sqlite3.connect(p)
    con.execute("CREATE TABLE counters_collection "
                "(kernel_name TEXT, counter_name TEXT, value REAL)")
    con.executemany("INSERT INTO counters_collection VALUES (?, ?, ?)",
                    [("copy", "FETCH_SIZE", 4.0), ("copy", "FETCH_SIZE", 2.0),
                     ("copy", "FETCH_SIZE", 262164.0), ("k", "FETCH_SIZE", 5.0)])
    con.commit()
    con.close()
    t = profiling.pmc_table(str(p))
    assert "| kernel | dispatches | FETCH_SIZE mean | FETCH_SIZE max |" in t
    assert "| `copy` | 3 | 87390 | 262164 |" in t
    assert "| `k` | 1 | 5 | 5 |" in t
