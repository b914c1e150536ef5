"""bench.py's bookkeeping, checked without a GPU.

* the HBM-traffic lookup takes a kernel by its whole name: profiler names
  hold single spaces ("k_visc_predict<true, true, true>"), and two template
  instances of one kernel must not be confused;
* every kernel kind the benchmark prices has its algorithmic bytes, its
  launches per step and, where the profiler names it differently, a symbol;
* the step's algorithmic bytes are 192 + 56 c per particle-step (one read of
  each phase's inputs and one write of its outputs, c = cells per particle);
* the autotuner keeps its recorded variant unless another is faster by more
  than the timing spread.
"""
import bench


def test_traffic_lookup_matches_whole_kernel_names(tmp_path, monkeypatch):
    f = tmp_path / "traffic.txt"
    f.write_text(
        "calibration: 256 MB device copy; scale FETCH_SIZE x1.9998, WRITE_SIZE x1.0000\n"
        "kernel                                     dispatches    FETCH_KiB    WRITE_KiB  "
        "HBM_MB_per_launch\n"
        "k_relax<true>                                     230      32018.9      26032.2"
        "             92.224\n"
        "k_visc_predict<true, false, true>                   1       8259.0      19691.9"
        "             37.077\n"
        "k_visc_predict<true, true, true>                  229      19730.9      18081.1"
        "             58.919\n")
    monkeypatch.setattr(bench, "TRAFFIC_FILE", str(f))
    assert bench.traffic_of("k_visc_predict<true, true, true>") == 58.919e6
    assert bench.traffic_of("k_visc_predict<true, false, true>") == 37.077e6
    assert bench.traffic_of("k_relax<true>") == 92.224e6
    assert bench.traffic_of("k_visc_predict<true,") is None   # a prefix is no name
    assert bench.traffic_of("k_density") is None


def test_committed_traffic_summary_covers_the_gather_kernels():
    for kind in ("visc_predict", "relax"):
        assert bench.traffic_of(bench.KERNEL_SYMBOLS[kind]) is not None, kind


def test_kernel_tables_agree():
    assert set(bench.KERNEL_LAUNCHES) == set(bench.KERNEL_BYTES)
    assert set(bench.KERNEL_SYMBOLS) <= set(bench.KERNEL_BYTES)
    # two sorts per step, one launch each of the gathers
    assert bench.KERNEL_LAUNCHES["scatter"] == bench.KERNEL_LAUNCHES["reorder"] == 2
    assert bench.KERNEL_LAUNCHES["density"] == bench.KERNEL_LAUNCHES["relax"] == 1


def test_step_bytes_per_particle_step():
    assert bench.B_ALG == (192, 56)
    c = 960000 / 1048576
    assert abs(bench.B_ALG[0] + bench.B_ALG[1] * c - 243.27) < 0.01


def test_autotuner_keeps_incumbent_within_noise():
    from sph_amd import autotune as at
    res = {"a.so": {"ms_per_step": 1.000}, "b.so": {"ms_per_step": 0.998},
           "c.so": {"ms_per_step": 1.100}}
    assert at.pick_best(res) == "b.so"
    assert at.pick_best(res, incumbent="a.so") == "a.so"        # 0.2% slower: kept
    assert at.pick_best(res, incumbent="c.so") == "b.so"        # 10% slower: replaced
    assert at.pick_best(res, incumbent="gone.so") == "b.so"     # not measured: fastest
    assert at.pick_best(res, incumbent="a.so", keep_within=0.001) == "b.so"


def test_counter_summary_means_and_derived_figures(tmp_path):
    import sqlite3
    db = tmp_path / "run_results.db"
    con = sqlite3.connect(db)
    con.execute("create table counters_collection "
                "(dispatch_id int, kernel_name text, counter_name text, value real)")
    rows = []
    for d, k in ((1, "k_a(int)"), (2, "k_a(int)"), (3, "void k_b<2>(float*)")):
        rows += [(d, k, "GRBM_GUI_ACTIVE", 1000.0), (d, k, "SQ_INSTS_VALU", 128000.0 * d),
                 (d, k, "SQ_WAVE_CYCLES", 4096000.0)]
    con.executemany("insert into counters_collection values (?,?,?,?)", rows)
    con.commit()
    out = bench.counter_summary(str(db), simds=1024, xcds=8).splitlines()
    a = next(l for l in out if l.startswith("k_a "))
    assert "dispatches=2" in a and "SQ_INSTS_VALU=192000" in a
    # busy cycles per XCD 1000/8; VALU 192000*4/(125*1024); waves 4096000*4/(125*1024)
    assert "VALU issue 6.000" in a and "waves/SIMD 128.00" in a
    b = next(l for l in out if l.startswith("k_b<2> "))
    assert "dispatches=1" in b
    last = bench.counter_summary(str(db), last=1).splitlines()
    assert len(last) == 1 and last[0].startswith("k_b<2> ")


def test_profile_summary_window(tmp_path):
    """--last N after --skip M: the library's dispatches before the final M,
    runtime blit kernels left out."""
    import sqlite3
    db = tmp_path / "run_results.db"
    con = sqlite3.connect(db)
    con.execute("create table kernels (name text, dispatch_id int, start int, end int, "
                "duration int)")
    rows, t = [], 0
    for d, (k, dur) in enumerate([("k_a(int)", 1000), ("__amd_rocclr_copyBuffer", 50),
                                  ("k_a(int)", 2000), ("k_b(int)", 3000), ("k_a(int)", 4000),
                                  ("k_b(int)", 5000)]):
        rows.append((k, d, t, t + dur, dur))
        t += dur + 10
    con.executemany("insert into kernels values (?,?,?,?,?)", rows)
    con.commit()
    out = bench.profile_summary(str(db), last=2, skip=2).splitlines()
    a = next(l for l in out if l.startswith("k_a "))
    b = next(l for l in out if l.startswith("k_b "))
    assert a.split()[1:3] == ["1", "2.00"] and b.split()[1:3] == ["1", "3.00"]
    assert out[-1].endswith("(2 dispatches)")
    tail = bench.profile_summary(str(db), last=2).splitlines()
    assert next(l for l in tail if l.startswith("k_a ")).split()[1:3] == ["1", "4.00"]
