"""The autotuner's tuning table on the host (no device calls): nearest-entry
lookup, save/load round trip, and a damaged file that merges nothing."""
import axon_amd


def write(path, lines):
    path.write_text("\n".join(lines) + "\n")


def test_table_round_trip_and_nearest(tmp_path):
    axon_amd.tune_clear()
    assert axon_amd.tune_lookup(4096, 128, 200) is None
    f = tmp_path / "tune.txt"
    write(f, ["axon-tune 1",
              "4096 128 200 4 32 3 12.5 0.25",
              "4096 256 35 8 0 3 5.25 0.125",
              "131072 8 35 16 16 3 40.0 1.0"])
    axon_amd.tune_load(str(f))
    assert len(axon_amd.tune_entries()) == 3
    e = axon_amd.tune_lookup(4096, 128, 200)  # exact key
    assert (e["tile"], e["waves_per_cu"], e["reps"], e["ms"]) == (4, 32, 3, 12.5)
    e = axon_amd.tune_lookup(4096, 256, 40)  # nearest: log2 distance 0 + 0.19 vs 1 + 2.32
    assert (e["probes"], e["T"], e["tile"]) == (256, 35, 8)
    e = axon_amd.tune_lookup(4096, 100, 150)  # nearest: 0.36 + 0.42 vs 1.36 + 2.1
    assert e["probes"] == 128
    assert axon_amd.tune_lookup(512, 128, 200) is None  # no entry for this ncol
    g = tmp_path / "again.txt"
    axon_amd.tune_save(str(g))
    axon_amd.tune_clear()
    assert axon_amd.tune_entries() == []
    axon_amd.tune_load(str(g))
    assert sorted((e["ncol"], e["probes"], e["T"], e["tile"], e["waves_per_cu"], e["ms"], e["spread_ms"])
                  for e in axon_amd.tune_entries()) == [(4096, 128, 200, 4, 32, 12.5, 0.25),
                                                        (4096, 256, 35, 8, 0, 5.25, 0.125),
                                                        (131072, 8, 35, 16, 16, 40.0, 1.0)]
    # a loaded key replaces the same key
    write(f, ["axon-tune 1", "4096 128 200 2 16 3 9.0 0.5"])
    axon_amd.tune_load(str(f))
    assert axon_amd.tune_lookup(4096, 128, 200)["tile"] == 2 and len(axon_amd.tune_entries()) == 3
    axon_amd.tune_clear()


def test_damaged_table_merges_nothing(tmp_path):
    axon_amd.tune_clear()
    f = tmp_path / "bad.txt"
    for lines in (["axon-tune 2", "4096 128 200 4 32 3 1.0 0.0"],          # unknown version
                  ["axon-tune 1", "4096 128 200 4 32 3 1.0 0.0", "4096 64 200 3 32 3 1.0 0.0"],  # tile 3
                  ["axon-tune 1", "4096 128 200 4 24 3 1.0 0.0"],          # 24 waves per CU
                  ["axon-tune 1", "4096 128 200 4 32 3 1.0"],              # short line
                  ["not a table"]):
        write(f, lines)
        try:
            axon_amd.tune_load(str(f))
            raise AssertionError(f"accepted {lines}")
        except axon_amd.AxonError as e:
            assert e.status == 8  # AXON_ERR_IO
        assert axon_amd.tune_entries() == []
    try:
        axon_amd.tune_load(str(tmp_path / "missing.txt"))
        raise AssertionError("accepted a missing file")
    except axon_amd.AxonError as e:
        assert e.status == 8
