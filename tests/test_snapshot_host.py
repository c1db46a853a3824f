"""Snapshot files that are not snapshots are refused on the host, before any
device work (these run without a GPU)."""
import pytest

import axon_amd


def test_missing_file_is_an_io_error(tmp_path):
    with pytest.raises(axon_amd.AxonError) as e:
        axon_amd.Net.load(str(tmp_path / "nope.axs"))
    assert e.value.status == 8  # AXON_ERR_IO


@pytest.mark.parametrize("blob", [b"", b"AXONSNA", b"NOTASNAP" + bytes(400), b"AXONSNAP" + bytes(16)],
                         ids=["empty", "short-magic", "foreign", "short-header"])
def test_foreign_or_short_files_are_bad_snapshots(tmp_path, blob):
    p = tmp_path / "bad.axs"
    p.write_bytes(blob)
    with pytest.raises(axon_amd.AxonError) as e:
        axon_amd.Net.load(str(p))
    assert e.value.status == 9  # AXON_ERR_BAD_SNAPSHOT
