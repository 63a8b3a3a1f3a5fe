"""Host side of factored-tree persistence: files that are not saved trees, or
of another format version, are refused before the handle changes (no device
call is made), and the saved header reads back in Python."""
import struct

import pytest

import robinsolver_amd as rsa

UNIT = [[0, 1], [0, 1], [0, 1]]


def _load(t, path):
    return rsa.lib().rs_tree_load(t.h, str(path).encode())


def test_missing_and_foreign_files_are_refused(tmp_path):
    t = rsa.GpuTree(2, 2, UNIT, 1.0 / 32, 5.0)
    assert _load(t, tmp_path / "none.rst") == 1
    assert "cannot open" in rsa.lib().rs_last_error(t.h).decode()
    junk = tmp_path / "junk.rst"
    junk.write_bytes(b"not a tree" * 40)
    assert _load(t, junk) == 1
    assert "is not a saved tree" in rsa.lib().rs_last_error(t.h).decode()
    with pytest.raises(rsa.RobinError):
        rsa.read_saved_header(str(junk))
    # the handle keeps its plan
    assert (t.n_leaves, t.nb) == (4, 64)


def test_other_format_version_is_refused(tmp_path):
    hdr = struct.pack("<8sI11i16d", b"RSTREE\x00\x01", 99, 2, 2, 0, 1, 0, 0, 4, 320, 256, 64, 0,
                      0, 1, 0, 1, 0, 1, 1.0 / 32, 5.0, 0, 5, 0, -5, 1, 0, 1, 0)
    p = tmp_path / "v99.rst"
    p.write_bytes(hdr)
    h = rsa.read_saved_header(str(p))
    assert h["version"] == 99 and h["d"] == 2 and h["kb"] == 5.0
    assert h["bndry_params"] == [5j, -5j, 1, 1]
    t = rsa.GpuTree(2, 2, UNIT, 1.0 / 32, 5.0)
    assert _load(t, p) == 1
    assert "format version 99" in rsa.lib().rs_last_error(t.h).decode()
