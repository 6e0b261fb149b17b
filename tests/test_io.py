"""OVF 2.0 and table files (mmx/io.py), after the reference's test/test_ovf.jl
and the formats of src/ovf2.jl and src/fileio.jl."""
import os

import numpy as np
import pytest

from mmx import io as mio

M = np.array([0.6, 0.8, 0.0, 0.6, 0.8, 0.0])  # test/test_ovf.jl:56


@pytest.mark.parametrize("kind", ["Float64", "Float32", "String"])
def test_ovf_round_trip(tmp_path, kind):
    fname = str(tmp_path / kind)
    mio.write_ovf(mio.OVF2(2, 1, 1, 1e-9, 1e-9, 1e-9, type=kind, name="saving", data=M), fname)
    ovf = mio.read_ovf_file(fname)
    assert len(ovf.data) == 6
    assert np.abs(ovf.data - M).max() < 1e-6
    assert (ovf.xnodes, ovf.ynodes, ovf.znodes) == (2, 1, 1)
    assert ovf.xstepsize == 1e-9 and ovf.name == "saving"
    # copying keeps the data (test_copying)
    mio.write_ovf(ovf, fname + "_copy")
    assert np.abs(mio.read_ovf_file(fname + "_copy").data - M).max() < 1e-6


def test_ovf_binary8_layout(tmp_path):
    """Header lines and the OOMMF check value precede the data (src/ovf2.jl:89-193)."""
    m = np.arange(3 * 6, dtype=np.float64) / 7
    fname = mio.mag2ovf(m, 3, 2, 1, dx=2e-9, dy=2e-9, dz=3e-9, fname=str(tmp_path / "b8"))
    raw = open(fname, "rb").read()
    assert raw.startswith(b"# OOMMF OVF 2.0\n# Segment count: 1\n# Begin: Segment\n# Begin: Header\n")
    for line in (b"# xnodes: 3\n", b"# ynodes: 2\n", b"# zstepsize: 3.0e-9\n", b"# xbase: 1.0e-9\n",
                 b"# valuelabels: m_x m_y m_z\n", b"# Begin: Data Binary 8\n"):
        assert line in raw
    i = raw.index(b"# Begin: Data Binary 8\n") + len(b"# Begin: Data Binary 8\n")
    assert np.frombuffer(raw[i:i + 8], "<f8")[0] == 123456789012345.0
    np.testing.assert_array_equal(np.frombuffer(raw[i + 8:i + 8 + 8 * m.size], "<f8"), m)
    assert raw.endswith(b"\n# End: Data Binary 8\n# End: Segment\n")


def test_ovf_rejects_other_files(tmp_path):
    p = tmp_path / "x.ovf"
    p.write_text("# OOMMF: rectangular mesh v1.0\n")
    with pytest.raises(ValueError):
        mio.read_ovf_file(str(p))


def test_table_round_trip(tmp_path):
    name = str(tmp_path / "t.txt")
    sv = mio.DataSaver(name)
    state = {"i": 0}
    sv.items.append(mio.SaverItem("step", "<unitless>", lambda o: state["i"]))
    sv.items.append(mio.SaverItem("time", "<s>", lambda o: state["i"] * 1e-12))
    sv.items.append(mio.SaverItem(("m_x", "m_y", "m_z"), ("<unitless>",) * 3, lambda o: (0.1, -0.2, 0.97)))
    for i in range(4):
        state["i"] = i
        sv.write(None)
    lines = open(name).read().splitlines()
    assert lines[0] == "#" + "%19s " % "step" + "%19s " % "time" + "%19s " % "m_x" + "%19s " % "m_y" + "%19s " % "m_z"
    assert lines[2] == "%18d " % 0 + "%+18.12e " % 0.0 + "%+18.12e " % 0.1 + "%+18.12e " % -0.2 + "%+18.12e " % 0.97
    data, units = mio.read_table(name)
    np.testing.assert_array_equal(data["step"], [0, 1, 2, 3])
    np.testing.assert_allclose(data["time"], [0, 1e-12, 2e-12, 3e-12])
    assert units["time"] == "s" and units["m_x"] == "unitless"
