"""Restart files (src/io/restart) through the C ABI, host only.

Fixtures under tests/golden/ref/bop/: pp-0.{bop,values} and
id-00.{bop,values} are the reference's post/stress/data and post/disp/data
files (ascii and iascii values); bop.ref.noflu.txt and
wall.ref.freeze.wall.t0.txt are its src/test_data headers of a solvent dump
with no particles (test noflu) and of the frozen wall template of 46,802
particles (test freeze.wall.t0, which compares the first four header lines).
The binary value layout (n x nvars little-endian 4-byte words) is the bop
library's (github.com/amlucas/bop, not vendored, no version pinned); no
binary file of the reference pins it, so it is checked by round trips and by
reading the values file back with numpy.
"""
import os

import numpy as np
import pytest

from ctc_amd import udx as U

BOP = os.path.join(os.path.dirname(__file__), "golden", "ref", "bop")


def text(path):
    with open(path) as f:
        return f.read()


def test_read_reference_ascii_files(tmp_path):
    d = tmp_path / "stress"
    d.mkdir()
    for name in ("pp-0.bop", "pp-0.values", "id-00.bop", "id-00.values"):
        (d / name).write_bytes(open(os.path.join(BOP, name), "rb").read())
    # the reference's files are named by hand (pp-0, id-00): read them through
    # ids the naming scheme can reach by linking them to %05d names
    os.symlink(d / "pp-0.bop", d / "00000.bop")
    pp = U.restart_read_pp(tmp_path, "stress", 0, 10)
    want = np.array([[0.1, 1.1, 1.2, 0, 0, 0], [1.1, 1.1, 0.2, 0, 0, 0], [2.1, 0.1, 0.2, 0, 0, 0],
                     [0.1, 2.1, 2.2, 0, 0, 0]], np.float32)
    assert np.array_equal(pp, want)
    os.symlink(d / "id-00.bop", d / "00001.bop")
    assert np.array_equal(U.restart_read_ii(tmp_path, "stress", 1, 10), [0, 1, 2])
    with pytest.raises(U.UdxError, match="capacity"):
        U.restart_read_pp(tmp_path, "stress", 0, 3)
    with pytest.raises(U.UdxError, match="6 variables, expected 1"):
        U.restart_read_ii(tmp_path, "stress", 0, 10)
    with pytest.raises(U.UdxError, match="could not open"):
        U.restart_read_pp(tmp_path, "stress", 7, 10)


def test_headers_match_reference(tmp_path):
    (tmp_path / "bop").mkdir()
    (tmp_path / "wall.pp").mkdir()
    # an empty solvent dump: the whole header of test noflu
    U.restart_write_pp(tmp_path, "bop", 1, np.zeros((0, 6), np.float32))
    got = text(tmp_path / "bop" / "00001.bop").replace("00001.values", "solvent-00001.values")
    assert got == text(os.path.join(BOP, "bop.ref.noflu.txt"))
    # the wall template: its first four lines, as test freeze.wall.t0 compares
    n = 46802
    U.restart_write_pp(tmp_path, "wall.pp", U.RESTART_TEMPL, np.zeros((n, 6), np.float32))
    head = "".join(text(tmp_path / "wall.pp" / "templ.bop").splitlines(True)[:4])
    assert head == text(os.path.join(BOP, "wall.ref.freeze.wall.t0.txt"))
    assert os.path.getsize(tmp_path / "wall.pp" / "templ.values") == n * 24


@pytest.mark.parametrize("id", [0, 42, 12345, U.RESTART_TEMPL, U.RESTART_FINAL])
def test_round_trip(tmp_path, id):
    (tmp_path / "flu.pp").mkdir()
    (tmp_path / "flu.ii").mkdir()
    rng = np.random.default_rng(id + 10)
    pp = rng.normal(0, 3, (1000, 6)).astype(np.float32)
    pp[0, 0] = np.float32(-0.0)
    pp[1, 1] = np.nextafter(np.float32(0), np.float32(1))   # a denormal survives
    ii = rng.integers(-2**31, 2**31, 1000, dtype=np.int64).astype(np.int32)
    U.restart_write_pp(tmp_path, "flu.pp", id, pp)
    U.restart_write_ii(tmp_path, "flu.ii", id, ii)
    name = {U.RESTART_TEMPL: "templ", U.RESTART_FINAL: "final"}.get(id, "%05d" % id)
    got = U.restart_read_pp(tmp_path, "flu.pp", id, 1000)
    assert np.array_equal(got.view(np.uint32), pp.view(np.uint32))
    assert np.array_equal(U.restart_read_ii(tmp_path, "flu.ii", id, 1000), ii)
    raw = np.fromfile(tmp_path / "flu.pp" / (name + ".values"), "<f4").reshape(-1, 6)
    assert np.array_equal(raw.view(np.uint32), pp.view(np.uint32))
    assert text(tmp_path / "flu.ii" / (name + ".bop")).splitlines()[:4] == [
        "1000", "DATA_FILE: %s.values" % name, "DATA_FORMAT: int", "VARIABLES: i"]


def test_errors(tmp_path):
    with pytest.raises(U.UdxError, match="could not open"):
        U.restart_write_pp(tmp_path, "missing", 0, np.zeros((1, 6), np.float32))
    with pytest.raises(U.UdxError, match="id -3"):
        U.restart_write_pp(tmp_path, "x", -3, np.zeros((1, 6), np.float32))
    (tmp_path / "x").mkdir()
    (tmp_path / "x" / "00000.bop").write_text("5\nDATA_FILE: 00000.values\nDATA_FORMAT: double\nVARIABLES: x y z vx vy vz\n")
    with pytest.raises(U.UdxError, match="not supported"):
        U.restart_read_pp(tmp_path, "x", 0, 10)
    (tmp_path / "x" / "00001.bop").write_text("5\nDATA_FILE: 00001.values\nDATA_FORMAT: float\nVARIABLES: x y z\n")
    with pytest.raises(U.UdxError, match="3 variables"):
        U.restart_read_pp(tmp_path, "x", 1, 10)
    (tmp_path / "x" / "00002.bop").write_text("5\nDATA_FILE: 00002.values\nDATA_FORMAT: float\nVARIABLES: x y z vx vy vz\n")
    (tmp_path / "x" / "00002.values").write_bytes(b"\0" * 24 * 4)
    with pytest.raises(U.UdxError, match="fewer than 30"):
        U.restart_read_pp(tmp_path, "x", 2, 10)
