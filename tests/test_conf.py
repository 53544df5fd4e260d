"""Configuration layer (host code of libudx.so, no GPU): the reference's own
conf/default.cfg and conf/rbc.test.cfg (tests/golden/ref/conf), the layer
order of src/conf/imp/get.h, libconfig's typing and syntax."""
import os

import numpy as np
import pytest

from ctc_amd import udx as U

CONF = os.path.join(os.path.dirname(__file__), "golden", "ref", "conf")
DEFAULT = os.path.join(CONF, "default.cfg")

# the rbc.u.simple test's conf.cfg (src/test/rbc/u/simple/conf.cfg)
SIMPLE = """time = {
    type = "const"
    dt = 5e-4
}

glb = {
    kBT = 0.0
    L = [16, 16, 16]
}

rbc = {
    cell = "rbc.off"
    ic = "rbcs-ic.txt"
    @include "rbc.test.cfg"
}"""


def test_default_cfg():
    c = U.Conf(default=DEFAULT)
    assert c.float("glb.kBT") == np.float32(0.0444302)
    assert c.int("glb.numdensity") == 10
    assert c.int3("glb.L") == (8, 8, 8)
    assert c.float("time.dt") == np.float32(1e-4)
    assert c.string("flu.self") == "pair_flu"
    assert c.bool("pair_flu.dpd") and not c.bool("wall.active")
    assert c.vfloat("fsi.a") == [np.float32(2.6666667)] * 3
    assert c.string("rbc.bending") == "kantor"
    p = c.pair("pair_flu", c.float("glb.kBT"), c.float("time.dt"))
    assert (p.a, p.g, p.spow) == (np.float32(2.6666667), 8.0, np.float32(c.float("pair_flu.spow")))
    assert p.s == np.float32(np.sqrt(np.float32(2 * np.float32(0.0444302) * 8.0) / np.float32(1e-4)))
    c.close()


def test_layers_and_include(tmp_path):
    """default < file < argument string < program setters; @include in a
    group reads from the include directory."""
    f = tmp_path / "conf.cfg"
    f.write_text(SIMPLE)
    c = U.Conf(default=DEFAULT, opt=str(f), args="rbc = { totArea = 80.0594 totVolume = 43.3586 }",
               include_dir=CONF)
    assert c.float("time.dt") == np.float32(5e-4)          # file over default
    assert c.int3("glb.L") == (16, 16, 16)
    assert c.float("glb.kBT") == 0.0
    assert c.string("flu.self") == "pair_flu"               # default where the file is silent
    r = c.rbc_params("rbc")
    want = dict(gammaC=30.0, gammaT=90.0, kBT=0.0444, kb=17.78, phi=6.97, ks=14.8, x0=0.5, mpow=2.0, ka=2450.0,
                kd=100.0, kv=3750.0, totArea=80.0594, totVolume=43.3586)
    for k, v in want.items():
        assert getattr(r, k) == np.float32(v), k
    assert c.string("rbc.cell") == "rbc.off" and c.int("rbc.seed") == 123456
    assert c.string("rbc.bending") == "kantor"              # from the default's rbc group
    c.set("time.dt", 1e-3)
    assert c.float("time.dt") == np.float32(1e-3)           # setter over everything
    c.close()


def test_types_and_errors():
    c = U.Conf(args='a = 2 b = 2.0 c = "x" d = [1, 2] e = true g = { h = 1e-3 }')
    assert c.int("a") == 2 and c.float("b") == 2.0 and c.float("g.h") == np.float32(1e-3)
    with pytest.raises(U.UdxError, match="wrong type"):
        c.float("a")
    with pytest.raises(U.UdxError, match="wrong type"):
        c.vfloat("d")
    with pytest.raises(U.UdxError, match="could not find"):
        c.int("zz")
    with pytest.raises(U.UdxError, match="could not find"):
        c.int("a.b")
    with pytest.raises(U.UdxError, match="int3"):
        c.int3("d")
    assert c.has("g.h") and not c.has("g.x")
    c.close()


def test_syntax():
    text = r"""
    // comment
    /* block
       comment */
    x : 0x1F;  # hex, colon assignment, semicolon
    s = "a\"b" "cd";          # escapes, adjacent strings join
    f = [1.5, -2e3, +3.];
    l = (1, "two", { q = 3 }, [4]);
    n = { m = { k = -7L } },
    b = FALSE
    """
    c = U.Conf(args=text)
    assert c.int("x") == 31
    assert c.string("s") == 'a"bcd'
    assert c.vfloat("f") == [1.5, -2000.0, 3.0]
    assert c.int("n.m.k") == -7
    assert c.bool("b") is False
    c.close()
    for bad in ("a = 1 a = 2", "a = [1, 2.0]", "a = {", "a = [1, {b = 1}]", "= 3"):
        with pytest.raises(U.UdxError):
            U.Conf(args=bad)


def test_wall_vel_conf():
    """wvel.type and its keys (src/wall/wvel/conf.cpp:10-47)"""
    v = U.Conf(args='wvel = { type = "constant" u = [0.5, 0.0, -1.0] }').wall_vel()
    assert v.type == U.WVEL_TYPES["constant"] and list(v.u) == [0.5, 0.0, -1.0]
    v = U.Conf(args='wvel = { type = "shear" gdot = 0.25 vdir = 0 gdir = 2 half = 1 }').wall_vel()
    assert (v.type, v.gdot, v.vdir, v.gdir, v.half) == (U.WVEL_TYPES["shear"], 0.25, 0, 2, 1)
    v = U.Conf(args='wvel = { type = "shear sin" gdot = 2.0 vdir = 1 gdir = 0 half = 0 w = 3.5 }').wall_vel()
    assert (v.type, v.gdot, v.vdir, v.gdir, v.half, v.w) == (U.WVEL_TYPES["shear sin"], 2.0, 1, 0, 0, 3.5)
    v = U.Conf(args='wvel = { type = "hele shaw" u = 0.125 h = 6.0 }').wall_vel()
    assert (v.type, v.hs_u, v.hs_h) == (U.WVEL_TYPES["hele shaw"], 0.125, 6.0)
    with pytest.raises(U.UdxError, match="unknown type <poiseuille>"):
        U.Conf(args='wvel = { type = "poiseuille" }').wall_vel()
    with pytest.raises(U.UdxError, match="could not find"):
        U.Conf(args='wvel = { type = "shear sin" gdot = 2.0 vdir = 1 gdir = 0 half = 0 }').wall_vel()
    with pytest.raises(U.UdxError, match="wrong type"):
        U.Conf(args='wvel = { type = "shear" gdot = 1 vdir = 0 gdir = 2 half = 0 }').wall_vel()


def test_wall_repulse_conf():
    """<obj>.repulsion names "none" or a group with lambda
    (src/sim/objects/imp/ini.h:69-75, src/wall/imp/repulse.h:8-12)"""
    assert U.Conf(args='rbc = { repulsion = "none" }').wall_repulse("rbc") == 0.0
    c = U.Conf(args='rbc = { repulsion = "wrep" } wrep = { lambda = 2.5 }')
    assert c.wall_repulse("rbc") == 2.5
    with pytest.raises(U.UdxError, match="could not find"):
        U.Conf(args='rbc = { repulsion = "wrep" }').wall_repulse("rbc")
    with pytest.raises(U.UdxError, match="could not find"):
        U.Conf(args='rbc = { }').wall_repulse("rbc")
