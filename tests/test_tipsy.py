"""Tipsy files through the C ABI (gg_tipsy_header, gg_read_tipsy,
gg_write_tipsy).  The layout is the one tipsydefs.h:9-45 defines and
pkdSeek (pkd.c:253-283) assumes: a 32-byte header (time, nbodies, ndim, nsph,
ndark, nstar, pad), then gas records of 12 floats, dark of 9, star of 11;
XDR big-endian in the standard form, the host's byte order otherwise.  The
fixtures are written here with numpy from that definition (the reference
ships no Tipsy files).

Bars: the header's numbers exactly; every field read equals the file's
float (times the unit factor); a file read and written back is the same
bytes (metals and formation times zero, which the store does not carry)."""
import numpy as np
import pytest

import gasoline_amd as gg

GAS = [("mass", "f4"), ("pos", "f4", (3,)), ("vel", "f4", (3,)), ("rho", "f4"), ("temp", "f4"),
       ("eps", "f4"), ("metals", "f4"), ("phi", "f4")]
DARK = [("mass", "f4"), ("pos", "f4", (3,)), ("vel", "f4", (3,)), ("eps", "f4"), ("phi", "f4")]
STAR = [("mass", "f4"), ("pos", "f4", (3,)), ("vel", "f4", (3,)), ("metals", "f4"),
        ("tform", "f4"), ("eps", "f4"), ("phi", "f4")]


def dt(fields, standard):
    e = ">" if standard else "<"
    return np.dtype([(f[0], e + f[1]) + f[2:] for f in fields])


def write_file(path, standard, ng, nd, ns, time=0.125, seed=0):
    rng = np.random.default_rng(seed)
    e = ">" if standard else "<"
    hdr = np.zeros(1, np.dtype([("time", e + "f8"), ("nbodies", e + "i4"), ("ndim", e + "i4"),
                                ("nsph", e + "i4"), ("ndark", e + "i4"), ("nstar", e + "i4"),
                                ("pad", e + "i4")]))
    hdr["time"], hdr["nbodies"], hdr["ndim"] = time, ng + nd + ns, 3
    hdr["nsph"], hdr["ndark"], hdr["nstar"] = ng, nd, ns
    recs = []
    for fields, k in ((GAS, ng), (DARK, nd), (STAR, ns)):
        a = np.zeros(k, dt(fields, standard))
        a["mass"] = rng.uniform(0.5, 1.5, k)
        a["pos"] = rng.uniform(-0.5, 0.5, (k, 3))
        a["vel"] = rng.normal(size=(k, 3))
        a["eps"] = rng.uniform(0.01, 0.02, k)
        a["phi"] = rng.normal(size=k)
        if fields is GAS:
            a["rho"] = rng.uniform(1, 2, k)
            a["temp"] = rng.uniform(1e3, 1e4, k)
        recs.append(a)
    with open(path, "wb") as fp:
        fp.write(hdr.tobytes())
        for a in recs:
            fp.write(a.tobytes())
    return recs


@pytest.mark.parametrize("standard", [True, False])
def test_header(tmp_path, standard):
    p = tmp_path / "h.std"
    write_file(p, standard, 5, 7, 3, time=0.375)
    assert gg.tipsy_header(p, standard) == (0.375, 15, 5, 7, 3)
    assert p.stat().st_size == 32 + 5 * 48 + 7 * 36 + 3 * 44   # pkdSeek's record sizes


@pytest.mark.gpu
@pytest.mark.parametrize("standard", [True, False])
def test_read_fields(gpu, tmp_path, standard):
    p = tmp_path / "r.bin"
    g, d, s = write_file(p, standard, 300, 500, 200)
    n = 1000
    with gpu.Pkd(n) as pk:
        assert pk.read_tipsy(p, standard, dvFac=2.0, dTuFac=1e-3) == 0.125
        r, v, vp = pk.get("r"), pk.get("v"), pk.get("vPred")
        m, h, pot, rho, u = (pk.get(k) for k in ("fMass", "fSoft", "fPot", "fDensity", "u"))
        bits = pk.get("iActive")
    def cat(k):
        return np.concatenate([g[k], d[k], s[k]]).astype(np.float64)

    assert np.array_equal(m, cat("mass"))
    assert np.array_equal(r, cat("pos"))
    assert np.array_equal(v, 2.0 * cat("vel"))
    assert np.array_equal(h, cat("eps"))
    assert np.array_equal(pot, cat("phi"))
    assert np.array_equal(rho[:300], g["rho"].astype(np.float64)) and (rho[300:] == 0).all()
    assert np.array_equal(u[:300], 1e-3 * g["temp"].astype(np.float64)) and (u[300:] == 0).all()
    assert np.array_equal(vp[:300], v[:300]) and (vp[300:] == 0).all()
    assert (bits == np.r_[np.full(300, 256), np.full(500, 512), np.full(200, 1024)]).all()


@pytest.mark.gpu
@pytest.mark.parametrize("standard", [True, False])
def test_read_write_same_bytes(gpu, tmp_path, standard):
    a, b = tmp_path / "a.bin", tmp_path / "b.bin"
    write_file(a, standard, 1000, 3000, 500, time=2.5, seed=3)
    with gpu.Pkd(4500) as pk:
        t = pk.read_tipsy(a, standard)
        pk.write_tipsy(b, t, standard)
    assert a.read_bytes() == b.read_bytes()


@pytest.mark.gpu
def test_write_groups_types_and_checks_bits(gpu, tmp_path):
    # types interleaved in the store come out gas, dark, star
    n = 9
    with gpu.Pkd(n) as pk:
        pk.put("r", np.arange(3 * n, dtype=float).reshape(n, 3))
        pk.put("fMass", np.arange(n, dtype=float) + 1)
        pk.put("iActive", np.array([512, 256, 1024, 256, 512, 1024, 256, 512, 512], np.uint32))
        pk.write_tipsy(tmp_path / "w.std", 1.0)
        assert gg.tipsy_header(tmp_path / "w.std", True) == (1.0, 9, 3, 4, 2)
        raw = (tmp_path / "w.std").read_bytes()[32:]
        gas = np.frombuffer(raw[:3 * 48], dt(GAS, True))
        dark = np.frombuffer(raw[3 * 48:3 * 48 + 4 * 36], dt(DARK, True))
        star = np.frombuffer(raw[3 * 48 + 4 * 36:], dt(STAR, True))
        assert gas["mass"].tolist() == [2, 4, 7]
        assert dark["mass"].tolist() == [1, 5, 8, 9]
        assert star["mass"].tolist() == [3, 6]
        pk.put("iActive", np.full(n, 7, np.uint32))      # no type bit
        with pytest.raises(gg.GGError, match="type bit"):
            pk.write_tipsy(tmp_path / "x.std", 1.0)


@pytest.mark.gpu
def test_ascii_output_arrays(gpu, tmp_path):
    # the count, then "%.14g" values; a vector component by component
    rng = np.random.default_rng(9)
    n = 1000
    a = rng.normal(size=(n, 3)) * 10.0 ** rng.integers(-20, 20, (n, 3))
    rho = rng.uniform(size=n)
    with gpu.Pkd(n) as pk:
        pk.put("a", a)
        pk.put("fDensity", rho)
        pk.out_array(tmp_path / "x.acc", "a")
        pk.out_array(tmp_path / "x.den", "fDensity")
    lines = (tmp_path / "x.acc").read_text().split("\n")
    assert lines[0] == str(n) and lines[-1] == ""
    assert lines[1:-1] == ["%.14g" % x for x in a.T.reshape(-1)]
    lines = (tmp_path / "x.den").read_text().split("\n")
    assert lines[0] == str(n) and lines[1:-1] == ["%.14g" % x for x in rho]
    with pytest.raises(gg.GGError, match="not an FP64 field"):
        with gpu.Pkd(n) as pk:
            pk.out_array(tmp_path / "x.iord", "iActive")


@pytest.mark.gpu
def test_empty_file(gpu, tmp_path):
    a, b = tmp_path / "e.std", tmp_path / "f.std"
    write_file(a, True, 0, 0, 0, time=0.5)
    with gpu.Pkd(10) as pk:
        assert pk.read_tipsy(a) == 0.5 and pk.n == 0
        pk.write_tipsy(b, 0.5)
    assert a.read_bytes() == b.read_bytes()
