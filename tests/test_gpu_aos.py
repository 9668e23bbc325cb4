"""gg_put_aos / gg_get_aos through the C ABI: fields in from and out to an
array of structs laid out like gasoline's PARTICLE (pkd.h:90-225, FLOAT =
double), here with an unaligned member as well.

Bars: every value bit for bit, across several staging chunks; members that
are not listed, and the bytes between members, are left as they were; a
gravity step fed and read through the struct array equals one through the
per-field copies bit for bit."""
import numpy as np
import pytest

from gasoline_amd import ics

pytestmark = pytest.mark.gpu

# iActive at byte 4, fMass at byte 13 (unaligned), vPred at the end
PARTICLE = np.dtype([("iOrder", "<i4"), ("iActive", "<u4"), ("iRung", "<i4"), ("flag", "u1"),
                     ("fMass", "<f8"), ("fSoft", "<f8"), ("r", "<f8", (3,)), ("v", "<f8", (3,)),
                     ("a", "<f8", (3,)), ("fPot", "<f8"), ("fBall2", "<f8"), ("fDensity", "<f8"),
                     ("dt", "<f8"), ("dtGrav", "<f8"), ("u", "<f8"), ("uPred", "<f8"),
                     ("spare", "<f8", (5,)), ("vPred", "<f8", (3,))])

IN = {"r": "r", "v": "v", "vPred": "vPred", "fMass": "fMass", "fSoft": "fSoft", "u": "u",
      "uPred": "uPred", "fBall2": "fBall2", "iActive": "iActive"}
OUT = {"a": "a", "fPot": "fPot", "fDensity": "fDensity", "dtGrav": "dtGrav"}


def random_particles(n, seed):
    raw = np.random.default_rng(seed).integers(0, 256, n * PARTICLE.itemsize, dtype=np.uint8)
    return raw.view(PARTICLE).copy()


def member_bytes(arr, mem):
    return np.ascontiguousarray(arr[mem]).view(np.uint8)


def test_roundtrip_bitexact_and_untouched_bytes(gpu):
    n = 1_000_003                     # four staging chunks of this record, the last ragged
    arr = random_particles(n, 1)
    with gpu.Pkd(n) as pk:
        pk.put_aos(arr, IN)
        for name, mem in IN.items():
            assert np.array_equal(member_bytes(arr, mem),
                                  np.ascontiguousarray(pk.get(name)).view(np.uint8)), name
        rng = np.random.default_rng(2)
        new = {"a": rng.normal(size=(n, 3)), "fPot": rng.normal(size=n),
               "fDensity": rng.uniform(size=n), "dtGrav": rng.uniform(size=n)}
        for name, v in new.items():
            pk.put(name, v)
        out = arr.copy()
        pk.get_aos(out, OUT)
        back = np.zeros_like(arr)
        pk.get_aos(back, IN)
    for name, mem in OUT.items():
        assert np.array_equal(out[mem].reshape(new[name].shape), new[name]), name
    # everything outside the listed members is as it was
    keep = np.ones(PARTICLE.itemsize, bool)
    for mem in OUT.values():
        dt, off = PARTICLE.fields[mem][:2]
        keep[off:off + dt.itemsize] = False
    rows_in, rows_out = arr.view(np.uint8).reshape(n, -1), out.view(np.uint8).reshape(n, -1)
    assert np.array_equal(rows_in[:, keep], rows_out[:, keep])
    for mem in IN.values():
        assert np.array_equal(member_bytes(back, mem), member_bytes(arr, mem)), mem


def test_gravity_step_through_the_struct_array(gpu):
    r, v, m = ics.plummer(20000, seed=9)
    n = len(m)
    arr = np.zeros(n, PARTICLE)
    arr["r"], arr["fMass"], arr["fSoft"], arr["iActive"] = r, m, 0.01, 7
    with gpu.Pkd(n) as pk:
        pk.put_aos(arr, {"r": "r", "fMass": "fMass", "fSoft": "fSoft", "iActive": "iActive"})
        pk.build_binary(16, 1, 0.7, 2)
        pk.grav_all(2)
        pk.get_aos(arr, {"a": "a", "fPot": "fPot", "dtGrav": "dtGrav"})
    with gpu.Pkd(n) as pk:
        pk.put("r", r)
        pk.put("fMass", m)
        pk.put("fSoft", np.full(n, 0.01))
        pk.build_binary(16, 1, 0.7, 2)
        pk.grav_all(2)
        a, pot, dtg = pk.get("a"), pk.get("fPot"), pk.get("dtGrav")
    assert np.array_equal(arr["a"], a)
    assert np.array_equal(arr["fPot"], pot) and np.array_equal(arr["dtGrav"], dtg)


def test_member_outside_the_record_is_rejected(gpu):
    from gasoline_amd import pkd as P
    arr = np.zeros(100, PARTICLE)
    with gpu.Pkd(100) as pk:
        tab = pk._aos_members(arr, {"r": "vPred"})
        tab[0].offset = PARTICLE.itemsize - 8          # 24 bytes from 8 before the end
        with pytest.raises(gpu.GGError, match="outside"):
            P._check(P.lib().gg_put_aos(pk.h, arr.ctypes.data_as(P._vp), PARTICLE.itemsize,
                                        tab, 1))
