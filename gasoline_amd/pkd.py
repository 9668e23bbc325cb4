"""ctypes binding of libgasoline_amd.so and a thin Pkd class over a GGPKD
handle.

Pkd mirrors the per-processor calls of gasoline that the C ABI replaces:
build_binary (pkdBuildBinary, pkd.c:2627), grav_all (pkdGravAll,
pkd.c:2869), build_local (pkdBuildLocal, pkd.c:2732), smooth (smInitialize +
smSmooth + smFinish, smooth.c:54/1439/505), resmooth (smReSmooth,
smooth.c:1716) and gas_pressure (pkdGasPressure, pkd.c:5750).  Every call
goes to the HIP library; a failure raises GGError with the library's message.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "libgasoline_amd.so")

GG_R, GG_V, GG_VPRED, GG_A, GG_MASS, GG_SOFT, GG_POT, GG_DTGRAV, GG_U, GG_UPRED, \
    GG_FBALL2, GG_DENSITY, GG_DIVV, GG_CURLV, GG_DVDS, GG_FDIVV_T, GG_BALSARA, GG_DIFF, \
    GG_POVERRHO2, GG_C, GG_DT, GG_UDOTPDV, GG_UDOTAV, GG_UDOTDIFF, GG_MUMAX, GG_DTNEW, \
    GG_IACTIVE, GG_IRUNG, GG_DTOLD, GG_FWEIGHT = range(30)

FIELDS = {
    "r": (GG_R, 3), "v": (GG_V, 3), "vPred": (GG_VPRED, 3), "a": (GG_A, 3),
    "fMass": (GG_MASS, 1), "fSoft": (GG_SOFT, 1), "fPot": (GG_POT, 1),
    "dtGrav": (GG_DTGRAV, 1), "u": (GG_U, 1), "uPred": (GG_UPRED, 1),
    "fBall2": (GG_FBALL2, 1), "fDensity": (GG_DENSITY, 1), "divv": (GG_DIVV, 1),
    "curlv": (GG_CURLV, 3), "dvds": (GG_DVDS, 1), "fDivv_t": (GG_FDIVV_T, 1),
    "BalsaraSwitch": (GG_BALSARA, 1), "diff": (GG_DIFF, 1),
    "PoverRho2": (GG_POVERRHO2, 1), "c": (GG_C, 1), "dt": (GG_DT, 1),
    "uDotPdV": (GG_UDOTPDV, 1), "uDotAV": (GG_UDOTAV, 1), "uDotDiff": (GG_UDOTDIFF, 1),
    "mumax": (GG_MUMAX, 1), "dtNew": (GG_DTNEW, 1), "iActive": (GG_IACTIVE, 1),
    "iRung": (GG_IRUNG, 1), "dtOld": (GG_DTOLD, 1),
    "fWeight": (GG_FWEIGHT, 1),
}
_INT_FIELDS = {"iActive": np.uint32, "iRung": np.int32}

PREC_FP64, PREC_FP32 = 0, 1
SMX_DENSITY, SMX_SPHPRESSURETERMS, SMX_DENDVDX = 1, 21, 22
TYPE_ACTIVE, TYPE_TREEACTIVE, TYPE_SMOOTHACTIVE, TYPE_SMOOTHDONE = 1, 2, 4, 8
TYPE_NbrOfACTIVE, TYPE_GAS, TYPE_DARK, TYPE_STAR = 32, 256, 512, 1024


def tipsy_header(path, standard=True):
    """(time, nbodies, nsph, ndark, nstar) of a Tipsy file (gg_tipsy_header)."""
    t = C.c_double()
    nb, ng, nd, ns = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    _check(lib().gg_tipsy_header(os.fsencode(path), int(bool(standard)), C.byref(t), C.byref(nb),
                                 C.byref(ng), C.byref(nd), C.byref(ns)))
    return t.value, nb.value, ng.value, nd.value, ns.value

# exported entry points (include/gasoline_amd.h)
ABI = ["gg_initialize", "gg_finish", "gg_last_error", "gg_set_local", "gg_local", "gg_set_period",
       "gg_tipsy_header", "gg_read_tipsy", "gg_write_tipsy", "gg_out_array",
       "gg_put", "gg_get", "gg_put_aos", "gg_get_aos", "gg_set_precision", "gg_build_binary", "gg_grav_all",
       "gg_grav_lists", "gg_grav_walk_checksums", "gg_tree_info", "gg_direct_grav",
       "gg_direct_grav_noalloc", "gg_direct_ewald", "gg_build_local", "gg_smooth", "gg_tune_sph", "gg_set_sph_lanes", "gg_tune_sph_lanes", "gg_resmooth",
       "gg_gas_pressure", "gg_smooth_lists", "gg_smooth_stats", "gg_timings",
       "gg_build_forest", "gg_grav_graph", "gg_grav_replay", "gg_tune_bucket",
       "gg_set_grav_schedule",
       "gg_set_seed_slack", "gg_reorder", "gg_init_accel", "gg_kick", "gg_drift",
       "gg_kick_vpred", "gg_init_dt", "gg_grav_step", "gg_accel_step", "gg_density_step",
       "gg_sph_step", "gg_dt_to_rung", "gg_active_mask_rung", "gg_active_type",
       "gg_active_type_rung", "gg_active_exact_type", "gg_set_type", "gg_reset_type",
       "gg_sph_curr_rung", "gg_emergency_adjust",
       "gg_hmlp_init", "gg_hmlp_record", "gg_hmlp_train", "gg_hmlp_predict",
       "gg_hmlp_params", "gg_hmlp_eval", "gg_hmlp_grad"]

HMLP_NW, HMLP_NP = 1249, 1257


class GGError(RuntimeError):
    pass


class AosField(C.Structure):
    """gg_aos_field: one member of the caller's particle struct."""
    _fields_ = [("field", C.c_int), ("pad", C.c_int), ("offset", C.c_size_t)]


class Smf(C.Structure):
    """gg_smf: SMF parameters of the density and pressure-force operators
    (smoothfcn.h:22-130)."""
    _fields_ = [("H", C.c_double), ("a", C.c_double), ("alpha", C.c_double),
                ("beta", C.c_double), ("gamma", C.c_double),
                ("dtFacCourant", C.c_double), ("dtFacDiffusion", C.c_double),
                ("dMetalDiffusionCoeff", C.c_double),
                ("dThermalDiffusionCoeff", C.c_double), ("dResolveJeans", C.c_double),
                ("bCannonical", C.c_int), ("iViscosityLimiter", C.c_int),
                ("bConstantDiffusion", C.c_int), ("pad", C.c_int)]

    @classmethod
    def default(cls, **kw):
        """H=0, a=1, alpha=1, beta=2, gamma=5/3, dEtaCourant=0.4
        (dtFacCourant = 0.4*2/1.6, pkd.c:5893), no viscosity limiter,
        no diffusion."""
        s = cls(H=0.0, a=1.0, alpha=1.0, beta=2.0, gamma=5.0 / 3.0,
                dtFacCourant=0.4 * 2 / 1.6, dtFacDiffusion=2 * 0.1,
                dMetalDiffusionCoeff=0.0, dThermalDiffusionCoeff=0.0, dResolveJeans=0.0,
                bCannonical=1, iViscosityLimiter=0, bConstantDiffusion=0, pad=0)
        for k, v in kw.items():
            setattr(s, k, v)
        return s


_LIB = None
_vp = C.c_void_p
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_i64p = C.POINTER(C.c_int64)


def lib():
    """Load libgasoline_amd.so; raises if it is not built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIBPATH):
        raise GGError("libgasoline_amd.so is not built: run __graft_entry__.build() "
                      "(make -C gasoline_amd/csrc)")
    L = C.CDLL(LIBPATH)
    L.gg_last_error.restype = C.c_char_p
    L.gg_initialize.argtypes = [C.POINTER(_vp), C.c_int, C.c_int]
    L.gg_finish.argtypes = [_vp]
    L.gg_set_local.argtypes = [_vp, C.c_int]
    L.gg_local.argtypes = [_vp]
    L.gg_set_period.argtypes = [_vp, _dp]
    L.gg_tipsy_header.argtypes = [C.c_char_p, C.c_int, _dp, _ip, _ip, _ip, _ip]
    L.gg_read_tipsy.argtypes = [_vp, C.c_char_p, C.c_int, C.c_double, C.c_double, _dp]
    L.gg_write_tipsy.argtypes = [_vp, C.c_char_p, C.c_int, C.c_double, C.c_double, C.c_double]
    L.gg_out_array.argtypes = [_vp, C.c_char_p, C.c_int]
    L.gg_put.argtypes = [_vp, C.c_int, _vp, C.c_int]
    L.gg_get.argtypes = [_vp, C.c_int, _vp, C.c_int]
    L.gg_put_aos.argtypes = [_vp, _vp, C.c_size_t, C.POINTER(AosField), C.c_int]
    L.gg_get_aos.argtypes = [_vp, _vp, C.c_size_t, C.POINTER(AosField), C.c_int]
    L.gg_set_precision.argtypes = [_vp, C.c_int]
    L.gg_build_binary.argtypes = [_vp, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int]
    L.gg_grav_all.argtypes = [_vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                              C.c_double, C.c_int, C.c_double, C.c_int, C.c_double, _dp,
                              _ip, _dp, _dp, _dp, _dp]
    L.gg_grav_lists.argtypes = [_vp, _ip, _i64p, _ip, _i64p, _ip, _ip]
    L.gg_grav_walk_checksums.argtypes = [_vp, C.c_int, _vp]
    L.gg_tree_info.argtypes = [_vp, _ip, _ip, _ip, _ip, _dp]
    L.gg_direct_grav.argtypes = [_vp, C.c_int, _ip, _dp, _dp]
    L.gg_direct_grav_noalloc.argtypes = [_vp, C.c_int, _vp, _vp, _vp]
    L.gg_direct_ewald.argtypes = [_vp, C.c_int, _ip, _dp, _dp]
    L.gg_set_grav_schedule.argtypes = [_vp, C.c_int]
    L.gg_reorder.argtypes = [_vp, _ip]
    L.gg_build_local.argtypes = [_vp, C.c_int, C.c_int, _dp]
    L.gg_tune_sph.argtypes = [_vp, C.POINTER(Smf), C.c_int, C.c_int, _dp, _ip, C.c_int, _dp, _ip]
    L.gg_set_sph_lanes.argtypes = [_vp, C.c_int]
    L.gg_tune_sph_lanes.argtypes = [_vp, C.POINTER(Smf), C.c_int, C.c_int, _ip, C.c_int, _dp, _ip]
    L.gg_smooth.argtypes = [_vp, C.POINTER(Smf), C.c_int, C.c_int, C.c_int, C.c_int, _ip]
    L.gg_resmooth.argtypes = [_vp, C.POINTER(Smf), C.c_int, C.c_int, C.c_int, C.c_int]
    L.gg_gas_pressure.argtypes = [_vp, C.POINTER(Smf)]
    L.gg_smooth_lists.argtypes = [_vp, C.c_int, _ip]
    L.gg_timings.argtypes = [_vp, _dp, C.c_int]
    L.gg_smooth_stats.argtypes = [_vp, _i64p, C.c_int]
    L.gg_build_forest.argtypes = [_vp, C.c_int, _ip, C.c_int, C.c_int, C.c_double, C.c_int]
    L.gg_grav_graph.argtypes = [_vp, C.c_int]
    L.gg_grav_replay.argtypes = [_vp, _ip, _dp, _dp, _dp, _dp]
    L.gg_tune_bucket.argtypes = [_vp, C.c_int, _ip, C.c_int, C.c_double, C.c_int, _ip, C.c_int,
                                 _dp, _ip]
    L.gg_set_seed_slack.argtypes = [_vp, C.c_double]
    L.gg_init_accel.argtypes = [_vp]
    L.gg_kick.argtypes = [_vp] + [C.c_double] * 6 + [C.c_int]
    L.gg_drift.argtypes = [_vp, C.c_double, _dp, C.c_int, _dp]
    L.gg_kick_vpred.argtypes = [_vp, C.c_double, C.c_double, C.c_double, C.c_int]
    L.gg_init_dt.argtypes = [_vp, C.c_double]
    L.gg_grav_step.argtypes = [_vp, C.c_double]
    L.gg_accel_step.argtypes = [_vp, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int,
                                C.c_int, C.c_double]
    L.gg_density_step.argtypes = [_vp, C.c_double, C.c_double]
    L.gg_sph_step.argtypes = [_vp] + [C.c_double] * 7 + [C.c_int, _dp]
    L.gg_dt_to_rung.argtypes = [_vp, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int, _ip, _ip,
                                _ip]
    L.gg_active_mask_rung.argtypes = [_vp, C.c_uint, C.c_int, C.c_int, _ip]
    L.gg_active_type.argtypes = [_vp, C.c_uint, C.c_uint, _ip]
    L.gg_active_type_rung.argtypes = [_vp, C.c_uint, C.c_uint, C.c_int, C.c_int, _ip]
    L.gg_active_exact_type.argtypes = [_vp, C.c_uint, C.c_uint, C.c_uint, _ip]
    L.gg_set_type.argtypes = [_vp, C.c_uint, C.c_uint, _ip]
    L.gg_reset_type.argtypes = [_vp, C.c_uint, C.c_uint, _ip]
    L.gg_sph_curr_rung.argtypes = [_vp, C.c_int, C.c_int, _ip]
    L.gg_emergency_adjust.argtypes = [_vp, C.c_int, C.c_int, C.c_double, C.c_double, _ip, _ip,
                                      _ip, _ip]
    _fp = C.POINTER(C.c_float)
    L.gg_hmlp_init.argtypes = [_vp, C.c_uint]
    L.gg_hmlp_record.argtypes = [_vp, C.c_double]
    L.gg_hmlp_train.argtypes = [_vp, C.c_int, C.c_int, C.c_double, C.c_uint, _dp, _dp, _ip]
    L.gg_hmlp_predict.argtypes = [_vp]
    L.gg_hmlp_params.argtypes = [_vp, _fp, C.c_int, C.c_int]
    L.gg_hmlp_eval.argtypes = [_vp, _fp, C.c_int, _fp]
    L.gg_hmlp_grad.argtypes = [_vp, _fp, _fp, C.c_int, _fp, _dp]
    _LIB = L
    return L


def _check(rc):
    if rc != 0:
        msg = lib().gg_last_error().decode(errors="replace")
        raise GGError(f"libgasoline_amd error {rc}: {msg}")


def _p(a, t=_dp):
    return a.ctypes.data_as(t)


class Pkd:
    """A GGPKD handle: device-resident particle store of nStore particles."""

    def __init__(self, n, device=0, prec=PREC_FP32):
        L = lib()
        h = _vp()
        _check(L.gg_initialize(C.byref(h), int(device), int(n)))
        self.h = h
        self._finish = L.gg_finish   # held so that close() works at interpreter exit
        self.n = int(n)
        self.set_precision(prec)

    def close(self):
        if getattr(self, "h", None):
            self._finish(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_precision(self, prec):
        _check(lib().gg_set_precision(self.h, int(prec)))

    # ---------------------------------------------------------- fields
    def put(self, name, arr):
        """Copy a field in, from a numpy array or a device tensor (torch)."""
        fid, w = FIELDS[name]
        if hasattr(arr, "data_ptr"):             # device tensor
            import torch
            # the handle's stream does not wait for torch's: the tensor's
            # values must be written before the copy is queued
            torch.cuda.current_stream(arr.device).synchronize()
            _check(lib().gg_put(self.h, fid, _vp(arr.data_ptr()), 1))
            return
        dt = _INT_FIELDS.get(name, np.float64)
        a = np.ascontiguousarray(arr, dtype=dt).reshape(-1)
        if a.size != self.n * w:
            raise ValueError(f"{name}: expected {self.n * w} values, got {a.size}")
        _check(lib().gg_put(self.h, fid, a.ctypes.data_as(_vp), 0))

    def get(self, name):
        fid, w = FIELDS[name]
        dt = _INT_FIELDS.get(name, np.float64)
        a = np.empty(self.n * w, dtype=dt)
        _check(lib().gg_get(self.h, fid, a.ctypes.data_as(_vp), 0))
        return a.reshape(self.n, w) if w > 1 else a

    @staticmethod
    def _aos_members(arr, members):
        """members: {field name: member name of arr's structured dtype}."""
        if arr.dtype.fields is None or not arr.flags.c_contiguous:
            raise ValueError("a C-contiguous structured array is required")
        tab = (AosField * max(len(members), 1))()
        for k, (name, mem) in enumerate(members.items()):
            tab[k].field = FIELDS[name][0]
            tab[k].offset = arr.dtype.fields[mem][1]
        return tab

    def put_aos(self, arr, members):
        """gg_put_aos: fields in from a structured array of nLocal records
        (gasoline's PARTICLE layout or any other)."""
        if len(arr) != self.n:
            raise ValueError(f"expected {self.n} records, got {len(arr)}")
        tab = self._aos_members(arr, members)
        _check(lib().gg_put_aos(self.h, arr.ctypes.data_as(_vp), arr.dtype.itemsize, tab,
                                len(members)))

    def get_aos(self, arr, members):
        """gg_get_aos: the listed members of arr overwritten in place."""
        if len(arr) != self.n:
            raise ValueError(f"expected {self.n} records, got {len(arr)}")
        tab = self._aos_members(arr, members)
        _check(lib().gg_get_aos(self.h, arr.ctypes.data_as(_vp), arr.dtype.itemsize, tab,
                                len(members)))

    # --------------------------------------------------------- gravity
    def build_binary(self, nBucket=16, iOpenType=1, dCrit=0.7, iOrder=2):
        _check(lib().gg_build_binary(self.h, int(nBucket), int(iOpenType), float(dCrit),
                                     int(iOrder), 0, 1))

    def read_tipsy(self, path, standard=True, dvFac=1.0, dTuFac=1.0):
        """gg_read_tipsy; returns the file's time."""
        t = C.c_double()
        _check(lib().gg_read_tipsy(self.h, os.fsencode(path), int(bool(standard)), dvFac, dTuFac,
                                   C.byref(t)))
        self.n = lib().gg_local(self.h)
        return t.value

    def write_tipsy(self, path, time, standard=True, dvFac=1.0, duTFac=1.0):
        _check(lib().gg_write_tipsy(self.h, os.fsencode(path), int(bool(standard)), time, dvFac,
                                    duTFac))

    def out_array(self, path, name):
        """gg_out_array: the field as the reference's ASCII output array."""
        _check(lib().gg_out_array(self.h, os.fsencode(path), FIELDS[name][0]))

    def set_period(self, period):
        L = np.ascontiguousarray(period, np.float64)
        _check(lib().gg_set_period(self.h, _p(L)))

    def grav_all(self, iOrder=2, nReps=0, periodic=False, bEwald=False, iEwOrder=4,
                 fEwCut=2.6, fEwhCut=2.8):
        """gg_grav_all; periodic uses the box of set_period."""
        na = C.c_int()
        ps, cs, ss, fl = C.c_double(), C.c_double(), C.c_double(), C.c_double()
        _check(lib().gg_grav_all(self.h, int(nReps), int(bool(periodic)), int(iOrder),
                                 int(bool(bEwald)), int(iEwOrder), float(fEwCut), float(fEwhCut),
                                 0, 0.0, 0, 0.0, None, C.byref(na), C.byref(ps), C.byref(cs),
                                 C.byref(ss), C.byref(fl)))
        return dict(nActive=na.value, dPartSum=ps.value, dCellSum=cs.value,
                    dSoftSum=ss.value, dFlop=fl.value)

    def build_forest(self, sysOff, nBucket=16, iOpenType=1, dCrit=0.7, iOrder=2):
        off = np.ascontiguousarray(sysOff, np.int32)
        _check(lib().gg_build_forest(self.h, len(off) - 1, _p(off, _ip), int(nBucket),
                                     int(iOpenType), float(dCrit), int(iOrder)))

    def set_grav_schedule(self, mode):
        """gg_set_grav_schedule: 0 one workgroup per bucket, 1 persistent
        queue (heaviest buckets first)."""
        _check(lib().gg_set_grav_schedule(self.h, int(mode)))

    def grav_graph(self, iOrder=2):
        _check(lib().gg_grav_graph(self.h, int(iOrder)))

    def grav_replay(self):
        na = C.c_int()
        ps, cs, ss, fl = C.c_double(), C.c_double(), C.c_double(), C.c_double()
        _check(lib().gg_grav_replay(self.h, C.byref(na), C.byref(ps), C.byref(cs),
                                    C.byref(ss), C.byref(fl)))
        return dict(nActive=na.value, dPartSum=ps.value, dCellSum=cs.value,
                    dSoftSum=ss.value, dFlop=fl.value)

    def tune_bucket(self, candidates=(8, 12, 16, 24, 32), sysOff=None, iOpenType=1, dCrit=0.7,
                    iOrder=2):
        cand = np.ascontiguousarray(candidates, np.int32)
        ms = np.zeros(len(cand))
        best = C.c_int()
        if sysOff is None:
            off, ns = None, 0
        else:
            off = np.ascontiguousarray(sysOff, np.int32)
            ns = len(off) - 1
        _check(lib().gg_tune_bucket(self.h, ns, None if off is None else _p(off, _ip),
                                    int(iOpenType), float(dCrit), int(iOrder), _p(cand, _ip),
                                    len(cand), _p(ms), C.byref(best)))
        return best.value, dict(zip(cand.tolist(), ms.tolist()))

    def grav_lists(self):
        nb = C.c_int()
        ne = C.c_int64()
        _check(lib().gg_grav_lists(self.h, C.byref(nb), C.byref(ne), None, None, None, None))
        bk = np.empty(max(nb.value, 1), np.int32)
        off = np.empty(3 * max(nb.value, 1), np.int64)
        cnt = np.empty(3 * max(nb.value, 1), np.int32)
        ent = np.empty(max(ne.value, 1), np.int32)
        _check(lib().gg_grav_lists(self.h, C.byref(nb), C.byref(ne), _p(bk, _ip),
                                   _p(off, _i64p), _p(cnt, _ip), _p(ent, _ip)))
        n = nb.value
        return bk[:n], off[:3 * n].reshape(n, 3), cnt[:3 * n].reshape(n, 3), ent[:ne.value]

    def walk_checksums(self, iOrder=2):
        """Per active bucket and list kind (PP, softened, Newtonian): the
        number and the sum of the members the fused walk evaluates and the
        sum of their splitmix64 mixes (gg_grav_walk_checksums), shape
        (nb, 3, 3), uint64."""
        nb = C.c_int()
        ne = C.c_int64()
        _check(lib().gg_grav_lists(self.h, C.byref(nb), C.byref(ne), None, None, None, None))
        out = np.zeros((max(nb.value, 1), 3, 3), np.uint64)
        _check(lib().gg_grav_walk_checksums(self.h, int(iOrder), out.ctypes.data_as(_vp)))
        return out[:nb.value]

    def direct_grav(self, targets):
        """FP64 direct-sum (a, pot) of the given caller indices (G = 1, spline
        softening h_i + h_j)."""
        t = np.ascontiguousarray(targets, dtype=np.int32)
        a = np.empty((len(t), 3))
        pot = np.empty(len(t))
        _check(lib().gg_direct_grav(self.h, len(t), _p(t, _ip), _p(a), _p(pot)))
        return a, pot

    def direct_ewald(self, targets):
        """FP64 periodic direct sum (a, pot) of the given caller indices in the
        cubic box set by set_period: Ewald-summed images with the uniform
        background, spline softening inside h_i + h_j."""
        t = np.ascontiguousarray(targets, dtype=np.int32)
        a = np.empty((len(t), 3))
        pot = np.empty(len(t))
        _check(lib().gg_direct_ewald(self.h, len(t), _p(t, _ip), _p(a), _p(pot)))
        return a, pot

    def tree_info(self):
        nn, root = C.c_int(), C.c_int()
        _check(lib().gg_tree_info(self.h, C.byref(nn), C.byref(root), None, None, None))
        perm = np.empty(self.n, np.int32)
        ints = np.empty((nn.value, 6), np.int32)
        dbls = np.empty((nn.value, 8))
        _check(lib().gg_tree_info(self.h, C.byref(nn), C.byref(root), _p(perm, _ip),
                                  _p(ints, _ip), _p(dbls)))
        return root.value, perm, ints, dbls

    # ------------------------------------------------------------- SPH
    def build_local(self, period=None, nBucket=32):
        if period is None:
            _check(lib().gg_build_local(self.h, int(nBucket), 0, None))
        else:
            L = np.ascontiguousarray(period, dtype=np.float64)
            _check(lib().gg_build_local(self.h, int(nBucket), 1, _p(L)))

    def smooth(self, smf, nSmooth=64, periodic=False, iSmoothType=SMX_DENDVDX, symmetric=False):
        """gg_smooth; symmetric selects DensitySym for SMX_DENSITY (DenDVDX
        is the same operator either way)."""
        ns = C.c_int()
        _check(lib().gg_smooth(self.h, C.byref(smf), int(nSmooth), int(bool(periodic)),
                               int(bool(symmetric)), int(iSmoothType), C.byref(ns)))
        return ns.value

    def tune_sph(self, smf, nSmooth=64, period=None, cand=(32, 24, 16)):
        """gg_tune_sph: (chosen nBucket, {candidate: ms})."""
        c = np.ascontiguousarray(cand, dtype=np.int32)
        ms = np.zeros(len(c))
        best = C.c_int()
        L = None if period is None else np.ascontiguousarray(period, dtype=np.float64)
        _check(lib().gg_tune_sph(self.h, C.byref(smf), int(nSmooth), int(period is not None),
                                 None if L is None else _p(L), _p(c, _ip), len(c), _p(ms),
                                 C.byref(best)))
        return best.value, {int(k): float(v) for k, v in zip(c, ms)}

    def set_sph_lanes(self, lanes):
        """gg_set_sph_lanes: 4, 8 or 16 lanes per particle in DenDVDX and
        the force pass."""
        _check(lib().gg_set_sph_lanes(self.h, int(lanes)))

    def tune_sph_lanes(self, smf, nSmooth=64, periodic=False, cand=(4, 8, 16)):
        """gg_tune_sph_lanes: (chosen lanes per particle, {candidate: ms})."""
        c = np.ascontiguousarray(cand, dtype=np.int32)
        ms = np.zeros(len(c))
        best = C.c_int()
        _check(lib().gg_tune_sph_lanes(self.h, C.byref(smf), int(nSmooth), int(bool(periodic)),
                                       _p(c, _ip), len(c), _p(ms), C.byref(best)))
        return best.value, {int(k): float(v) for k, v in zip(c, ms)}

    def resmooth(self, smf, nSmooth=64, periodic=False, iSmoothType=SMX_SPHPRESSURETERMS):
        _check(lib().gg_resmooth(self.h, C.byref(smf), int(nSmooth), int(bool(periodic)), 1,
                                 int(iSmoothType)))

    def gas_pressure(self, smf):
        _check(lib().gg_gas_pressure(self.h, C.byref(smf)))

    def smooth_lists(self, nSmooth=64):
        out = np.empty((self.n, nSmooth), np.int32)
        _check(lib().gg_smooth_lists(self.h, int(nSmooth), _p(out, _ip)))
        return out

    def smooth_stats(self):
        s = np.zeros(10, np.int64)
        _check(lib().gg_smooth_stats(self.h, _p(s, _i64p), 10))
        return dict(retried=int(s[0]), ties=int(s[1]), extras=int(s[2]), queries=int(s[3]),
                    groups_overflowed=int(s[4]), short_at_RL=int(s[5]), unsettled=int(s[6]),
                    max_buckets=int(s[7]), extra_passes=int(s[8]), candidates=int(s[9]))

    def reorder(self, want_perm=True):
        """Permute the store into the SPH structure's order (gg_reorder);
        returns perm with new[k] = old[perm[k]] (None if not wanted: the
        copy to the host is then skipped)."""
        perm = np.empty(self.n, np.int32) if want_perm else None
        _check(lib().gg_reorder(self.h, _p(perm, _ip) if want_perm else None))
        return perm

    def set_seed_slack(self, slack):
        _check(lib().gg_set_seed_slack(self.h, float(slack)))

    # --- kick, drift, accel reset (pkdInitAccel, pkdKick, pkdDrift)
    def init_accel(self):
        _check(lib().gg_init_accel(self.h))

    def kick(self, dvFacOne, dvFacTwo, dvPredFacOne, dvPredFacTwo, duDelta, duPredDelta,
             bIsothermal=False):
        _check(lib().gg_kick(self.h, dvFacOne, dvFacTwo, dvPredFacOne, dvPredFacTwo, duDelta,
                             duPredDelta, int(bIsothermal)))

    def drift(self, dDelta, period=None, center=(0.0, 0.0, 0.0)):
        if period is None:
            _check(lib().gg_drift(self.h, dDelta, None, 0, None))
        else:
            c = np.ascontiguousarray(center, np.float64)
            L = np.ascontiguousarray(period, np.float64)
            _check(lib().gg_drift(self.h, dDelta, _p(c), 1, _p(L)))

    def kick_vpred(self, dvFacOne, dvFacTwo, duDelta, bIsothermal=False):
        _check(lib().gg_kick_vpred(self.h, dvFacOne, dvFacTwo, duDelta, int(bIsothermal)))

    # --- timesteps and rungs (pkdInitDt, pkdGravStep, pkdAccelStep,
    # pkdDensityStep, pkdSphStep, pkdDtToRung)
    def init_dt(self, dDelta):
        _check(lib().gg_init_dt(self.h, dDelta))

    def grav_step(self, dEta):
        _check(lib().gg_grav_step(self.h, dEta))

    def accel_step(self, dEta, dVelFac=1.0, dAccFac=1.0, bDoGravity=True, bEpsAcc=True,
                   bSqrtPhi=False, dhMinOverSoft=0.0):
        _check(lib().gg_accel_step(self.h, dEta, dVelFac, dAccFac, int(bDoGravity), int(bEpsAcc),
                                   int(bSqrtPhi), dhMinOverSoft))

    def density_step(self, dEta, dRhoFac):
        _check(lib().gg_density_step(self.h, dEta, dRhoFac))

    def sph_step(self, dCosmoFac, dEtaCourant, dEtauDot=0.0, duMinDt=0.0, dDiffCoeff=0.0,
                 dEtaDiffusion=0.1, dResolveJeans=0.0, bViscosityLimitdt=False):
        """Returns dtMinGas."""
        m = C.c_double()
        _check(lib().gg_sph_step(self.h, dCosmoFac, dEtaCourant, dEtauDot, duMinDt, dDiffCoeff,
                                 dEtaDiffusion, dResolveJeans, int(bViscosityLimitdt),
                                 C.byref(m)))
        return m.value

    def active_rung(self, iRung, bGreater, mask=TYPE_ACTIVE):
        """gg_active_mask_rung; returns the number of particles set."""
        n = C.c_int()
        _check(lib().gg_active_mask_rung(self.h, int(mask), int(iRung), int(bool(bGreater)),
                                         C.byref(n)))
        return n.value

    def active_type(self, iTestMask, iSetMask, iRung=None, bGreater=True, iFilterMask=None):
        """gg_active_type (pkdActiveType); with iRung gg_active_type_rung,
        with iFilterMask gg_active_exact_type.  Returns the matching count."""
        n = C.c_int()
        if iFilterMask is not None:
            _check(lib().gg_active_exact_type(self.h, int(iFilterMask), int(iTestMask),
                                              int(iSetMask), C.byref(n)))
        elif iRung is not None:
            _check(lib().gg_active_type_rung(self.h, int(iTestMask), int(iSetMask), int(iRung),
                                             int(bool(bGreater)), C.byref(n)))
        else:
            _check(lib().gg_active_type(self.h, int(iTestMask), int(iSetMask), C.byref(n)))
        return n.value

    def set_type(self, iTestMask, iSetMask, reset=False):
        """gg_set_type / gg_reset_type (pkdSetType / pkdResetType)."""
        n = C.c_int()
        fn = lib().gg_reset_type if reset else lib().gg_set_type
        _check(fn(self.h, int(iTestMask), int(iSetMask), C.byref(n)))
        return n.value

    def sph_curr_rung(self, iRung, bGreater=True):
        c = C.c_int()
        _check(lib().gg_sph_curr_rung(self.h, int(iRung), int(bool(bGreater)), C.byref(c)))
        return bool(c.value)

    def emergency_adjust(self, iRung, iMaxRung, dDelta, dDeltaThresh):
        """gg_emergency_adjust; returns (nUn, iMaxRungOut, nMaxRung, iMaxRungIdeal)."""
        v = [C.c_int() for _ in range(4)]
        _check(lib().gg_emergency_adjust(self.h, int(iRung), int(iMaxRung), float(dDelta),
                                         float(dDeltaThresh), *[C.byref(x) for x in v]))
        return v[0].value, v[3].value, v[2].value, v[1].value

    def dt_to_rung(self, iRung, dDelta, iMaxRung, bAll=True, bDiagExceed=False):
        """Returns (iMaxRungOut, nMaxRung, iMaxRungIdeal)."""
        nm, ideal, top = C.c_int(), C.c_int(), C.c_int()
        _check(lib().gg_dt_to_rung(self.h, int(iRung), dDelta, int(iMaxRung), int(bAll),
                                   int(bDiagExceed), C.byref(nm), C.byref(ideal), C.byref(top)))
        return top.value, nm.value, ideal.value

    # --- h-proposal MLP
    def hmlp_init(self, seed=1):
        _check(lib().gg_hmlp_init(self.h, int(seed)))

    def hmlp_record(self, dt):
        _check(lib().gg_hmlp_record(self.h, float(dt)))

    def hmlp_train(self, nIter, nBatch=4096, lr=1e-3, seed=1):
        l0, l1, ns = C.c_double(), C.c_double(), C.c_int()
        _check(lib().gg_hmlp_train(self.h, int(nIter), int(nBatch), float(lr), int(seed),
                                   C.byref(l0), C.byref(l1), C.byref(ns)))
        return l0.value, l1.value, ns.value

    def hmlp_predict(self):
        _check(lib().gg_hmlp_predict(self.h))

    def hmlp_params(self, w=None):
        fp = C.POINTER(C.c_float)
        if w is None:
            out = np.zeros(HMLP_NP, np.float32)
            _check(lib().gg_hmlp_params(self.h, out.ctypes.data_as(fp), HMLP_NP, 0))
            return out
        w = np.ascontiguousarray(w, np.float32)
        _check(lib().gg_hmlp_params(self.h, w.ctypes.data_as(fp), len(w), 1))

    def hmlp_eval(self, x):
        """x: (n, 4) features -> y (n,)."""
        fp = C.POINTER(C.c_float)
        xs = np.ascontiguousarray(np.asarray(x, np.float32).T)
        y = np.zeros(xs.shape[1], np.float32)
        _check(lib().gg_hmlp_eval(self.h, xs.ctypes.data_as(fp), xs.shape[1],
                                  y.ctypes.data_as(fp)))
        return y

    def hmlp_grad(self, x, t):
        """Gradient of mean (y - t)^2 over samples x (n, 4): (g[HMLP_NW], loss)."""
        fp = C.POINTER(C.c_float)
        xs = np.ascontiguousarray(np.asarray(x, np.float32).T)
        ts = np.ascontiguousarray(t, np.float32)
        g = np.zeros(HMLP_NW, np.float32)
        loss = C.c_double()
        _check(lib().gg_hmlp_grad(self.h, xs.ctypes.data_as(fp), ts.ctypes.data_as(fp),
                                  xs.shape[1], g.ctypes.data_as(fp), C.byref(loss)))
        return g, loss.value

    def timings(self):
        t = np.zeros(11)
        _check(lib().gg_timings(self.h, _p(t), 11))
        return dict(zip(["walk", "interact", "sph_tree", "knn", "density", "lists",
                         "revlists", "hydro", "other", "knn_lists", "knn_retry"],
                        (float(x) for x in t)))


# ---------------------------------------------------- C host sequences
class StepParam(C.Structure):
    """ggStepParam (gasoline_amd/host/gg_msr.h): the msr->param fields the
    block-timestep recursion reads."""
    _fields_ = [("dDelta", C.c_double), ("iMaxRung", C.c_int), ("bAdjust", C.c_int),
                ("dEta", C.c_double),
                ("dEtaCourant", C.c_double), ("dEtauDot", C.c_double), ("duMinDt", C.c_double),
                ("bGravStep", C.c_int), ("bAccelStep", C.c_int), ("bEpsAccStep", C.c_int),
                ("bSqrtPhiStep", C.c_int), ("bSphStep", C.c_int), ("bViscosityLimitdt", C.c_int),
                ("bDoGravity", C.c_int), ("bDoGas", C.c_int), ("nBucket", C.c_int),
                ("iOrder", C.c_int), ("dTheta", C.c_double), ("nSmooth", C.c_int),
                ("nBucketSph", C.c_int), ("smf", Smf), ("bPeriodic", C.c_int),
                ("fPeriod", C.c_double * 3), ("fCenter", C.c_double * 3)]


class StepState(C.Structure):
    _fields_ = [("iCurrMaxRung", C.c_int), ("nForce", C.c_int), ("nLogCap", C.c_int),
                ("logKickRung", _ip), ("logActive", _ip), ("nEmergency", C.c_int),
                ("dtMinGas", C.c_double)]


_MSR = None


def msr_lib():
    """gasoline_amd/host/libggmsr.so, the C host sequences (built with the
    library)."""
    global _MSR
    if _MSR is None:
        path = os.path.join(_HERE, "host", "libggmsr.so")
        if not os.path.exists(path):
            raise RuntimeError("gasoline_amd/host/libggmsr.so is not built (run make in csrc)")
        lib()
        L = C.CDLL(path)
        L.ggMsrTopStepKDK.argtypes = [_vp, C.POINTER(StepParam), C.POINTER(StepState),
                                      C.c_double, C.c_double, C.c_double, C.c_int, C.c_int,
                                      C.c_int]
        L.ggMsrForces.argtypes = [_vp, C.POINTER(StepParam), C.POINTER(StepState), C.c_int]
        _MSR = L
    return _MSR


class BlockSteps:
    """Drives ggMsrTopStepKDK (msrTopStepKDK, master.c:8242) of the C host
    over a Pkd: p holds the parameters, the state keeps iCurrMaxRung and a
    log of the force evaluations."""

    def __init__(self, pk, dDelta, iMaxRung=8, dEta=0.2, dEtaCourant=0.4, smf=None,
                 bDoGravity=True, bDoGas=True, bGravStep=True, bSphStep=True, dTheta=0.7,
                 iOrder=2, nBucket=16, nSmooth=32, nBucketSph=32, period=None, log_cap=4096,
                 bAdjust=True):
        smf = smf if smf is not None else Smf.default()
        smf.dtFacCourant = dEtaCourant * 2 / 1.6
        self.pk = pk
        self.p = StepParam(dDelta=dDelta, iMaxRung=iMaxRung, bAdjust=int(bAdjust), dEta=dEta,
                           dEtaCourant=dEtaCourant,
                           dEtauDot=0.25, duMinDt=0.0, bGravStep=int(bGravStep), bAccelStep=0,
                           bEpsAccStep=0, bSqrtPhiStep=0, bSphStep=int(bSphStep),
                           bViscosityLimitdt=0, bDoGravity=int(bDoGravity), bDoGas=int(bDoGas),
                           nBucket=nBucket, iOrder=iOrder, dTheta=dTheta, nSmooth=nSmooth,
                           nBucketSph=nBucketSph, smf=smf, bPeriodic=int(period is not None))
        if period is not None:
            for k in range(3):
                self.p.fPeriod[k] = period[k]
        self._rung = np.zeros(log_cap, np.int32)
        self._act = np.zeros(log_cap, np.int32)
        self.s = StepState(iCurrMaxRung=0, nForce=0, nLogCap=log_cap,
                           logKickRung=_p(self._rung, _ip), logActive=_p(self._act, _ip),
                           nEmergency=0, dtMinGas=0.0)
        self.dTime = 0.0

    def forces(self, iKickRung=0):
        rc = msr_lib().ggMsrForces(self.pk.h, C.byref(self.p), C.byref(self.s), int(iKickRung))
        _check(rc)

    def top_step(self):
        rc = msr_lib().ggMsrTopStepKDK(self.pk.h, C.byref(self.p), C.byref(self.s), 0.0,
                                      self.dTime, self.p.dDelta, 0, 0, 1)
        _check(rc)
        self.dTime += self.p.dDelta

    def force_log(self):
        n = min(self.s.nForce, self.s.nLogCap)
        return list(zip(self._rung[:n].tolist(), self._act[:n].tolist()))
