"""gg_direct_grav_noalloc, the entry point that allocates nothing (its
partial sums in a fixed device array), through the C ABI: the same results
as gg_direct_grav bit for bit, on two calls in a row, the second with fewer
targets (the array keeps the first call's contents); a target outside the
store gives NaN; more targets than the array holds are refused."""
import ctypes as C

import numpy as np
import pytest

from gasoline_amd import ics

pytestmark = pytest.mark.gpu

MAX = 4096      # GG_DIRECT_NOALLOC_MAX


def test_direct_grav_noalloc_equals_direct_grav(gpu):
    import torch
    L = gpu.lib()
    n = 20000
    r, v, m = ics.plummer(n, seed=51)
    r[17] = r[5]                                   # a coincident pair
    soft = np.where(np.arange(n) % 2 == 0, 0.01, 0.03)
    rng = np.random.default_rng(52)
    with gpu.Pkd(n, prec=gpu.PREC_FP64) as pk:
        pk.put("r", r)
        pk.put("fMass", m)
        pk.put("fSoft", soft)
        for nt in (MAX, 700):
            idx = np.sort(rng.choice(n, nt, replace=False)).astype(np.int32)
            idx[0] = 5
            a_ref, p_ref = pk.direct_grav(idx)
            t = torch.from_numpy(idx).cuda()
            a = torch.full((nt, 3), float("nan"), dtype=torch.float64, device="cuda")
            p = torch.full((nt,), float("nan"), dtype=torch.float64, device="cuda")
            torch.cuda.synchronize()
            rc = L.gg_direct_grav_noalloc(pk.h, nt, C.c_void_p(t.data_ptr()),
                                          C.c_void_p(a.data_ptr()), C.c_void_p(p.data_ptr()))
            assert rc == 0
            torch.cuda.synchronize()
            assert np.array_equal(a.cpu().numpy(), a_ref), nt
            assert np.array_equal(p.cpu().numpy(), p_ref), nt
        # a target outside the store: NaN there, the others unchanged
        idx = np.array([3, n + 7, 11], np.int32)
        a_ref, p_ref = pk.direct_grav(idx[[0, 2]])
        t = torch.from_numpy(idx).cuda()
        a = torch.zeros((3, 3), dtype=torch.float64, device="cuda")
        p = torch.zeros(3, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        assert L.gg_direct_grav_noalloc(pk.h, 3, C.c_void_p(t.data_ptr()),
                                        C.c_void_p(a.data_ptr()), C.c_void_p(p.data_ptr())) == 0
        torch.cuda.synchronize()
        a, p = a.cpu().numpy(), p.cpu().numpy()
        assert np.isnan(a[1]).all() and np.isnan(p[1])
        assert np.array_equal(a[[0, 2]], a_ref) and np.array_equal(p[[0, 2]], p_ref)
        # more targets than the fixed array holds
        big = torch.zeros(MAX + 1, dtype=torch.int32, device="cuda")
        out = torch.zeros(4 * (MAX + 1), dtype=torch.float64, device="cuda")
        assert L.gg_direct_grav_noalloc(pk.h, MAX + 1, C.c_void_p(big.data_ptr()),
                                        C.c_void_p(out.data_ptr()),
                                        C.c_void_p(out.data_ptr())) == -1     # GG_EINVAL
        assert b"GG_DIRECT_NOALLOC_MAX" in L.gg_last_error()
