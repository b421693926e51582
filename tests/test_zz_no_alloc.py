"""dgx_calc_sensors_noalloc: the allocation-free variant of Lcalc_sensors
(its scratch is a fixed per-device array, not hipMalloc).  Two calls in a
row, the second on a smaller mesh, each equal bit for bit to Lcalc_sensors
and to the oracle's sensor on the same cell averages; sizes beyond the fixed
scratch are refused on the host with an error message."""
import ctypes

import numpy as np
import pytest
import torch

import dg_amd
import oracle
from dg_amd import MULTIFLUID, ROE, Config, problems
from tests.helpers import small_mesh

pytestmark = pytest.mark.gpu
vp = ctypes.c_void_p
MAX_ELEMENTS = 1 << 21
TH = (0.01, 0.01, 0.01)


def _lib():
    lib = dg_amd.lib()
    i, d, b = ctypes.c_int, ctypes.c_double, ctypes.c_bool
    args = [i, i, b, d, b, d, b, d, vp, vp, vp]
    lib.Lcalc_sensors.argtypes = args
    lib.Lcalc_sensors.restype = None
    lib.dgx_calc_sensors_noalloc.argtypes = args
    lib.dgx_calc_sensors_noalloc.restype = ctypes.c_int
    lib.dgx_last_error.restype = ctypes.c_char_p
    return lib


def _case(device, n, seed):
    mesh = small_mesh(2, 2, MULTIFLUID, n=n)
    t = mesh.tables()
    prob = oracle.Problem(t, MULTIFLUID, ROE)
    U = problems.riemann_state(mesh, MULTIFLUID, seed=seed) * (1 + 0.01 * np.sin(np.arange(prob.state_size)))
    sens_ref, uavg = prob.sensing(U, TH)
    assert (sens_ref != 0).any()
    nb = torch.from_numpy(np.ascontiguousarray(t["neighbors"])).to(device)
    ua = torch.from_numpy(uavg).to(device)
    return t["N_E"], t["N_N"], nb, ua, sens_ref


def _run(fn, NE, NN, nb, ua, device):
    sd = torch.zeros(NE, dtype=torch.int32, device=device)  # SENSOR::sensing starts from zero flags
    rc = fn(NE, NN, True, TH[0], True, TH[1], True, TH[2], vp(nb.data_ptr()), vp(ua.data_ptr()), vp(sd.data_ptr()))
    torch.cuda.synchronize()
    return rc, sd.cpu().numpy()


def test_noalloc_sensors_equal_lcalc_sensors_twice(device):
    lib = _lib()
    assert lib.dgx_set_config(ctypes.byref(Config(2, 2, MULTIFLUID, ROE))) == 0
    for n, seed in (((96, 48), 31), ((40, 24), 32)):  # the second call is smaller
        NE, NN, nb, ua, sens_ref = _case(device, n, seed)
        rc, got = _run(lib.dgx_calc_sensors_noalloc, NE, NN, nb, ua, device)
        assert rc == 0, lib.dgx_last_error()
        _, want = _run(lib.Lcalc_sensors, NE, NN, nb, ua, device)
        print(f"mesh {n}: {int((got != 0).sum())} flagged of {NE}", flush=True)
        assert np.array_equal(got, want)
        assert np.array_equal(got, sens_ref)


def test_noalloc_sensors_refuse_sizes_beyond_scratch(device):
    lib = _lib()
    assert lib.dgx_set_config(ctypes.byref(Config(2, 2, MULTIFLUID, ROE))) == 0
    stand_in = torch.zeros(4, dtype=torch.int32, device=device)  # never read: the size is refused first
    p = vp(stand_in.data_ptr())
    rc = lib.dgx_calc_sensors_noalloc(MAX_ELEMENTS + 1, 4, True, 0.01, True, 0.01, True, 0.01, p, p, p)
    assert rc != 0
    assert b"exceed the fixed scratch" in lib.dgx_last_error()
