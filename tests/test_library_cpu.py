"""The C-ABI library without a GPU: it loads, exports every entry point
include/*.h declares, and its host-side logic (colour pipeline,
camera placement, argument validation) matches the oracle.  No kernel runs."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import magik_amd as M
from magik_amd import abi
from oracle import oracle as O

INCLUDE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")


def declared_symbols():
    text = "".join(open(os.path.join(INCLUDE, f)).read() for f in sorted(os.listdir(INCLUDE)) if f.endswith(".h"))
    return sorted(set(re.findall(r"MGK_API\s+[\w\s\*]+?\b(\w+)\s*\(", text)))


def test_library_loads_and_exports_every_declared_symbol():
    lib = abi.load()
    names = declared_symbols()
    assert len(names) >= 12
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(abi.SIGNATURES), "abi.py must mirror every declared entry point"


def test_version():
    maj, mi, rev = C.c_uint32(), C.c_uint32(), C.c_uint32()
    s = C.c_char_p()
    assert abi.load().magik_get_version(C.byref(maj), C.byref(mi), C.byref(rev), C.byref(s)) == abi.MAGIK_SUCCESS
    assert (maj.value, mi.value) == (0, 7) and b"magik" in s.value


def test_color_weights_match_oracle_bitexact():
    for n_bin, lo, hi in ((24, 360.0, 830.0), (16, 380.0, 780.0), (64, 360.0, 830.0)):
        spec = M.spectral(n_bin=n_bin, bin_min=lo, bin_max=hi)
        W, Mx = M.color_weights(spec)
        ref = O.bin_weights(n_bin, lo, hi, O.cie_table())
        assert np.array_equal(W, ref)
        assert float(W[:, 1].sum()) * (hi - lo) / n_bin == pytest.approx(1.0, rel=1e-5)


@pytest.mark.parametrize("T", [3200.0, 5600.0, 6500.0])
def test_xyz_to_scene_matrix_white_balances_the_blackbody(T):
    # the white-balance CAT maps the blackbody white point onto the display white:
    # its XYZ must come out neutral, R = G = B, in linear Rec.709
    _, Mx = M.color_weights(M.spectral(white_K=T))
    tab = O.cie_table().astype(np.float64)
    lam = 360.0 + np.arange(471)
    B = np.array([O.lib().orc_planck(O.F64, l, T) for l in lam])
    xyz = (tab * B[:, None]).sum(axis=0)
    rgb = Mx @ (xyz / xyz[1])
    assert np.allclose(rgb, rgb[1], rtol=2e-3), rgb


def test_argument_validation_without_gpu():
    lib = abi.load()
    steps = (abi.Step * 4)()
    cont = abi.Container(capacity=123, n_allocated_step=4, n_written=0, steps=steps)
    mp = abi.MetricParameters(metric_type=abi.KERR, kerr_mass=6.7e26, kerr_spin=0.5, event_horizon_scalar=1.01)
    ip = abi.IntegratorParameters(56, 1e-5, 1e-5)
    ic = abi.InitialConditions()
    assert lib.magik_trace_geodesic(None, ic, mp, ip) == abi.MAGIK_ERROR_INVALID_POINTER
    assert lib.magik_trace_geodesic(C.byref(cont), ic, mp, ip) == \
        abi.MAGIK_ERROR_HOST_TRACE_ALLOCATED_STEPS_AND_CAPACITY_MISMATCH
    cont.capacity = C.sizeof(abi.Step) * 4
    mp.metric_type = 7
    assert lib.magik_trace_geodesic(C.byref(cont), ic, mp, ip) == abi.MAGIK_ERROR_UNKNOWN
    cam = abi.Camera()
    assert lib.mgk_camera_boyer_lindquist(50.0, 0.0, 0.9, 6.7e26, 25.0, 64, 36, C.byref(cam)) == \
        abi.MGK_ERROR_INVALID_ARGUMENT   # on the spin axis: no up vector
    assert lib.mgk_camera_boyer_lindquist(-1.0, 80.0, 0.9, 6.7e26, 25.0, 64, 36, C.byref(cam)) == \
        abi.MGK_ERROR_INVALID_ARGUMENT
    assert lib.mgk_color_weights(C.byref(M.spectral(n_bin=65)), None, None) == abi.MGK_ERROR_INVALID_ARGUMENT
    h = C.c_void_p()
    assert lib.mgk_context_create(-1, C.byref(h)) == abi.MAGIK_INVALID_CUDA_DEVICE
    assert lib.mgk_context_destroy(None) == abi.MAGIK_ERROR_INVALID_POINTER


def test_workload_definitions():
    integ, sc, spec, cam, spp = M.workload("kerr_disk_1080p")
    assert (cam.x_resolution, cam.y_resolution, spp, spec.n_hero) == (1920, 1080, 32, 4)
    assert sc.disk_r_in == pytest.approx(2.3209, abs=1e-4)        # r_ISCO(a = 0.9)
    assert integ.spin == pytest.approx(0.9) and integ.abs_tol == pytest.approx(1e-7)
    integ, sc, spec, cam, spp = M.workload("schwarzschild_disk")
    assert (cam.x_resolution, spp, sc.n_star) == (512, 16, 100000)
    integ, sc, spec, cam, spp = M.workload("minkowski_sphere")
    assert (cam.x_resolution, spp, integ.metric_type) == (256, 4, M.MINKOWSKI)
