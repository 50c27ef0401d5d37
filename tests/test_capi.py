"""CPU checks of the C ABI: the library loads, exports every function that
include/*.h declares, and sia_params_default() equals the pism_config.cdl
defaults (no GPU work is launched)."""
import ctypes
import os
import re
import subprocess

import pytest

import sia_types as T

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(REPO, "pismv0.7.3_amd", "lib", "libpism_sia.so")


def declared_functions():
    names = set()
    inc = os.path.join(REPO, "include")
    for fn in sorted(os.listdir(inc)):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(inc, fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b(sia_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        subprocess.check_call(["make", "-s", "-C", os.path.join(REPO, "pismv0.7.3_amd")])
    return ctypes.CDLL(LIB)


def test_header_declares_the_boundary():
    names = declared_functions()
    for n in ("sia_create", "sia_destroy", "sia_update", "sia_diffusive_flux", "sia_max_diffusivity",
              "sia_velocity_u", "sia_velocity_v", "sia_velocity_w", "sia_mass_continuity_step",
              "sia_step", "sia_last_error", "sia_params_default", "sia_workload_generate"):
        assert n in names


def test_library_exports_every_declared_symbol(lib):
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.check_output(["nm", "-D", "--defined-only", LIB]).decode()
    exported = set(re.findall(r" T (sia_\w+)", out))
    assert names <= exported


def test_python_wrapper_covers_the_abi():
    import pism_sia
    assert set(pism_sia.SIGNATURES) == declared_functions()


def test_params_default_matches_config_defaults(lib):
    p = T.SiaParams()
    lib.sia_params_default(ctypes.byref(p))
    d = T.params_dict(p)
    for k, v in T.DEFAULTS.items():
        assert d[k] == v, k
    assert ctypes.sizeof(T.SiaParams) == 4 * 8 + 8 * len(T._PARAM_DOUBLE_FIELDS)


def test_create_rejects_invalid_grids_before_any_device_call(lib):
    """sia_create checks the grid and the parameter selections before its
    first HIP call and returns a status: at least 3 columns each way, at least
    2 and at most 4096 levels, strictly increasing levels, known flow laws."""
    create = lib.sia_create
    create.argtypes = [ctypes.POINTER(T.SiaGrid), ctypes.POINTER(T.SiaParams), ctypes.POINTER(ctypes.c_void_p)]
    create.restype = ctypes.c_int
    h = ctypes.c_void_p()

    def status(Mx, My, z, params=None):
        g = T.make_grid(Mx, My, 1000.0, 1000.0, z)
        p = params if params is not None else T.make_params()
        return create(ctypes.byref(g), ctypes.byref(p), ctypes.byref(h))

    z = T.vertical_levels(4000.0, 11)
    assert status(5, 5, T.vertical_levels(4000.0, 4097)) == T.SIA_ERR_ARGUMENT
    assert status(2, 5, z) == T.SIA_ERR_ARGUMENT
    assert status(5, 2, z) == T.SIA_ERR_ARGUMENT
    assert status(5, 5, z[:1]) == T.SIA_ERR_ARGUMENT
    flat = z.copy()
    flat[5] = flat[4]
    assert status(5, 5, flat) == T.SIA_ERR_ARGUMENT
    assert status(5, 5, z, T.make_params(sia_flow_law=99)) == T.SIA_ERR_CONFIG
    assert status(5, 5, z, T.make_params(surface_gradient_method=7)) == T.SIA_ERR_CONFIG
    assert h.value is None


def test_library_is_built_for_gfx950(lib):
    data = open(LIB, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in data
