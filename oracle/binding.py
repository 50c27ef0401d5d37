"""ctypes binding of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and the CPU-baseline leg of bench.py may
import this module; the HIP product never does. It loads
oracle/liboracle_sia.so (built by oracle/Makefile) and exposes numpy-level
wrappers around the restated PISM functions in oracle/sia_oracle.c.
"""
import ctypes as C
import os
import subprocess
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_REPO = os.path.dirname(_HERE)
_PKG = os.path.join(_REPO, "pismv0.7.3_amd")
if _PKG not in sys.path:
    sys.path.insert(0, _PKG)

import sia_types as T  # noqa: E402

LIB_PATH = os.path.join(_HERE, "liboracle_sia.so")

_dp = C.POINTER(C.c_double)


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def _load():
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    P = C.POINTER(T.SiaParams)
    G = C.POINTER(T.SiaGrid)
    d = C.c_double
    sig = {
        "or_ec_pressure": (d, [P, d]),
        "or_ec_melting_temperature": (d, [P, d]),
        "or_ec_enthalpy_cts": (d, [P, d]),
        "or_ec_temperature": (d, [P, d, d]),
        "or_ec_pressure_adjusted_temperature": (d, [P, d, d]),
        "or_ec_water_fraction": (d, [P, d, d]),
        "or_ec_enthalpy": (d, [P, d, d, d]),
        "or_ec_is_temperate": (C.c_int, [P, d, d]),
        "or_flow": (d, [P, d, d, d, d]),
        "or_flow_uses_grain_size": (C.c_int, [P]),
        "or_grain_size_vostok": (d, [P, d]),
        "or_kbelow": (C.c_int, [_dp, C.c_int, d]),
        "or_bed_smoother": (C.c_int, [G, P, _dp, _dp, _dp, _dp, _dp, _dp,
                                      C.POINTER(C.c_int), C.POINTER(C.c_int)]),
        "or_theta": (C.c_int, [G, P, C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
        "or_thksmooth": (C.c_int, [G, _dp, _dp, _dp, _dp, _dp, _dp]),
        "or_surface_gradient": (C.c_int, [G, P, _dp, _dp, _dp, _dp, _dp, _dp]),
        "or_diffusive_flux": (C.c_int, [G, P, _dp, _dp, _dp, _dp, _dp, _dp, C.c_int,
                                        _dp, _dp, _dp, _dp, _dp]),
        "or_compute_I": (C.c_int, [G, _dp, _dp, _dp, _dp, _dp]),
        "or_horizontal_velocity": (C.c_int, [G, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
        "or_vertical_velocity": (C.c_int, [G, _dp, _dp, _dp, _dp, _dp]),
        "or_hardness": (d, [P, d, d]),
        "or_regional_gradient": (C.c_int, [G, _dp, _dp, _dp, _dp]),
        "or_sia_sliding": (C.c_int, [G, P, C.c_void_p, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                     _dp, _dp, _dp, _dp, _dp]),
        "or_strain_heating": (C.c_int, [G, P, _dp, _dp, _dp, _dp, _dp, _dp]),
        "or_sia_update": (C.c_int, [G, P, C.c_void_p, C.c_int]),
        "or_update_mask_surface": (C.c_int, [G, P, _dp, _dp, _dp, _dp]),
        "or_mass_continuity": (C.c_int, [G, P, d, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
        "or_max_timestep": (d, [G, P, d, d, _dp, _dp, _dp]),
        "or_last_error": (C.c_char_p, []),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


class OracleError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("oracle status %d: %s" % (code, msg))
        self.code = code


def _check(rc):
    if rc != 0:
        raise OracleError(rc, lib().or_last_error().decode())


def _ptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


class OrFields(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in (
        "H", "b", "h", "mask", "E", "age", "ub", "vb", "bmr",
        "topgsmooth", "maxtl", "C2", "C3", "C4",
        "theta", "thksmooth", "h_x", "h_y", "flux", "D_stag",
        "delta0", "delta1", "I0", "I1", "u", "v", "w")] + [
        ("D_max", C.c_double), ("Nx", C.c_int), ("Ny", C.c_int), ("Sigma", C.c_void_p),
        ("no_model_mask", C.c_void_p), ("usurfstore", C.c_void_p)]


# ---------------------------------------------------------------- scalars
def flow(params, stress, E, p, gs=1e-3):
    return lib().or_flow(C.byref(params), stress, E, p, gs)


def kbelow(z, height):
    z = np.ascontiguousarray(z, dtype=np.float64)
    return lib().or_kbelow(_ptr(z), len(z), height)


class EC:
    """Enthalpy converter of the oracle (enthalpyConverter.cc:56-357)."""

    def __init__(self, params):
        self.p = params

    def pressure(self, depth):
        return lib().or_ec_pressure(C.byref(self.p), depth)

    def melting_temperature(self, P):
        return lib().or_ec_melting_temperature(C.byref(self.p), P)

    def enthalpy_cts(self, P):
        return lib().or_ec_enthalpy_cts(C.byref(self.p), P)

    def temperature(self, E, P):
        return lib().or_ec_temperature(C.byref(self.p), E, P)

    def pressure_adjusted_temperature(self, E, P):
        return lib().or_ec_pressure_adjusted_temperature(C.byref(self.p), E, P)

    def water_fraction(self, E, P):
        return lib().or_ec_water_fraction(C.byref(self.p), E, P)

    def enthalpy(self, T, omega, P):
        return lib().or_ec_enthalpy(C.byref(self.p), T, omega, P)

    def is_temperate(self, E, P):
        return bool(lib().or_ec_is_temperate(C.byref(self.p), E, P))


# ---------------------------------------------------------------- fields
def bed_smoother(grid, params, topg):
    n2 = grid.Mx * grid.My
    out = {k: np.zeros(n2) for k in ("topgsmooth", "maxtl", "C2", "C3", "C4")}
    Nx, Ny = C.c_int(0), C.c_int(0)
    _check(lib().or_bed_smoother(C.byref(grid), C.byref(params), _ptr(np.ascontiguousarray(topg.ravel())),
                                 _ptr(out["topgsmooth"]), _ptr(out["maxtl"]), _ptr(out["C2"]),
                                 _ptr(out["C3"]), _ptr(out["C4"]), C.byref(Nx), C.byref(Ny)))
    out["Nx"], out["Ny"] = Nx.value, Ny.value
    return out


def theta(grid, params, sm, usurf):
    th = np.zeros(grid.Mx * grid.My)
    _check(lib().or_theta(C.byref(grid), C.byref(params), sm["Nx"], sm["Ny"],
                          _ptr(np.ascontiguousarray(usurf.ravel())), _ptr(sm["topgsmooth"]),
                          _ptr(sm["maxtl"]), _ptr(sm["C2"]), _ptr(sm["C3"]), _ptr(sm["C4"]), _ptr(th)))
    return th


def thksmooth(grid, usurf, thk, mask, maxtl, topgsmooth):
    """BedSmoother::get_smoothed_thk (PISMBedSmoother.cc:286-336) pointwise."""
    arr = [np.ascontiguousarray(a, dtype=np.float64).ravel() for a in (usurf, thk, mask, maxtl, topgsmooth)]
    out = np.zeros(grid.Mx * grid.My)
    _check(lib().or_thksmooth(C.byref(grid), *[_ptr(a) for a in arr], _ptr(out)))
    return out


def surface_gradient(grid, params, H, b, h, mask):
    n2 = grid.Mx * grid.My
    hx, hy = np.zeros(2 * n2), np.zeros(2 * n2)
    _check(lib().or_surface_gradient(C.byref(grid), C.byref(params), _ptr(H), _ptr(b), _ptr(h),
                                     _ptr(mask), _ptr(hx), _ptr(hy)))
    return hx, hy


def update_mask_surface(grid, params, bed, H):
    n2 = grid.Mx * grid.My
    mask, h = np.zeros(n2), np.zeros(n2)
    _check(lib().or_update_mask_surface(C.byref(grid), C.byref(params), _ptr(bed), _ptr(H),
                                        _ptr(mask), _ptr(h)))
    return mask, h


def sia_update(grid, params, H, b, h, mask, E, full=True, age=None, ub=None, vb=None, bmr=None,
               no_model_mask=None, usurfstore=None):
    """One SIAFD::update (+ w when full). Arrays are flat float64 in PISM layout.

    Returns a dict of all outputs and intermediates.
    """
    Mx, My, Mz = grid.Mx, grid.My, grid.Mz
    n2, n3 = Mx * My, Mx * My * Mz
    keep = {}

    def inp(name, a):
        if a is None:
            return None
        a = np.ascontiguousarray(a, dtype=np.float64).ravel()
        keep[name] = a
        return a.ctypes.data

    out = {k: np.zeros(n2) for k in ("topgsmooth", "maxtl", "C2", "C3", "C4", "theta", "thksmooth")}
    out.update({k: np.zeros(2 * n2) for k in ("h_x", "h_y", "flux", "D_stag")})
    n3w = n3 if full else Mz  # oracle touches delta only when full
    out.update({k: np.zeros(n3w) for k in ("delta0", "delta1", "I0", "I1", "u", "v", "w", "Sigma")})
    f = OrFields()
    f.H, f.b, f.h, f.mask, f.E = inp("H", H), inp("b", b), inp("h", h), inp("mask", mask), inp("E", E)
    f.age, f.ub, f.vb, f.bmr = inp("age", age), inp("ub", ub), inp("vb", vb), inp("bmr", bmr)
    f.no_model_mask, f.usurfstore = inp("no_model_mask", no_model_mask), inp("usurfstore", usurfstore)
    for k in ("topgsmooth", "maxtl", "C2", "C3", "C4", "theta", "thksmooth", "h_x", "h_y",
              "flux", "D_stag", "delta0", "delta1", "I0", "I1", "u", "v", "w"):
        setattr(f, k, out[k].ctypes.data)
    f.Sigma = out["Sigma"].ctypes.data if full else None
    _check(lib().or_sia_update(C.byref(grid), C.byref(params), C.byref(f), 1 if full else 0))
    out["D_max"] = f.D_max
    out["Nx"], out["Ny"] = f.Nx, f.Ny
    if not full:
        for k in ("delta0", "delta1", "I0", "I1", "u", "v", "w", "Sigma"):
            del out[k]
    return out


class OrSliding(C.Structure):
    _fields_ = [("mode", C.c_int), ("mu_sliding", C.c_double), ("minimum_temperature_for_sliding", C.c_double)]


SLIDING_MODES = {"pressure_melting": 0, "test_E": 1, "eismint2_G": 2, "eismint2_H": 3}


def sia_sliding(grid, params, mode, x, y, H, h, b, mask, E, mu_sliding=0.0, min_T=273.0):
    """SIA_Sliding::update: (u_b, v_b, basal frictional heating, h_x, h_y)."""
    n2 = grid.Mx * grid.My
    arr = [np.ascontiguousarray(a, dtype=np.float64).ravel() for a in (x, y, H, h, b, mask, E)]
    hx, hy = np.zeros(2 * n2), np.zeros(2 * n2)
    ub, vb, heat = np.zeros(n2), np.zeros(n2), np.zeros(n2)
    sp = OrSliding(SLIDING_MODES[mode], mu_sliding, min_T)
    _check(lib().or_sia_sliding(C.byref(grid), C.byref(params), C.byref(sp), *[_ptr(a) for a in arr],
                                _ptr(hx), _ptr(hy), _ptr(ub), _ptr(vb), _ptr(heat)))
    return ub, vb, heat, hx, hy


def hardness(params, E, p):
    """FlowLaw::hardness_parameter of the selected flow law."""
    return lib().or_hardness(C.byref(params), E, p)


def strain_heating(grid, params, H, mask, E, u, v):
    """StressBalance::compute_volumetric_strain_heating from given u, v."""
    arr = [np.ascontiguousarray(a, dtype=np.float64).ravel() for a in (H, mask, E, u, v)]
    out = np.zeros(grid.Mx * grid.My * grid.Mz)
    _check(lib().or_strain_heating(C.byref(grid), C.byref(params), *[_ptr(a) for a in arr], _ptr(out)))
    return out


def mass_continuity(grid, params, dt, H, flux, mask, smb=None, bmr=None, ub=None, vb=None):
    n2 = grid.Mx * grid.My
    Hn = np.zeros(n2)
    acc = np.zeros(8)
    a = [np.ascontiguousarray(x, dtype=np.float64).ravel() if x is not None else None
         for x in (H, flux, ub, vb, mask, smb, bmr)]
    _check(lib().or_mass_continuity(C.byref(grid), C.byref(params), dt, *[_ptr(x) for x in a],
                                    _ptr(Hn), _ptr(acc)))
    return Hn, acc


def max_timestep(grid, params, D_max, time_to_end=0.0, ub=None, vb=None, mask=None):
    return lib().or_max_timestep(C.byref(grid), C.byref(params), D_max, time_to_end,
                                 _ptr(ub), _ptr(vb), _ptr(mask))


def step(grid, params, state, E, full=True, smb=None, bmr=None, time_to_end=0.0, ub=None, vb=None):
    """One SIA-only IceModel::step: update, dt, mass continuity, mask/surface.

    ``state`` is a dict with H, b, h, mask (flat arrays), updated in place.
    ub, vb: the sliding (SSA) velocity given to the update, the 2-D CFL and
    the upwind flux of mass continuity. Returns (dt, sia_outputs, accounting).
    """
    out = sia_update(grid, params, state["H"], state["b"], state["h"], state["mask"], E,
                     full=full, bmr=bmr, ub=ub, vb=vb)
    dt = max_timestep(grid, params, out["D_max"], time_to_end, ub=ub, vb=vb, mask=state["mask"])
    Hn, acc = mass_continuity(grid, params, dt, state["H"], out["flux"], state["mask"],
                              smb=smb, bmr=bmr, ub=ub, vb=vb)
    mask, h = update_mask_surface(grid, params, state["b"], Hn)
    state["H"], state["mask"], state["h"] = Hn, mask, h
    return dt, out, acc
