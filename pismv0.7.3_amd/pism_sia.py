"""Python host side of the HIP SIA library (ctypes over include/pism_sia.h).

``SIAFD`` mirrors PISM's SSB_Modifier/SIAFD interface
(src/base/stressbalance/SSB_Modifier.hh:37-67, sia/SIAFD.hh:51-116):
``update(fast)``, ``diffusive_flux()``, ``max_diffusivity()``,
``velocity_u()/velocity_v()`` plus ``velocity_w()`` (StressBalance). Device
memory and streams come from PyTorch (ROCm build); every computation runs in
lib/libpism_sia.so. There is no CPU fallback: importing this module on a
machine without the built library raises.
"""
import ctypes as C
import os

import numpy as np

import sia_types as T

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libpism_sia.so")

STATUS = {0: "SIA_OK", 1: "SIA_ERR_ARGUMENT", 2: "SIA_ERR_HIP", 3: "SIA_ERR_NEGATIVE_THICKNESS",
          4: "SIA_ERR_THICKNESS_ABOVE_LZ", 5: "SIA_ERR_OMEGA_NEGATIVE", 6: "SIA_ERR_CONFIG",
          7: "SIA_ERR_OUT_OF_MEMORY", 8: "SIA_ERR_STATE"}

_vp = C.c_void_p
_PP = C.POINTER(C.c_void_p)


class SiaWorkloadSpec(C.Structure):
    _fields_ = [("Mx_full", C.c_int32), ("My_full", C.c_int32), ("i0", C.c_int32), ("j0", C.c_int32),
                ("n_bumps", C.c_int32), ("reserved", C.c_int32), ("seed", C.c_uint64),
                ("dx", C.c_double), ("Hmax", C.c_double)]


# exported function -> (restype, argtypes); the ABI declared in include/*.h
SIGNATURES = {
    "sia_params_default": (None, [C.POINTER(T.SiaParams)]),
    "sia_create": (C.c_int, [C.POINTER(T.SiaGrid), C.POINTER(T.SiaParams), _PP]),
    "sia_destroy": (C.c_int, [_vp]),
    "sia_set_stream": (C.c_int, [_vp, _vp]),
    "sia_set_sync_check": (C.c_int, [_vp, C.c_int]),
    "sia_check_errors": (C.c_int, [_vp]),
    "sia_update": (C.c_int, [_vp, C.POINTER(T.SiaInputs), C.c_int]),
    "sia_diffusive_flux": (C.c_int, [_vp, _PP]),
    "sia_max_diffusivity": (C.c_int, [_vp, C.POINTER(C.c_double)]),
    "sia_max_diffusivity_device": (C.c_int, [_vp, _PP]),
    "sia_velocity_u": (C.c_int, [_vp, _PP]),
    "sia_velocity_v": (C.c_int, [_vp, _PP]),
    "sia_velocity_w": (C.c_int, [_vp, _PP]),
    "sia_volumetric_strain_heating": (C.c_int, [_vp, _PP]),
    "sia_sliding_update": (C.c_int, [_vp, C.POINTER(T.SiaInputs), C.c_void_p, _vp, _vp, _vp]),
    "sia_diffusivity": (C.c_int, [_vp, _vp]),
    "sia_surface_values": (C.c_int, [_vp, _vp, _vp, _vp]),
    "sia_surface_gradient": (C.c_int, [_vp, _PP, _PP]),
    "sia_diffusivity_staggered": (C.c_int, [_vp, _PP]),
    "sia_theta_thksmooth": (C.c_int, [_vp, _PP, _PP]),
    "sia_bed_smoother_fields": (C.c_int, [_vp, _PP, _PP, _PP, _PP, _PP, C.POINTER(C.c_int32),
                                          C.POINTER(C.c_int32)]),
    "sia_max_timestep": (C.c_int, [_vp, C.c_double, C.POINTER(C.c_double)]),
    "sia_mass_continuity_step": (C.c_int, [_vp, C.POINTER(T.SiaGeometry), C.c_double,
                                           C.POINTER(T.SiaMassAccounting)]),
    "sia_step": (C.c_int, [_vp, C.POINTER(T.SiaGeometry), _vp, C.c_int, C.c_double, _vp, _vp]),
    "sia_last_error": (C.c_char_p, [_vp]),
    "sia_flux_update_noalloc": (C.c_int, [C.POINTER(T.SiaGrid), C.POINTER(T.SiaParams), C.POINTER(T.SiaInputs),
                                          _vp, _vp, _vp, _vp]),
    "sia_profiling": (C.c_int, [_vp, C.c_int]),
    "sia_profiling_read": (C.c_int, [_vp, C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
    "sia_event_name": (C.c_char_p, [C.c_int]),
    "sia_workload_generate": (C.c_int, [C.POINTER(SiaWorkloadSpec), C.POINTER(T.SiaGrid),
                                        C.POINTER(T.SiaParams), _vp, _vp, _vp, _vp, _vp, _vp]),
    # learned surrogate (include/sia_surrogate.h)
    "sia_surrogate_create": (C.c_int, [C.c_uint64, C.c_int64, _PP]),
    "sia_surrogate_destroy": (C.c_int, [_vp]),
    "sia_surrogate_set_stream": (C.c_int, [_vp, _vp]),
    "sia_surrogate_collect": (C.c_int, [_vp, _vp, C.POINTER(T.SiaInputs), C.c_int, C.POINTER(C.c_int64)]),
    "sia_surrogate_fit_normalization": (C.c_int, [_vp]),
    "sia_surrogate_train": (C.c_int, [_vp, C.c_int64, C.c_int, C.c_double, C.c_double, C.c_uint64,
                                      C.POINTER(C.c_double)]),
    "sia_surrogate_get_params": (C.c_int, [_vp, C.POINTER(C.c_float), C.c_size_t]),
    "sia_surrogate_set_params": (C.c_int, [_vp, C.POINTER(C.c_float), C.c_size_t]),
    "sia_surrogate_grad": (C.c_int, [_vp, C.c_int, _vp, _vp, _vp, _vp, C.POINTER(C.c_double),
                                     C.POINTER(C.c_float)]),
    "sia_surrogate_save": (C.c_int, [_vp, _vp, C.POINTER(C.c_size_t)]),
    "sia_surrogate_load": (C.c_int, [_vp, _vp, C.c_size_t]),
    "sia_surrogate_samples": (C.c_int, [_vp, _PP, C.POINTER(C.c_int64)]),
    "sia_use_surrogate": (C.c_int, [_vp, _vp]),
    "sia_surrogate_counts": (C.c_int, [_vp, C.POINTER(C.c_int64)]),
    "sia_surrogate_calibrate_fallback": (C.c_int, [_vp, _vp, C.POINTER(T.SiaInputs), C.c_double,
                                                   C.POINTER(C.c_double)]),
    "sia_surrogate_set_fallback_score": (C.c_int, [_vp, C.c_double]),
    # batched ensembles (include/sia_ensemble.h)
    "sia_ensemble_create": (C.c_int, [C.POINTER(_vp), C.POINTER(T.SiaGeometry), C.POINTER(_vp),
                                      C.POINTER(C.c_double), C.c_int, _PP]),
    "sia_ensemble_destroy": (C.c_int, [_vp]),
    "sia_ensemble_set_stream": (C.c_int, [_vp, _vp]),
    "sia_ensemble_set_tile": (C.c_int, [_vp, C.c_int]),
    "sia_ensemble_autotune": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double)]),
    "sia_ensemble_step": (C.c_int, [_vp, C.c_int, C.c_int]),
    "sia_ensemble_time": (C.c_int, [_vp, C.c_int, _PP, _PP]),
    "sia_ensemble_check_errors": (C.c_int, [_vp]),
    "sia_ensemble_last_error": (C.c_char_p, [_vp]),
}

N_EVENTS = 9  # SIA_EV_COUNT
SLIDING_MODES = {"pressure_melting": 0, "test_E": 1, "eismint2_G": 2, "eismint2_H": 3}


class SiaSlidingParams(C.Structure):
    _fields_ = [("mode", C.c_int32), ("reserved", C.c_int32), ("mu_sliding", C.c_double),
                ("minimum_temperature_for_sliding", C.c_double), ("x", C.c_void_p), ("y", C.c_void_p)]
SG_NPARAM = 58144
SG_KIN = 64
SG_SAMPLE_DTYPE = np.dtype([("f", "<f4", 36), ("r", "<f4", 32), ("share", "<f4", 32), ("has", "<u4"),
                            ("col", "<i4"), ("dref", "<f8")])  # SgSample, 416 bytes

_lib = None


def lib():
    """Load lib/libpism_sia.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("HIP SIA library not built: %s (run __graft_entry__.build())" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


class SIAError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (%d): %s" % (STATUS.get(code, "?"), code, msg))
        self.code = code


def default_params():
    p = T.SiaParams()
    lib().sia_params_default(C.byref(p))
    return p


# ------------------------------------------------------------ device memory
def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("no GPU visible to PyTorch; the HIP SIA library needs one")
    return torch


def to_device(a):
    torch = _torch()
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64).ravel()).to("cuda")


def empty_device(n):
    torch = _torch()
    return torch.empty(int(n), dtype=torch.float64, device="cuda")


def ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def device_to_numpy(p, n):
    """Copy n doubles from a raw device pointer (returned by the library) to numpy."""
    torch = _torch()
    out = np.empty(int(n), dtype=np.float64)
    torch.cuda.synchronize()
    rc = _hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(p), C.c_size_t(8 * int(n)), 2)
    if rc != 0:
        raise RuntimeError("hipMemcpy failed: %d" % rc)
    return out


_hiplib = None


def _hip():
    global _hiplib
    if _hiplib is None:
        _hiplib = C.CDLL("libamdhip64.so")
        _hiplib.hipMemcpy.restype = C.c_int
        _hiplib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return _hiplib


# ------------------------------------------------------------ SIAFD mirror
class SIAFD:
    """SSB_Modifier/SIAFD over the C ABI; fields are torch CUDA float64 tensors
    (flat, PISM layout) owned by the caller (inputs) or the library (outputs)."""

    def __init__(self, grid, params, stream=None):
        self.grid = grid
        self.params = params
        h = C.c_void_p()
        self._check(lib().sia_create(C.byref(grid), C.byref(params), C.byref(h)), None)
        self.h = h
        if stream is not None:
            self._check(lib().sia_set_stream(self.h, C.c_void_p(stream)))
        self.n2 = grid.Mx * grid.My
        self.n3 = self.n2 * grid.Mz

    def _check(self, rc, h="self"):
        if rc != 0:
            hh = self.h if h == "self" else None
            msg = lib().sia_last_error(hh).decode() if hh is not None else ""
            raise SIAError(rc, msg)

    def close(self):
        if getattr(self, "h", None):
            lib().sia_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_sync_check(self, on):
        self._check(lib().sia_set_sync_check(self.h, 1 if on else 0))

    def check_errors(self):
        """Raises SIAError for the device-side error flags of every update and
        step since the last check (ParallelSection::check, error_handling.cc:153-192),
        and clears them."""
        self._check(lib().sia_check_errors(self.h))

    def update(self, H, b, h, mask, E, fast=False, age=None, ub=None, vb=None, bmr=None,
               bed_state_counter=0, no_model_mask=None, usurfstore=None):
        """SIAFD::update(vel_input, fast) (SIAFD.cc:119-148) + strain heating and
        w when not fast; SIAFD_Regional's gradient when no_model_mask and
        usurfstore are given."""
        inp = T.SiaInputs(ptr(H), ptr(b), ptr(h), ptr(mask), ptr(E), ptr(age), ptr(ub), ptr(vb),
                          ptr(bmr), bed_state_counter, ptr(no_model_mask), ptr(usurfstore))
        self._keep = (H, b, h, mask, E, age, ub, vb, bmr, no_model_mask, usurfstore)
        self._check(lib().sia_update(self.h, C.byref(inp), 1 if fast else 0))

    def _get(self, fn):
        p = C.c_void_p()
        self._check(fn(self.h, C.byref(p)))
        return p.value

    def diffusive_flux(self):
        return device_to_numpy(self._get(lib().sia_diffusive_flux), 2 * self.n2)

    def max_diffusivity(self):
        d = C.c_double()
        self._check(lib().sia_max_diffusivity(self.h, C.byref(d)))
        return d.value

    def velocity_u(self):
        return device_to_numpy(self._get(lib().sia_velocity_u), self.n3)

    def velocity_v(self):
        return device_to_numpy(self._get(lib().sia_velocity_v), self.n3)

    def velocity_w(self):
        return device_to_numpy(self._get(lib().sia_velocity_w), self.n3)

    def volumetric_strain_heating(self):
        """StressBalance::volumetric_strain_heating() of the last full update, W m-3."""
        return device_to_numpy(self._get(lib().sia_volumetric_strain_heating), self.n3)

    def diffusivity_staggered(self):
        return device_to_numpy(self._get(lib().sia_diffusivity_staggered), 2 * self.n2)

    def surface_gradient(self):
        a, b = C.c_void_p(), C.c_void_p()
        self._check(lib().sia_surface_gradient(self.h, C.byref(a), C.byref(b)))
        return device_to_numpy(a.value, 2 * self.n2), device_to_numpy(b.value, 2 * self.n2)

    def theta_thksmooth(self):
        a, b = C.c_void_p(), C.c_void_p()
        self._check(lib().sia_theta_thksmooth(self.h, C.byref(a), C.byref(b)))
        return device_to_numpy(a.value, self.n2), device_to_numpy(b.value, self.n2)

    def bed_smoother_fields(self):
        ps = [C.c_void_p() for _ in range(5)]
        Nx, Ny = C.c_int32(), C.c_int32()
        self._check(lib().sia_bed_smoother_fields(self.h, *[C.byref(p) for p in ps], C.byref(Nx), C.byref(Ny)))
        out = {k: device_to_numpy(p.value, self.n2) for k, p in zip(("topgsmooth", "maxtl", "C2", "C3", "C4"), ps)}
        out["Nx"], out["Ny"] = Nx.value, Ny.value
        return out

    def max_timestep(self, time_to_end=0.0):
        d = C.c_double()
        self._check(lib().sia_max_timestep(self.h, time_to_end, C.byref(d)))
        return d.value

    def mass_continuity_step(self, geom, dt, accounting=True):
        acc = T.SiaMassAccounting()
        self._check(lib().sia_mass_continuity_step(self.h, C.byref(geom), dt,
                                                   C.byref(acc) if accounting else None))
        return np.array([getattr(acc, n) for n, _ in T.SiaMassAccounting._fields_])

    def profiling(self, on):
        self._check(lib().sia_profiling(self.h, 1 if on else 0))

    def profiling_read(self):
        """{event name: (total ms, launches)} since the last read."""
        ms = (C.c_double * N_EVENTS)()
        cnt = (C.c_int64 * N_EVENTS)()
        self._check(lib().sia_profiling_read(self.h, ms, cnt))
        return {lib().sia_event_name(e).decode(): (ms[e], cnt[e]) for e in range(N_EVENTS)}

    def step(self, geom, E, fast=False, t_end=0.0, t_dev=None, dt_dev=None):
        self._check(lib().sia_step(self.h, C.byref(geom), ptr(E), 1 if fast else 0, t_end,
                                   ptr(t_dev), ptr(dt_dev)))

    def sliding(self, H, b, h, mask, E, mode="pressure_melting", x=None, y=None, mu_sliding=0.0,
                min_T=273.0):
        """SIA_Sliding::update (SIA_Sliding.cc:106-206): (u_b, v_b, basal
        frictional heating) as device tensors; mode one of SLIDING_MODES."""
        torch = _torch()
        n2 = self.grid.Mx * self.grid.My
        ub, vb, heat = (torch.empty(n2, dtype=torch.float64, device="cuda") for _ in range(3))
        inp = T.SiaInputs(ptr(H), ptr(b), ptr(h), ptr(mask), ptr(E), None, None, None, None, 0, None, None)
        xs = np.ascontiguousarray(x, dtype=np.float64) if x is not None else None
        ys = np.ascontiguousarray(y, dtype=np.float64) if y is not None else None
        sp = SiaSlidingParams(SLIDING_MODES[mode], 0, mu_sliding, min_T,
                              xs.ctypes.data if xs is not None else None,
                              ys.ctypes.data if ys is not None else None)
        self._check(lib().sia_sliding_update(self.h, C.byref(inp), C.byref(sp), ptr(ub), ptr(vb), ptr(heat)))
        return ub, vb, heat

    def diffusivity(self):
        """SIAFD's "diffusivity" diagnostic (regular grid), numpy [Mx*My]."""
        torch = _torch()
        out = torch.empty(self.n2, dtype=torch.float64, device="cuda")
        self._check(lib().sia_diffusivity(self.h, ptr(out)))
        return out.cpu().numpy()

    def surface_values(self, field3d, thickness):
        """IceModelVec3::getSurfaceValues of a device 3-D field (tensor or raw
        device pointer) at z = thickness."""
        torch = _torch()
        out = torch.empty(self.n2, dtype=torch.float64, device="cuda")
        fp = C.c_void_p(field3d) if isinstance(field3d, int) else ptr(field3d)
        self._check(lib().sia_surface_values(self.h, fp, ptr(thickness), ptr(out)))
        return out

    def velocity_u_device(self):
        """Device pointer of u (owned by the library, valid until the next update)."""
        return self._get(lib().sia_velocity_u)

    def velocity_v_device(self):
        return self._get(lib().sia_velocity_v)

    def use_surrogate(self, sg):
        """Fast updates use the learned surrogate `sg` (None: FP64 again)."""
        self._check(lib().sia_use_surrogate(self.h, sg.h if sg is not None else None))
        self._sg = sg

    def surrogate_counts(self):
        """(zero, surrogate, fallback) staggered columns of the last surrogate update."""
        c = (C.c_int64 * 3)()
        self._check(lib().sia_surrogate_counts(self.h, c))
        return tuple(c)


class Surrogate:
    """The learned column surrogate (include/sia_surrogate.h): training samples
    from the FP64 path, training on the GPU, save/load, gradient hook."""

    def __init__(self, seed=1, capacity=0, stream=None):
        h = C.c_void_p()
        rc = lib().sia_surrogate_create(seed, capacity, C.byref(h))
        if rc != 0:
            raise SIAError(rc, "sia_surrogate_create")
        self.h = h
        if stream is not None:
            self._check(lib().sia_surrogate_set_stream(self.h, C.c_void_p(stream)))

    def _check(self, rc, what=""):
        if rc != 0:
            raise SIAError(rc, what)

    def close(self):
        """Frees the model; SIAError(SIA_ERR_STATE) while a handle is still
        bound to it (the model is then kept)."""
        if getattr(self, "h", None):
            rc = lib().sia_surrogate_destroy(self.h)
            if rc != 0:
                raise SIAError(rc, "sia_surrogate_destroy: a handle is still bound to the model")
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def collect(self, sia, H, b, h, mask, E, stride=1, bed_state_counter=0):
        """Appends the samples of the FP64 fast update of these fields on handle
        sia. As in update(), the bed is smoothed again only when
        bed_state_counter exceeds the last one the handle saw: a handle reused
        for another bed needs a larger counter."""
        inp = T.SiaInputs(ptr(H), ptr(b), ptr(h), ptr(mask), ptr(E), None, None, None, None, bed_state_counter)
        n = C.c_int64()
        rc = lib().sia_surrogate_collect(self.h, sia.h, C.byref(inp), stride, C.byref(n))
        if rc != 0:
            raise SIAError(rc, lib().sia_last_error(sia.h).decode())
        return n.value

    def calibrate_fallback(self, sia, H, b, h, mask, E, fraction, bed_state_counter=0):
        """Sets the score fallback threshold tau from one surrogate update of
        these fields (training fields, not the ones it will be evaluated on): about
        `fraction` of the in-range columns then fall back to FP64. Returns tau."""
        inp = T.SiaInputs(ptr(H), ptr(b), ptr(h), ptr(mask), ptr(E), None, None, None, None, bed_state_counter)
        tau = C.c_double()
        rc = lib().sia_surrogate_calibrate_fallback(self.h, sia.h, C.byref(inp), float(fraction), C.byref(tau))
        if rc != 0:
            raise SIAError(rc, lib().sia_last_error(sia.h).decode())
        return tau.value

    def set_fallback_score(self, tau):
        """Score threshold of the FP64 fallback (float("inf"): none)."""
        self._check(lib().sia_surrogate_set_fallback_score(self.h, float(tau)), "set_fallback_score")

    def fit_normalization(self):
        self._check(lib().sia_surrogate_fit_normalization(self.h), "fit_normalization")

    def train(self, steps, batch=1024, lr_start=2e-3, lr_end=1e-5, seed=7):
        loss = np.zeros(int(steps))
        self._check(lib().sia_surrogate_train(self.h, int(steps), int(batch), lr_start, lr_end, seed,
                                              loss.ctypes.data_as(C.POINTER(C.c_double))), "train")
        return loss

    def get_params(self):
        w = np.zeros(SG_NPARAM, dtype=np.float32)
        self._check(lib().sia_surrogate_get_params(self.h, w.ctypes.data_as(C.POINTER(C.c_float)), SG_NPARAM))
        return w

    def set_params(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        assert w.size == SG_NPARAM
        self._check(lib().sia_surrogate_set_params(self.h, w.ctypes.data_as(C.POINTER(C.c_float)), SG_NPARAM))

    def grad(self, x, r, share, has):
        """Loss and FP32 gradient of one batch; arrays are host float32:
        x [B][64] normalized inputs, r/share/has [B][32] (targets in output units)."""
        torch = _torch()
        B = x.shape[0]
        dev = [torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to("cuda") for a in (x, r, share, has)]
        loss = C.c_double()
        g = np.zeros(SG_NPARAM, dtype=np.float32)
        torch.cuda.synchronize()
        self._check(lib().sia_surrogate_grad(self.h, B, *[ptr(t) for t in dev], C.byref(loss),
                                             g.ctypes.data_as(C.POINTER(C.c_float))), "grad")
        return loss.value, g

    def save(self):
        n = C.c_size_t()
        self._check(lib().sia_surrogate_save(self.h, None, C.byref(n)), "save")
        buf = (C.c_char * n.value)()
        self._check(lib().sia_surrogate_save(self.h, buf, C.byref(n)), "save")
        return bytes(buf)

    def load(self, blob):
        buf = (C.c_char * len(blob)).from_buffer_copy(blob)
        self._check(lib().sia_surrogate_load(self.h, buf, len(blob)), "load")

    def samples(self):
        """The training samples as a numpy structured array (copied to the host)."""
        p, n = C.c_void_p(), C.c_int64()
        self._check(lib().sia_surrogate_samples(self.h, C.byref(p), C.byref(n)), "samples")
        out = np.zeros(n.value, dtype=SG_SAMPLE_DTYPE)
        if n.value:
            _torch().cuda.synchronize()
            rc = _hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(p.value),
                                  C.c_size_t(out.nbytes), 2)
            if rc != 0:
                raise RuntimeError("hipMemcpy failed: %d" % rc)
        return out


class Ensemble:
    """Many SIAFD members advanced together (include/sia_ensemble.h): one kernel
    launch per stage for all members, optional HIP-graph capture of multi-step
    runs. members: SIAFD objects whose bed smoother is ready; geoms: their
    geometry() structs (kept alive here); Es: their enthalpy tensors."""

    def __init__(self, members, geoms, Es, t_end=None, stream=None):
        n = len(members)
        self.members, self.geoms, self.Es = list(members), list(geoms), list(Es)
        hs = (_vp * n)(*[m.h.value for m in members])
        gs = (T.SiaGeometry * n)(*geoms)
        es = (_vp * n)(*[E.data_ptr() for E in Es])
        te = (C.c_double * n)(*t_end) if t_end is not None else None
        h = C.c_void_p()
        rc = lib().sia_ensemble_create(hs, gs, es, te, n, C.byref(h))
        if rc != 0:
            raise SIAError(rc, "sia_ensemble_create")
        self.h = h
        if stream is not None:
            self._check(lib().sia_ensemble_set_stream(self.h, C.c_void_p(stream)))

    def _check(self, rc):
        if rc != 0:
            raise SIAError(rc, lib().sia_ensemble_last_error(self.h).decode())

    def set_tile(self, cols_per_wave):
        self._check(lib().sia_ensemble_set_tile(self.h, cols_per_wave))

    def autotune(self, steps=3):
        """Times each tile width, keeps the fastest; returns (chosen, {width: ms per step})."""
        chosen = C.c_int()
        ms = (C.c_double * 3)()
        self._check(lib().sia_ensemble_autotune(self.h, steps, C.byref(chosen), ms))
        return chosen.value, {8: ms[0], 16: ms[1], 32: ms[2]}

    def step(self, n_steps=1, graph=False):
        self._check(lib().sia_ensemble_step(self.h, n_steps, 1 if graph else 0))

    def time(self, member):
        """(t, dt) of a member in seconds (synchronises)."""
        t, dt = C.c_void_p(), C.c_void_p()
        self._check(lib().sia_ensemble_time(self.h, member, C.byref(t), C.byref(dt)))
        return device_to_numpy(t.value, 1)[0], device_to_numpy(dt.value, 1)[0]

    def check_errors(self):
        self._check(lib().sia_ensemble_check_errors(self.h))

    def close(self):
        if getattr(self, "h", None):
            lib().sia_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def geometry(H, h, mask, b, smb=None, bmr=None):
    return T.SiaGeometry(ptr(H), ptr(h), ptr(mask), ptr(b), ptr(smb), ptr(bmr))


def generate_greenland_like(grid, params, i0=0, j0=0, Mx_full=1501, My_full=2801, seed=20261203,
                            dx=1000.0, Hmax=3200.0, n_bumps=200, with_E=True):
    """Device generator of the polythermal benchmark sheet (include/sia_workload.h)."""
    torch = _torch()
    n2 = grid.Mx * grid.My
    H, b, h, mask = (torch.empty(n2, dtype=torch.float64, device="cuda") for _ in range(4))
    E = torch.empty(n2 * grid.Mz, dtype=torch.float64, device="cuda") if with_E else None
    spec = SiaWorkloadSpec(Mx_full, My_full, i0, j0, n_bumps, 0, seed, dx, Hmax)
    torch.cuda.synchronize()
    rc = lib().sia_workload_generate(C.byref(spec), C.byref(grid), C.byref(params), ptr(H), ptr(b),
                                     ptr(h), ptr(mask), ptr(E), None)
    if rc != 0:
        raise SIAError(rc, "sia_workload_generate")
    return dict(H=H, b=b, h=h, mask=mask, E=E)


# largest grid (Mx * My) of sia_flux_update_noalloc (include/pism_sia.h)
NOALLOC_MAX_COLUMNS = 524288


def flux_update_noalloc(grid, params, H, b, h, mask, E, flux, D_max, err_flags=None, age=None,
                        no_model_mask=None, usurfstore=None, stream=None):
    """sia_flux_update_noalloc: the flux-only update without a handle, written
    into the caller's device tensors flux [Mx*My*2], D_max [1] (float64) and
    err_flags [1] (int32, optional). Returns the status code (0 = OK)."""
    inp = T.SiaInputs(ptr(H), ptr(b), ptr(h), ptr(mask), ptr(E), ptr(age), None, None, None, 0,
                      ptr(no_model_mask), ptr(usurfstore))
    return lib().sia_flux_update_noalloc(C.byref(grid), C.byref(params), C.byref(inp), ptr(flux), ptr(D_max),
                                         ptr(err_flags), C.c_void_p(stream) if stream else None)
