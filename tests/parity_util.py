"""Helpers shared by the GPU parity tests: run one workload through the HIP
library and through the CPU oracle, and measure disagreement with the
tolerances the tests state."""
import sys

import numpy as np


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run_gpu(w, full=True, ub=None, vb=None, bmr=None, age=None, no_model_mask=None, usurfstore=None):
    import pism_sia as S
    f = w.fields
    dev = {k: S.to_device(v) for k, v in f.items()}
    extra = {k: (S.to_device(v) if v is not None else None) for k, v in
             (("ub", ub), ("vb", vb), ("bmr", bmr), ("age", age), ("no_model_mask", no_model_mask),
              ("usurfstore", usurfstore))}
    sia = S.SIAFD(w.grid, w.params)
    sia.update(dev["H"], dev["b"], dev["h"], dev["mask"], dev["E"], fast=not full, **extra)
    out = dict(flux=sia.diffusive_flux(), D_max=sia.max_diffusivity(), D_stag=sia.diffusivity_staggered())
    out["h_x"], out["h_y"] = sia.surface_gradient()
    out["theta"], out["thksmooth"] = sia.theta_thksmooth()
    out.update({k: v for k, v in sia.bed_smoother_fields().items()})
    if full:
        out["u"], out["v"], out["w"] = sia.velocity_u(), sia.velocity_v(), sia.velocity_w()
        out["Sigma"] = sia.volumetric_strain_heating()
        out["_heat_inputs"] = (w.grid, w.params, f["H"], f["mask"], f["E"])
    sia.close()
    return out


def run_oracle(oracle, w, full=True, ub=None, vb=None, bmr=None, age=None, no_model_mask=None, usurfstore=None):
    f = w.fields
    return oracle.sia_update(w.grid, w.params, f["H"], f["b"], f["h"], f["mask"], f["E"], full=full,
                             ub=ub, vb=vb, bmr=bmr, age=age, no_model_mask=no_model_mask,
                             usurfstore=usurfstore)


def rel_elementwise(a, b, floor=1e-300):
    """max |a - b| / |b| over entries with |b| > floor (entries with |b| <= floor
    must agree to floor absolutely)."""
    a, b = np.asarray(a), np.asarray(b)
    big = np.abs(b) > floor
    r = 0.0
    if big.any():
        r = float(np.max(np.abs(a[big] - b[big]) / np.abs(b[big])))
    small = ~big
    if small.any():
        assert np.max(np.abs(a[small] - b[small])) <= floor
    return r


def uv_scale(grid, o, ub=None, vb=None):
    """Size of the summands of u, v (SIAFD.cc:949-956): the natural scale of
    their rounding error, since the four terms can cancel."""
    Mx, My, Mz = grid.Mx, grid.My, grid.Mz
    I0 = o["I0"].reshape(Mx, My, Mz)
    I1 = o["I1"].reshape(Mx, My, Mz)
    hx = o["h_x"].reshape(Mx, My, 2)
    hy = o["h_y"].reshape(Mx, My, 2)
    Iw = np.roll(I0, 1, axis=0)
    Is = np.roll(I1, 1, axis=1)
    hxw = np.roll(hx[:, :, 0], 1, axis=0)[..., None]
    hxs = np.roll(hx[:, :, 1], 1, axis=1)[..., None]
    hyw = np.roll(hy[:, :, 0], 1, axis=0)[..., None]
    hys = np.roll(hy[:, :, 1], 1, axis=1)[..., None]
    su = 0.25 * (np.abs(I0 * hx[:, :, 0:1]) + np.abs(Iw * hxw) + np.abs(I1 * hx[:, :, 1:2]) + np.abs(Is * hxs))
    sv = 0.25 * (np.abs(I0 * hy[:, :, 0:1]) + np.abs(Iw * hyw) + np.abs(I1 * hy[:, :, 1:2]) + np.abs(Is * hys))
    if ub is not None:
        su = su + np.abs(ub.reshape(Mx, My, 1))
        sv = sv + np.abs(vb.reshape(Mx, My, 1))
    return su.ravel(), sv.ravel()


def w_scale(grid, o, bmr=None):
    """Rounding scale of w: |w_0| + cumulative trapezoid of an upper bound of
    |u_x| + |v_y| built from |u|, |v| of the column and its neighbours."""
    Mx, My, Mz = grid.Mx, grid.My, grid.Mz
    u = np.abs(o["u"].reshape(Mx, My, Mz))
    v = np.abs(o["v"].reshape(Mx, My, Mz))
    A = (2 * u + np.roll(u, 1, 0) + np.roll(u, -1, 0)) / grid.dx + \
        (2 * v + np.roll(v, 1, 1) + np.roll(v, -1, 1)) / grid.dy
    z = np.array([grid.z[k] for k in range(Mz)])
    dz = np.diff(z)
    inc = 0.5 * dz[None, None, :] * (A[:, :, 1:] + A[:, :, :-1])
    S = np.concatenate([np.zeros((Mx, My, 1)), np.cumsum(inc, axis=2)], axis=2)
    if bmr is not None:
        S = S + np.abs(bmr.reshape(Mx, My, 1))
    return S.ravel()


def check_against_oracle(g, o, grid, full=True, tol=1e-12, ub=None, vb=None, bmr=None, label=""):
    """Assert the HIP outputs g agree with the oracle outputs o; returns the errors."""
    errs = {}
    # 2-D fields: same formulas in the same order; only libm ulps may differ
    for k in ("topgsmooth", "maxtl", "C2", "C3", "C4", "theta", "thksmooth", "h_x", "h_y"):
        errs[k] = rel_elementwise(g[k], o[k], floor=1e-300)
        assert errs[k] <= tol, (label, k, errs[k])
    assert g["Nx"] == o["Nx"] and g["Ny"] == o["Ny"]
    # D, flux, D_max: relative to each value (all terms of D are >= 0)
    for k in ("D_stag", "flux"):
        errs[k] = rel_elementwise(g[k], o[k], floor=1e-300)
        assert errs[k] <= tol, (label, k, errs[k])
    errs["D_max"] = abs(g["D_max"] - o["D_max"]) / max(o["D_max"], 1e-300)
    assert errs["D_max"] <= tol, (label, "D_max", errs["D_max"])
    if full:
        su, sv = uv_scale(grid, o, ub, vb)
        for k, s in (("u", su), ("v", sv)):
            d = np.abs(g[k] - o[k])
            errs[k] = float(np.max(d / np.maximum(s, 1e-300)))
            assert np.all(d <= tol * s + 1e-300), (label, k, errs[k])
        sw = w_scale(grid, o, bmr)
        d = np.abs(g["w"] - o["w"])
        errs["w"] = float(np.max(d / np.maximum(sw, 1e-300)))
        assert np.all(d <= tol * sw + 1e-300), (label, "w", errs["w"])
        if "Sigma" in g:
            # strain heating of the library's own u, v against the oracle's
            # restatement on the same u, v: only libm ulps may differ
            import binding as oracle
            gr, P, H, mask, E = g["_heat_inputs"]
            ref = oracle.strain_heating(gr, P, H, mask, E, g["u"], g["v"])
            errs["Sigma"] = rel_elementwise(g["Sigma"], ref, floor=1e-300)
            assert errs["Sigma"] <= tol, (label, "Sigma", errs["Sigma"])
            assert np.all(g["Sigma"][ref == 0.0] == 0.0)
    return errs
