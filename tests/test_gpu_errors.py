"""GPU tests of the device-side error reporting (the reference's
RuntimeError cases, IceGrid.cc:381-389 and PISMBedSmoother.cc:312-314,
395-398) and of the surrogate's lifetime rule, through the C ABI.

* Flags raised by an update that is not checked at once survive later
  updates and steps and are reported by the next sia_check_errors, which
  clears them (the reference throws at the first one, SIAFD.cc:654).
* A staggered column thicker than Lz reports SIA_ERR_THICKNESS_ABOVE_LZ
  whether or not a learned surrogate is bound, also where the surface is
  flat (kBelowHeight is called for any thk != 0, SIAFD.cc:654).
* A surrogate cannot be destroyed while a handle is bound to it.
"""
import numpy as np
import pytest

import sia_types as T
import workloads as W
from parity_util import say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def S():
    import pism_sia
    import torch
    assert torch.cuda.is_available()
    return pism_sia


def thick_dome():
    """The dome on 21 x 21 x 11 levels to 4000 m, with one interior column
    (and its surface) 600 m above the top of the grid."""
    w = W.dome(Mx=21, My=21, Mz=11)
    f = {k: v.copy() for k, v in w.fields.items()}
    q = 10 * 21 + 10
    f["H"][q] = 4600.0
    f["h"][q] = 4600.0
    return w, f


def test_flags_accumulate_across_unchecked_updates(S):
    say("[gpu] error flags of an unchecked update survive the next updates")
    w, thick = thick_dome()
    ok = w.fields
    dev_ok = {k: S.to_device(v) for k, v in ok.items()}
    dev_bad = {k: S.to_device(v) for k, v in thick.items()}
    sia = S.SIAFD(w.grid, w.params)
    sia.set_sync_check(False)
    d = dev_bad
    sia.update(d["H"], d["b"], d["h"], d["mask"], d["E"], fast=True)
    for _ in range(2):
        d = dev_ok
        sia.update(d["H"], d["b"], d["h"], d["mask"], d["E"], fast=False)
    with pytest.raises(S.SIAError) as e:
        sia.check_errors()
    assert e.value.code == T.SIA_ERR_THICKNESS_ABOVE_LZ
    sia.check_errors()  # cleared by the check that reported them
    sia.close()


def test_flags_accumulate_across_device_steps(S):
    say("[gpu] error flags of the first of three sia_step calls")
    import torch
    w, thick = thick_dome()
    dev = {k: S.to_device(v) for k, v in thick.items()}
    sia = S.SIAFD(w.grid, w.params)
    sia.set_sync_check(False)
    geom = S.geometry(dev["H"], dev["h"], dev["mask"], dev["b"])
    t = torch.zeros(1, dtype=torch.float64, device="cuda")
    dt = torch.zeros(1, dtype=torch.float64, device="cuda")
    sia.step(geom, dev["E"], fast=True, t_end=0.0, t_dev=t, dt_dev=dt)
    # the thick column is gone before the next two steps
    torch.cuda.synchronize()
    for k in ("H", "h", "mask"):
        dev[k].copy_(S.to_device(w.fields[k]))
    for _ in range(2):
        sia.step(geom, dev["E"], fast=True, t_end=0.0, t_dev=t, dt_dev=dt)
    with pytest.raises(S.SIAError) as e:
        sia.check_errors()
    assert e.value.code == T.SIA_ERR_THICKNESS_ABOVE_LZ
    sia.close()


def _tiny_surrogate(S, w):
    """A surrogate with a normalization (a few training steps on w): enough
    to bind it; its accuracy does not matter here."""
    dev = {k: S.to_device(v) for k, v in w.fields.items()}
    sg = S.Surrogate(seed=3, capacity=200_000)
    c = S.SIAFD(w.grid, w.params)
    sg.collect(c, dev["H"], dev["b"], dev["h"], dev["mask"], dev["E"], stride=1)
    c.close()
    sg.fit_normalization()
    sg.train(20, batch=256)
    return sg


def test_flat_column_above_Lz_with_surrogate(S):
    """A 6 x 6 plateau with a flat surface at 4100 m in a 401-level window to
    4000 m, no bed smoothing: ice 3990 m thick on its rim and 4100 m in its
    2 x 2 centre. Only the faces at the centre are thicker than Lz, and their
    surface slope is exactly 0 (so the surrogate's zero-slope shortcut would
    skip them): the surrogate-bound fast update reports the thickness error
    as the FP64 update does."""
    say("[gpu] flat column above Lz, FP64 and surrogate")
    w = W.greenland_like(700, 740, 1180, 1220, bed_smoother_range=0.0)
    sg = _tiny_surrogate(S, w)
    Mx, My = w.grid.Mx, w.grid.My
    f = {k: v.copy() for k, v in w.fields.items()}
    H, b, h = (f[k].reshape(Mx, My) for k in ("H", "b", "h"))
    H[15:21, 15:21] = 3990.0
    H[17:19, 17:19] = 4100.0
    h[15:21, 15:21] = 4100.0
    b[15:21, 15:21] = h[15:21, 15:21] - H[15:21, 15:21]
    dev = {k: S.to_device(v) for k, v in f.items()}
    for bound in (False, True):
        sia = S.SIAFD(w.grid, w.params)
        if bound:
            sia.use_surrogate(sg)
        with pytest.raises(S.SIAError) as e:
            sia.update(dev["H"], dev["b"], dev["h"], dev["mask"], dev["E"], fast=True)
        assert e.value.code == T.SIA_ERR_THICKNESS_ABOVE_LZ, bound
        hx, hy = sia.surface_gradient()
        hx, hy = hx.reshape(Mx, My, 2), hy.reshape(Mx, My, 2)
        assert np.all(hx[16:19, 16:19] == 0.0) and np.all(hy[16:19, 16:19] == 0.0)
        sia.close()
    sg.close()


def test_surrogate_destroy_refused_while_bound(S):
    say("[gpu] surrogate lifetime")
    w = W.greenland_like(700, 732, 1180, 1212)
    sg = _tiny_surrogate(S, w)
    a = S.SIAFD(w.grid, w.params)
    b = S.SIAFD(w.grid, w.params)
    a.use_surrogate(sg)
    b.use_surrogate(sg)
    with pytest.raises(S.SIAError) as e:
        sg.close()
    assert e.value.code == T.SIA_ERR_STATE
    a.use_surrogate(None)
    with pytest.raises(S.SIAError):
        sg.close()
    b.close()  # destroying a bound handle unbinds it
    sg.close()
    assert sg.h is None
    a.close()
