"""GPU test of sia_flux_update_noalloc (include/pism_sia.h), the flux-only
update that allocates nothing: through the C ABI, its flux, D_max and error
flags must be bitwise those of sia_update(fast = 1) on a handle with the same
grid and parameters, on two calls in a row with the second input smaller
than the first (the scratch array keeps the first call's contents), and for
the regional gradient override and the error flags; grids beyond the
scratch array's size are refused."""
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


def handle_fast(S, w, dev, **kw):
    sia = S.SIAFD(w.grid, w.params)
    sia.set_sync_check(False)
    sia.update(dev["H"], dev["b"], dev["h"], dev["mask"], dev["E"], fast=True, **kw)
    flux = sia.diffusive_flux()
    try:
        sia.check_errors()
        code = 0
    except S.SIAError as e:
        code = e.code
    import ctypes as C
    p = C.c_void_p()
    assert S.lib().sia_max_diffusivity_device(sia.h, C.byref(p)) == 0
    D_max = S.device_to_numpy(p.value, 1)[0]
    sia.close()
    return flux, D_max, code


def noalloc_fast(S, w, dev, **kw):
    import torch
    n2 = w.grid.Mx * w.grid.My
    flux = torch.full((2 * n2,), np.nan, dtype=torch.float64, device="cuda")
    D_max = torch.full((1,), np.nan, dtype=torch.float64, device="cuda")
    err = torch.full((1,), -1, dtype=torch.int32, device="cuda")
    rc = S.flux_update_noalloc(w.grid, w.params, dev["H"], dev["b"], dev["h"], dev["mask"], dev["E"], flux, D_max,
                               err, **kw)
    assert rc == 0, rc
    torch.cuda.synchronize()
    return flux.cpu().numpy(), D_max.item(), int(err.item())


def same(a, b):
    return np.array_equal(a.view(np.uint64), b.view(np.uint64))


@pytest.mark.parametrize("win", [(640, 736, 1150, 1262, 401), (150, 190, 1380, 1428, 201)])
def test_noalloc_equals_sia_update(S, win):
    """Called for the larger window first, then the smaller one."""
    i0, i1, j0, j1, Mz = win
    say("[gpu] no-alloc flux update, window %dx%dx%d" % (i1 - i0, j1 - j0, Mz))
    w = W.greenland_like(i0, i1, j0, j1, Mz=Mz)
    dev = {k: S.to_device(v) for k, v in w.fields.items()}
    f0, d0, e0 = handle_fast(S, w, dev)
    f1, d1, e1 = noalloc_fast(S, w, dev)
    assert e0 == 0 and e1 == 0
    assert same(f1, f0) and d1 == d0 and d0 > 0
    # twice in a row on the same inputs: the same bits
    f2, d2, _ = noalloc_fast(S, w, dev)
    assert same(f2, f0) and d2 == d0


def test_noalloc_regional_and_no_smoothing(S):
    say("[gpu] no-alloc flux update: regional override, bed_smoother_range = 0")
    w = W.greenland_like(700, 740, 1180, 1224, bed_smoother_range=0.0)
    Mx, My = w.grid.Mx, w.grid.My
    nmm = np.zeros((Mx, My))
    nmm[:5, :] = 1.0
    hst = w.fields["h"].reshape(Mx, My) + np.random.default_rng(4).uniform(-20.0, 20.0, (Mx, My))
    dev = {k: S.to_device(v) for k, v in w.fields.items()}
    kw = dict(no_model_mask=S.to_device(nmm), usurfstore=S.to_device(hst))
    f0, d0, _ = handle_fast(S, w, dev, **kw)
    f1, d1, _ = noalloc_fast(S, w, dev, **kw)
    assert same(f1, f0) and d1 == d0


def test_noalloc_error_flags_and_size_limit(S):
    say("[gpu] no-alloc flux update: error flags, size limit")
    w = W.dome(Mx=21, My=21, Mz=11)
    f = {k: v.copy() for k, v in w.fields.items()}
    f["H"][10 * 21 + 10] = 4600.0  # above Lz = 4000 m
    f["h"][10 * 21 + 10] = 4600.0
    dev = {k: S.to_device(v) for k, v in f.items()}
    f0, d0, e0 = handle_fast(S, w, dev)
    f1, d1, e1 = noalloc_fast(S, w, dev)
    assert e0 == T.SIA_ERR_THICKNESS_ABOVE_LZ and e1 == 2  # the device flag of that case
    assert same(f1, f0) and d1 == d0
    big = T.make_grid(1024, 513, 1e3, 1e3, T.vertical_levels(4000.0, 11))
    assert big.Mx * big.My > S.NOALLOC_MAX_COLUMNS
    import torch
    x = torch.zeros(1, dtype=torch.float64, device="cuda")
    rc = S.flux_update_noalloc(big, w.params, x, x, x, x, x, x, x)
    assert rc == T.SIA_ERR_ARGUMENT
