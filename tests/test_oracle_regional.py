"""CPU checks of the oracle's SIAFD_Regional gradient override
(oracle/sia_oracle.c: or_regional_gradient, restating regional.cc:45-109).

Known answers used: with a planar stored surface usurfstore = a x + b y + c
every overridden staggered gradient equals (a, b) exactly (one-sided and
four-point forms alike), except where the stencil leaves the domain (0);
every point whose stencil does not touch the no-model mask keeps the
ordinary SIAFD gradient. The set of overridden points is computed here
independently from the stencils with numpy.
"""
import numpy as np

import sia_types as T
import workloads as W


def stencil_hits(nmm):
    """Boolean [Mx, My] per component: does the component's stencil touch nmm > 0.5."""
    m = nmm > 0.5
    sh = lambda di, dj: np.roll(np.roll(m, -di, axis=0), -dj, axis=1)  # value at (i+di, j+dj)
    hx0 = m | sh(1, 0)
    hx1 = sh(-1, 1) | sh(1, 1) | sh(-1, 0) | sh(1, 0)
    hy0 = sh(0, 1) | sh(1, 1) | sh(0, -1) | sh(1, -1)
    hy1 = m | sh(0, 1)
    return hx0, hx1, hy0, hy1


def test_regional_override_closed_form(oracle):
    w = W.dome(Mx=31, My=27, Mz=11)
    Mx, My = w.grid.Mx, w.grid.My
    f = w.fields
    base = oracle.sia_update(w.grid, w.params, f["H"], f["b"], f["h"], f["mask"], f["E"], full=False)
    nmm = np.zeros((Mx, My))
    nmm[:6, :] = 1.0          # a strip along i = 0 (touches the domain edge)
    nmm[20:24, 10:15] = 1.0   # an interior block
    nmm[:, 25:] = 1.0         # a strip along j = My - 1
    x = np.arange(Mx) * w.grid.dx
    y = np.arange(My) * w.grid.dy
    a, b = 2.5e-3, -1.5e-3
    hst = a * x[:, None] + b * y[None, :] + 100.0
    reg = oracle.sia_update(w.grid, w.params, f["H"], f["b"], f["h"], f["mask"], f["E"], full=False,
                            no_model_mask=nmm.ravel(), usurfstore=hst.ravel())
    hx = reg["h_x"].reshape(Mx, My, 2)
    hy = reg["h_y"].reshape(Mx, My, 2)
    bx = base["h_x"].reshape(Mx, My, 2)
    by = base["h_y"].reshape(Mx, My, 2)
    hit = stencil_hits(nmm)
    I, J = np.meshgrid(np.arange(Mx), np.arange(My), indexing="ij")
    # where the stencil leaves the domain (regional.cc:64, 73, 83, 93)
    out = [I + 1 > Mx - 1, (I - 1 < 0) | (J + 1 > My - 1) | (I + 1 > Mx - 1),
           (J + 1 > My - 1) | (I + 1 > Mx - 1) | (J - 1 < 0), J + 1 > My - 1]
    comps = [(hx[..., 0], bx[..., 0], a), (hx[..., 1], bx[..., 1], a), (hy[..., 0], by[..., 0], b),
             (hy[..., 1], by[..., 1], b)]
    for c, ((g, base_c, exact), h, o) in enumerate(zip(comps, hit, out)):
        assert np.array_equal(g[~h], base_c[~h]), c                     # untouched points keep SIAFD's
        assert np.all(g[h & o] == 0.0), c                               # stencil outside the domain
        np.testing.assert_allclose(g[h & ~o], exact, rtol=1e-12, err_msg=str(c))
        assert (h & ~o).sum() > 10 and (h & o).sum() > 0               # both cases occur


def test_zero_mask_changes_nothing(oracle):
    w = W.greenland_like(700, 724, 1180, 1204, Mz=51)
    f = w.fields
    n2 = w.grid.Mx * w.grid.My
    a = oracle.sia_update(w.grid, w.params, f["H"], f["b"], f["h"], f["mask"], f["E"], full=True)
    b = oracle.sia_update(w.grid, w.params, f["H"], f["b"], f["h"], f["mask"], f["E"], full=True,
                          no_model_mask=np.zeros(n2), usurfstore=np.full(n2, 7.0))
    for k in ("h_x", "h_y", "flux", "u", "v", "w", "Sigma"):
        assert np.array_equal(a[k], b[k]), k
