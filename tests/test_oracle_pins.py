"""Independent pins of oracle branches that the reference's golden vectors and
the closed forms elsewhere leave open. Every expected value here is worked
out by hand from the reference's rules for a small, specially built input
(or comes from a golden number of the reference's own tests), so a dropped
branch, a swapped index or a wrong comparison in the oracle fails a check:

* haseloff's two-point averages and its zero-weight branch
  (SIAFD.cc:441-499), on a grid of floating, ocean, grounded and cliff rows,
  and the same grid transposed for the y-derivative;
* get_smoothed_thk with an active smoother over a non-flat bed
  (PISMBedSmoother.cc:286-336): test_22's bed with usurf = 1000 m, whose
  smoothed-bed extremes (test/regression/test_22.sh) fix the extremes of
  thksmooth, plus the maxtl >= H tie, the floating branch, H = 0 and H < 0;
* adjust_flow's floating and ocean cases with cell_interface_fluxes' SSA
  face velocities and the upwind thickness (iMgeometry.cc:172-415, 637-650);
* GeometryCalculator's flotation criterion (Mask.hh:82-118);
* the surrogate's MLP forward pass against torch's own linear and
  tanh-GELU (oracle/surrogate_oracle.py).
"""
import math

import numpy as np
import pytest

import sia_types as T


# ---------------------------------------------------------------- haseloff
def _marine_rows(Mx=6):
    """Rows j of a periodic 6 x 6 grid, identical along i:
    j=0 grounded ice  h = 1000 + 10 i          b = 0
    j=1 floating      h = 50 + 2 i             b = -1000
    j=2 ice-free ocean h = 0                   b = -1000
    j=3 floating      h = 60 + 3 i             b = -1000
    j=4 grounded ice  h = 500 + 5 i            b = 0
    j=5 ice-free land h = b = 900              (a cliff above row 4, not above row 0)
    """
    My = 6
    i = np.arange(Mx, dtype=float)[:, None] * np.ones((1, My))
    mask = np.zeros((Mx, My))
    h = np.zeros((Mx, My))
    b = np.zeros((Mx, My))
    mask[:, 0], h[:, 0], b[:, 0] = T.MASK_GROUNDED, 1000 + 10 * i[:, 0], 0.0
    mask[:, 1], h[:, 1], b[:, 1] = T.MASK_FLOATING, 50 + 2 * i[:, 1], -1000.0
    mask[:, 2], h[:, 2], b[:, 2] = T.MASK_ICE_FREE_OCEAN, 0.0, -1000.0
    mask[:, 3], h[:, 3], b[:, 3] = T.MASK_FLOATING, 60 + 3 * i[:, 3], -1000.0
    mask[:, 4], h[:, 4], b[:, 4] = T.MASK_GROUNDED, 500 + 5 * i[:, 4], 0.0
    mask[:, 5], h[:, 5], b[:, 5] = T.MASK_ICE_FREE_BEDROCK, 900.0, 900.0
    H = np.where((mask == 2) | (mask == 3), 100.0, 0.0)  # not read by haseloff
    return H, b, h, mask


def _haseloff(oracle, H, b, h, mask, dx, dy):
    Mx, My = mask.shape
    grid = T.make_grid(Mx, My, dx, dy, T.vertical_levels(4000.0, 11))
    P = T.make_params(surface_gradient_method="haseloff")
    hx, hy = oracle.surface_gradient(grid, P, *(np.ascontiguousarray(a).ravel() for a in (H, b, h, mask)))
    return hx.reshape(Mx, My, 2), hy.reshape(Mx, My, 2)


# Hand values of h_x on the j-offset (north) faces, i = 1..4 (away from the
# periodic wrap of the linear rows), dx = 100:
#  j=0: w_j = 1 (grounded -> floating is not a margin): four faces
#       (0.1 + 0.1 + 0.02 + 0.02) / 4 = 0.06
#  j=1: floating, w_j = 0 (floating -> ocean), icy: two faces of the own row
#       (0.02 + 0.02) / 2 = 0.02        (four faces would give 0.01)
#  j=2: ocean, w_j = 0 (ocean -> floating), not icy: the two faces of row 3
#       (0.03 + 0.03) / 2 = 0.03        (four faces would give 0.015)
#  j=3: w_j = 1 (floating -> grounded): (0.03 + 0.03 + 0.05 + 0.05) / 4 = 0.04
#  j=4: grounded, w_j = 0 (cliff: b(i,5) = 900 > h(i,4)), icy: (0.05 + 0.05) / 2 = 0.05
#  j=5: w_j = 1 (land -> grounded row 0, b = 900 < h(i,0)): (0 + 0 + 0.1 + 0.1) / 4 = 0.05
HX_NORTH = {0: 0.06, 1: 0.02, 2: 0.03, 3: 0.04, 4: 0.05, 5: 0.05}
# h_y on those north faces, dy = 200: (h(i,j+1) - h(i,j)) / dy, or 0 at the margins
def _hy_north(i, j):
    vals = {0: ((50 + 2 * i) - (1000 + 10 * i)) / 200.0, 1: 0.0, 2: 0.0,
            3: ((500 + 5 * i) - (60 + 3 * i)) / 200.0, 4: 0.0, 5: ((1000 + 10 * i) - 900.0) / 200.0}
    return vals[j]


def test_haseloff_two_point_branches_x(oracle):
    H, b, h, mask = _marine_rows()
    hx, hy = _haseloff(oracle, H, b, h, mask, 100.0, 200.0)
    for i in range(1, 5):
        for j in range(6):
            assert hx[i, j, 1] == pytest.approx(HX_NORTH[j], rel=1e-13, abs=1e-15), (i, j)
            assert hy[i, j, 1] == pytest.approx(_hy_north(i, j), rel=1e-13, abs=1e-15), (i, j)
    # the i-offset (east) faces: every east face joins cells of one type (default case)
    slopes = {0: 0.1, 1: 0.02, 2: 0.0, 3: 0.03, 4: 0.05, 5: 0.0}
    for i in range(5):
        for j in range(6):
            assert hx[i, j, 0] == pytest.approx(slopes[j], rel=1e-13, abs=1e-15), (i, j)
    # at the periodic wrap the east face of i = 5 joins h(5,j) to h(0,j):
    # row 1 gives (50 - 60) / 100 = -0.1, so h_x(0,1) on the north face is the
    # two-point (0.02 - 0.1) / 2 = -0.04; row 3 wraps (60 - 75) / 100 = -0.15,
    # so the ocean cells (0,2) and (5,2) give (0.03 - 0.15) / 2 = -0.06
    assert hx[0, 1, 1] == pytest.approx(-0.04, rel=1e-13)
    assert hx[0, 2, 1] == pytest.approx(-0.06, rel=1e-13)
    assert hx[5, 2, 1] == pytest.approx(-0.06, rel=1e-13)


def test_haseloff_two_point_branches_y(oracle):
    """The same rows turned into columns: h_y on the east faces must take the
    values h_x took on the north faces (dx and dy swap with the grid)."""
    H, b, h, mask = (np.ascontiguousarray(a.T) for a in _marine_rows())
    hx, hy = _haseloff(oracle, H, b, h, mask, 200.0, 100.0)
    for j in range(1, 5):
        for i in range(6):
            assert hy[i, j, 0] == pytest.approx(HX_NORTH[i], rel=1e-13, abs=1e-15), (i, j)
            assert hx[i, j, 0] == pytest.approx(_hy_north(j, i), rel=1e-13, abs=1e-15), (i, j)


def test_haseloff_zero_weight_branch(oracle):
    """A floating cell with ice-free ocean on every side: all four of its
    faces are marine margins, so w_i(i,j) = w_i(i-1,j) = w_j(i,j) = 0 and the
    icy branch has W = 0: both cross components are 0 there. The ocean's
    surface is given a slope (7 per cell in i, 11 in j) so that any average
    that wrongly included an ocean face would be non-zero. The ocean cell
    south of it sees w_j = 0 and is not icy: its h_x on the north face
    averages the two east faces of the floating cell's row, one of which is
    a margin (weight 0) and one an ocean-ocean face (slope 7 / dx):
    7 / dx / 1."""
    Mx = My = 5
    ii, jj = np.meshgrid(np.arange(Mx, dtype=float), np.arange(My, dtype=float), indexing="ij")
    h = 7.0 * ii + 11.0 * jj
    b = np.full((Mx, My), -1000.0)
    mask = np.full((Mx, My), float(T.MASK_ICE_FREE_OCEAN))
    mask[2, 2] = T.MASK_FLOATING
    h[2, 2] = 40.0
    H = np.where(mask == 3, 300.0, 0.0)
    dx, dy = 100.0, 200.0
    hx, hy = _haseloff(oracle, H, b, h, mask, dx, dy)
    assert hx[2, 2, 0] == 0.0 and hx[1, 2, 0] == 0.0 and hy[2, 2, 1] == 0.0 and hy[2, 1, 1] == 0.0
    assert hx[2, 2, 1] == 0.0  # w_j = 0, icy, W = w_i(2,2) + w_i(1,2) = 0
    assert hy[2, 2, 0] == 0.0  # w_i = 0, icy, W = w_j(2,2) + w_j(2,1) = 0
    # ocean cell (2,1): w_j(2,1) = 0 (ocean -> floating), not icy:
    # (h_x(1,2,0) + h_x(2,2,0)) / (w_i(2,2) + w_i(1,2)) over the row j = 2, both margins: W = 0 -> 0
    assert hx[2, 1, 1] == 0.0
    # ocean cell (1,1): w_j(1,1) = 1 (ocean -> ocean): four faces (1,1),(0,1),(0,2),(1,2);
    # (1,2)'s east face is the margin (0, weight 0): (7 + 7 + 7 + 0) / dx / 3
    assert hx[1, 1, 1] == pytest.approx(21.0 / dx / 3.0, rel=1e-14)


# ---------------------------------------------------------------- thksmooth
def _test22_grid():
    Mx = My = 81
    L = 1200e3
    x = T.horizontal_coordinates(L, Mx)
    y = T.horizontal_coordinates(L, My)
    dx = T.horizontal_spacing(L, Mx)
    X, Y = np.meshgrid(x, y, indexing="ij")
    topg = 400.0 * np.sin(2.0 * math.pi * X / 600.0e3) + 100.0 * np.sin(2.0 * math.pi * (X + 1.5 * Y) / 40.0e3)
    P = T.make_params(bed_smoother_range=50.0e3, sia_Glen_exponent=3.0, sia_flow_law="isothermal_glen",
                      is_dry_simulation=1)
    grid = T.make_grid(Mx, My, dx, dx, T.vertical_levels(4000.0, 31))
    return grid, P, topg


def test_thksmooth_active_smoother_test22_bed(oracle):
    """test_22's sinusoidal bed under a flat surface usurf = 1000 m with
    H = usurf - b, grounded everywhere: maxtl < H at every point (checked), so
    thksmooth = max(usurf - topgsmooth, 0), whose extremes follow from the
    golden smoothed-bed extremes -+372.992474 m: 627.007526 and 1372.992474.
    A call site that passed topg instead of topgsmooth would give H, whose
    extremes are 500 and 1500."""
    grid, P, topg = _test22_grid()
    Mx, My = grid.Mx, grid.My
    usurf = np.full((Mx, My), 1000.0)
    H = usurf - topg
    mask = np.full((Mx, My), float(T.MASK_GROUNDED))
    E = np.zeros(Mx * My * grid.Mz)
    out = oracle.sia_update(grid, P, H.ravel(), topg.ravel(), usurf.ravel(), mask.ravel(), E, full=False)
    assert np.all(out["maxtl"] < H.ravel())
    ts = out["thksmooth"]
    assert "%.6f" % ts.min() == "627.007526"
    assert "%.6f" % ts.max() == "1372.992474"
    assert abs(H.min() - 500.0) < 1e-6 and abs(H.max() - 1500.0) < 1e-6


def test_thksmooth_branches_by_hand(oracle):
    """get_smoothed_thk point by point (PISMBedSmoother.cc:312-330) on chosen
    inputs: H = 0 -> 0; maxtl == H exactly -> H (the comparison is >=);
    maxtl slightly below H, grounded -> usurf - topgsmooth; the same floating
    -> H; usurf below topgsmooth, grounded -> 0; H < 0 -> the error."""
    grid = T.make_grid(6, 1, 1e3, 1e3, T.vertical_levels(4000.0, 11))
    usurf = np.array([700.0, 700.0, 700.0, 700.0, 100.0, 700.0])
    thk = np.array([0.0, 250.0, 250.0, 250.0, 50.0, 300.0])
    maxtl = np.array([10.0, 250.0, 249.5, 249.5, 20.0, 400.0])
    topgs = np.array([0.0, 420.0, 420.0, 420.0, 120.0, 500.0])
    mask = np.array([0.0, 2.0, 2.0, 3.0, 2.0, 2.0])
    expected = [0.0, 250.0, 280.0, 250.0, 0.0, 300.0]
    out = oracle.thksmooth(grid, usurf, thk, mask, maxtl, topgs)
    assert list(out) == expected
    with pytest.raises(oracle.OracleError) as e:
        oracle.thksmooth(grid, usurf, np.array([0.0, -1.0, 0, 0, 0, 0]), mask, maxtl, topgs)
    assert e.value.code == T.SIA_ERR_NEGATIVE_THICKNESS


# ---------------------------------------------------------------- adjust_flow
def test_adjust_flow_floating_and_ocean_by_hand(oracle):
    """One explicit step on a row of cells, identical along j (My = 3):
       i:    0         1         2         3          4         5
       mask  grounded  floating  floating  ice-free   ice-free  grounded
                                           ocean      land
    East faces and what adjust_flow + cell_interface_fluxes leave of them
    (iMgeometry.cc:172-281, 309-415):
       (0,1) grounded-floating  SIA kept,   SSA 0.5 (u0 + u1)
       (1,2) floating-floating  SIA zeroed, SSA 0.5 (u1 + u2)
       (2,3) floating-ocean     SIA zeroed, SSA u2 (the icy side's velocity)
       (3,4) ocean-land         SIA zeroed, SSA zeroed
       (4,5) land-grounded      SIA kept,   SSA u5
       (5,0) grounded-grounded  SIA kept,   SSA 0.5 (u5 + u0)
    North and south faces join equal cells with equal fluxes and cancel.
    The SSA flux is upwind: V H_west if V > 0, else V H_east."""
    Mx, My = 6, 3
    dx, dy, dt = 1000.0, 500.0, 1.0e7
    grid = T.make_grid(Mx, My, dx, dy, T.vertical_levels(4000.0, 11))
    P = T.make_params(is_dry_simulation=0)
    mrow = np.array([2.0, 3.0, 3.0, 4.0, 0.0, 2.0])
    Hrow = np.array([1000.0, 400.0, 300.0, 0.0, 0.0, 800.0])
    Qe = np.array([1.0, 2.0, 3.0, 4.0, 5.0, 6.0]) * 1e-3     # east-face SIA fluxes, m2 s-1
    Qn = np.array([0.5, 0.25, 0.125, 1.0, 2.0, 0.75]) * 1e-3  # north-face fluxes (cancel)
    u = np.array([1.0, -2.0, 3.0, 4.0, 5.0, -6.0]) * 1e-6     # sliding u, m s-1
    mask = np.repeat(mrow[:, None], My, 1)
    H = np.repeat(Hrow[:, None], My, 1)
    flux = np.zeros((Mx, My, 2))
    flux[:, :, 0] = Qe[:, None]
    flux[:, :, 1] = Qn[:, None]
    ub = np.repeat(u[:, None], My, 1)
    vb = np.zeros((Mx, My))
    Hn, acc = oracle.mass_continuity(grid, P, dt, H.ravel(), flux.ravel(), mask.ravel(), ub=ub.ravel(),
                                     vb=vb.ravel())
    Hn = Hn.reshape(Mx, My)
    # face values after adjustment (east face of cell i)
    Qf = [Qe[0], 0.0, 0.0, 0.0, Qe[4], Qe[5]]
    Vf = [0.5 * (u[0] + u[1]), 0.5 * (u[1] + u[2]), u[2], 0.0, u[5], 0.5 * (u[5] + u[0])]
    def up(f):  # upwind thickness of east face f (between f and f+1)
        return Vf[f] * (Hrow[f] if Vf[f] > 0 else Hrow[(f + 1) % Mx])
    expected = []
    for i in range(Mx):
        w = (i - 1) % Mx
        divQ_SIA = (Qf[i] - Qf[w]) / dx
        divQ_SSA = (up(i) - up(w)) / dx
        Hi = Hrow[i] + dt * (0.0 - 0.0 - (divQ_SIA + divQ_SSA))
        expected.append(max(Hi, 0.0))
    # by hand, for the record: cell 1 loses the SIA flux of face (0,1) to its
    # west and gains nothing across the floating face to its east; cell 3
    # (ocean) receives u2 H2 from the floating cell and has SMB forced to 0
    assert expected[3] == pytest.approx(dt * u[2] * Hrow[2] / dx, rel=1e-15)
    for j in range(My):
        np.testing.assert_allclose(Hn[:, j], expected, rtol=1e-14, atol=1e-12)
    # the negative cells were clipped and counted by the non-negativity rule
    raw = [Hrow[i] - dt * ((Qf[i] - Qf[(i - 1) % Mx]) / dx + (up(i) - up((i - 1) % Mx)) / dx)
           for i in range(Mx)]
    nonneg = sum(-r for r in raw if r < 0) * My * dx * dy * P.ice_density
    assert acc[1] == pytest.approx(nonneg, rel=1e-12, abs=1e-6)


# ---------------------------------------------------------------- flotation
def test_geometry_flotation_by_hand(oracle):
    """GeometryCalculator::compute (Mask.hh:82-118), sea level 0,
    rho_i / rho_w = 910 / 1028, floating if sea_level + (1 - rho_i/rho_w) H >
    b + H + 1 m, ice-free if H < 0.01 m:
      b = -1000, H = 500   -> floating, h = (1 - 910/1028) 500 = 57.392996...
      b = -1000, H = 0.005 -> ice-free ocean, h = 0.005 (1 - 910/1028)
      b = -1000, H = 2000  -> grounded (229.57 < 1000 + 1), h = 1000
      b = 10,    H = 0     -> ice-free land, h = 10
      b = -100,  H = 111   -> floating (12.74 > 12), h = 12.7412...
    and with is_dry_simulation every cell is grounded or ice-free land."""
    grid = T.make_grid(5, 1, 1e3, 1e3, T.vertical_levels(4000.0, 11))
    b = np.array([-1000.0, -1000.0, -1000.0, 10.0, -100.0])
    H = np.array([500.0, 0.005, 2000.0, 0.0, 111.0])
    a = 1.0 - 910.0 / 1028.0
    P = T.make_params(is_dry_simulation=0, sea_level=0.0)
    mask, h = oracle.update_mask_surface(grid, P, b, H)
    assert list(mask) == [3.0, 4.0, 2.0, 0.0, 3.0]
    np.testing.assert_allclose(h, [a * 500.0, a * 0.005, 1000.0, 10.0, a * 111.0], rtol=1e-15)
    Pd = T.make_params(is_dry_simulation=1, sea_level=0.0)
    mask, h = oracle.update_mask_surface(grid, Pd, b, H)
    assert list(mask) == [2.0, 0.0, 2.0, 0.0, 2.0]
    np.testing.assert_allclose(h, b + H, rtol=1e-15)


# ---------------------------------------------------------------- surrogate MLP
def test_surrogate_forward_matches_torch():
    """The oracle's MLP (34 -> 4 x 128 tanh-GELU -> 32) against torch's
    F.linear and F.gelu(approximate="tanh") on random weights, the flat
    parameter vector sliced in its documented order W1[128][34] b1 W2 b2 W3
    b3 W4 b4 W5[32][128] b5. Without rounding the two agree to FP64 rounding;
    with the BF16 rounding of the GPU kernels, against torch's own bfloat16
    casts (round to nearest even) at every layer."""
    torch = pytest.importorskip("torch")
    import surrogate_oracle as SO
    rng = np.random.default_rng(7)
    w = rng.normal(0.0, 0.15, SO.NPARAM).astype(np.float32)
    X = rng.normal(0.0, 1.0, (40, 64)).astype(np.float32)
    sizes = [(128, 34), (128, 128), (128, 128), (128, 128), (32, 128)]
    assert sum(o * i + o for o, i in sizes) == SO.NPARAM
    tw = torch.from_numpy(w.astype(np.float64))
    layers, off = [], 0
    for o, i in sizes:
        W = tw[off:off + o * i].reshape(o, i)
        off += o * i
        bb = tw[off:off + o]
        off += o
        layers.append((W, bb))

    def torch_forward(x, bf16):
        R = (lambda t: t.float().to(torch.bfloat16).double()) if bf16 else (lambda t: t)
        a = R(torch.from_numpy(x[:, :34].astype(np.float64)))
        for l, (W, bb) in enumerate(layers):
            bias = bb.float().double() if bf16 else bb
            z = torch.nn.functional.linear(a, R(W), bias)
            if l < 4:
                a = R(torch.nn.functional.gelu(z, approximate="tanh"))
        return z.numpy()

    out, _ = SO.forward(w, X, rounding=False)
    ref = torch_forward(X, False)
    np.testing.assert_allclose(out, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
    out, _ = SO.forward(w, X, rounding=True)
    ref = torch_forward(X, True)
    # torch rounds GELU's FP64 value to bfloat16 through float32 as the oracle
    # does; summation order may flip a rare half-ulp tie
    np.testing.assert_allclose(out, ref, rtol=0, atol=2e-2 * np.abs(ref).max())
    assert np.median(np.abs(out - ref)) <= 1e-12 * np.abs(ref).max()
    # the BF16 rounding itself: bitwise torch's round-to-nearest-even
    v = rng.normal(0.0, 3.0, 10000).astype(np.float32)
    assert np.array_equal(SO.bf16(v), torch.from_numpy(v).to(torch.bfloat16).double().numpy())
