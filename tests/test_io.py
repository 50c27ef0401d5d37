"""CPU checks of the PISM-compatible output (pismv0.7.3_amd/pism_io.py):
getValZ's known answers (iceModelVec3.cc:131-155), and a NetCDF-3 round trip
with PISM's names, dimension order and glaciological units."""
import os

import numpy as np

import pism_io as IO
import sia_types as T


def test_get_val_z_known_answers():
    z = T.vertical_levels(1000.0, 11, "quadratic", 4.0)
    col = 3.0 + 0.5 * z                       # linear in z: interpolation is exact
    for zq in (0.0, 1.0, 123.4, z[3], 999.0):
        assert IO.get_val_z(col, z, zq) == np.float64(3.0 + 0.5 * zq) or \
            abs(IO.get_val_z(col, z, zq) - (3.0 + 0.5 * zq)) < 1e-12
    assert IO.get_val_z(col, z, z[3]) == col[3]          # on a level: that level
    assert IO.get_val_z(col, z, 5000.0) == col[-1]       # above the column: top level
    assert IO.get_val_z(col, z, -1.0) == col[0]          # below: bottom level
    sq = z * z                                           # between levels: the chord, not the curve
    zq = 0.5 * (z[2] + z[3])
    assert IO.get_val_z(sq, z, zq) == sq[2] + 0.5 * (sq[3] - sq[2])


def test_netcdf_round_trip(tmp_path):
    Mx, My, Mz = 5, 4, 3
    z = T.vertical_levels(100.0, Mz)
    grid = T.make_grid(Mx, My, 1000.0, 1000.0, z)
    x = np.arange(Mx) * 1000.0
    y = np.arange(My) * 1000.0
    rng = np.random.default_rng(0)
    f = {"uvel": rng.normal(size=Mx * My * Mz), "wvel_rel": rng.normal(size=Mx * My * Mz),
         "strainheat": rng.uniform(size=Mx * My * Mz), "diffusivity": rng.uniform(size=Mx * My),
         "h_x_i": rng.normal(size=Mx * My), "thksmooth": rng.uniform(size=Mx * My)}
    path = os.path.join(tmp_path, "sia.nc")
    IO.write_sia_netcdf(path, grid, x, y, f, time=7.0)
    r = IO.read_netcdf(path)
    u, a, dims = r["uvel"]
    assert dims == ("time", "x", "y", "z")               # PISM's default "xyz" order
    assert a["units"] == "m year-1" and a["standard_name"] == "land_ice_x_velocity"
    np.testing.assert_array_equal(u[0], (f["uvel"] * IO.SECONDS_PER_YEAR).reshape(Mx, My, Mz))
    s, a, _ = r["strainheat"]
    assert a["units"] == "W m-3" and a["pism_intent"] == "internal"   # not written in mW m-3
    np.testing.assert_array_equal(s[0], f["strainheat"].reshape(Mx, My, Mz))
    d, a, dims = r["diffusivity"]
    assert dims == ("time", "x", "y") and a["units"] == "m2 s-1"
    np.testing.assert_array_equal(d[0], f["diffusivity"].reshape(Mx, My))
    assert r["time"][0][0] == 7.0 and np.array_equal(r["z"][0], z)


def test_names_are_the_reference_outputs():
    """SSB_Modifier (uvel, vvel, strainheat), StressBalance (wvel_rel) and the
    SIAFD diagnostics (SIAFD_diagnostics.cc:26-35, staggered components _i/_j)."""
    names = {"uvel", "vvel", "wvel_rel", "strainheat", "diffusivity", "diffusivity_i", "diffusivity_j",
             "h_x_i", "h_x_j", "h_y_i", "h_y_j", "schoofs_theta", "thksmooth", "topgsmooth"}
    assert set(IO.META) == names
