"""GPU checks of the device diagnostics behind the PISM-compatible output:
the regular-grid diffusivity (SIAFD.cc:736-744, iceModelVec2.cc:679-689) and
getSurfaceValues (iceModelVec3.cc:131-155, 206-222) against their numpy
restatements on the library's own fields (bitwise), and a NetCDF file of a
full update read back."""
import os

import numpy as np
import pytest

import pism_io as IO
import workloads as W
from parity_util import say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def S():
    import pism_sia
    import torch
    assert torch.cuda.is_available()
    return pism_sia


def test_diffusivity_and_surface_values(S, tmp_path):
    say("[gpu] diffusivity, getSurfaceValues, NetCDF output")
    w = W.greenland_like(700, 740, 1180, 1216, Mz=201)
    Mx, My, Mz = w.grid.Mx, w.grid.My, w.grid.Mz
    dev = {k: S.to_device(v) for k, v in w.fields.items()}
    sia = S.SIAFD(w.grid, w.params)
    us, vs = IO.compute_sia_surface_velocities(sia, dev["H"], dev["b"], dev["h"], dev["mask"], dev["E"])
    z = np.array([w.grid.z[k] for k in range(Mz)])
    H = w.fields["H"]
    u = sia.velocity_u().reshape(Mx * My, Mz)
    v = sia.velocity_v().reshape(Mx * My, Mz)
    ref_u = np.array([IO.get_val_z(u[q], z, H[q]) for q in range(Mx * My)])
    ref_v = np.array([IO.get_val_z(v[q], z, H[q]) for q in range(Mx * My)])
    assert np.array_equal(us, ref_u) and np.array_equal(vs, ref_v)
    assert np.abs(us).max() > 0
    D = sia.diffusivity_staggered().reshape(Mx, My, 2)
    ref = 0.25 * (D[:, :, 0] + D[:, :, 1] + np.roll(D[:, :, 1], 1, axis=1) + np.roll(D[:, :, 0], 1, axis=0))
    assert np.array_equal(sia.diffusivity().reshape(Mx, My), ref)
    # a PISM-style output file of the whole update
    path = os.path.join(tmp_path, "sia_out.nc")
    fields = IO.sia_fields(sia)
    IO.write_sia_netcdf(path, w.grid, np.arange(Mx) * w.grid.dx, np.arange(My) * w.grid.dy, fields)
    r = IO.read_netcdf(path)
    np.testing.assert_array_equal(r["wvel_rel"][0][0], (fields["wvel_rel"] * IO.SECONDS_PER_YEAR).reshape(Mx, My, Mz))
    np.testing.assert_array_equal(r["h_x_j"][0][0], fields["h_x_j"].reshape(Mx, My))
    sia.close()
