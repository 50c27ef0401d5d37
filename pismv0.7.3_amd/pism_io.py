"""PISM-compatible output of the SIA fields and the sia.py harness semantics.

* write_sia_netcdf: a NetCDF-3 file (PISM's default output_format "netcdf3",
  pism_config.cdl:1349) with PISM's variable names, dimensions in the default
  "xyz" order (time, x, y[, z]) (pism_config.cdl:1355, io_helpers.cc:570-598),
  and PISM's metadata:
  - uvel, vvel: SSB_Modifier.cc:34-44;
  - wvel_rel: PISMStressBalance.cc:43-49;
  - strainheat: SSB_Modifier.cc:46-50;
  - the SIAFD diagnostics diffusivity, diffusivity_i/_j, h_x_i/_j, h_y_i/_j,
    schoofs_theta, thksmooth, topgsmooth: SIAFD_diagnostics.cc:26-210.
  Fields flagged write_in_glaciological_units are stored in their
  glaciological units (m year-1) as PISM writes them.
* get_val_z: IceModelVec3D::getValZ (iceModelVec3.cc:131-155) on host arrays.
* compute_sia_surface_velocities: PISM.sia.computeSIASurfaceVelocities
  (site-packages/PISM/sia.py:24-55) on the GPU: a full update with zero
  sliding, then u and v at the ice surface (IceModelVec3::getSurfaceValues).

NetCDF-3 files are written with scipy.io.netcdf_file (64-bit offsets).
"""
import numpy as np

SECONDS_PER_YEAR = 3.15569259747e7  # UDUNITS "year"

# name: (long_name, standard_name, units, glaciological units or None, pism_intent, written in glaciological units)
META = {
    "uvel": ("horizontal velocity of ice in the X direction", "land_ice_x_velocity", "m s-1", "m year-1",
             "diagnostic", True),
    "vvel": ("horizontal velocity of ice in the Y direction", "land_ice_y_velocity", "m s-1", "m year-1",
             "diagnostic", True),
    "wvel_rel": ("vertical velocity of ice, relative to base of ice directly below", "", "m s-1", "m year-1",
                 "diagnostic", True),
    "strainheat": ("rate of strain heating in ice (dissipation heating)", "", "W m-3", "mW m-3", "internal", False),
    "diffusivity": ("diffusivity of SIA mass continuity equation", "", "m2 s-1", "m2 s-1", "diagnostic", False),
    "diffusivity_i": ("diffusivity of SIA mass continuity equation on the staggered grid (i-offset)", "",
                      "m2 s-1", "m2 s-1", "diagnostic", True),
    "diffusivity_j": ("diffusivity of SIA mass continuity equation on the staggered grid (j-offset)", "",
                      "m2 s-1", "m2 s-1", "diagnostic", True),
    "h_x_i": ("the x-component of the surface gradient, i-offset", "", "", None, "diagnostic", True),
    "h_x_j": ("the x-component of the surface gradient, j-offset", "", "", None, "diagnostic", True),
    "h_y_i": ("the y-component of the surface gradient, i-offset", "", "", None, "diagnostic", True),
    "h_y_j": ("the y-component of the surface gradient, j-offset", "", "", None, "diagnostic", True),
    "schoofs_theta": ("multiplier 'theta' in Schoof's (2003) theory of bed roughness in SIA", "", "1", None,
                      "diagnostic", False),
    "thksmooth": ("thickness relative to smoothed bed elevation in Schoof's (2003) theory of bed roughness in SIA",
                  "", "m", "m", "diagnostic", False),
    "topgsmooth": ("smoothed bed elevation in Schoof's (2003) theory of bed roughness in SIA", "", "m", "m",
                   "diagnostic", False),
}
THREE_D = ("uvel", "vvel", "wvel_rel", "strainheat")

_TO_GLACIOLOGICAL = {("m s-1", "m year-1"): SECONDS_PER_YEAR, ("W m-3", "mW m-3"): 1e3,
                     ("m2 s-1", "m2 s-1"): 1.0, ("m", "m"): 1.0}


def sia_fields(sia, full=True):
    """The outputs of a SIAFD object's last update as {PISM name: numpy array}
    (2-D [Mx*My], 3-D [Mx*My*Mz], staggered components split into _i/_j)."""
    out = {}
    D = sia.diffusivity_staggered()
    out["diffusivity_i"], out["diffusivity_j"] = D[0::2], D[1::2]
    out["diffusivity"] = sia.diffusivity()
    hx, hy = sia.surface_gradient()
    out["h_x_i"], out["h_x_j"] = hx[0::2], hx[1::2]
    out["h_y_i"], out["h_y_j"] = hy[0::2], hy[1::2]
    out["schoofs_theta"], out["thksmooth"] = sia.theta_thksmooth()
    out["topgsmooth"] = sia.bed_smoother_fields()["topgsmooth"]
    if full:
        out["uvel"], out["vvel"], out["wvel_rel"] = sia.velocity_u(), sia.velocity_v(), sia.velocity_w()
        out["strainheat"] = sia.volumetric_strain_heating()
    return out


def write_sia_netcdf(path, grid, x, y, fields, time=0.0):
    """Write fields ({PISM name: flat array in PISM layout}) to a NetCDF-3 file."""
    from scipy.io import netcdf_file
    Mx, My, Mz = grid.Mx, grid.My, grid.Mz
    z = np.array([grid.z[k] for k in range(Mz)])
    with netcdf_file(path, "w", version=2) as nc:
        nc.createDimension("time", None)
        nc.createDimension("x", Mx)
        nc.createDimension("y", My)
        nc.createDimension("z", Mz)
        for name, vals, units, axis in (("x", x, "m", "X"), ("y", y, "m", "Y"), ("z", z, "m", "Z")):
            v = nc.createVariable(name, "d", (name,))
            v[:] = np.asarray(vals, dtype=np.float64)
            v.units, v.axis = units, axis
        t = nc.createVariable("time", "d", ("time",))
        t[0] = time
        t.units = "seconds since 1-1-1"
        t.axis = "T"
        for name, arr in fields.items():
            long_name, standard_name, units, glac, intent, glac_write = META[name]
            dims = ("time", "x", "y", "z") if name in THREE_D else ("time", "x", "y")
            shape = (1, Mx, My, Mz) if name in THREE_D else (1, Mx, My)
            data = np.asarray(arr, dtype=np.float64).reshape(shape)
            out_units = units
            if glac_write and glac is not None and glac != units:
                data = data * _TO_GLACIOLOGICAL[(units, glac)]
                out_units = glac
            v = nc.createVariable(name, "d", dims)
            v[:] = data
            v.long_name = long_name
            if standard_name:
                v.standard_name = standard_name
            v.units = out_units
            if glac is not None:
                v.glaciological_units = glac
            v.pism_intent = intent


def read_netcdf(path):
    """{name: (array, {attribute: value})} of every variable (a copy)."""
    from scipy.io import netcdf_file
    out = {}
    with netcdf_file(path, "r", mmap=False) as nc:
        for name, v in nc.variables.items():
            attrs = {k: (val.decode() if isinstance(val, bytes) else val) for k, val in v._attributes.items()}
            out[name] = (np.array(v[:]), attrs, v.dimensions)
    return out


def get_val_z(column, z_levels, z):
    """IceModelVec3D::getValZ (iceModelVec3.cc:131-155) of one column."""
    zl = np.asarray(z_levels)
    if z >= zl[-1]:
        return column[-1]
    if z <= zl[0]:
        return column[0]
    m = int(np.searchsorted(zl, z, side="right") - 1)
    m = min(m, len(zl) - 2)
    incr = (z - zl[m]) / (zl[m + 1] - zl[m])
    valm = column[m]
    return valm + incr * (column[m + 1] - valm)


def compute_sia_surface_velocities(sia, H, b, h, mask, E):
    """PISM.sia.computeSIASurfaceVelocities (sia.py:24-55): full update with
    zero basal sliding, then u and v at the ice surface. Device tensors in,
    numpy (u_s, v_s) out."""
    sia.update(H, b, h, mask, E, fast=False)
    us = sia.surface_values(sia.velocity_u_device(), H)
    vs = sia.surface_values(sia.velocity_v_device(), H)
    return us.cpu().numpy(), vs.cpu().numpy()
