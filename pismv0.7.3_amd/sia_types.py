"""ctypes mirrors of the structs in include/pism_sia.h and PISM's defaults.

The default values are those of PISM v0.7.3's src/pism_config.cdl (line
numbers given per entry), already converted to SI units the way
Config::get_double(name, units) would convert them. ``ice_grain_size`` is read
by SIAFD without a unit argument (SIAFD.cc:565), so its raw value 1.0 is kept.
"""
import ctypes as C

import numpy as np

# UDUNITS "year" used by PISM for "years" -> "seconds" (iceCompModel.cc:53)
SECONDS_PER_YEAR = 3.15569259747e7

FLOW_LAWS = {  # sia_flow_law keywords, cdl:1093; flowlaw_factory.cc:77-83
    "isothermal_glen": 0, "pb": 1, "gpbld": 2, "hooke": 3,
    "arr": 4, "arrwarm": 5, "gk": 6,
}
GRADIENT_METHODS = {"haseloff": 0, "mahaffy": 1, "eta": 2}  # cdl:1272; SIAFD.cc:187-207
ENTHALPY_CONVERTERS = {"default": 0, "cold": 1, "kirchhoff": 2}

# status codes of include/pism_sia.h (enum sia_status)
SIA_OK, SIA_ERR_ARGUMENT, SIA_ERR_HIP, SIA_ERR_NEGATIVE_THICKNESS, SIA_ERR_THICKNESS_ABOVE_LZ, \
    SIA_ERR_OMEGA_NEGATIVE, SIA_ERR_CONFIG, SIA_ERR_OUT_OF_MEMORY, SIA_ERR_STATE = range(9)

# mask codes, Mask.hh:28-34
MASK_ICE_FREE_BEDROCK = 0
MASK_GROUNDED = 2
MASK_FLOATING = 3
MASK_ICE_FREE_OCEAN = 4


class SiaGrid(C.Structure):
    _fields_ = [("Mx", C.c_int32), ("My", C.c_int32), ("Mz", C.c_int32),
                ("reserved", C.c_int32), ("dx", C.c_double), ("dy", C.c_double),
                ("z", C.POINTER(C.c_double))]


_PARAM_INT_FIELDS = [
    "sia_flow_law", "surface_gradient_method", "enthalpy_converter",
    "compute_grain_size_using_age", "do_age", "is_dry_simulation",
    "include_bmr_in_continuity", "reserved",
]
_PARAM_DOUBLE_FIELDS = [
    "sia_enhancement_factor", "sia_Glen_exponent", "bed_smoother_range", "ice_grain_size",
    "ice_softness", "Paterson_Budd_A_cold", "Paterson_Budd_A_warm", "Paterson_Budd_Q_cold",
    "Paterson_Budd_Q_warm", "Paterson_Budd_critical_temperature", "gpbld_water_frac_coeff",
    "gpbld_water_frac_observed_limit", "Hooke_A", "Hooke_Q", "Hooke_C", "Hooke_k", "Hooke_Tr",
    "ideal_gas_constant", "beta_CC", "ice_specific_heat_capacity",
    "water_specific_heat_capacity", "water_latent_heat_fusion", "surface_pressure",
    "ice_density", "standard_gravity", "water_melting_point_temperature",
    "enthalpy_converter_reference_temperature", "sea_water_density", "sea_level",
    "mask_icefree_thickness_standard", "mask_is_floating_thickness_standard",
    "adaptive_timestepping_ratio", "maximum_time_step", "seconds_per_year",
]


class SiaParams(C.Structure):
    _fields_ = ([(n, C.c_int32) for n in _PARAM_INT_FIELDS] +
                [(n, C.c_double) for n in _PARAM_DOUBLE_FIELDS])


# pism_config.cdl defaults (line numbers in src/pism_config.cdl)
DEFAULTS = {
    "sia_flow_law": FLOW_LAWS["gpbld"],                      # cdl:1094
    "surface_gradient_method": GRADIENT_METHODS["haseloff"],  # cdl:1272
    "enthalpy_converter": ENTHALPY_CONVERTERS["default"],     # enthalpy_converter_from_options
    "compute_grain_size_using_age": 0,                        # cdl:309
    "do_age": 0,                                              # cdl:34
    "is_dry_simulation": 0,                                   # cdl:136
    "include_bmr_in_continuity": 1,                           # cdl:148
    "reserved": 0,
    "sia_enhancement_factor": 1.0,                            # cdl:281
    "sia_Glen_exponent": 3.0,                                 # cdl:1009
    "bed_smoother_range": 5.0e3,                              # cdl:560
    "ice_grain_size": 1.0,                                    # cdl:304
    "ice_softness": 3.1689e-24,                               # cdl:1020
    "Paterson_Budd_A_cold": 3.61e-13,                         # cdl:1060
    "Paterson_Budd_A_warm": 1.73e3,                           # cdl:1065
    "Paterson_Budd_Q_cold": 6.0e4,                            # cdl:1070
    "Paterson_Budd_Q_warm": 13.9e4,                           # cdl:1075
    "Paterson_Budd_critical_temperature": 263.15,             # cdl:1080
    "gpbld_water_frac_coeff": 181.25,                         # cdl:1115
    "gpbld_water_frac_observed_limit": 0.01,                  # cdl:1120
    "Hooke_A": 4.42165e-9,                                    # cdl:1025
    "Hooke_Q": 7.88e4,                                        # cdl:1030
    "Hooke_C": 0.16612,                                       # cdl:1035
    "Hooke_k": 1.17,                                          # cdl:1040
    "Hooke_Tr": 273.39,                                       # cdl:1045
    "ideal_gas_constant": 8.31441,                            # cdl:1193
    "beta_CC": 7.9e-8,                                        # cdl:963
    "ice_specific_heat_capacity": 2009.0,                     # cdl:1003
    "water_specific_heat_capacity": 4170.0,                   # cdl:988
    "water_latent_heat_fusion": 3.34e5,                       # cdl:983
    "surface_pressure": 0.0,                                  # cdl:968
    "ice_density": 910.0,                                     # cdl:993
    "standard_gravity": 9.81,                                 # cdl:1188
    "water_melting_point_temperature": 273.15,                # cdl:973
    "enthalpy_converter_reference_temperature": 223.15,       # cdl:978
    "sea_water_density": 1028.0,                              # cdl:1140
    "sea_level": 0.0,
    "mask_icefree_thickness_standard": 0.01,                  # cdl:1225
    "mask_is_floating_thickness_standard": 1.0,               # cdl:1230
    "adaptive_timestepping_ratio": 0.12,                      # cdl:330
    "maximum_time_step": 60.0 * SECONDS_PER_YEAR,             # cdl:341 (years)
    "seconds_per_year": SECONDS_PER_YEAR,
}


def make_params(**overrides):
    """Return a SiaParams filled with PISM defaults plus overrides.

    Keyword values for the selection fields may be given as the cdl keyword
    strings (e.g. ``sia_flow_law="arr"``, ``surface_gradient_method="eta"``,
    ``enthalpy_converter="cold"``).
    """
    vals = dict(DEFAULTS)
    for k, v in overrides.items():
        if k not in vals:
            raise KeyError("unknown SIA parameter %r" % k)
        if k == "sia_flow_law" and isinstance(v, str):
            v = FLOW_LAWS[v]
        elif k == "surface_gradient_method" and isinstance(v, str):
            v = GRADIENT_METHODS[v]
        elif k == "enthalpy_converter" and isinstance(v, str):
            v = ENTHALPY_CONVERTERS[v]
        vals[k] = v
    p = SiaParams()
    for k, v in vals.items():
        setattr(p, k, v)
    return p


def params_dict(p):
    return {name: getattr(p, name) for name, _ in SiaParams._fields_}


def vertical_levels(Lz, Mz, spacing="equal", lam=4.0):
    """IceGrid::compute_vertical_levels (IceGrid.cc:333-376): EQUAL spacing
    (352-360) or QUADRATIC with slope lambda at the base (362-369); the top
    level is exactly Lz."""
    if spacing == "equal":
        dz = Lz / (float(Mz) - 1)
        z = [dz * float(k) for k in range(Mz - 1)]
    elif spacing == "quadratic":
        z = []
        for k in range(Mz - 1):
            zeta = float(k) / (float(Mz) - 1)
            z.append(Lz * ((zeta / lam) * (1.0 + (lam - 1.0) * zeta)))
    else:
        raise ValueError("spacing must be 'equal' or 'quadratic'")
    return np.array(z + [Lz], dtype=np.float64)


def horizontal_spacing(half_width, M):
    """compute_horizontal_spacing, non-periodic (IceGrid.cc:540-546)."""
    return 2.0 * half_width / (M - 1)


def horizontal_coordinates(half_width, M, center=0.0):
    """compute_coordinates, non-periodic (IceGrid.cc:549-569)."""
    delta = horizontal_spacing(half_width, M)
    v_min, v_max = center - half_width, center + half_width
    x = np.array([v_min + i * delta for i in range(M)], dtype=np.float64)
    x[M - 1] = v_max
    return x


def make_grid(Mx, My, dx, dy, z):
    """Build a SiaGrid; keeps a reference to the z array on the struct."""
    z = np.ascontiguousarray(z, dtype=np.float64)
    g = SiaGrid(Mx, My, len(z), 0, dx, dy, z.ctypes.data_as(C.POINTER(C.c_double)))
    g._z_keepalive = z
    return g


class SiaMassAccounting(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "grounded_basal_ice_flux", "nonneg_rule_flux", "sub_shelf_ice_flux",
        "surface_ice_flux", "sum_divQ_SIA", "sum_divQ_SSA", "Href_to_H_flux",
        "H_to_Href_flux")]


class SiaInputs(C.Structure):
    _fields_ = [("land_ice_thickness", C.c_void_p), ("bedrock_altitude", C.c_void_p),
                ("surface_altitude", C.c_void_p), ("mask", C.c_void_p),
                ("enthalpy", C.c_void_p), ("age", C.c_void_p),
                ("vel_input_u", C.c_void_p), ("vel_input_v", C.c_void_p),
                ("basal_melt_rate", C.c_void_p), ("bed_state_counter", C.c_int64),
                ("no_model_mask", C.c_void_p), ("usurfstore", C.c_void_p)]


class SiaGeometry(C.Structure):
    _fields_ = [("land_ice_thickness", C.c_void_p), ("surface_altitude", C.c_void_p),
                ("mask", C.c_void_p), ("bedrock_altitude", C.c_void_p),
                ("climatic_mass_balance", C.c_void_p), ("basal_melt_rate", C.c_void_p)]
