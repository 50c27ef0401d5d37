"""CPU oracle of the learned column surrogate (numpy, float64). TEST
INFRASTRUCTURE ONLY: imported by tests/ and nothing else; the product path is
the HIP library.

What it restates:
* the discrete column integral of SIAFD::compute_diffusive_flux
  (src/base/stressbalance/sia/SIAFD.cc:620-721) regrouped per level:
  D = C sum_k W_k p_k^n A_k with C = 2 e theta alpha^(n-1) and
  W_k = d_k (dz_k [k >= 1] + dz_{k+1} [k < ks]) / 2 + d_ks^2 / 2 [k == ks]
  (the terms of SIAFD.cc:677-684 that multiply delta_k);
* GPBLD softness (flowlaws.cc:209-221, 90-95) and the default enthalpy
  converter (enthalpyConverter.cc:92-234), vectorised over levels;
* the surrogate's cell features, physics baseline and targets, and the MLP
  (34 -> 4 x 128 GELU(tanh form) -> 32) forward and backward, optionally with
  the BF16 rounding of weights and activations that the GPU kernels apply;
* the score of the surrogate's FP64 fallback (fallback_score).
"""
import numpy as np

NCELL = 32
NIN = 34
KIN = 64
WIDTH = 128
NOUT = 32
T_REF = 263.15

# flat parameter layout (row = output unit), as the library's FP32 master array
LAYOUT = []
_off = 0
for _name, _shape in (("W1", (WIDTH, NIN)), ("b1", (WIDTH,)), ("W2", (WIDTH, WIDTH)), ("b2", (WIDTH,)),
                      ("W3", (WIDTH, WIDTH)), ("b3", (WIDTH,)), ("W4", (WIDTH, WIDTH)), ("b4", (WIDTH,)),
                      ("W5", (NOUT, WIDTH)), ("b5", (NOUT,))):
    _n = int(np.prod(_shape))
    LAYOUT.append((_name, _off, _shape))
    _off += _n
NPARAM = _off  # 58144


def unpack(w):
    return {name: np.asarray(w[off:off + int(np.prod(shape))], dtype=np.float64).reshape(shape)
            for name, off, shape in LAYOUT}


def pack(p):
    w = np.zeros(NPARAM, dtype=np.float64)
    for name, off, shape in LAYOUT:
        w[off:off + int(np.prod(shape))] = np.asarray(p[name]).ravel()
    return w


# ---------------------------------------------------------------- bf16
def bf16(x):
    """float32 -> bfloat16 (round to nearest even) -> float64, elementwise."""
    f = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


# ---------------------------------------------------------------- GPBLD + EC (vectorised)
def ec_pressure(P, depth):
    """EnthalpyConverter::pressure (enthalpyConverter.cc:92-98)."""
    return np.where(depth > 0.0, P.surface_pressure + P.ice_density * P.standard_gravity * depth,
                    P.surface_pressure)


def ec_melting_temperature(P, p):
    return P.water_melting_point_temperature - P.beta_CC * p  # enthalpyConverter.cc:128-130


def ec_enthalpy_cts(P, p):
    return P.ice_specific_heat_capacity * (ec_melting_temperature(P, p)
                                           - P.enthalpy_converter_reference_temperature)  # :141-143


def softness_pb(P, T_pa):
    """FlowLaw::softness_parameter_paterson_budd (flowlaws.cc:90-95)."""
    cold = T_pa < P.Paterson_Budd_critical_temperature
    A = np.where(cold, P.Paterson_Budd_A_cold, P.Paterson_Budd_A_warm)
    Q = np.where(cold, P.Paterson_Budd_Q_cold, P.Paterson_Budd_Q_warm)
    return A * np.exp(-Q / (P.ideal_gas_constant * T_pa))


def stress_pow(n, s):
    """s^(n-1) with the integer-power path of the library (n - 1 in {1, 2, 3})."""
    nm1 = n - 1.0
    if nm1 == 1.0:
        return s
    if nm1 == 2.0:
        return s * s
    if nm1 == 3.0:
        return s * s * s
    return np.power(s, nm1)


# ---------------------------------------------------------------- column decomposition
def kbelow(z, h):
    """IceGrid::kBelowHeight clamped to Mz-2 (IceGrid.cc:379-391)."""
    return np.clip(np.searchsorted(z, h, side="right") - 1, 0, len(z) - 2)


def columns(P, z, thk, alpha, theta_l, E, targets=True):
    """Surrogate decomposition of N staggered columns.

    thk, alpha, theta_l: [N]; E: [N, Mz] enthalpy averaged over the two cells
    (SIAFD.cc:660-661). Returns a dict with the features F [N, 34] (float32
    values as the GPU stores them), the baseline C Qb [N, 32], C S_t [N], the
    cold-cell mask, and with targets the exact per-cell sums, r, shares and
    D_ref (the reference's D from the per-level regrouping)."""
    z = np.asarray(z, dtype=np.float64)
    N, Mz = E.shape
    n = P.sia_Glen_exponent
    ks = kbelow(z, thk)
    k = np.arange(Mz)[None, :]
    d = thk[:, None] - z[None, :]
    inside = k <= ks[:, None]
    dzk = np.concatenate([[0.0], np.diff(z)])[None, :]
    dzk1 = np.concatenate([np.diff(z), [0.0]])[None, :]
    W = 0.5 * d * (dzk + np.where(k < ks[:, None], dzk1, 0.0))
    W = W + np.where(k == ks[:, None], 0.5 * d * d, 0.0)
    p = ec_pressure(P, np.where(inside, d, 0.0))
    G = np.where(inside, W * stress_pow(n, p) * p, 0.0)
    Es = ec_enthalpy_cts(P, p)
    cold = inside & (E < Es)
    temp = inside & ~(E < Es)
    T = E / P.ice_specific_heat_capacity + P.enthalpy_converter_reference_temperature
    Tpa = T - ec_melting_temperature(P, p) + P.water_melting_point_temperature  # :197-199, cold branch
    omega = np.minimum((E - Es) / P.water_latent_heat_fusion, P.gpbld_water_frac_observed_limit)
    A_T0 = float(softness_pb(P, np.array(P.water_melting_point_temperature)))
    St = np.where(temp, G * (A_T0 * (1.0 + P.gpbld_water_frac_coeff * omega)), 0.0).sum(axis=1)
    Cf = P.sia_enhancement_factor * theta_l * 2.0 * stress_pow(n, alpha)
    s32 = NCELL / thk
    cell = np.minimum((z[None, :] * s32[:, None]).astype(np.int64), NCELL - 1)
    x = Tpa - T_REF
    g0 = np.zeros((N, NCELL))
    g1 = np.zeros((N, NCELL))
    g2 = np.zeros((N, NCELL))
    qref = np.zeros((N, NCELL))
    Acold = softness_pb(P, np.where(cold, Tpa, 260.0))
    for m in range(NCELL):
        sel = cold & (cell == m)
        g0[:, m] = np.where(sel, G, 0.0).sum(axis=1)
        g1[:, m] = np.where(sel, G * x, 0.0).sum(axis=1)
        g2[:, m] = np.where(sel, G * x * x, 0.0).sum(axis=1)
        if targets:
            qref[:, m] = np.where(sel, G * Acold, 0.0).sum(axis=1)
    has = g0 > 0.0
    g0s = np.where(has, g0, 1.0)
    mx = g1 / g0s
    Tbar = T_REF + mx
    var = np.maximum(g2 / g0s - mx * mx, 0.0)
    Qa = np.where(Tbar < P.Paterson_Budd_critical_temperature, P.Paterson_Budd_Q_cold, P.Paterson_Budd_Q_warm)
    kk = Qa / (P.ideal_gas_constant * Tbar * Tbar)
    Qb = np.where(has, g0 * softness_pb(P, Tbar) * (1.0 + 0.5 * (kk * kk - 2.0 * kk / Tbar) * var), 0.0)
    F = np.zeros((N, NIN), dtype=np.float32)
    F[:, :NCELL] = np.where(has, (Tbar - 250.0) / 15.0, 2.0).astype(np.float32)
    F[:, 32] = (thk / 1000.0).astype(np.float32)
    F[:, 33] = (np.log10(alpha) + 2.0).astype(np.float32)
    out = dict(F=F, CQb=(Cf[:, None] * Qb).astype(np.float32), CSt=Cf * St, has=has, Cf=Cf, St=St, Qb=Qb, ks=ks)
    if targets:
        Dref = Cf * (St + qref.sum(axis=1))
        out["qref"] = qref
        out["Dref"] = Dref
        out["r"] = np.where(has, np.log(np.where(has, qref, 1.0) / np.where(has, Qb, 1.0)), 0.0)
        out["share"] = np.where(Dref[:, None] > 0, Cf[:, None] * qref / np.where(Dref > 0, Dref, 1.0)[:, None], 0.0)
    return out


def staggered_inputs(grid, fields, thks, theta, hx, hy):
    """Per staggered column (o = 0, 1 of every map column): thk, alpha,
    theta_l and the averaged enthalpy (SIAFD.cc:629-661), flattened in the
    library's staggered order s = 2 (i My + j) + o."""
    Mx, My, Mz = grid.Mx, grid.My, grid.Mz
    ts = thks.reshape(Mx, My)
    th = theta.reshape(Mx, My)
    E = fields["E"].reshape(Mx, My, Mz)
    hx = hx.reshape(Mx, My, 2)
    hy = hy.reshape(Mx, My, 2)
    out = {}
    parts = []
    for o, (di, dj) in enumerate(((1, 0), (0, 1))):
        ts2 = np.roll(ts, (-di, -dj), axis=(0, 1))
        th2 = np.roll(th, (-di, -dj), axis=(0, 1))
        E2 = np.roll(E, (-di, -dj), axis=(0, 1))
        parts.append(dict(thk=0.5 * (ts + ts2), theta=0.5 * (th + th2), E=0.5 * (E + E2),
                          alpha=np.sqrt(hx[:, :, o] ** 2 + hy[:, :, o] ** 2),
                          slope=hx[:, :, o] if o == 0 else hy[:, :, o]))
    for key in ("thk", "theta", "alpha", "slope"):
        out[key] = np.stack([parts[0][key], parts[1][key]], axis=-1).reshape(-1)
    out["E"] = np.stack([parts[0]["E"], parts[1]["E"]], axis=2).reshape(-1, Mz)
    return out


def fallback_score(CQb, CSt, out, rscale):
    """The library's score fallback: sum_m t_m |r_m| / D with r_m = out_m rscale
    (the network's log-correction), t_m = C Qb_m exp(r_m) (cell m's predicted
    part of D) and D = C S_t + sum_m t_m. Cells without cold levels have
    CQb_m = 0 and add nothing. (No counterpart in the reference: it decides
    which columns the FP64 kernel computes instead of the surrogate.)"""
    r = np.asarray(out, dtype=np.float64) * rscale
    t = np.asarray(CQb, dtype=np.float64) * np.exp(r)
    D = np.asarray(CSt, dtype=np.float64) + t.sum(axis=1)
    return (t * np.abs(r)).sum(axis=1) / D


# ---------------------------------------------------------------- MLP
def gelu(x):
    u = 0.7978845608028654 * (x + 0.044715 * x ** 3)
    return 0.5 * x * (1.0 + np.tanh(u))


def gelu_grad(x):
    u = 0.7978845608028654 * (x + 0.044715 * x ** 3)
    th = np.tanh(u)
    du = 0.7978845608028654 * (1.0 + 3.0 * 0.044715 * x * x)
    return 0.5 * (1.0 + th) + 0.5 * x * (1.0 - th * th) * du


def forward(w, X, rounding=True):
    """X [B, 64] normalized inputs (only the first 34 columns matter).
    Returns the outputs [B, 32] and the per-layer (input, pre-activation)."""
    p = unpack(w)
    R = bf16 if rounding else (lambda a: np.asarray(a, dtype=np.float64))
    a = R(np.asarray(X, dtype=np.float32))[:, :NIN]
    cache = []
    for l in range(1, 6):
        Wl = R(p["W%d" % l])
        b = p["b%d" % l].astype(np.float32).astype(np.float64) if rounding else p["b%d" % l]
        z = a @ Wl.T + b
        cache.append((a, z))
        if l < 5:
            a = R(gelu(z).astype(np.float32)) if rounding else gelu(z)
    return z, cache


def loss_grad(w, X, Tt, S, Hm, rscale, lam=1e-3, rounding=True):
    """Loss of the library's training step and its gradient w.r.t. the flat
    parameters. Tt: targets in output units; S: shares; Hm: cold-cell masks.
    With rounding, the backward also uses BF16 copies of the weights and of
    dZ, as the GPU GEMMs do."""
    p = unpack(w)
    R = bf16 if rounding else (lambda a: np.asarray(a, dtype=np.float64))
    out, cache = forward(w, X, rounding)
    B = out.shape[0]
    dd = (out - Tt) * rscale
    e = (S * dd).sum(axis=1)
    loss = float(np.mean(e * e + lam * (Hm * dd * dd).sum(axis=1)))
    dZ = 2.0 / B * rscale * (e[:, None] * S + lam * Hm * dd)
    g = {}
    for l in range(5, 0, -1):
        a_in, z = cache[l - 1]
        dZr = R(dZ.astype(np.float32)) if rounding else dZ
        g["W%d" % l] = dZr.T @ a_in
        g["b%d" % l] = dZ.sum(axis=0)
        if l > 1:
            dA = dZr @ R(p["W%d" % l])
            dZ = dA * gelu_grad(cache[l - 2][1])
    return loss, pack(g)
