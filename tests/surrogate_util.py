"""Helpers of the surrogate GPU tests: the FP64 and surrogate fast updates of
one set of fields through the C ABI, the model blob's normalization, and the
oracle's inference (oracle/surrogate_oracle.py) of the same columns."""
import struct

import numpy as np

import surrogate_oracle as SO


def plain_fast(S, w, dev):
    sia = S.SIAFD(w.grid, w.params)
    sia.update(dev["H"], dev["b"], dev["h"], dev["mask"], dev["E"], fast=True)
    out = dict(D=sia.diffusivity_staggered(), flux=sia.diffusive_flux(), D_max=sia.max_diffusivity())
    out["theta"], out["thks"] = sia.theta_thksmooth()
    out["hx"], out["hy"] = sia.surface_gradient()
    sia.close()
    return out


def surrogate_fast(S, sg, w, dev):
    sia = S.SIAFD(w.grid, w.params)
    sia.use_surrogate(sg)
    sia.update(dev["H"], dev["b"], dev["h"], dev["mask"], dev["E"], fast=True)
    out = dict(D=sia.diffusivity_staggered(), flux=sia.diffusive_flux(), D_max=sia.max_diffusivity(),
               counts=sia.surrogate_counts())
    sia.close()
    return out


def eligible(w, st):
    z = np.array([w.grid.z[k] for k in range(w.grid.Mz)])
    return (st["thk"] > 0) & (st["alpha"] > 0) & (st["thk"] >= z[1]) & (st["thk"] <= z[-1] + 1e-6)


def parse_norm(blob):
    hdr = struct.calcsize("8sii d ii")
    vals = np.frombuffer(blob, dtype=np.float32, count=4 * 34 + 1, offset=hdr)
    mean, inv_std, lo, hi = (vals[34 * i:34 * (i + 1)] for i in range(4))
    w = np.frombuffer(blob, dtype=np.float32, offset=hdr + 4 * (4 * 34 + 1))
    return dict(mean=mean, inv_std=inv_std, lo=lo, hi=hi, rscale=float(vals[-1])), w


def oracle_inference(sg, w, ref):
    """The oracle's surrogate D of every eligible staggered column of w (the
    FP64 update ref gives thksmooth, theta and the gradient): (column index,
    D, in-range flag)."""
    norm, wts = parse_norm(sg.save())
    st = SO.staggered_inputs(w.grid, w.fields, ref["thks"], ref["theta"], ref["hx"], ref["hy"])
    sel = np.nonzero(eligible(w, st))[0]
    z = np.array([w.grid.z[k] for k in range(w.grid.Mz)])
    c = SO.columns(w.params, z, st["thk"][sel], st["alpha"][sel], st["theta"][sel], st["E"][sel], targets=False)
    F = c["F"].astype(np.float32)
    inside = np.all((F >= norm["lo"]) & (F <= norm["hi"]), axis=1)
    X = np.zeros((len(sel), SO.KIN), dtype=np.float32)
    X[:, :34] = (F - norm["mean"]) * norm["inv_std"]
    out, _ = SO.forward(wts.astype(np.float64), X, rounding=True)
    Dor = c["CSt"] + (c["CQb"].astype(np.float64) * np.exp(out * norm["rscale"])).sum(axis=1)
    return sel, Dor, inside
