"""GPU parity of the time step and explicit mass-continuity path
(iMadaptive.cc:157-171, 214-388; iMgeometry.cc:56-149, 497-822), the
device-side step used by hipGraph capture, and the C++ host driver.

Tolerances: H after a step agrees with the oracle to 1e-12 of max(H) per
step; dt to 1e-12 relative; the accounting sums to 1e-12 of their absolute
size; mass is conserved to 14 digits (test_12.sh:33-44).
"""
import math
import os
import subprocess

import numpy as np
import pytest

import sia_types as T
import workloads as W
from parity_util import say

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def S():
    import pism_sia
    return pism_sia


def test_mass_continuity_step_matches_oracle(S, oracle):
    say("[gpu] dome: 10 steps of update + dt + mass continuity vs oracle")
    w = W.dome(Mx=41, My=41, Mz=31)
    M0 = W.vialov_smb(w.params)
    steps_vs_oracle(S, oracle, w, M0 * w.params.ice_density, 10)


def test_mass_continuity_on_rectangular_cells(S, oracle):
    """dx = 1 km, dy = 2.5 km on a margin window (a 1501x1121 sheet of the
    same extent): the flux divergence, the diffusivity dt limit
    (iMadaptive.cc:157-171) and the interface fluxes each take their own
    spacing, which a grid with dx = dy cannot tell apart."""
    say("[gpu] rectangular cells: 6 steps of update + dt + mass continuity vs oracle")
    w = W.greenland_like(160, 200, 570, 606, My=1121, Mz=101, dx=1000.0, dy=2500.0)
    assert w.grid.dx != w.grid.dy
    steps_vs_oracle(S, oracle, w, 0.3 / T.SECONDS_PER_YEAR * w.params.ice_density, 6)


def steps_vs_oracle(S, oracle, w, smb_value, nsteps):
    """nsteps of update (full, fast alternately) + dt + mass continuity on the
    GPU and in the oracle from the same state, with a uniform SMB (kg m-2 s-1)."""
    st = {k: v.copy() for k, v in w.fields.items()}
    E = st.pop("E")
    dev = {k: S.to_device(v) for k, v in st.items()}
    dE = S.to_device(E)
    smb = np.full(w.grid.Mx * w.grid.My, smb_value)
    dsmb = S.to_device(smb)
    sia = S.SIAFD(w.grid, w.params)
    geom = S.geometry(dev["H"], dev["h"], dev["mask"], dev["b"], smb=dsmb)
    for step in range(nsteps):
        sia.update(dev["H"], dev["b"], dev["h"], dev["mask"], dE, fast=(step % 2 == 1))
        dt_g = sia.max_timestep()
        acc_g = sia.mass_continuity_step(geom, dt_g)
        dt_o, _, acc_o = oracle.step(w.grid, w.params, st, E, full=(step % 2 == 0), smb=smb)
        assert abs(dt_g - dt_o) <= 1e-12 * dt_o
        Hg = dev["H"].cpu().numpy()
        assert np.max(np.abs(Hg - st["H"])) <= 1e-12 * st["H"].max() * (step + 1)
        assert np.array_equal(dev["mask"].cpu().numpy(), st["mask"])
        np.testing.assert_allclose(dev["h"].cpu().numpy(), st["h"], rtol=0, atol=1e-12 * st["H"].max() * (step + 1))
        for a, b in zip(acc_g, acc_o):
            assert abs(a - b) <= 1e-12 * (abs(b) + np.abs(acc_o).max())
        st["H"], st["h"], st["mask"] = Hg, dev["h"].cpu().numpy(), dev["mask"].cpu().numpy()


def test_mass_conservation_and_device_step(S):
    """sia_step (dt on the device) conserves mass to 14 digits with zero SMB."""
    say("[gpu] device-side steps conserve mass")
    import torch
    w = W.dome(Mx=61, My=61, Mz=31)
    dev = {k: S.to_device(v) for k, v in w.fields.items()}
    sia = S.SIAFD(w.grid, w.params)
    sia.set_sync_check(False)
    geom = S.geometry(dev["H"], dev["h"], dev["mask"], dev["b"])
    t = torch.zeros(1, dtype=torch.float64, device="cuda")
    dt = torch.zeros(1, dtype=torch.float64, device="cuda")
    area = w.grid.dx * w.grid.dy
    ivol = [dev["H"].sum().item() * area]
    for _ in range(20):
        sia.step(geom, dev["E"], fast=False, t_end=1e12, t_dev=t, dt_dev=dt)
        ivol.append(dev["H"].sum().item() * area)
    torch.cuda.synchronize()
    ivol = np.array(ivol)
    threshold = 10 ** (math.floor(math.log10(ivol.max())) - 14)
    assert np.abs(np.diff(ivol)).max() < threshold
    assert t.item() == pytest.approx(20 * 1.63 * T.SECONDS_PER_YEAR, rel=0.25)


def test_graph_captured_steps_replay_exactly(S):
    """Ten captured steps replayed twice from the same state give the same
    bitwise result as ten eager steps (no memset nodes in the graph)."""
    say("[gpu] hipGraph capture of steps")
    import torch
    w = W.greenland_like(700, 764, 1180, 1260, Mz=401)
    base = {k: S.to_device(v) for k, v in w.fields.items()}

    def fresh():
        return {k: v.clone() for k, v in base.items()}

    # eager reference
    st = fresh()
    s1 = torch.cuda.Stream()
    sia = S.SIAFD(w.grid, w.params, stream=s1.cuda_stream)
    sia.set_sync_check(False)
    t = torch.zeros(1, dtype=torch.float64, device="cuda")
    dt = torch.zeros(1, dtype=torch.float64, device="cuda")
    geom = S.geometry(st["H"], st["h"], st["mask"], st["b"])
    with torch.cuda.stream(s1):
        for _ in range(10):
            sia.step(geom, st["E"], t_end=0.0, t_dev=t, dt_dev=dt)
    torch.cuda.synchronize()
    H_eager = st["H"].cpu().numpy()
    # captured
    st2 = fresh()
    geom2 = S.geometry(st2["H"], st2["h"], st2["mask"], st2["b"])
    t2 = torch.zeros(1, dtype=torch.float64, device="cuda")
    sia2 = S.SIAFD(w.grid, w.params, stream=s1.cuda_stream)
    sia2.set_sync_check(False)
    with torch.cuda.stream(s1):
        sia2.step(geom2, st2["E"], t_end=0.0, t_dev=t2, dt_dev=dt)  # warm-up: allocations, bed smoother
    torch.cuda.synchronize()
    for k in ("H", "h", "mask"):
        st2[k].copy_(base[k])
    t2.zero_()
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=s1):
        for _ in range(10):
            sia2.step(geom2, st2["E"], t_end=0.0, t_dev=t2, dt_dev=dt)
    for rep in range(2):
        for k in ("H", "h", "mask"):
            st2[k].copy_(base[k])
        t2.zero_()
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert np.array_equal(st2["H"].cpu().numpy(), H_eager), rep
        assert t2.item() == t.item()


def test_cpp_host_driver():
    """pismv0.7.3_amd/bin/siafd_hip_test drives the C++ class SIAFD_HIP."""
    say("[gpu] C++ host driver")
    exe = os.path.join(REPO, "pismv0.7.3_amd", "bin", "siafd_hip_test")
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    say(out.stdout)
    assert out.returncode == 0 and "PASSED" in out.stdout


@pytest.mark.parametrize("Mx,key", [(31, "Mx31"), (41, "Mx41")])
def test_gpu_reproduces_test_15_table(S, Mx, key):
    """The HIP path reproduces every printed digit of test_15's error table
    (test/regression/test_15.sh:19-28): 5000 years of verification test C."""
    import json
    import verification as V
    say("[gpu] verification test C", Mx)
    gold = json.load(open(os.path.join(REPO, "tests", "golden", "regression_outputs.json")))["test_C"]
    errs, steps, _ = V.run_test_c_gpu(Mx, Mx)
    say("  errors", ["%.6f" % e for e in errs], "golden", gold[key])
    assert ["%.6f" % e for e in errs] == gold[key]
    assert steps == 84


@pytest.mark.parametrize("Mx,key", [(21, "Mx21"), (41, "Mx41")])
def test_gpu_reproduces_test_14_table(S, Mx, key):
    """The HIP path (SIA_Sliding + SIAFD with sliding + mass continuity)
    reproduces every printed digit of test_14's error tables
    (test/regression/test_14.sh:17-35): 100 years of verification test E."""
    import json
    import verification as V
    say("[gpu] verification test E", Mx)
    gold = json.load(open(os.path.join(REPO, "tests", "golden", "regression_outputs.json")))["test_E"][key]
    geo, vel, steps = V.run_test_e_gpu(Mx)
    got_geo = ["%.6f" % e for e in geo]
    got_vel = ["%.4f" % vel[0], "%.5f" % vel[1], "%.5f" % vel[2], "%.4f" % vel[3], "%.4f" % vel[4]]
    say("  geometry", got_geo, "golden", gold["geometry"])
    say("  base vels", got_vel, "golden", gold["base_vels"])
    assert got_geo == gold["geometry"] and got_vel == gold["base_vels"]


def test_gpu_test_12_mass_conservation(S, oracle):
    """test_12.sh on the HIP path: test B (Halfar), 31x31x31, years 1000-6000,
    max_dt 25 a, ice volume every 25 a; diff(ivol).max() below the 14th digit
    of max(ivol). The final thickness agrees with the oracle's run to 1e-9 of
    max(H) (FP64 rounding over 200 steps)."""
    say("[gpu] test_12: test B mass conservation, 200 steps")
    import verification as V
    from test_verification_B import run_oracle_test_b
    ivol, H, steps = V.run_test_b_gpu(31)
    assert len(ivol) == 201 and steps >= 200
    threshold = 10.0 ** (math.floor(math.log10(ivol.max())) - 14)
    say("  diff(ivol) in [%.2e, %.2e] m^3, threshold %.1e" % (np.diff(ivol).min(), np.diff(ivol).max(), threshold))
    assert np.diff(ivol).max() < threshold
    ivol_o, H_o, _ = run_oracle_test_b(oracle)
    assert np.abs(H - H_o).max() <= 1e-9 * H_o.max()
    assert np.allclose(ivol, ivol_o, rtol=1e-12, atol=0.0)
