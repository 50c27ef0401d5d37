"""GPU tests of batched ensembles (include/sia_ensemble.h).

Each member is a seeded polythermal sheet of its own size, enhancement factor
and Glen exponent (GPBLD). The ensemble advances all members with one launch
per stage; the bar is bitwise equality with stepping every member alone
through sia_step (fast), whatever the tile width, eagerly or through a
captured HIP graph replayed several times.
"""
import numpy as np
import pytest

import sia_types as T
import workloads as W
from parity_util import say

pytestmark = pytest.mark.gpu

# (Mx, My, Mz, dx, seed, e, n)
MEMBERS = [(41, 41, 31, 5000.0, 11, 1.0, 3.0), (61, 53, 41, 4000.0, 12, 2.5, 4.0),
           (81, 81, 51, 3000.0, 13, 5.0, 3.0), (45, 71, 36, 4500.0, 14, 3.7, 4.0)]


@pytest.fixture(scope="module")
def S():
    import pism_sia
    import torch
    assert torch.cuda.is_available()
    return pism_sia


def member_workload(Mx, My, Mz, dx, seed, e, n):
    return W.greenland_like(Mx=Mx, My=My, Mz=Mz, dx=dx, seed=seed, Hmax=2400.0 + 200.0 * (seed % 4),
                            sia_enhancement_factor=e, sia_Glen_exponent=n)


class Member:
    """One member: its handle, device geometry, enthalpy and time."""

    def __init__(self, S, w):
        import torch
        self.w = w
        self.dev = {k: S.to_device(v) for k, v in w.fields.items()}
        self.sia = S.SIAFD(w.grid, w.params)
        self.sia.set_sync_check(False)
        self.geom = S.geometry(self.dev["H"], self.dev["h"], self.dev["mask"], self.dev["b"])
        self.t = torch.zeros(1, dtype=torch.float64, device="cuda")
        self.dt = torch.zeros(1, dtype=torch.float64, device="cuda")
        # bed smoother + allocations (the state is restored afterwards)
        d = self.dev
        self.sia.update(d["H"], d["b"], d["h"], d["mask"], d["E"], fast=True)

    def step_alone(self, k):
        for _ in range(k):
            self.sia.step(self.geom, self.dev["E"], fast=True, t_end=0.0, t_dev=self.t, dt_dev=self.dt)

    def state(self):
        import torch
        torch.cuda.synchronize()
        out = {k: self.dev[k].cpu().numpy() for k in ("H", "h", "mask")}
        out["flux"] = self.sia.diffusive_flux()
        out["D"] = self.sia.diffusivity_staggered()
        return out


def make_members(S):
    return [Member(S, member_workload(*m)) for m in MEMBERS]


def assert_same(a, b, what):
    for k in a:
        assert np.array_equal(a[k], b[k]), (what, k)


@pytest.fixture(scope="module")
def reference(S):
    """Every member stepped alone: states after 3 and after 6 steps, and times."""
    say("[gpu] ensemble: members stepped alone")
    ms = make_members(S)
    after3, after6, t3, t6 = [], [], [], []
    for m in ms:
        m.step_alone(3)
        after3.append(m.state())
        t3.append(m.t.item())
        m.step_alone(3)
        after6.append(m.state())
        t6.append(m.t.item())
    for m in ms:
        m.sia.close()
    return after3, after6, t3, t6


@pytest.mark.parametrize("tile", [8, 16, 32])
def test_ensemble_equals_members_alone(S, reference, tile):
    say("[gpu] ensemble eager steps, tile %d" % tile)
    after3, after6, t3, t6 = reference
    ms = make_members(S)
    ens = S.Ensemble([m.sia for m in ms], [m.geom for m in ms], [m.dev["E"] for m in ms])
    ens.set_tile(tile)
    ens.step(3)
    ens.check_errors()
    for i, m in enumerate(ms):
        assert_same(m.state(), after3[i], "member %d after 3" % i)
        assert ens.time(i)[0] == t3[i]
    ens.step(3)
    for i, m in enumerate(ms):
        assert_same(m.state(), after6[i], "member %d after 6" % i)
        assert ens.time(i)[0] == t6[i]
    ens.close()
    for m in ms:
        m.sia.close()


def test_ensemble_graph_replays(S, reference):
    """Three captured steps launched twice = six steps alone; the graph is
    built once and replayed (the member states carry over between launches)."""
    say("[gpu] ensemble hipGraph")
    after3, after6, t3, t6 = reference
    ms = make_members(S)
    ens = S.Ensemble([m.sia for m in ms], [m.geom for m in ms], [m.dev["E"] for m in ms])
    ens.step(3, graph=True)
    for i, m in enumerate(ms):
        assert_same(m.state(), after3[i], "graph member %d after 3" % i)
    ens.step(3, graph=True)
    for i, m in enumerate(ms):
        assert_same(m.state(), after6[i], "graph member %d after 6" % i)
        assert ens.time(i)[0] == t6[i]
    ens.check_errors()
    ens.close()
    for m in ms:
        m.sia.close()


def test_autotune_restores_state(S, reference):
    say("[gpu] ensemble autotuner")
    after3, _, t3, _ = reference
    ms = make_members(S)
    ens = S.Ensemble([m.sia for m in ms], [m.geom for m in ms], [m.dev["E"] for m in ms])
    chosen, ms_per = ens.autotune(steps=2)
    say("  tile times", ms_per, "chosen", chosen)
    assert chosen in (8, 16, 32) and all(v > 0 for v in ms_per.values())
    assert ms_per[chosen] == min(ms_per.values())
    for i in range(len(ms)):
        assert ens.time(i) == (0.0, 0.0)
    ens.step(3, graph=True)
    for i, m in enumerate(ms):
        assert_same(m.state(), after3[i], "after autotune, member %d" % i)
    ens.close()
    for m in ms:
        m.sia.close()


def test_ensemble_errors(S):
    say("[gpu] ensemble argument checks")
    w1 = member_workload(*MEMBERS[0])
    a = Member(S, w1)
    # a member whose bed smoother never ran
    fresh = S.SIAFD(w1.grid, w1.params)
    with pytest.raises(S.SIAError) as e:
        S.Ensemble([a.sia, fresh], [a.geom, a.geom], [a.dev["E"], a.dev["E"]])
    assert e.value.code == 8
    # members with different flow laws
    w2 = W.dome(Mx=21, My=21, Mz=11)
    b = Member(S, w2)
    with pytest.raises(S.SIAError) as e:
        S.Ensemble([a.sia, b.sia], [a.geom, b.geom], [a.dev["E"], b.dev["E"]])
    assert e.value.code == 6
    ens = S.Ensemble([a.sia], [a.geom], [a.dev["E"]])
    with pytest.raises(S.SIAError):
        ens.set_tile(12)
    ens.close()
    for h in (a.sia, b.sia, fresh):
        h.close()
