"""GPU tests of the learned surrogate inside batched ensembles
(include/sia_ensemble.h with members bound by sia_use_surrogate).

Members bound to a surrogate run its stages (classification and list, feature
pass, MLP, FP64 fallback) with one launch per stage for all members of that
surrogate; the others run the FP64 column stage. The bar is the ensemble's:
bitwise equality of every member's geometry, flux, D, time and surrogate
counts with stepping it alone through sia_step (fast) with the same binding,
eagerly and through a HIP graph replayed twice.
"""
import numpy as np
import pytest

from parity_util import say
from test_gpu_ensemble import MEMBERS, Member, assert_same, member_workload

pytestmark = pytest.mark.gpu

# member index -> surrogate (by Glen exponent): member 3 stays FP64 in BOUND;
# ALL_BOUND leaves the column stage without blocks
BOUND = {0: 3.0, 1: 4.0, 2: 3.0}
ALL_BOUND = {0: 3.0, 1: 4.0, 2: 3.0, 3: 4.0}


@pytest.fixture(scope="module")
def S():
    import pism_sia
    import torch
    assert torch.cuda.is_available()
    return pism_sia


@pytest.fixture(scope="module")
def surrogates(S):
    """One briefly trained surrogate per Glen exponent, from the members' own columns."""
    say("[gpu] ensemble surrogate: collect + train (n = 3, 4)")
    out = {}
    for n in (3.0, 4.0):
        sg = S.Surrogate(seed=int(n), capacity=60000)
        for m in MEMBERS:
            if m[6] != n:
                continue
            w = member_workload(*m)
            dev = {k: S.to_device(v) for k, v in w.fields.items()}
            sia = S.SIAFD(w.grid, w.params)
            sg.collect(sia, dev["H"], dev["b"], dev["h"], dev["mask"], dev["E"], stride=1)
            sia.close()
        sg.fit_normalization()
        sg.train(300, batch=512, lr_start=2e-3, lr_end=1e-4, seed=3)
        out[n] = sg
    return out


def make_bound(S, surrogates, bound=BOUND):
    ms = [Member(S, member_workload(*m)) for m in MEMBERS]
    for i, n in bound.items():
        ms[i].sia.use_surrogate(surrogates[n])
    return ms


def state(m, bound):
    st = m.state()
    if bound:
        st["counts"] = np.array(m.sia.surrogate_counts())
    return st


def run_alone(S, surrogates, bound):
    say("[gpu] ensemble surrogate: members stepped alone, bound %s" % sorted(bound))
    ms = make_bound(S, surrogates, bound)
    after3, after6, t6 = [], [], []
    for i, m in enumerate(ms):
        m.step_alone(3)
        after3.append(state(m, i in bound))
        m.step_alone(3)
        after6.append(state(m, i in bound))
        t6.append(m.t.item())
    for i in bound:
        c = after6[i]["counts"]
        assert c[1] > 0, "member %d: no surrogate columns" % i
    for m in ms:
        m.sia.close()
    return after3, after6, t6


@pytest.fixture(scope="module")
def reference(S, surrogates):
    return run_alone(S, surrogates, BOUND)


@pytest.mark.parametrize("graph", [False, True])
def test_surrogate_ensemble_equals_members_alone(S, surrogates, reference, graph):
    say("[gpu] ensemble with surrogate groups, graph=%s" % graph)
    after3, after6, t6 = reference
    ms = make_bound(S, surrogates)
    ens = S.Ensemble([m.sia for m in ms], [m.geom for m in ms], [m.dev["E"] for m in ms])
    ens.step(3, graph=graph)
    ens.check_errors()
    for i, m in enumerate(ms):
        assert_same(state(m, i in BOUND), after3[i], "member %d after 3" % i)
    ens.step(3, graph=graph)
    for i, m in enumerate(ms):
        assert_same(state(m, i in BOUND), after6[i], "member %d after 6" % i)
        assert ens.time(i)[0] == t6[i]
    ens.close()
    for m in ms:
        m.sia.close()


def test_binding_change_is_refused(S, surrogates):
    say("[gpu] ensemble surrogate: binding changed after creation")
    ms = make_bound(S, surrogates)
    ens = S.Ensemble([m.sia for m in ms], [m.geom for m in ms], [m.dev["E"] for m in ms])
    ms[0].sia.use_surrogate(None)
    with pytest.raises(S.SIAError):
        ens.step(1)
    ms[0].sia.use_surrogate(surrogates[3.0])
    ens.step(1)
    ens.check_errors()
    ens.close()
    for m in ms:
        m.sia.close()


def test_every_member_bound(S, surrogates):
    """All members bound (the column stage has no blocks), from a graph."""
    after3, after6, t6 = run_alone(S, surrogates, ALL_BOUND)
    say("[gpu] ensemble with every member bound")
    ms = make_bound(S, surrogates, ALL_BOUND)
    ens = S.Ensemble([m.sia for m in ms], [m.geom for m in ms], [m.dev["E"] for m in ms])
    ens.step(3, graph=True)
    ens.check_errors()
    for i, m in enumerate(ms):
        assert_same(state(m, True), after3[i], "member %d after 3" % i)
    ens.step(3, graph=True)
    for i, m in enumerate(ms):
        assert_same(state(m, True), after6[i], "member %d after 6" % i)
        assert ens.time(i)[0] == t6[i]
    ens.close()
    for m in ms:
        m.sia.close()
