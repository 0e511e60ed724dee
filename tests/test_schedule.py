"""Host-side turn scheduling of the fused engine (no GPU needed): the
fixed-point shortcut gives exactly the programs of the full per-turn
recurrence, and the host scalars end where the element-by-element path
leaves them."""
import numpy as np

from oracle import workloads as W
from tests.machines import nonlinear_machine, linear_machine


class _Scalars(object):
    pass


def _beam(ob):
    b = _Scalars()
    b.mass, b.charge, b.circumference = ob.mass, ob.charge, ob.circumference
    b.macroparticlenumber, b.particlenumber_per_mp = ob.macroparticlenumber, ob.particlenumber_per_mp
    b._gamma, b._beta, b._betagamma, b._p0 = ob._gamma, ob._beta, ob._betagamma, ob._p0
    return b


def _full(eng, beam, n):
    from pyheadtail_amd.particles.particles import ScalarState
    st = ScalarState(beam)
    keys = [b"".join(bytes(memoryview(q).cast("B")) for q in eng._program(st))
            for _ in range(n)]
    return keys, st


def test_schedule_shortcut_equals_full_recurrence():
    from pyheadtail_amd.engine import FusedOneTurnMap
    ob = W.nonlinear_bunch(100, seed=1)
    for mach in (nonlinear_machine(), linear_machine(), nonlinear_machine(n_turns_wake=8)):
        eng = FusedOneTurnMap(mach)
        b1, b2 = _beam(ob), _beam(ob)
        progs, table = eng._schedule(b1, 300)
        keys, st = _full(eng, b2, 300)
        assert [progs[i] for i in table] == keys
        assert (b1._gamma, b1._p0) == (st._gamma, st._p0)


def test_rf_kick_p0_round_trip_matches_oracle():
    """Kick.track sets p0 through the Particles.p0 setter (gamma round
    trip); the host scalars follow the oracle's bit for bit."""
    from pyheadtail_amd.engine import FusedOneTurnMap
    from oracle import refpath as R
    ob = W.nonlinear_bunch(100, seed=2)
    eng = FusedOneTurnMap(nonlinear_machine(wake=False))
    b = _beam(ob)
    eng._schedule(b, 5)
    for _ in range(5):
        R.RFSystemsOracle(W.CIRCUMFERENCE, [W.H_RF], [W.V_RF], [0.], [W.ALPHA_C], W.GAMMA,
                          ob.charge).track(ob)
    assert (b._gamma, b._beta, b._p0) == (ob.gamma, ob.beta, ob.p0)
