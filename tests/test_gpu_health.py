"""GPU: failure detection in the run loop, forced from the host.

The reference has none (it only exits on bad inputs).  Here a loop body
that leaves a depth non-finite or below -1e-4 m reports an infinite maximum
depth, and the next body is not run; neither is a body whose flow timestep
(the Courant limit at the last maximum depth) would fall below the floor.
Both stop the run with HC_ERR_UNSTABLE, in the device-resident loop (the
control kernel halts it; hc_loop_sync reports), in the host-stepped driver
and per member of a batched ensemble.  The conditions are planted in the
uploaded depth field (NaN, +inf, a negative depth, a depth of 1e9 m) or by a
floor above the run's timestep; every kernel runs to its end as usual.
"""
import numpy as np
import pytest

import hail_caesar_amd.inputs as inputs
from hail_caesar_amd import hc
from hail_caesar_amd.model import Ensemble, Model

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def wet_members():
    return inputs.ensemble_members(3, sizes=[64, 80, 96], hours=1, wet_depth=0.1)


def _model(member, device_loop, prec=hc.HC_FP64, **kw):
    arrays, params = member
    fp32 = prec == hc.HC_FP32
    return Model(None, None, precision=prec, arrays=arrays, params=params, device_loop=device_loop,
                 use_graph=device_loop, elev_fixed=fp32, **kw)


def _plant(m, value):
    """Set the deepest active interior cell's depth to `value`; returns it."""
    h = m.download("water_depth")
    mask = m.download("mask").astype(bool)
    inner = np.zeros_like(mask)
    inner[2:-2, 2:-2] = True
    cand = np.where(mask & inner, h, -np.inf)
    x, y = np.unravel_index(int(np.argmax(cand)), cand.shape)
    assert mask[x, y] and np.isfinite(cand[x, y])
    h[x, y] = value
    m.context().upload("water_depth", h)
    return x, y


@pytest.mark.parametrize("device_loop", [True, False])
@pytest.mark.parametrize("prec", [hc.HC_FP64, hc.HC_FP32])
@pytest.mark.parametrize("value", [np.nan, np.inf, -1.0])
def test_bad_depth_stops_the_run(wet_members, device_loop, prec, value):
    m = _model(wet_members[0], device_loop, prec)
    tt, _, _ = m.iterate(20)  # the active mask is first built at counter 5
    assert len(tt) == 20 and np.isfinite(m.scalar("maxdepth"))
    _plant(m, value)
    with pytest.raises(hc.HCError) as ei:
        m.iterate(20)
    assert ei.value.code == hc.HC_ERR_UNSTABLE, str(ei.value)
    assert "non-finite" in str(ei.value)
    # the body that met the planted depth ran; the next one did not
    assert len(ei.value.results[0]) == 1
    assert m.scalar("maxdepth") == np.inf
    # the run stays stopped
    with pytest.raises(hc.HCError) as ei2:
        m.iterate(5)
    assert ei2.value.code == hc.HC_ERR_UNSTABLE
    m.close()


@pytest.mark.parametrize("device_loop", [True, False])
@pytest.mark.parametrize("prec", [hc.HC_FP64, hc.HC_FP32])
def test_huge_depth_trips_the_timestep_floor(wet_members, device_loop, prec):
    """A finite 1e9 m depth: the next Courant limit, 0.7 * 5 / sqrt(9.81e9)
    = 3.5e-5 s, is below the default floor of 1e-4 s."""
    m = _model(wet_members[0], device_loop, prec)
    m.iterate(20)
    _plant(m, 1e9)
    with pytest.raises(hc.HCError) as ei:
        m.iterate(20)
    assert ei.value.code == hc.HC_ERR_UNSTABLE, str(ei.value)
    assert "floor" in str(ei.value)
    assert len(ei.value.results[0]) == 1
    md = m.scalar("maxdepth")
    assert np.isfinite(md) and md > 1e8
    m.close()


@pytest.mark.parametrize("device_loop", [True, False])
def test_forced_floor_stops_before_the_first_body(wet_members, device_loop):
    """A floor of 60 s is above this run's Courant limit (0.7 * 5 /
    sqrt(9.81 * 0.1) = 3.5 s at the 0.1 m minimum depth): no body runs, and
    the same model with the floor off steps normally."""
    m = _model(wet_members[0], device_loop, dt_floor=60.0)
    c0, t0 = m.scalar("counter"), m.scalar("cycle")
    with pytest.raises(hc.HCError) as ei:
        m.iterate(10)
    assert ei.value.code == hc.HC_ERR_UNSTABLE, str(ei.value)
    assert "floor" in str(ei.value)
    assert len(ei.value.results[0]) == 0
    assert (m.scalar("counter"), m.scalar("cycle")) == (c0, t0)  # nothing of the body ran
    m.close()
    ok = _model(wet_members[0], device_loop, dt_floor=-1.0)
    tt, _, dt = ok.iterate(10)
    assert len(tt) == 10 and np.all(dt < 60.0)
    ok.close()


def test_sound_runs_match_with_and_without_the_tests(wet_members):
    """The tests change nothing in a sound run: the same 60 device-loop
    bodies with the default floor and with the timestep test off agree bit
    for bit."""
    out = []
    for fl in (0.0, -1.0):
        m = _model(wet_members[1], True, dt_floor=fl)
        tt, cy, dt = m.iterate(60)
        out.append((tt, cy, dt, m.download("water_depth")))
        m.close()
    for a, b in zip(out[0], out[1]):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("prec", [hc.HC_FP64, hc.HC_FP32])
def test_ensemble_member_stops_alone(wet_members, prec):
    """A NaN planted in member 1 of 3: that member halts after the body that
    met it; the other two run all their steps, with the records of a run
    without the fault."""
    ref = [_model(mb, True, prec) for mb in wet_members]
    ens_ref = Ensemble(ref)
    ens_ref.iterate(20)
    want = ens_ref.iterate(20)
    ms = [_model(mb, True, prec) for mb in wet_members]
    ens = Ensemble(ms)
    ens.iterate(20)
    _plant(ms[1], np.nan)
    with pytest.raises(hc.HCError) as ei:
        ens.iterate(20)
    assert ei.value.code == hc.HC_ERR_UNSTABLE and ei.value.member == 1, str(ei.value)
    got = ei.value.results
    assert len(got[1][0]) == 1
    for k in (0, 2):
        for a, b in zip(got[k], want[k]):
            np.testing.assert_array_equal(a, b)
    for e in (ens, ens_ref):
        e.close()
    for m in ms + ref:
        m.close()
