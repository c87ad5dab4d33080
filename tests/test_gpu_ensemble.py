"""GPU: the batched ensemble (include/hc.h hc_ens_*) against each member
stepped alone by the device-resident loop.

Members are uneven (33..100 cells a side), with their own DEM, Manning n,
rain and adaptive dt, so they take different numbers of steps and finish at
different ensemble steps.  Bar: every member's depth/discharge fields, clock
and dt series bit-identical to its standalone run; outlet discharge and the
output series to 1e-12 (the edge-sum grouping follows the tile height).
"""
import numpy as np
import pytest

import hail_caesar_amd.inputs as inputs
from hail_caesar_amd import hc
from hail_caesar_amd.model import Ensemble, Model

pytestmark = pytest.mark.gpu

SIZES = [40, 70, 33, 100, 57]


def _members(wet):
    specs = inputs.ensemble_members(len(SIZES), hours=1, sizes=SIZES, seed0=4100,
                                    wet_depth=0.05 if wet else None)
    return specs


# precision forms: FP64; FP32 with FP64 elevations; FP32 with fixed-point elevations
FIXED = "fp32_fixed"


def _model(spec, prec):
    arrays, params = spec
    if prec == FIXED:
        return Model(None, None, precision=hc.HC_FP32, arrays=arrays, params=params,
                     device_loop=True, use_graph=False, elev_fixed=True)
    return Model(None, None, precision=prec, arrays=arrays, params=params, device_loop=True,
                 use_graph=False, flat_dz=-1.0 if prec == hc.HC_FP32 else 0.0)


def _run_alone(spec, prec):
    m = _model(spec, prec)
    parts = []
    while True:
        parts.append(m.iterate(500))
        if m.finished:
            break
    return m, [np.concatenate([p[k] for p in parts]) for k in range(3)]


def _run_ensemble(specs, prec, chunk, tune=False):
    ms = [_model(s, prec) for s in specs]
    ens = Ensemble(ms)
    if tune:
        ens.tune()
    acc = [[[], [], []] for _ in ms]
    for _ in range(10000):
        res = ens.iterate(chunk, use_graph=True)
        for a, r in zip(acc, res):
            for k in range(3):
                a[k].append(r[k])
        if all(m.finished for m in ms):
            break
    return ms, ens, [[np.concatenate(a[k]) for k in range(3)] for a in acc]


@pytest.mark.parametrize("prec,wet", [(hc.HC_FP64, False), (hc.HC_FP64, True), (hc.HC_FP32, True),
                                      (FIXED, True)])
def test_ensemble_equals_members_alone(prec, wet):
    specs = _members(wet)
    alone = [_run_alone(s, prec) for s in specs]
    ms, ens, series = _run_ensemble(specs, prec, chunk=250)
    steps = [len(s[0]) for s in series]
    assert len(set(steps)) > 1, "members should finish at different steps"
    for (ma, (tt, cy, dt)), mb, (tt2, cy2, dt2) in zip(alone, ms, series):
        assert len(tt2) == len(tt)
        assert np.array_equal(cy2, cy) and np.array_equal(dt2, dt)
        np.testing.assert_allclose(tt2, tt, rtol=1e-12, atol=1e-15)
        for f in ("water_depth", "qx", "qy"):
            assert np.array_equal(mb.download(f), ma.download(f)), f
        np.testing.assert_allclose(mb.series(), ma.series(), rtol=1e-12, atol=1e-14)


def test_ensemble_tuned_tiles_same_results():
    """The autotuner may change tile rows and the persistent grid; the fields
    do not depend on either."""
    specs = _members(True)
    alone = [_run_alone(s, hc.HC_FP32) for s in specs]
    ms, ens, series = _run_ensemble(specs, hc.HC_FP32, chunk=400, tune=True)
    for (ma, (tt, cy, dt)), mb, (tt2, cy2, dt2) in zip(alone, ms, series):
        assert np.array_equal(cy2, cy)
        np.testing.assert_allclose(tt2, tt, rtol=1e-12, atol=1e-15)
        for f in ("water_depth", "qx", "qy"):
            assert np.array_equal(mb.download(f), ma.download(f)), f


@pytest.mark.parametrize("form", ["fp32_z32_flat", "fp32_z64", FIXED])
def test_wide_members_aligned_equals_overlapping(form):
    """Members wider than one four-wavefront tile column (300 and 520 cells):
    stepped alone they use the aligned FP32 tiles, whose edge columns come
    from the neighbouring wavefronts through LDS and, on a block's sides, from
    one-row-per-lane column loads, with the east faces recomputed lane-batched;
    in the ensemble they use the overlapping tiles, whose halo lanes compute
    every face themselves.  Both must give the same bits, rain input at the
    halo columns included."""
    specs = inputs.ensemble_members(2, hours=1, sizes=[300, 520], seed0=4200, wet_depth=0.05)

    def model(spec):
        arrays, params = spec
        if form == FIXED:
            return Model(None, None, precision=hc.HC_FP32, arrays=arrays, params=params,
                         device_loop=True, use_graph=False, elev_fixed=True)
        return Model(None, None, precision=hc.HC_FP32, arrays=arrays, params=params,
                     device_loop=True, use_graph=False,
                     flat_dz=1e-3 if form == "fp32_z32_flat" else -1.0)

    alone = []
    for spec in specs:
        m = model(spec)
        parts = []
        while True:
            parts.append(m.iterate(1000))
            if m.finished:
                break
        alone.append((m, [np.concatenate([p[k] for p in parts]) for k in range(3)]))
    ms = [model(s) for s in specs]
    ens = Ensemble(ms)
    acc = [[[], [], []] for _ in ms]
    for _ in range(10000):
        for a, r in zip(acc, ens.iterate(500, use_graph=True)):
            for k in range(3):
                a[k].append(r[k])
        if all(m.finished for m in ms):
            break
    for (ma, (tt, cy, dt)), mb, a in zip(alone, ms, acc):
        cy2, dt2 = np.concatenate(a[1]), np.concatenate(a[2])
        assert np.array_equal(cy2, cy) and np.array_equal(dt2, dt)
        for f in ("water_depth", "qx", "qy"):
            assert np.array_equal(mb.download(f), ma.download(f)), f


@pytest.mark.parametrize("pa,pb", [(hc.HC_FP64, hc.HC_FP32), (hc.HC_FP32, FIXED)])
def test_ensemble_rejects_mixed_modes(pa, pb):
    specs = _members(True)[:2]
    a = _model(specs[0], pa)
    b = _model(specs[1], pb)
    with pytest.raises(hc.HCError):
        Ensemble([a, b])
