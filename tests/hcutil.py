"""Shared helpers for the parity tests: build oracle states and mirror them
onto a device context."""
from __future__ import annotations

import os

import numpy as np

import hail_caesar_amd.inputs as inputs
from pyoracle import OracleModel, orc_params

STATE_FIELDS = ("water_depth", "elev", "qx", "qy")


def hc_params_from(p: orc_params, flat_dz: float = 0.0, elev_fixed: bool = False):
    from hail_caesar_amd import hc
    return hc.default_params(
        courant_number=p.courant_number, froude_limit=p.froude_limit,
        hflow_threshold=p.hflow_threshold, edgeslope=p.edgeslope, mannings=p.mannings,
        spatially_var_mannings=p.spatially_var_mannings,
        water_depth_erosion_threshold=p.water_depth_erosion_threshold,
        ERODEFACTOR=p.ERODEFACTOR, active=p.active, Csuspmax=p.Csuspmax, max_vel=p.max_vel,
        transport_law=1 if p.wilcock else 2, d=list(p.d), dprop=list(p.dprop),
        fallVelocity=list(p.fallVelocity), bedrock_erosion_threshold=p.bedrock_erosion_threshold,
        bedrock_erosion_rate=p.bedrock_erosion_rate, lateral_constant=p.lateral_constant,
        chann_lateral_erosion=p.chann_lateral_erosion, veg_lat_restriction=p.veg_lat_restriction,
        flat_dz=flat_dz, elev_fixed=1 if elev_fixed else 0)


def wet_flow_state(tmp, size=64, steps=400, seed=1001, spatial_n=False):
    """Config-1-style workload advanced `steps` iterations by the oracle."""
    w = inputs.flow_only_256(os.path.join(str(tmp), f"flow{size}"), size=size, hours=2, seed=seed)
    if spatial_n:
        rng = np.random.Generator(np.random.PCG64(seed + 7))
        n = inputs.lognormal_mannings(size, size, 0.035, 0.3, rng)
        w = inputs.make_workload(w.path, w.dem, w.dx, np.array([10.0, 0.0]), 60, 2,
                                 mannings=n, name="flow_n")
    o = OracleModel.from_workload(w)
    if steps:
        o.iterate(steps)
    return w, o


def to_context(o: OracleModel, precision=0, flat_dz: float = 0.0, sediment: bool = False,
               elev_fixed: bool = False):
    from hail_caesar_amd import hc
    p = o.p
    ctx = hc.Context(p.imax, p.jmax, p.DX, hc_params_from(p, flat_dz, elev_fixed), precision,
                     sediment)
    for f in STATE_FIELDS:
        ctx.upload(f, o.field(f))
    if p.spatially_var_mannings:
        ctx.upload("mannings", o.field("mannings"))
    ctx.upload("mask", o.field_int("mask"))
    return ctx


def rel_err(a, b, floor=1e-300):
    d = np.abs(a - b)
    s = np.maximum(np.abs(b), floor)
    return float(np.max(d / s)) if d.size else 0.0
