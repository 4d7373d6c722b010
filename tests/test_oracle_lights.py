"""The oracle's sphere, spot and triangle lights and its light grid against
the closed forms of light_cases.py, on the CPU: one light over a plane,
ReSTIR DI reuse off, four frames of a still camera (the grid's weight
converges as 1 - 9^-k inside it)."""
import numpy as np
import pytest

import light_cases as lc
from raytracedgl1_amd import ffi
from oracle_lib import OracleFramebuffers, OraclePathTracer
from test_oracle_pathtrace import params_for, W, H

P = ffi.PLANE


@pytest.mark.parametrize("name", sorted(lc.CASES))
def test_oracle_light_closed_form(oracle, name):
    lights, over = lc.CASES[name]
    scene = lc.plane_scene(lights)
    fb = OracleFramebuffers(oracle, W, H)
    pt = OraclePathTracer(oracle, scene, W, H)
    cell = 1.0
    seen_in = False
    for f in range(4):
        if f:
            fb.swap()
        p = params_for(lc.CAM, f, restirTemporalSamples=0, restirSpatialSamples=0, cellWorldSize=cell, **over)
        pt.frame(p, fb, threads=8)
        o = fb.cur
        m_in, _ = lc.check(name, o[P["UNFILTERED_DIRECT"]], o[P["SURFACE_POSITION"]], o[P["IS_SKY"]], f, p.cameraPosition, cell,
                           p.polyLightSpotlightFactor, "oracle ")
        seen_in = seen_in or m_in == m_in
    assert seen_in
