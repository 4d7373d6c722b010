"""The oracle's reflection / refraction chain (oracle/oracle_pathtrace.cpp,
ReflectRefract) on the CPU against the same closed forms as the GPU path
(reflrefr_cases.py): mirror law and path lengths, both Schlick Fresnel
forms, Snell's law and total internal reflection, Beer-Lambert media, the
checkerboard split and the depth limit; and the refracting range skipped by
shadow rays."""
import pytest

from raytracedgl1_amd.ffi import PLANE
from oracle_lib import OracleFramebuffers, OraclePathTracer
import reflrefr_cases as rc


def _run(oracle, scene, flags, params, emissive=None):
    fb = OracleFramebuffers(oracle, rc.W, rc.H)
    pt = OraclePathTracer(oracle, scene, rc.W, rc.H, flags=flags, emissive=emissive)
    pt.frame(params, fb, threads=8)
    return fb.cur


@pytest.mark.parametrize("case", list(rc.CASES))
def test_oracle_chain_closed_forms(oracle, case):
    flags, settings, fovy, check, *emissive = rc.CASES[case]
    p = rc.frame_params(rc.camera(fovy), **settings)
    cur = _run(oracle, rc._planes_scene(), flags, p, emissive[0] if emissive else None)
    out = {k: cur[PLANE[k]] for k in rc.NAMES}
    out["params"] = p
    check(out)


@pytest.mark.parametrize("pane_flags, lit", rc.SHADOW_CASES)
def test_oracle_refracting_geometry_casts_no_shadow(oracle, pane_flags, lit):
    scene, flags = rc.sun_scene(pane_flags)
    cur = _run(oracle, scene, flags, rc.frame_params(rc.SUN_CAMERA))
    rc.check_shadow(cur[PLANE["UNFILTERED_DIRECT"]], cur[PLANE["SURFACE_POSITION"]], pane_flags, lit)
