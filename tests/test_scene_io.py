"""Scene-description I/O through the C ABI (vfd_read_descriptions / vfd_write_description; no GPU needed).

The fixture is VFD's own DFSPH example scene (tests/golden/scene_default.json, copied by make_golden.py): an entt
snapshot in cereal JSON whose simulation component holds one description (DFSPHSimulationComponent.h:7-39). The
expected values come from reading the same text with Python's json module, independently of the C reader."""
import json
import os

import numpy as np
import pytest

from vfd_amd import VfdError, read_descriptions, write_description
from vfd_amd.simulation import DFSPHSimulationDescription

HERE = os.path.dirname(os.path.abspath(__file__))
SCENE = os.path.join(HERE, "golden", "scene_default.json")

# cereal name -> vfd_desc field (DFSPHSimulationComponent.h:11-37)
NAMES = {
    "timeStepSize": "TimeStepSize", "minTimeStepSize": "MinTimeStepSize", "maxTimeStepSize": "MaxTimeStepSize",
    "frameLength": "FrameLength", "frameCount": "FrameCount",
    "minPressureSolverIterations": "MinPressureSolverIterations",
    "maxPressureSolverIterations": "MaxPressureSolverIterations", "maxPressureSolverError": "MaxPressureSolverError",
    "enableDivergenceSolverError": "EnableDivergenceSolverError",
    "minDivergenceSolverIterations": "MinDivergenceSolverIterations",
    "maxDivergenceSolverIterations": "MaxDivergenceSolverIterations",
    "maxDivergenceSolverError": "MaxDivergenceSolverError", "enableViscositySolver": "EnableViscositySolver",
    "minViscositySolverIterations": "MinViscositySolverIterations",
    "maxViscositySolverIterations": "MaxViscositySolverIterations",
    "maxViscositySolverError": "MaxViscositySolverError", "viscosity": "Viscosity",
    "boundaryViscosity": "BoundaryViscosity", "tangentialDistanceFactor": "TangentialDistanceFactor",
    "enableSurfaceTensionSolver": "EnableSurfaceTensionSolver",
    "surfaceTensionSmoothPassCount": "SurfaceTensionSmoothPassCount", "surfaceTension": "SurfaceTension",
    "temporalSmoothing": "TemporalSmoothing", "CSDFix": "CSDFix", "CSD": "CSD", "particleRadius": "ParticleRadius",
}


def _fields(desc):
    out = {}
    for name, _ in type(desc)._fields_:
        v = getattr(desc, name)
        out[name] = tuple(np.float32(v[k]) for k in range(3)) if name == "Gravity" else v
    return out


def _find_descriptions(node, found):
    if isinstance(node, dict):
        for k, v in node.items():
            if k == "description" and isinstance(v, dict):
                found.append(v)
            else:
                _find_descriptions(v, found)
    elif isinstance(node, list):
        for v in node:
            _find_descriptions(v, found)


def test_reads_the_reference_example_scene():
    descs = read_descriptions(SCENE)
    expected = []
    _find_descriptions(json.load(open(SCENE)), expected)
    assert len(descs) == len(expected) == 1
    d, e = descs[0], expected[0]
    assert set(e) == set(NAMES) | {"gravity"}
    for cname, field in NAMES.items():
        got = getattr(d, field)
        want = e[cname]
        if isinstance(want, bool):
            assert got == int(want), cname
        elif isinstance(got, float):
            assert np.float32(got) == np.float32(want), cname
        else:
            assert got == want, cname
    assert tuple(np.float32(d.Gravity[k]) for k in range(3)) == tuple(
        np.float32(e["gravity"][a]) for a in "xyz")
    # the scene differs from the defaults in the fields its author changed
    assert d.FrameLength == np.float32(0.016) and d.EnableSurfaceTensionSolver == 0


def test_write_then_read_gives_the_same_description():
    d = DFSPHSimulationDescription(ParticleRadius=0.004, Viscosity=0.01, MaxPressureSolverError=0.01,
                                   MinDivergenceSolverIterations=2, CSDFix=17, EnableViscositySolver=0,
                                   Gravity=(0.5, -9.81, 1.0 / 3.0))
    text = write_description(d)
    assert json.loads(text)["description"]["particleRadius"] == float(np.float32(0.004))
    back = read_descriptions(text)
    assert len(back) == 1 and _fields(back[0]) == _fields(d)


def test_missing_fields_keep_the_defaults_and_unknown_ones_are_skipped():
    text = '{"a": [1, {"b": null}], "x": {"description": {"viscosity": 0.5, "someFutureField": {"k": [1, 2]}}},' \
           ' "y": {"description": {}}}'
    descs = read_descriptions(text)
    assert len(descs) == 2
    ref = _fields(DFSPHSimulationDescription())
    first = _fields(descs[0])
    assert first.pop("Viscosity") == np.float32(0.5)
    assert first == {k: v for k, v in ref.items() if k != "Viscosity"}
    assert _fields(descs[1]) == ref


@pytest.mark.parametrize("text", ['{"description": {"frameCount": -3}}', '{"description": {"viscosity": }}',
                                  '{"description": {"viscosity": 1.0', '{"description": {"gravity": 3}}',
                                  '[1, 2] trailing'])
def test_malformed_text_is_an_invalid_argument(text):
    with pytest.raises(VfdError, match="INVALID_ARGUMENT"):
        read_descriptions(text)


def test_scene_file_entities_become_mesh_inputs():
    """The example scene's entities (Scene.cpp:10-40 snapshot order; editor bake, ComponentPanel.cpp:608-656): a
    Cone.obj fluid object (medium density, 20^3 SDF, flipped and lifted by 3) and a Cube.obj rigid body scaled
    (2, 0.2, 2)."""
    from vfd_amd import scenes
    d, fluids, bodies = scenes.load_scene_file(SCENE, os.path.join(HERE, "golden", "models"))
    assert np.float32(d.ParticleRadius) == np.float32(0.025)
    assert len(fluids) == 1 and len(bodies) == 1
    fo, rb = fluids[0], bodies[0]
    assert fo.mesh.triangles.shape == (62, 3) and fo.sample_mode == 1 and fo.resolution == (20, 20, 20)
    assert fo.transform[1, 3] == np.float32(3.0) and fo.transform[1, 1] == np.float32(-1.0)
    assert rb.is_mesh and rb.mesh.triangles.shape == (12, 3) and rb.resolution == (20, 20, 20)
    assert np.array_equal(np.diag(rb.transform), np.array([2.0, 0.2, 2.0, 1.0], np.float32))
    assert not rb.inverted and rb.padding == 0.0


@pytest.mark.parametrize("case", ["index", "empty", "mode", "radius"])
def test_mesh_sampler_rejects_bad_input(case):
    """Argument checks of vfd_sample_mesh_volume run before any GPU work."""
    from vfd_amd.mesh import TriangleMesh
    from vfd_amd.simulation import sample_mesh_volume
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], np.float32)
    t = np.array([[0, 2, 1], [0, 1, 3], [0, 3, 2], [1, 2, 3]], np.uint32)
    mesh, radius, mode = TriangleMesh(v, t), 0.1, 0
    if case == "index":
        mesh.triangles = np.array([[0, 2, 7]], np.uint32)
    elif case == "empty":
        mesh.triangles = np.zeros((0, 3), np.uint32)
    elif case == "mode":
        mode = 3
    else:
        radius = 0.0
    with pytest.raises(VfdError, match="INVALID_ARGUMENT"):
        sample_mesh_volume(mesh, radius, (8, 8, 8), False, mode)
