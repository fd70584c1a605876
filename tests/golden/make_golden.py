"""Writes the golden vectors under tests/golden/ from data tables held in the reference's source tree.

The reference (Goubermouche/VFD) has no tests or fixtures of its own for the DFSPH path. The only reference-held
data that pin parts of the path are two numeric tables, read here as text (nothing is compiled or run):

  - VFD/Source/Simulation/DFSPH/HaltonVec323.cuh: 16,384 unit-sphere sample points (49,152 floats) used by the
    Monte-Carlo surface tension classifier (DFSPHKernels.cu:917-924).
  - VFD/Source/Core/Math/GaussQuadrature.cpp: Gauss-Legendre abscissae/weights; row p = 30 (16 points) is the one
    the volume-map build uses (RigidBody.cu:71, GaussQuadrature.cpp:5621-5655).

  - VFD/Resources/Scenes/New/DFSPH/default.json: the DFSPH example scene (cereal JSON of an entt snapshot), whose
    simulation component stores a DFSPHSimulationDescription (DFSPHSimulationComponent.h:7-39).

  - VFD/Resources/Models/{Cube,Sphere,Cone,Torus}.obj: triangle meshes (Wavefront OBJ) the example scenes use as
    rigid bodies and fluid objects; Cube.obj and Cone.obj are the two meshes of the DFSPH example scene.

Outputs (small fixtures, committed):
  halton_samples.json    every 97th point of the table plus the first and last 64 (index -> xyz)
  halton_table.npz       the whole table: 49,152 values as written in the source (float64 "values")
  gauss_legendre_16.json the 16 abscissae and weights of row p = 30
  scene_default.json     the example scene file, byte for byte (input of the scene-description reader)
  models/*.obj           the four meshes, byte for byte (inputs of the mesh distance and sampler tests)

Run in a container where the reference is mounted:  python tests/golden/make_golden.py [/root/reference]
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def _numbers(text):
    return [float(t) for t in re.findall(r"[-+]?\d+\.\d+(?:[eE][-+]?\d+)?", text)]


def halton(ref):
    path = os.path.join(ref, "VFD/Source/Simulation/DFSPH/HaltonVec323.cuh")
    text = open(path, encoding="utf-8", errors="replace").read()
    body = text[text.index("{") + 1: text.rindex("}")]
    vals = _numbers(body)
    assert len(vals) == 49152, len(vals)
    idx = sorted(set(list(range(64)) + list(range(0, 16384, 97)) + list(range(16384 - 64, 16384))))
    return {
        "source": "VFD/Source/Simulation/DFSPH/HaltonVec323.cuh",
        "count": 16384,
        "points": {str(i): vals[3 * i: 3 * i + 3] for i in idx},
    }


def halton_table(ref):
    path = os.path.join(ref, "VFD/Source/Simulation/DFSPH/HaltonVec323.cuh")
    text = open(path, encoding="utf-8", errors="replace").read()
    vals = _numbers(text[text.index("{") + 1: text.rindex("}")])
    assert len(vals) == 49152, len(vals)
    return vals


def gauss_legendre(ref):
    path = os.path.join(ref, "VFD/Source/Core/Math/GaussQuadrature.cpp")
    lines = open(path, encoding="utf-8", errors="replace").read().splitlines()

    def table_row(decl, p):
        start = next(i for i, l in enumerate(lines) if decl in l)
        marker = next(i for i in range(start, len(lines)) if lines[i].strip() == "// p = %d" % p)
        row = []
        for l in lines[marker + 1:]:
            row += _numbers(l)
            if "}" in l:
                break
        return row

    x = table_row("GaussianAbscissae1", 30)
    w = table_row("GaussianWeights1", 30)
    assert len(x) == 16 and len(w) == 16, (len(x), len(w))
    return {"source": "VFD/Source/Core/Math/GaussQuadrature.cpp (p = 30)", "abscissae": x, "weights": w}


def scene(ref):
    with open(os.path.join(ref, "VFD/Resources/Scenes/New/DFSPH/default.json"), "rb") as f:
        return f.read()


MODELS = ("Cube.obj", "Sphere.obj", "Cone.obj", "Torus.obj")


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    with open(os.path.join(HERE, "scene_default.json"), "wb") as f:
        f.write(scene(ref))
    os.makedirs(os.path.join(HERE, "models"), exist_ok=True)
    for name in MODELS:
        with open(os.path.join(ref, "VFD/Resources/Models", name), "rb") as f:
            data = f.read()
        with open(os.path.join(HERE, "models", name), "wb") as f:
            f.write(data)
    with open(os.path.join(HERE, "halton_samples.json"), "w") as f:
        json.dump(halton(ref), f, separators=(",", ":"))
    import numpy as np
    np.savez_compressed(os.path.join(HERE, "halton_table.npz"), values=np.array(halton_table(ref), np.float64))
    with open(os.path.join(HERE, "gauss_legendre_16.json"), "w") as f:
        json.dump(gauss_legendre(ref), f, indent=1)
    print("wrote halton_samples.json, halton_table.npz, gauss_legendre_16.json, scene_default.json, models/")


if __name__ == "__main__":
    main()
