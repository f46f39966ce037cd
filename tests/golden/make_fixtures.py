"""Regenerate the golden fixtures under tests/golden/ from the n2p2 tree.

Run in a container that has the reference tree (default /root/reference):

    python tests/golden/make_fixtures.py [/path/to/n2p2]

What it writes (data only -- numbers and the example's input files):

* ``H2O_RPBE-D3/`` -- the model and structure files of
  ``examples/nnp-predict/H2O_RPBE-D3`` (input.nn, scaling.data,
  weights.001.data, weights.008.data, input.data), copied verbatim, and
  ``expected.json`` with the energy and the 1080 force vectors that
  ``test/cpp/ExamplePrediction.h:34-1117`` holds for them (checked by the
  reference at 1000*DBL_EPSILON, ``test/cpp/test_Prediction.cpp:25,44-56``).
* ``Cu2S_PBE/`` -- the same for ``examples/nnp-predict/Cu2S_PBE`` and
  ``ExamplePrediction.h:1118-1265`` (SF types 2, 3 and 9, CT_POLY2 cutoff,
  SCALESIGMA scaling, softplus networks).
* ``Anisole_SCAN/``, ``DMABN_SCAN/``, ``Ethylbenzene_SCAN/`` -- the same for
  the compact-function models (SF types 20 and 22, polynomial and asymmetric
  polynomial subtypes, normalised units; ``ExamplePrediction.h:1266-1525,
  1526-1550, 1551-1846``).
* ``LJ_checkdw/`` -- the input files of ``examples/nnp-checkdw/LJ`` (101 Ne
  dimers, radial functions, a 5-3-1 network), the example the reference's
  nnp-checkdw tool test runs (``test/cpp/Example_nnp_checkdw.h:21-24``);
  ``examples/nnp-checkf/LJ`` holds the same files.
* ``H2O_RPBE-D3_checkf/`` -- the input files of
  ``examples/nnp-checkf/H2O_RPBE-D3`` (three 48-atom periodic water cells,
  normalised units), run by the reference's nnp-checkf tool test
  (``test/cpp/Example_nnp_checkf.h:18-20``).
* ``sf_kat.json`` -- the symmetry-function known answers (types 2, 3, 9, 12,
  13, 20-25) of
  ``test/cpp/fixtures.h:151-258`` with the cluster geometries of
  ``fixtures.h:11-149`` (values for atom 0, CT_TANHU, no scaling, 1000*eps).
* ``cutoff_kat.json`` -- f and df of all cutoff types at rc=3, alpha=0.2,
  r=2.4 from ``test/cpp/test_CutoffFunction.cpp:17-50`` (100*eps).
"""
import json
import os
import re
import shutil
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def floats_in(text):
    return [float(x) for x in re.findall(r"[-+]?\d+\.\d*(?:[eE][-+]?\d+)?|[-+]?\d+[eE][-+]?\d+", text)]


def block(text, name):
    """Body of a C++ initializer `name = { ... };`."""
    m = re.search(re.escape(name) + r"\s*=\s*\{(.*?)\};", text, re.S)
    if not m:
        raise RuntimeError("initializer %s not found" % name)
    return m.group(1)


def example(ref, name, count):
    """Copy an nnp-predict example's files and write its golden energy and
    forces (the ExamplePrediction.h block for `name`)."""
    ex = os.path.join(ref, "examples", "nnp-predict", name)
    out = os.path.join(HERE, name)
    os.makedirs(out, exist_ok=True)
    for fn in sorted(os.listdir(ex)):
        if fn in ("input.nn", "scaling.data", "input.data") or re.match(r"weights\.\d{3}\.data$", fn):
            shutil.copyfile(os.path.join(ex, fn), os.path.join(out, fn))
    text = open(os.path.join(ref, "test", "cpp", "ExamplePrediction.h")).read()
    start = text.index('ExamplePrediction("%s")' % name)
    stop = text.find("examples.push_back", start + 10)
    part = text[start:stop if stop > 0 else len(text)]
    line0 = text[:start].count("\n") + 1
    energy = float(re.search(r"e->energy\s*=\s*([-+0-9.eE]+);", part).group(1))
    forces = [floats_in(m) for m in re.findall(r"Vec3D\(([^)]*)\)", part)]
    assert len(forces) == count and all(len(f) == 3 for f in forces), (name, len(forces))
    json.dump({"source": "test/cpp/ExamplePrediction.h:%d-%d" % (line0, line0 + part.count("\n") - 1),
               "tolerance": "1000*DBL_EPSILON absolute (test/cpp/test_Prediction.cpp:25)",
               "energy": energy, "forces": forces},
              open(os.path.join(out, "expected.json"), "w"), indent=0)


def tool_example(ref, tool, name):
    """Copy an nnp-checkdw / nnp-checkf example's input files (no expected
    values: the reference's tool tests only check that the output files
    exist)."""
    ex = os.path.join(ref, "examples", "nnp-" + tool, name)
    out = os.path.join(HERE, name + "_" + tool)
    os.makedirs(out, exist_ok=True)
    for fn in sorted(os.listdir(ex)):
        if fn in ("input.nn", "scaling.data", "input.data") or re.match(r"weights\.\d{3}\.data$", fn):
            shutil.copyfile(os.path.join(ex, fn), os.path.join(out, fn))


def main(ref):
    example(ref, "H2O_RPBE-D3", 1080)
    example(ref, "Cu2S_PBE", 144)
    example(ref, "Anisole_SCAN", 256)
    example(ref, "DMABN_SCAN", 21)
    example(ref, "Ethylbenzene_SCAN", 288)
    tool_example(ref, "checkdw", "LJ")
    tool_example(ref, "checkf", "H2O_RPBE-D3")

    fx = open(os.path.join(ref, "test", "cpp", "fixtures.h")).read()
    kat = {"source": "test/cpp/fixtures.h:11-258; test/cpp/test_SymmetryFunctions.cpp:36,68-74",
           "cutoff_type": 2, "cutoff_alpha": 0.0, "neighbor_cutoff": 10.0,
           "tolerance": "1000*DBL_EPSILON absolute", "sets": {}}
    geoms = {
        # fixtures.h:11-149 (positions, element indices in ElementMap order)
        "Mono": {"elements": ["H"], "pos": [[0, 0, 0], [1, 2, 1], [0, 3, -1]], "elem": [0, 0, 0]},
        "Dual": {"elements": ["H", "He"], "pos": [[0, 0, 0], [1, 2, 1], [0, 3, -1]], "elem": [0, 1, 1]},
        "Mixed": {"elements": ["S", "Cu", "U"],
                  "pos": [[0, 0, 0], [1, 2, 1], [0, 3, -1], [-2, 1, -0.5]], "elem": [0, 0, 1, 2]},
    }
    for name in ("Mono", "Dual", "Mixed"):
        types = [int(x) for x in re.findall(r"\d+", block(fx, "types" + name))]
        lines = re.findall(r'"([^"]*)"', block(fx, "setupLines" + name))
        vals = floats_in(block(fx, "values" + name))
        assert len(types) == len(lines) == len(vals)
        rows = [{"type": t, "line": l, "value": v} for t, l, v in zip(types, lines, vals) if t in (2, 3, 9, 12, 13, 20, 21, 22, 23, 24, 25)]
        kat["sets"][name] = dict(geoms[name], functions=rows)
    json.dump(kat, open(os.path.join(HERE, "sf_kat.json"), "w"), indent=1)

    tc = open(os.path.join(ref, "test", "cpp", "test_CutoffFunction.cpp")).read()
    f = floats_in(block(tc, "vector<double> const f"))
    df = floats_in(block(tc, "vector<double> const df"))
    json.dump({"source": "test/cpp/test_CutoffFunction.cpp:17-50",
               "rc": 3.0, "alpha": 0.2, "r": 2.4, "tolerance": "100*DBL_EPSILON",
               "types": ["CT_HARD", "CT_COS", "CT_TANHU", "CT_TANH", "CT_EXP",
                         "CT_POLY1", "CT_POLY2", "CT_POLY3", "CT_POLY4"],
               "f": f, "df": df},
              open(os.path.join(HERE, "cutoff_kat.json"), "w"), indent=1)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
