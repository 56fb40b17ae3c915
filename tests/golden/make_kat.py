"""Extracts the reference's InterpolateArray known-answer data into a JSON fixture.

Reads LinearBasis/tests/InterpolateArray/test.cpp of the reference checkout as
text (init(): the 59 input coordinates; check(): the 118 expected values,
printed there to 6 significant digits and compared with EPSILON = 0.001) and
copies the binary fixture surplus.plt it evaluates. Run once in the build
container: python tests/golden/make_kat.py /root/reference
"""
import json
import os
import re
import shutil
import sys

ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
src = os.path.join(ref, "LinearBasis", "tests", "InterpolateArray")
text = open(os.path.join(src, "test.cpp")).read()
x = {int(m.group(1)): float(m.group(2)) for m in re.finditer(r"input\[(\d+)\]\s*=\s*([-0-9.eE+]+);", text)}
y = {int(m.group(1)): float(m.group(2)) for m in re.finditer(r"expected\[(\d+)\]\s*=\s*([-0-9.eE+]+);", text)}
eps = float(re.search(r"#define EPSILON\s+([0-9.eE+-]+)", text).group(1))
out = {
    "source": "LinearBasis/tests/InterpolateArray/test.cpp init() and check(); fixture surplus.plt; hddm-solver.cfg n_agents 59",
    "nagents": 59,
    "epsilon": eps,
    "x": [x[i] for i in range(len(x))],
    "expected": [y[i] for i in range(len(y))],
}
here = os.path.dirname(os.path.abspath(__file__))
json.dump(out, open(os.path.join(here, "interpolate_array_kat.json"), "w"), indent=1)
shutil.copyfile(os.path.join(src, "surplus.plt"), os.path.join(here, "surplus.plt"))
print("x:", len(out["x"]), "expected:", len(out["expected"]), "epsilon:", eps)
