"""Writes tests/golden/reference_tables.json: the small lookup tables the
reference declares as data (LUTparent / LUTchild, base/Constants.h:5-24;
trianglesCount and the triangles[256][16] triangulation,
mesh/ComputeTriangleIndices.cu:41-315; childrenVertexKind and edgeVertex,
:31-38; parentFaceKind, mesh/BuildMeshGeometry.cu:19-26).  Run in the build
container, where the reference sources are readable; the output is a fixture
of numbers only.
"""
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"


def grab(path, name, count):
    src = open(os.path.join(REF, path), encoding="latin-1").read()
    m = re.search(name + r"[^=]*=\s*\{(.*?)\};", src, re.S)
    nums = [int(x) for x in re.findall(r"-?\d+", m.group(1))]
    assert len(nums) == count, (name, len(nums))
    return nums


out = {
    "LUTparent": grab("base/Constants.h", r"LUTparent_Host\[8\]\[27\]", 216),
    "LUTchild": grab("base/Constants.h", r"LUTchild_Host\[8\]\[27\]", 216),
    "trianglesCount": grab("mesh/ComputeTriangleIndices.cu", r"trianglesCount\[256\]", 256),
    "triangles": grab("mesh/ComputeTriangleIndices.cu", r"int triangles\[256\]\[16\]", 4096),
    "childrenVertexKind": grab("mesh/ComputeTriangleIndices.cu", r"childrenVertexKind\[8\]", 8),
    "edgeVertex": grab("mesh/ComputeTriangleIndices.cu", r"edgeVertex\[12\]\[2\]", 24),
    "parentFaceKind": grab("mesh/BuildMeshGeometry.cu", r"__constant__ int parentFaceKind\[8\]\[6\]", 48),
}
here = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(here, "reference_tables.json"), "w") as f:
    json.dump(out, f)
print("wrote", {k: len(v) for k, v in out.items()})
