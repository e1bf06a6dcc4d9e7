"""Extract the numeric tables the reference holds for this path into a
fixture (reference_tables.json).  Run in a container that has the reference
checkout; the fixture is committed and the tests never read the reference.

Tables (data only):
  * ke_a, ke_nu: the 300-entry integer vectors whose combination
    (A + nu*B) / ((1+nu)(2nu-1)144) fills the lower triangle of Ke0,
    column-major (Cantilever/omp/top3D125mod.m:50-76).
  * p1: the 8x27 trilinear prolongation table P1 of one coarse element
    (Cantilever/omp/mgcg1.c:1185-1192).
"""
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def _numbers(text):
    return [float(t) for t in re.findall(r"[-+]?\d+(?:\.\d+)?", text)]


def ke_tables():
    src = open(os.path.join(REF, "Cantilever/omp/top3D125mod.m")).read()
    start = src.index("Ke = 1/(1+nu)/(2*nu-1)/144")
    body = src[start:src.index("Ke0( tril", start)]
    # drop MATLAB comments, keep the two bracketed vectors
    body = "\n".join(line.split("%")[0] for line in body.splitlines())
    vecs = re.findall(r"\[(.*?)\]", body, flags=re.S)
    assert len(vecs) == 2, len(vecs)
    a = [int(v) for v in _numbers(vecs[0].replace("...", ""))]
    b = [int(v) for v in _numbers(vecs[1].replace("...", ""))]
    assert len(a) == 300 and len(b) == 300, (len(a), len(b))
    return a, b


def p1_table():
    src = open(os.path.join(REF, "Cantilever/omp/mgcg1.c")).read()
    start = src.index("double P1[8][27]")
    body = src[start + len("double P1[8][27]"):src.index(";", start)]
    vals = _numbers(body)
    assert len(vals) == 8 * 27, len(vals)
    return [vals[27 * r:27 * (r + 1)] for r in range(8)]


def main():
    a, b = ke_tables()
    out = {
        "source": {
            "ke": "Cantilever/omp/top3D125mod.m:50-76",
            "p1": "Cantilever/omp/mgcg1.c:1185-1192",
        },
        "ke_a": a,
        "ke_nu": b,
        "p1": p1_table(),
    }
    with open(os.path.join(HERE, "reference_tables.json"), "w") as f:
        json.dump(out, f)
    print("wrote reference_tables.json")


if __name__ == "__main__":
    main()
