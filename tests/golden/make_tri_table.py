"""Writes tri_table.json: the 256 x 16 marching-cubes triangulation table as the
reference holds it (marching_cubes.h:98-355), read as data from the reference checkout,
so that the GPU mesher can be checked on the very table the reference passes to it.

Run in the build container, where the reference is mounted:
    python tests/golden/make_tri_table.py /root/reference/marching_cubes.h
"""
import json
import os
import re
import sys


def main(path):
    text = open(path).read()
    body = text[text.index("TRIANGLE_TABLE[256][16]"):]
    body = body[body.index("{") + 1:body.index("};")]
    vals = [int(v) for v in re.findall(r"-?\d+", body)]
    assert len(vals) == 256 * 16, len(vals)
    rows = [vals[16 * i:16 * i + 16] for i in range(256)]
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tri_table.json")
    with open(out, "w") as f:
        json.dump({"source": "marching_cubes.h:98-355 TRIANGLE_TABLE", "rows": rows}, f)
    print("wrote", out)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/marching_cubes.h")
