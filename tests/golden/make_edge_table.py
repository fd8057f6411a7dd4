"""Writes edge_table.json: the 256-entry marching-cubes edge table as the reference
holds it (marching_cubes.h:61-96), read as data from the reference checkout.

Run in the build container, where the reference is mounted:
    python tests/golden/make_edge_table.py /root/reference/marching_cubes.h
"""
import json
import os
import re
import sys


def main(path):
    text = open(path).read()
    body = text[text.index("EDGE_TABLE[256]"):]
    body = body[body.index("{") + 1:body.index("};")]
    vals = [int(v, 16) for v in re.findall(r"0x[0-9a-fA-F]+", body)]
    assert len(vals) == 256, len(vals)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "edge_table.json")
    with open(out, "w") as f:
        json.dump({"source": "marching_cubes.h:61-96 EDGE_TABLE", "values": vals}, f)
    print("wrote", out)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/marching_cubes.h")
