"""Writes tests/golden/mc_tables.json: SHA-256 fingerprints of the reference's
marching-cubes tables, read as text from the reference tree (run in a
container that has it; the fixture is committed, the tables are not):

  triTable[256 * 16]       traversehp.cl:74   (edge ids, -1 padded), as int8
  nrOfTriangles[256]       classifycubes.cl:37, as uint8
  offsets3[72]             traversehp.cl:35   (edge endpoints), as int8

tests/test_surface_cpu.py compares the library's tables (rtps_mc_case_table,
rtps_mc_edge_table) against these fingerprints.

    python3 tests/golden/make_mc_digest.py /root/reference
"""
import hashlib
import json
import os
import re
import sys

import numpy as np


def table(text: str, name: str) -> list[int]:
    m = re.search(r"\b" + name + r"\s*\[\s*\d*\s*\]\s*=\s*\{(.*?)\};", text, re.S)
    if not m:
        raise SystemExit(f"{name} not found")
    body = re.sub(r"//[^\n]*", "", m.group(1))
    return [int(v) for v in re.findall(r"-?\d+", body)]


def main(ref: str) -> None:
    cl = os.path.join(ref, "rtpslib", "system", "common", "cl_src")
    trav = open(os.path.join(cl, "traversehp.cl")).read()
    cls = open(os.path.join(cl, "classifycubes.cl")).read()
    tri = np.array(table(trav, "triTable"), np.int8)
    off = np.array(table(trav, "offsets3"), np.int8)
    ntri = np.array(table(cls, "nrOfTriangles"), np.uint8)
    assert tri.size == 4096 and off.size == 72 and ntri.size == 256
    out = {
        "triTable_int8_sha256": hashlib.sha256(tri.tobytes()).hexdigest(),
        "nrOfTriangles_uint8_sha256": hashlib.sha256(ntri.tobytes()).hexdigest(),
        "offsets3_int8_sha256": hashlib.sha256(off.tobytes()).hexdigest(),
        "triangles_total": int(ntri.astype(int).sum()),
        "source": "rtpslib/system/common/cl_src/traversehp.cl:35,74; classifycubes.cl:37",
    }
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "mc_tables.json"), "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")
    print(out)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
