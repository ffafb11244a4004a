"""Writes the reference's polygon fixtures as JSON data (run in the build container only).

Reads crates/tessellation/src/fuzz_tests.rs and earcut_tests.rs from the lyon source tree
as text and keeps only their coordinate data:

* fuzz_polygons.json — every `fuzzing_test_case_*`: sub-paths of begin/line_to points
  with their close flag (fuzz_tests.rs:54-992).  The reference requires each to
  tessellate without error under EvenOdd and NonZero at tolerance 0.05.
* earcut_polygons.json — every `earcut_test`/`earcut_test_f32` call: rings of points.
  Replays must reproduce the harness quirk of earcut_tests.rs:20184-20219, whose
  sub-paths *begin at (start[0], start[0])* and then line_to every ring point.

Usage: python tests/golden/extract_fixtures.py /root/reference
"""
import json
import os
import re
import sys

NUM = r"-?\d+(?:\.\d*)?(?:[eE][-+]?\d+)?"


def fuzz(src):
    out = {}
    for m in re.finditer(r"fn (fuzzing_test_case_\w+)\(\)\s*\{(.*?)\n\}", src, re.S):
        name, body = m.group(1), m.group(2)
        subs = []
        cur = None
        for tok in re.finditer(r"builder\.(begin|line_to|end|close)\((.*?)\);", body, re.S):
            op, arg = tok.group(1), tok.group(2)
            if op in ("begin", "line_to"):
                xy = [float(v) for v in re.findall(NUM, arg)[:2]]
                if op == "begin":
                    cur = {"points": [xy], "closed": False}
                    subs.append(cur)
                else:
                    cur["points"].append(xy)
            elif op == "end":
                cur["closed"] = "true" in arg
            else:
                cur["closed"] = True
        out[name] = subs
    return out


def earcut(src):
    out = {}
    for m in re.finditer(r"fn (\w+)\(\)\s*\{\s*(earcut_test(?:_f32)?)\(&\[(.*?)\]\);\s*\}", src, re.S):
        name, body = m.group(1), m.group(3)
        rings = []
        for rm in re.finditer(r"&\[(.*?)\]\s*,?\s*(?=&\[|$)", body, re.S):
            pts = [[float(a), float(b)]
                   for a, b in re.findall(r"\[\s*(" + NUM + r")\s*,\s*(" + NUM + r")\s*\]", rm.group(1))]
            rings.append(pts)
        out[name] = rings
    return out


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    here = os.path.dirname(os.path.abspath(__file__))
    tdir = os.path.join(ref, "crates", "tessellation", "src")
    f = fuzz(open(os.path.join(tdir, "fuzz_tests.rs")).read())
    e = earcut(open(os.path.join(tdir, "earcut_tests.rs")).read())
    json.dump(f, open(os.path.join(here, "fuzz_polygons.json"), "w"))
    json.dump(e, open(os.path.join(here, "earcut_polygons.json"), "w"))
    print(f"{len(f)} fuzz cases, {len(e)} earcut cases "
          f"({sum(len(p) for r in e.values() for p in r)} earcut points)")


if __name__ == "__main__":
    main()
