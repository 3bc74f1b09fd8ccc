"""Writes tests/golden/camera_glm.json: camera-to-world matrices of LookAt
cameras as the reference computes them, inverse(glm::lookAtLH(eye, center,
up)) (SceneLoading/Parsing.cu:196-204,257), from the glm 0.9.8 headers the
reference vendors, compiled by oracle/Makefile (target ref_glm) into
oracle/_ref/glm_camera. Inputs: axis-aligned and textbook cases, the
benchmark cameras, and seeded random ones. Run from the repository root in a
container that has the reference; the fixture is committed."""
import json
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))


def main() -> int:
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "ref_glm"], check=True, capture_output=True)
    exe = os.path.join(ROOT, "oracle", "_ref", "glm_camera")
    cases = [
        ((0, 0, 0), (0, 0, 1), (0, 1, 0)),
        ((0, 0, 0), (1, 0, 0), (0, 1, 0)),
        ((0, 0, 0), (0, 0, -1), (0, 1, 0)),
        ((0, 0, 0), (0, -1, 0), (0, 0, 1)),
        ((0.5, 0.5, -1.3), (0.5, 0.5, 0.5), (0, 1, 0)),   # the Cornell box camera
        ((5, 2, 0.3), (5, 1.2, 5), (0, 1, 0)),            # the interior camera
        ((1, 2, 3), (4, 5, 6), (0, 0, 1)),
        ((-2, 1, 7), (0, 0, 0), (0.3, 1, -0.2)),
    ]
    rng = np.random.default_rng(20261203)
    for _ in range(120):
        e = rng.uniform(-10, 10, 3)
        c = e + rng.normal(size=3) * rng.uniform(0.1, 5)
        u = rng.normal(size=3)
        cases.append((tuple(e), tuple(c), tuple(u)))
    vals = [np.float32(x) for case in cases for v in case for x in v]
    lines = []
    for k in range(len(cases)):
        lines.append(" ".join("%.9g" % float(vals[9 * k + j]) for j in range(9)))
    out = subprocess.run([exe], input="\n".join(lines) + "\n", capture_output=True, text=True, check=True)
    rows = out.stdout.split("\n")
    fixture = []
    for k, line in enumerate(lines):
        inp = [int(np.float32(vals[9 * k + j]).view(np.uint32)) for j in range(9)]
        fixture.append({"eye_center_up_bits": ["%08x" % b for b in inp], "cam_to_world_bits": rows[k].split()})
    path = os.path.join(ROOT, "tests", "golden", "camera_glm.json")
    with open(path, "w") as f:
        f.write('{"source": "inverse(glm::lookAtLH(eye, center, up)), glm 0.9.8 headers of the reference, '
                'g++ -O2 -ffp-contract=off",\n "cases": [\n')
        f.write(",\n".join(json.dumps(c, separators=(",", ":")) for c in fixture))
        f.write("\n]}\n")
    print(f"wrote {len(fixture)} cases to {path}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
