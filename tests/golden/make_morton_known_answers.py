"""Writes morton_known_answers.json: hand-derived Morton codes.

Each case is a point (a degenerate triangle whose box centroid is the point)
and a scene box. The expected codes are written out from the bit positions
worked out by hand below -- x bit i lands at 3i+2, y at 3i+1, z at 3i; a
normalised coordinate c in (-0.5, 0.5] contributes floor(c * 2^B) with
B = 21 (63-bit) or 10 (30-bit); c <= 0 and NaN give 0 (saturating cast);
+inf saturates and is then clamped to 2^(B-1) (BVH/BVH.cu:61-62).
Run: python tests/golden/make_morton_known_answers.py
"""
import json
import os

UNIT = [0, 0, 0, 1, 1, 1]
DENORM = 1.401298464324817e-45  # 2^-149, smallest float32 denormal


def bits(*positions):
    v = 0
    for p in positions:
        v |= 1 << p
    return f"{v:#x}"


CASES = [
    # c = (0.5, 0.5, 0.5): 2^20 / 2^9 on every axis
    {"point": [1, 1, 1], "scene": UNIT, "code63": bits(62, 61, 60), "code30": bits(29, 28, 27)},
    # c < 0 on every axis: all zero
    {"point": [0.25, 0.25, 0.25], "scene": UNIT, "code63": bits(), "code30": bits()},
    # c = (0.25, 0, -0.25): x = 2^19 / 2^8 only
    {"point": [0.75, 0.5, 0.25], "scene": UNIT, "code63": bits(59), "code30": bits(26)},
    # c = (0.375, 0.125, 2^-10): x = 2^19+2^18 / 2^8+2^7, y = 2^18 / 2^7, z = 2^11 / 1
    {"point": [0.875, 0.625, 0.5009765625], "scene": UNIT,
     "code63": bits(59, 56, 55, 33), "code30": bits(26, 23, 22, 0)},
    # zero x extent: c.x = 0/0 = NaN -> 0; y, z at 0.5
    {"point": [0, 1, 1], "scene": [0, 0, 0, 0, 1, 1], "code63": bits(61, 60),
     "code30": bits(28, 27)},
    # x extent 2^-149: c.x = +inf -> saturates -> clamped to 2^20 / 2^9;
    # y, z: c = 0.25 -> 2^19 / 2^8
    {"point": [1, 0.75, 0.75], "scene": [0, 0, 0, DENORM, 1, 1],
     "code63": bits(62, 58, 57), "code30": bits(29, 25, 24)},
    # negative zero coordinates behave as zero: c = -0.5 -> 0 on x
    {"point": [-0.0, 1, 0.5], "scene": UNIT, "code63": bits(61), "code30": bits(28)},
]

if __name__ == "__main__":
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "morton_known_answers.json")
    with open(out, "w") as f:
        json.dump(CASES, f, indent=1)
    print("wrote", out)
