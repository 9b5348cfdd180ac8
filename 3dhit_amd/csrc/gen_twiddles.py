"""Writes twiddles.h: exp(-2 pi i j / 1024), j < 1024, and cos/sin(2 pi j / 64),
j < 64, as float64 literals (repr round-trips exactly). The float tables are
the float64 values rounded to nearest. Run: python gen_twiddles.py"""

import math
import os

NT = 1024
here = os.path.dirname(os.path.abspath(__file__))


def cos_sin(j, n):
    """cos, sin of 2 pi j / n with exact zeros/ones and exact octant symmetry:
    reduce to r in [0, n/8] and rotate by whole quadrants."""
    j %= n
    quad, r = divmod(4 * j, n)          # 4j = quad*n + r, angle = (quad + r/n) pi/2
    if 2 * r <= n:
        c, s = math.cos(math.pi * r / (2 * n)), math.sin(math.pi * r / (2 * n))
    else:
        s, c = math.cos(math.pi * (n - r) / (2 * n)), math.sin(math.pi * (n - r) / (2 * n))
    if r == 0:
        c, s = 1.0, 0.0
    for _ in range(quad):
        c, s = -s, c
    return c + 0.0, s + 0.0


def lit(x):
    r = repr(float(x))
    return r if ("e" in r or "." in r) else r + ".0"


lines = [
    "// Generated by gen_twiddles.py: do not edit.",
    "#pragma once",
    f"#define HIT3D_TW_N {NT}",
    "namespace h3 {",
    "// cos(2 pi j / 64), sin(2 pi j / 64)",
    "constexpr double kCos64[64] = {" + ", ".join(lit(cos_sin(j, 64)[0]) for j in range(64)) + "};",
    "constexpr double kSin64[64] = {" + ", ".join(lit(cos_sin(j, 64)[1]) for j in range(64)) + "};",
    f"// exp(-2 pi i j / {NT}) as (re, im) pairs",
]
vals = [(cos_sin(j, NT)[0], -cos_sin(j, NT)[1] + 0.0) for j in range(NT)]
body = ",\n".join(f"{{{lit(c)}, {lit(s)}}}" for c, s in vals)
lines.append(f"static __constant__ cpx<double> kTwD[{NT}] = {{\n{body}}};")
lines.append(f"static __constant__ cpx<float> kTwF[{NT}] = {{\n" + ",\n".join(
    f"{{{lit(c)}f, {lit(s)}f}}" for c, s in vals) + "};")
lines.append("}  // namespace h3")
with open(os.path.join(here, "twiddles.h"), "w") as f:
    f.write("\n".join(lines) + "\n")
