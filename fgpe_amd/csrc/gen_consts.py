"""Writes radix_consts.hpp: cos/sin(2 pi m / R) for the in-register DFT
codelets, rounded from 50-digit mpmath values.  Run once; the header is
committed."""
import os

import mpmath

mpmath.mp.dps = 50
RADICES = (3, 6, 8, 12, 16)

def rnd(x):
    v = float(x)
    return 0.0 if abs(v) < 1e-30 else v


lines = ["// Generated by gen_consts.py: cos/sin(2*pi*m/R), m < R, correctly rounded.",
         "#pragma once", "namespace fgpe {"]
for R in RADICES:
    cs = ", ".join(repr(rnd(mpmath.cos(2 * mpmath.pi * m / R))) for m in range(R))
    sn = ", ".join(repr(rnd(mpmath.sin(2 * mpmath.pi * m / R))) for m in range(R))
    lines.append(f"constexpr double kCos{R}[{R}] = {{{cs}}};")
    lines.append(f"constexpr double kSin{R}[{R}] = {{{sn}}};")
lines.append("}  // namespace fgpe")
out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "radix_consts.hpp")
with open(out, "w") as f:
    f.write("\n".join(lines) + "\n")
print(out)
