"""Writes twiddles.h: W_256^j = exp(-2*pi*i*j/256), j = 0..255, rounded once from float64 to float32.

Every inter-stage twiddle of the 2-D FFTs (N = 32, 64, 128, 256) is a power of W_256, so one
table serves all sizes. Run: python quantem_amd/csrc/gen_twiddles.py
"""
import math
import os

import numpy as np

rows = []
for j in range(256):
    c = np.float32(math.cos(2.0 * math.pi * j / 256.0))
    s = np.float32(-math.sin(2.0 * math.pi * j / 256.0))
    rows.append("    {%.9ef, %.9ef}," % (float(c), float(s)))
text = (
    "// Generated by gen_twiddles.py: W_256^j = exp(-2 pi i j / 256) as float32 (cos, -sin).\n"
    "#pragma once\n"
    "namespace qem {\n"
    "__device__ __constant__ const float2 kTw256[256] = {\n" + "\n".join(rows) + "\n};\n"
    "}  // namespace qem\n"
)
out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "twiddles.h")
with open(out, "w") as f:
    f.write(text)
