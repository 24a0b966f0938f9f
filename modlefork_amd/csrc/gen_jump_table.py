"""Writes det_jump.h: the Xoshiro256++ state transition raised to the power JUMP (2048, the
words one batch of the wave-parallel engine ring holds), as 256 columns of four 64-bit words.

The engine's state update (s1 << 17, the xor cascade and rotl(s3, 45)) is linear over GF(2), so
the state after J steps is M^J applied to the state. Column b of M^J is the state reached after
J steps from the basis state with only bit b set (bit b % 64 of word b / 64); a lane applies the
table as y = xor of the columns of the bits set in its state. The columns are produced here by
stepping all 256 basis states J times.
"""
import sys

import numpy as np

JUMP = 2048


def step(s0, s1, s2, s3):
    t = s1 << np.uint64(17)
    s2 = s2 ^ s0
    s3 = s3 ^ s1
    s1 = s1 ^ s2
    s0 = s0 ^ s3
    s2 = s2 ^ t
    s3 = (s3 << np.uint64(45)) | (s3 >> np.uint64(19))
    return s0, s1, s2, s3


def columns(jump=JUMP):
    s = [np.zeros(256, dtype=np.uint64) for _ in range(4)]
    for b in range(256):
        s[b // 64][b] = np.uint64(1) << np.uint64(b % 64)
    for _ in range(jump):
        s = list(step(*s))
    return np.stack(s, axis=1)  # [256][4]


def main(out):
    cols = columns()
    rows = [", ".join(f"0x{int(v):016x}ULL" for v in c) for c in cols]
    with open(out, "w") as fh:
        fh.write("// Generated by gen_jump_table.py: Xoshiro256++ transition to the power "
                 f"{JUMP}, column b = image of state bit b.\n#pragma once\n#include <stdint.h>\n"
                 f"constexpr uint64_t kDetJumpSteps = {JUMP};\n"
                 f"__device__ __constant__ uint64_t kDetJump[256][4] = {{\n    {{"
                 + "},\n    {".join(rows) + "}};\n")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "det_jump.h")
