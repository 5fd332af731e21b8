"""Writes hv_twiddles.inc: exp(-2 pi i k / 4096), k < 4096, as float pairs
evaluated in double and rounded once.  Run by holovibes_amd/build.py."""
import math
import sys

N = 4096


def main(path: str) -> None:
    lines = ["// generated by gen_twiddles.py: exp(-2 pi i k / 4096), rounded from double", "#pragma once",
             "static constexpr int kTwN = %d;" % N,
             "__device__ static constexpr float kTw[%d][2] = {" % N]
    for k in range(N):
        a = -2.0 * math.pi * k / N
        c, s = math.cos(a), math.sin(a)
        lines.append("    {%sf, %sf}," % (repr(c), repr(s)))
    lines.append("};")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


if __name__ == "__main__":
    main(sys.argv[1])
