"""Regenerate tests/golden/fed_tau_4x4.txt from the reference's own FED code.

The reference's src/lib/fed.cpp is compiled in place by oracle/Makefile into
oracle/_ref/libfed_ref.so (target `ref`).  Inputs follow AKAZE.cpp:176-197 for
the default options (soffset 1.6, 4 octaves x 4 sublevels): etime of level k
is 0.5*esigma^2 with esigma = 1.6 * 2^(k/4), and level k >= 1 gets
fed_tau_by_process_time(etime[k] - etime[k-1], 1, 0.25, reordering=true).
Values are printed with %.9g, which round-trips a float exactly.
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))


def main():
    lib = ctypes.CDLL(os.path.join(REPO, "oracle", "_ref", "libfed_ref.so"))
    f = lib.fedref_tau_by_process_time
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_float, ctypes.c_int, ctypes.c_float, ctypes.c_int,
                  ctypes.POINTER(ctypes.c_float), ctypes.c_int]
    et = []
    for k in range(16):
        o, s = divmod(k, 4)
        esigma = np.float32(1.6) * np.float32(2.0) ** np.float32(np.float32(s) / np.float32(4) + np.float32(o))
        et.append(np.float32(0.5 * float(np.float32(esigma * esigma))))
    lines = []
    buf = (ctypes.c_float * 256)()
    for k in range(1, 16):
        t = np.float32(et[k] - et[k - 1])
        n = f(float(t), 1, 0.25, 1, buf, 256)
        taus = " ".join("%.9g" % buf[i] for i in range(n))
        lines.append("lvl %2d ttime=%f n=%-2d : %s" % (k, t, n, taus))
    out = os.path.join(HERE, "fed_tau_4x4.txt")
    if len(sys.argv) > 1 and sys.argv[1] == "--check":
        with open(out) as fh:
            ok = fh.read().split() == "\n".join(lines).split()
        print("match" if ok else "MISMATCH")
        return 0 if ok else 1
    with open(out, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
