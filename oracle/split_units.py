"""Splits a fixed-form Fortran source into one file per program unit
(SUBROUTINE, FUNCTION, MODULE) so that a build recipe can compile the units
one by one and leave out those it does not need.  Test infrastructure: used
by the reference-PES recipe in oracle/Makefile, writing only under
oracle/_ref/.

    python3 split_units.py SOURCE.f OUTDIR   -> OUTDIR/NN_<name>.f
"""
import os
import re
import sys

UNIT = re.compile(r"^\s+(?:(?:double\s+precision|real\*8|real|integer|logical)\s+)?"
                  r"(subroutine|function|module|block\s+data)\s+(\w+)", re.I)


def split(src, outdir):
    os.makedirs(outdir, exist_ok=True)
    units, cur, name = [], [], None
    for ln in open(src).read().split("\n"):
        comment = ln[:1] in ("c", "C", "*", "!")
        m = None if comment else UNIT.match(ln)
        if m and name is not None:
            units.append((name, cur))
            cur = []
        if m:
            name = m.group(2)
        cur.append(ln)
    units.append((name, cur))
    for i, (n, lines) in enumerate(units):
        with open(os.path.join(outdir, "%02d_%s.f" % (i, n)), "w") as f:
            f.write("\n".join(lines) + "\n")
    return [n for n, _ in units]


if __name__ == "__main__":
    print(" ".join(split(sys.argv[1], sys.argv[2])))
