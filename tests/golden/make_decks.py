"""Writes tests/golden/decks/*.in: the settings of each input deck the
reference ships (input/*.in), re-emitted as an &invar namelist in this
repository's own layout (one key per line, sorted).  The decks are data for
the deck reader and for the time-dependent and steady parity cases; nothing
at test time reads the reference.

Run from the repository root:  python tests/golden/make_decks.py
"""
import glob
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "legion-snap_amd"))

from snap_amd.deck import _FLOATS, parse_namelist  # noqa: E402

REF = "/root/reference/input"


def main():
    out_dir = os.path.join(HERE, "decks")
    os.makedirs(out_dir, exist_ok=True)
    for path in sorted(glob.glob(os.path.join(REF, "*.in"))):
        with open(path) as f:
            vals = parse_namelist(f.read())
        lines = ["&invar"]
        for k in sorted(vals):
            v = vals[k]
            lines.append(f"  {k} = {v!r}" if k in _FLOATS else f"  {k} = {int(v)}")
        lines.append("/")
        name = os.path.basename(path)
        with open(os.path.join(out_dir, name), "w") as f:
            f.write("\n".join(lines) + "\n")
        print(name)


if __name__ == "__main__":
    main()
