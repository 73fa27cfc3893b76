"""Extract the per-connection transmissibility lists of the reference's
legacy fixture rs/data/test/mpfa_quad (a 10x10 unit-permeability MPFA-O
quadrilateral mesh) into tests/golden/mpfa_quad.json.

Run from the repository root; reads the fixture from /root/reference.
The JSON holds data only: {"connections": [[[elements...], [t...]], ...]}.
"""
import json
import os
import re
import sys

SRC = "/root/reference/rs/data/test/mpfa_quad"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "mpfa_quad.json")


def main():
    txt = open(SRC).read()
    block = txt[txt.index("begin abspermtransmissibilities"):txt.index("begin rockdata")]
    out = []
    for m in re.finditer(r"begin (\d+)\s+array transmissibilities(.*?)end\s+array elements"
                         r"(.*?)end", block, re.S):
        t = [float(v) for v in m.group(2).split()]
        e = [int(v) for v in m.group(3).split()]
        assert len(t) == len(e)
        out.append([int(m.group(1)), e, t])
    json.dump({"source": "rs/data/test/mpfa_quad", "grid": [10, 10], "dx": 0.1,
               "connections": out}, open(DST, "w"))
    print(f"{len(out)} connections -> {DST}")


if __name__ == "__main__":
    sys.exit(main())
