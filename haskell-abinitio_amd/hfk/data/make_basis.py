"""Regenerate basis_sets.json from the NWChem-format basis text files.

Run in a container that has the reference tree mounted:

    python haskell-abinitio_amd/hfk/data/make_basis.py /root/reference/data/basis_text

The text files are EMSL Basis Set Exchange exports (data/basis_text/*.txt in
the reference).  Only the elements the workloads use are kept.  SP blocks are
parsed with their real exponent (the reference's cereal binaries carry the
exponent 1.0 for every SP primitive, ParsecTools/ParserBasis.hs:116-118, so
they are not used).
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ELEMENTS = ("H", "He", "C", "N", "O")
FILES = {"sto-3g": "sto-3g.txt", "6-31g": "6-31g.txt", "6-31g*": "6-31gs.txt"}


def main(src_dir):
    sys.path.insert(0, os.path.dirname(HERE))
    from basis import parse_nwchem  # the package's own parser

    out = {}
    for name, fname in FILES.items():
        with open(os.path.join(src_dir, fname)) as f:
            parsed = parse_nwchem(f.read())
        out[name] = {el: parsed[el] for el in ELEMENTS if el in parsed}
    with open(os.path.join(HERE, "basis_sets.json"), "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/data/basis_text")
