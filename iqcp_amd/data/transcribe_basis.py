"""Transcribe the published basis-set tables that qc-core ships into basis_sets.json.

The exponents and contraction coefficients are data (PySCF 2.11 / Basis Set
Exchange values), transcribed with provenance from
crates/qc-core/src/basis/builtin.rs (get_element_basis :97, get_sto3g :155,
get_631g :701, get_631gs :1032, get_631pgs :1072).  Shell order per element is
kept exactly as listed there, because the basis-function order of every matrix
follows it (BasisSet::build, basis/mod.rs:164-193).

Run once in a container that has the reference checkout:
    python iqcp_amd/data/transcribe_basis.py /root/reference
"""
import json
import re
import sys
from pathlib import Path

ELEMENTS = {1: "h", 2: "he", 3: "li", 4: "be", 5: "b", 6: "c", 7: "n", 8: "o", 9: "f", 10: "ne"}
L_OF = {"S": 0, "P": 1, "D": 2}
# 6-31G* d exponents for Z = 3..10 (builtin.rs:1038-1050): appended to 6-31G.
D631GS = {3: 0.2, 4: 0.4, 5: 0.6, 6: 0.8, 7: 0.8, 8: 0.8, 9: 0.8, 10: 0.8}


def parse_functions(src: str):
    """Return {fn_name: [(l, [(exp, coef), ...]), ...]} for every `fn x() -> Vec<ShellData>`."""
    out = {}
    for m in re.finditer(r"fn (\w+)\(\) -> Vec<ShellData> \{(.*?)\n\}", src, re.S):
        name, body = m.group(1), m.group(2)
        shells = []
        # Each shell starts with AngularMomentum::X and is followed by its vec![...]
        parts = re.split(r"AngularMomentum::([SPD])", body)
        for k in range(1, len(parts), 2):
            l = L_OF[parts[k]]
            seg = parts[k + 1]
            # primitives are (exp, coef) tuples up to the closing of the vec!
            end = seg.find("]")
            prims = re.findall(r"\(\s*([-+0-9.eE]+)\s*,\s*([-+0-9.eE]+)\s*\)", seg[:end])
            shells.append((l, [(float(a), float(b)) for a, b in prims]))
        out[name] = shells
    return out


def main(ref_root: str):
    src = Path(ref_root, "crates/qc-core/src/basis/builtin.rs").read_text()
    fns = parse_functions(src)
    data = {"source": "ExaPsi/IQCP crates/qc-core/src/basis/builtin.rs (PySCF 2.11 tables)",
            "bases": {}}
    for basis, prefix in (("sto-3g", "sto3g"), ("6-31g", "b631g"), ("6-31+g*", "b631pgs")):
        table = {}
        for z, sym in ELEMENTS.items():
            table[str(z)] = [[l, prims] for l, prims in fns[f"{prefix}_{sym}"]]
        data["bases"][basis] = table
    gs = {}
    for z, sym in ELEMENTS.items():
        shells = [[l, prims] for l, prims in fns[f"b631g_{sym}"]]
        if z in D631GS:
            shells.append([2, [(D631GS[z], 1.0)]])
        gs[str(z)] = shells
    data["bases"]["6-31g*"] = gs
    out = Path(__file__).with_name("basis_sets.json")
    out.write_text(json.dumps(data, indent=1))
    print("wrote", out)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
