"""Copies the wall-normal face files (yu.txt) of the reference's four example
cases into this folder, records their grid parameters from params.dat, and
copies the grit-blasted rough-surface modes (fortran_surf_files/mnab_coeffs.dat).

Run in a container that has the reference checkout; the fixtures it writes
(reference_grids.json, yu_*.txt) are data and are committed."""
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
CASES = {
    "nicoud2000_tr4": "input_Nicoud_2000_Tr_4",
    "peeters_sandham_ret180_k30": "input_Peeters_Sandham_Ret180_Pr1_k30_gritblasted_as_varprop",
    "ret540_k45": "input_Ret540_Pr1_k45_gritblasted_incomp",
    "wan2025_case_a": "input_Wan_2025_Case_A_as_roughsurf",
}


def params(path):
    out = {}
    for line in open(path):
        m = re.match(r"\s*(nx|nz|ny|Lx|Lz|Ly)\s*=\s*([-+0-9.eEdD]+)", line)
        if m:
            v = m.group(2).replace("d", "e").replace("D", "e")
            out[m.group(1)] = int(v) if m.group(1) in ("nx", "nz", "ny") else float(v)
    return out


def main():
    manifest = {}
    for name, folder in CASES.items():
        src = os.path.join(REF, "examples", folder)
        text = open(os.path.join(src, "yu.txt")).read()
        with open(os.path.join(HERE, f"yu_{name}.txt"), "w") as fw:
            fw.write(text)
        manifest[name] = dict(params(os.path.join(src, "params.dat")), source=f"examples/{folder}/yu.txt")
    with open(os.path.join(HERE, "reference_grids.json"), "w") as fw:
        json.dump(manifest, fw, indent=1, sort_keys=True)
    # the grit-blasted rough surface (Fourier modes m, n, a, b) of two examples
    src = os.path.join(REF, "examples", CASES["peeters_sandham_ret180_k30"], "fortran_surf_files", "mnab_coeffs.dat")
    with open(os.path.join(HERE, "mnab_gritblasted.dat"), "w") as fw:
        fw.write(open(src).read())


if __name__ == "__main__":
    main()
