"""Writes props_wan2025.json: the piecewise-cubic property laws of the
reference's supercritical-CO2 example,
examples/input_Wan_2025_Case_A_as_roughsurf/expressions_rho_mu_cond_buoyancy_cp_div_jacobian.f90,
as data (breakpoints and coefficients), for the property-update tests.

The file holds, for each of enth, rho, mu, cond, cp and drho_dT, a binary
tree of `if (T_ext_ijk<=(hi)) then ... else ... end if` whose leaves are
`q_ijk = (c0) + (c1)*((T_ext_ijk - x0)) + (c2)*(((T_ext_ijk - x0))**2) + (c3)*(((T_ext_ijk - x0))**3)`,
then div_jacobian and buoyancy as formulas of rho, cp and drho_dT.  The
in-order leaves of a tree with the thresholds between them are the sorted
segments [(hi, x0, c0, c1, c2, c3), ...] (last hi = +inf).

Run in this container (reads /root/reference):  python tests/golden/make_property_tables.py
"""
import json
import os
import re

SRC = "/root/reference/examples/input_Wan_2025_Case_A_as_roughsurf/expressions_rho_mu_cond_buoyancy_cp_div_jacobian.f90"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "props_wan2025.json")
NUM = r"\(?\s*([-+]?\d+\.\d+D[-+]\d+)\s*\)?"


def fnum(s):
    return float(s.replace("D", "e"))


def statements(text):
    """Joins Fortran continuation lines, drops comments and blank lines."""
    out, cur = [], ""
    for line in text.split("\n"):
        line = line.split("!")[0].rstrip()
        if not line.strip():
            continue
        if cur:
            cur += " " + line.strip()
        else:
            cur = line.strip()
        if cur.endswith("&"):
            cur = cur[:-1].rstrip()
            continue
        out.append(cur)
        cur = ""
    return out


LEAF = re.compile(r"^(\w+)_ijk\s*=\s*" + NUM + r"\s*\+\s*" + NUM + r"\*\(\(T_ext_ijk - (\S+?)\)\)\s*\+\s*" + NUM +
                  r"\*\(\(\(T_ext_ijk - (\S+?)\)\)\*\*2\)\s*\+\s*" + NUM + r"\*\(\(\(T_ext_ijk - (\S+?)\)\)\*\*3\)$")
IF = re.compile(r"^if \(T_ext_ijk<=\(([-+]?\d+\.\d+D[-+]\d+)\)\) then$")


def parse_tree(st, i):
    """Returns (leaves, thresholds, next index) of the tree starting at st[i]."""
    m = IF.match(st[i])
    if m:
        hi = fnum(m.group(1))
        left, lthr, i = parse_tree(st, i + 1)
        assert st[i] == "else", st[i]
        right, rthr, i = parse_tree(st, i + 1)
        assert st[i] == "end if", st[i]
        return left + right, lthr + [hi] + rthr, i + 1
    m = LEAF.match(st[i].replace(" ", "").replace("T_ext_ijk-", "T_ext_ijk - ").replace("_ijk=", "_ijk = "))
    assert m, st[i]
    name = m.group(1)
    c0, c1, x0, c2, x0b, c3, x0c = m.groups()[1:]
    assert x0 == x0b == x0c
    return [(name, fnum(x0), fnum(c0), fnum(c1), fnum(c2), fnum(c3))], [], i + 1


def main():
    st = statements(open(SRC).read())
    st = [s for s in st if not s.startswith("T_ext_ijk = T_ext(i,j,k)")]
    laws, i = {}, 0
    while i < len(st) and IF.match(st[i]):
        leaves, thr, i = parse_tree(st, i)
        names = {lf[0] for lf in leaves}
        assert len(names) == 1, names
        assert thr == sorted(thr) and len(thr) == len(leaves) - 1
        segs = [[thr[s] if s < len(thr) else None, *leaves[s][1:]] for s in range(len(leaves))]
        laws[names.pop()] = segs
    rest = " ".join(st[i:])
    assert "div_jacobian_arr(i,j,k) = -1_dp/((rho_ijk**2)*cp_ijk)*drho_dT_ijk" in rest
    m = re.search(r"buoyancy_arr\(i,j,k\)\s*=\s*([-+]?\d+\.\d+)_dp\*\(rho_ijk-(\d+\.\d+)_dp\)", rest)
    assert m
    out = {"source": "examples/input_Wan_2025_Case_A_as_roughsurf/expressions_rho_mu_cond_buoyancy_cp_div_jacobian.f90",
           "laws": laws, "divjac": "formula", "buoyancy": [float(m.group(1)), float(m.group(2))]}
    assert sorted(laws) == sorted(["enth", "rho", "mu", "cond", "cp", "drho_dT"]), sorted(laws)
    json.dump(out, open(OUT, "w"), indent=1)
    print({k: len(v) for k, v in laws.items()}, "->", OUT)


if __name__ == "__main__":
    main()
