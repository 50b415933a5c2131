"""Extract the SpMV and triangular-solve known answers of the reference's
torture tests as data.

Reads test/torture/psb_{d,s}_mvsv_tester.f90 from the reference checkout and
writes tests/golden/torture_usmv.json (every usmv case with trans='n'),
tests/golden/torture_usmv_t.json (the usmv cases with trans 't' and 'c') and
tests/golden/torture_ussv.json (every ussv case, trans 'n', 't' and 'c', with
the triangle flags the case sets on its matrix): the 1-based triplets (IA,
JA, VA), x, alpha, beta, y before (bcy) and the expected y after (cy).  Only
numbers and flags are kept; no reference source text.

usage: python tests/golden/make_torture_fixture.py [/root/reference]
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def _nums(s):
    return [float(t) for t in re.findall(r"[-+]?\d+(?:\.\d*)?(?:[eEdD][-+]?\d+)?", s)]


def parse(path, kind, op="usmv", transposed=False):
    text = open(path).read()
    cases = []
    for m in re.finditer(r"subroutine (\w_" + op + r"_\w+)\(res,afmt,ctxt\)(.*?)end subroutine",
                         text, re.S):
        name, body = m.group(1), m.group(2)
        tr = re.search(r"character::transa='(\w)'", body).group(1)
        if op == "usmv" and (tr != "n") != transposed:
            continue

        def arr(var):
            mm = re.search(r"::\s*" + var + r"\(\d+\)=\(/([^/]*)/\)", body)
            return _nums(mm.group(1))

        def scal(var):
            mm = re.search(r"::\s*" + var + r"\s*=\s*([-+]?\d+(?:\.\d*)?)", body)
            return float(mm.group(1))

        m_ = int(scal("m"))
        k_ = int(scal("k"))
        case = {
            "name": name, "kind": kind, "m": m_, "k": k_,
            "alpha": scal("alpha"), "beta": scal("beta"),
            "ia": [int(v) for v in arr("IA")], "ja": [int(v) for v in arr("JA")],
            "va": arr("VA"), "x": arr("x"), "bcy": arr("bcy"), "cy": arr("cy"),
        }
        if op == "ussv" or transposed:
            case["trans"] = tr
        if op == "ussv":
            case["lower"] = "a%set_upper()" not in body
            case["unit"] = re.search(r"a%set_unit\(\.true\.\)", body) is not None
        cases.append(case)
    return cases


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    for op, label, tag in (("usmv", "usmv, trans=n", "usmv"), ("usmv", "usmv, trans t/c", "usmv_t"),
                           ("ussv", "ussv, trans n/t/c", "ussv")):
        out = []
        for kind in ("d", "s"):
            p = os.path.join(ref, "test", "torture", f"psb_{kind}_mvsv_tester.f90")
            out += parse(p, kind, op, transposed=tag == "usmv_t")
        dst = os.path.join(HERE, f"torture_{tag}.json")
        with open(dst, "w") as f:
            json.dump({"source": f"test/torture/psb_{{d,s}}_mvsv_tester.f90 ({label})",
                       "cases": out}, f, indent=1)
        print(f"{len(out)} cases -> {dst}")


if __name__ == "__main__":
    main()
