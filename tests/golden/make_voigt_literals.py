"""Write tests/golden/voigt_literals.json: the numeric literals of Korg's harris_series and
voigt_hjerting (src/line_absorption.jl:251-310), as data for tests/test_voigt_literals.py.

Per function: "fitted", the signed coefficients with six or more significant digits in the order
they appear (the sign is that of the term: a literal after a minus sign is negative), and
"thresholds", the constants its branch conditions compare against, in order.

    python tests/golden/make_voigt_literals.py /root/reference
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
NUM = re.compile(r"(?<![\w.])(\d+\.\d+|\d+)(?![\w.])")
CMP = re.compile(r"(<=|>=|<|>)\s*\(?\s*(\d+(?:\.\d+)?)")


def body(src, name):
    """The text of `function name(...)` up to its closing `end` (indentation 0)."""
    i = src.index("function " + name + "(")
    j = src.index("\nend", i)
    return src[i:j]


def signed_literals(text):
    out = []
    for m in NUM.finditer(text):
        digits = m.group(1).replace(".", "").lstrip("0")
        if "." not in m.group(1) or len(digits) < 6:
            continue
        k = m.start() - 1
        while k >= 0 and text[k] in " (\n":
            k -= 1
        out.append(-float(m.group(1)) if k >= 0 and text[k] == "-" else float(m.group(1)))
    return out


def main(ref):
    src = open(os.path.join(ref, "src", "line_absorption.jl")).read()
    data = {}
    for name in ("harris_series", "voigt_hjerting"):
        b = body(src, name)
        code = "\n".join(l.split("#")[0] for l in b.splitlines())  # comments off
        data[name] = {"fitted": signed_literals(code),
                      "thresholds": [float(m.group(2)) for m in CMP.finditer(code)]}
    data["source"] = "src/line_absorption.jl:251-310"
    with open(os.path.join(HERE, "voigt_literals.json"), "w") as f:
        json.dump(data, f, indent=1)
    print(json.dumps(data))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
