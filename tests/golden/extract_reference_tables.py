"""Writes tests/golden/reference_tables.json: the linear tables of the
reference solver as coded, extracted from its Fortran source read as text.

Extracted (all exact rationals):
  * the MRT moment matrices M (m(r) = sum_a M[r][a] f(a)) and their coded
    inverses (f_post(a) = sum_r Minv[a][r] m_post(r)) of collision_U
    (natural-convection_2dim_mono.f90:445-453, :493-509;
     natural-convection_3dim_mono.f90:460-520, :589-655);
  * the thermal matrices N and N^-1 of collision_T (2dim_mono.f90:531-535,
    :553-557; 3dim_mono.f90:681-687, :709-715);
  * the relaxation-rate assignments s(r) and Q(r) (names or constants);
  * the wall assignments of bounceback_u / bounceback_T
    (2dim_mono.f90:662-724; 3dim_mono.f90:878-986): target population,
    source population, the node-index expressions and, for the thermal walls,
    the sign and wall temperature.
Each right-hand side is linear in the populations; it is evaluated on unit
vectors with exact fractions by a small expression walker that accepts only
numbers, + - * / and the array references named above.  Nothing from the
reference is executed or copied into the repository: the output is a table
of numbers.  Run it where the reference is mounted:

    python3 tests/golden/extract_reference_tables.py /root/reference
"""
import ast
import json
import os
import re
import sys
from fractions import Fraction

FILES = {"2d": "natural-convection_2dim_mono.f90",
         "3d": "natural-convection_3dim_mono.f90"}


def logical_lines(text):
    """Fortran free-form lines with '&' continuations joined."""
    out, cur = [], ""
    for raw in text.splitlines():
        line = raw.split("!")[0].rstrip()
        s = line.strip()
        if s.startswith("&"):
            s = s[1:]
        if s.endswith("&"):
            cur += s[:-1]
            continue
        cur += s
        if cur.strip():
            out.append(cur.strip())
        cur = ""
    return out


def subroutine_body(text, name):
    m = re.search(r"subroutine\s+%s\s*\(\)(.*?)end\s+subroutine" % name, text,
                  flags=re.S | re.I)
    if not m:
        raise KeyError(name)
    return logical_lines(m.group(1))


class Linear:
    """A linear form sum_k c_k x_k over the symbols of one array."""

    def __init__(self, terms=None, const=Fraction(0)):
        self.t = dict(terms or {})
        self.c = const

    def __add__(self, o):
        o = lift(o)
        t = dict(self.t)
        for k, v in o.t.items():
            t[k] = t.get(k, 0) + v
        return Linear(t, self.c + o.c)

    def __neg__(self):
        return Linear({k: -v for k, v in self.t.items()}, -self.c)

    def __sub__(self, o):
        return self + (-lift(o))

    def scale(self, s):
        return Linear({k: v * s for k, v in self.t.items()}, self.c * s)


def lift(x):
    return x if isinstance(x, Linear) else Linear(const=Fraction(x))


def evaluate(expr, array):
    """Evaluates a linear Fortran expression in array(index, ...) terms."""
    src = re.sub(r"(\d+\.\d*|\d+)d0", r"\1", expr)
    src = re.sub(r"\b%s\((\d+)(,[^()]*)?\)" % array, r"__v[\1]", src)
    tree = ast.parse(src, mode="eval")

    def walk(n):
        if isinstance(n, ast.Expression):
            return walk(n.body)
        if isinstance(n, ast.Constant) and isinstance(n.value, (int, float)):
            return Fraction(str(n.value))
        if isinstance(n, ast.Subscript) and isinstance(n.value, ast.Name) \
                and n.value.id == "__v":
            idx = n.slice.value if isinstance(n.slice, ast.Constant) else None
            return Linear({int(idx): Fraction(1)})
        if isinstance(n, ast.UnaryOp) and isinstance(n.op, ast.USub):
            v = walk(n.operand)
            return -v if isinstance(v, Linear) else -v
        if isinstance(n, ast.BinOp):
            a, b = walk(n.left), walk(n.right)
            if isinstance(n.op, ast.Add):
                return lift(a) + lift(b) if isinstance(a, Linear) or isinstance(b, Linear) else a + b
            if isinstance(n.op, ast.Sub):
                return lift(a) - lift(b) if isinstance(a, Linear) or isinstance(b, Linear) else a - b
            if isinstance(n.op, ast.Mult):
                if isinstance(a, Linear):
                    return a.scale(b)
                if isinstance(b, Linear):
                    return b.scale(a)
                return a * b
            if isinstance(n.op, ast.Div) and not isinstance(b, Linear):
                return a.scale(1 / b) if isinstance(a, Linear) else a / b
        raise ValueError("unsupported expression: " + expr)

    return lift(walk(tree))


def table(lines, lhs, array, q):
    rows = {}
    for line in lines:
        m = re.match(r"%s\((\d+)(,[^=]*)?\)\s*=\s*(.*)$" % lhs, line)
        if m:
            lin = evaluate(m.group(3), array)
            rows[int(m.group(1))] = [str(lin.t.get(a, Fraction(0))) for a in range(q)]
    return [rows[r] for r in range(q)]


def rates(lines, name):
    out = {}
    for line in lines:
        m = re.match(r"%s\((\d+)\)\s*=\s*(\S+)$" % name, line)
        if m:
            out[int(m.group(1))] = m.group(2).replace("d0", "")
    return [out[k] for k in sorted(out)]


def walls(lines, target, source):
    out = []
    pat = re.compile(r"%s\((\d+),([^)]*)\)\s*=\s*(-?)%s\((\d+),([^)]*)\)(.*)$" % (target, source))
    for line in lines:
        m = pat.match(line.replace(" ", ""))
        if m:
            extra = m.group(6)
            temp = None
            if "Thot" in extra:
                temp = "Thot"
            elif "Tcold" in extra:
                temp = "Tcold"
            out.append({"to": int(m.group(1)), "from": int(m.group(4)),
                        "node": m.group(2).split(","), "sign": -1 if m.group(3) else 1,
                        "wall_temperature": temp})
    return out


def main(ref_dir):
    data = {}
    for dim, fname in FILES.items():
        text = open(os.path.join(ref_dir, fname)).read()
        q, qg = (9, 5) if dim == "2d" else (19, 7)
        cu = subroutine_body(text, "collision_U")
        ct = subroutine_body(text, "collision_T")
        data[dim] = {
            "M": table(cu, "m", "f", q),
            "Minv": table(cu, "f_post", "m_post", q),
            "s": rates(cu, "s"),
            "N": table(ct, "n", "g", qg),
            "Ninv": table(ct, "g_post", "n_post", qg),
            "Q": rates(ct, "Q"),
            "walls_u": walls(subroutine_body(text, "bounceback_u"), "f", "f_post"),
            "walls_T": walls(subroutine_body(text, "bounceback_T"), "g", "g_post"),
        }
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_tables.json")
    with open(out, "w") as fh:
        json.dump(data, fh, indent=1)
    print("wrote", out)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
