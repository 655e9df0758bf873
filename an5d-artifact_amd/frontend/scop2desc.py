#!/usr/bin/env python3
"""Stencil-spec front end: parse a ``#pragma scop`` time loop into a tap table.

Each AN5D benchmark source (e.g. ``j2d5pt.c:14-21``, ``star3d1r.c:15-25``,
``box3d4r.c:15-758``) holds exactly one statement of the form

    A[(t+1)%2][i][j]... = <expr over A[t%2][i+a][j+b]...>;

inside ``#pragma scop`` / ``#pragma endscop``, plus ``BENCH_DIM``,
``BENCH_FPP`` and ``BENCH_RAD`` macros (``j2d5pt.c:1-3``).  This module
tokenises and parses that C expression with C precedence and associativity,
then classifies it:

* ``linear``: a left-associated chain ``c0*A[..] (+|-) c1*A[..] ...``,
  optionally parenthesised and divided by an integer literal.  The tap table
  keeps the source term order, sign, float literal text and offsets, because
  the reference evaluates the sum left to right without contraction and the
  kernels must reproduce that rounding sequence bit for bit.
* ``nonlinear``: anything else (only ``gradient2d.c:18-22``); those are
  emitted with their access offsets and evaluated by hand-written code.

``python scop2desc.py --src <dir with *.c> --out stencils_gen.h`` writes the
C/C++/HIP header used by the engine, its CPU path and the oracle.  The header
is committed so that nothing reads the stencil sources at run time.
"""
from __future__ import annotations

import argparse
import os
import re
import sys
from dataclasses import dataclass, field

# Stencil set evaluated by the artifact (compiled/run.sh and the generated
# code under compiled/; 3d125pt.c excluded: it is in no run script and has no
# generated code).
STENCILS = [
    "star2d1r", "star2d2r", "star2d3r", "star2d4r",
    "box2d1r", "box2d2r", "box2d3r", "box2d4r",
    "j2d5pt", "j2d9pt", "j2d9pt-gol", "gradient2d",
    "star3d1r", "star3d2r", "star3d3r", "star3d4r",
    "box3d1r", "box3d2r", "box3d3r", "box3d4r",
    "j3d27pt",
]

_TOKEN = re.compile(r"""
    (?P<num>(?:\d+\.\d*|\.\d+|\d+)(?:[eE][+-]?\d+)?[fF]?)
  | (?P<id>[A-Za-z_]\w*)
  | (?P<op>\(|\)|\[|\]|\+|-|\*|/|%)
  | (?P<ws>\s+)
""", re.VERBOSE)


class ParseError(Exception):
    pass


def tokenize(s: str):
    pos = 0
    out = []
    while pos < len(s):
        m = _TOKEN.match(s, pos)
        if not m:
            raise ParseError(f"bad character {s[pos]!r} at {pos}")
        pos = m.end()
        if m.lastgroup == "ws":
            continue
        out.append((m.lastgroup, m.group()))
    return out


# AST nodes: ('num', text) | ('ref', (offsets...)) | ('neg', x)
#            | ('bin', op, lhs, rhs) | ('call', name, arg)
class _Parser:
    def __init__(self, toks, loop_vars):
        self.t = toks
        self.i = 0
        self.loop_vars = loop_vars

    def peek(self):
        return self.t[self.i] if self.i < len(self.t) else (None, None)

    def take(self, val=None):
        tok = self.peek()
        if val is not None and tok[1] != val:
            raise ParseError(f"expected {val!r}, got {tok[1]!r}")
        self.i += 1
        return tok

    # expr := term (('+'|'-') term)*        (left associative)
    def expr(self):
        node = self.term()
        while self.peek()[1] in ("+", "-"):
            op = self.take()[1]
            node = ("bin", op, node, self.term())
        return node

    # term := unary (('*'|'/') unary)*
    def term(self):
        node = self.unary()
        while self.peek()[1] in ("*", "/"):
            op = self.take()[1]
            node = ("bin", op, node, self.unary())
        return node

    def unary(self):
        if self.peek()[1] == "-":
            self.take()
            return ("neg", self.unary())
        if self.peek()[1] == "+":
            self.take()
            return self.unary()
        return self.primary()

    def primary(self):
        kind, val = self.peek()
        if val == "(":
            self.take()
            node = self.expr()
            self.take(")")
            return node
        if kind == "num":
            self.take()
            return ("num", val)
        if kind == "id" and val == "A":
            return self.ref()
        if kind == "id":
            self.take()
            self.take("(")
            arg = self.expr()
            self.take(")")
            return ("call", val, arg)
        raise ParseError(f"unexpected token {val!r}")

    def ref(self):
        self.take("A")
        self.take("[")
        # time index must be t%2 on the right-hand side
        if [self.take()[1] for _ in range(3)] != ["t", "%", "2"]:
            raise ParseError("right-hand side must read A[t%2]")
        self.take("]")
        offs = []
        for var in self.loop_vars:
            self.take("[")
            if self.take()[1] != var:
                raise ParseError(f"index must be {var}")
            off = 0
            if self.peek()[1] in ("+", "-"):
                sign = 1 if self.take()[1] == "+" else -1
                kind, val = self.take()
                if kind != "num":
                    raise ParseError("offset must be an integer literal")
                off = sign * int(val)
            self.take("]")
            offs.append(off)
        return ("ref", tuple(offs))


@dataclass
class Tap:
    sign: int          # +1: acc = acc + c*A ; -1: acc = acc - c*A  (first tap: +1)
    coef: str          # float literal text exactly as written, e.g. "5.1f", "-3.240f"
    off: tuple         # (d_outer, d_mid, d_fast); 2D sources map (i, j) -> (i, 0, j)


@dataclass
class Stencil:
    name: str
    dim: int
    fpp: int
    rad: int
    kind: str                       # "linear" | "nonlinear"
    taps: list = field(default_factory=list)
    divisor: int = 0                # 0: no division
    points: list = field(default_factory=list)  # nonlinear: distinct offsets
    src_lines: tuple = (0, 0)


def _lit_text(node):
    """Coefficient literal as C text, with a folded unary minus."""
    if node[0] == "num":
        return node[1]
    if node[0] == "neg" and node[1][0] == "num":
        return "-" + node[1][1]
    return None


def _as_product(node):
    """Return (coef_text, offsets) for `lit * A[..]` or `A[..] * lit`, else None."""
    if node[0] != "bin" or node[1] != "*":
        return None
    lhs, rhs = node[2], node[3]
    if lhs[0] != "ref" and rhs[0] == "ref" and _lit_text(lhs) is not None:
        return _lit_text(lhs), rhs[1]
    if lhs[0] == "ref" and _lit_text(rhs) is not None:
        return _lit_text(rhs), lhs[1]
    return None


def _flatten_chain(node):
    """Unroll a left-associated +/- chain into [(sign, product_node), ...]."""
    terms = []
    while node[0] == "bin" and node[1] in ("+", "-"):
        terms.append((1 if node[1] == "+" else -1, node[3]))
        node = node[2]
    terms.append((1, node))
    terms.reverse()
    return terms


def _collect_refs(node, acc):
    if node[0] == "ref":
        if node[1] not in acc:
            acc.append(node[1])
    elif node[0] in ("neg",):
        _collect_refs(node[1], acc)
    elif node[0] == "bin":
        _collect_refs(node[2], acc)
        _collect_refs(node[3], acc)
    elif node[0] == "call":
        _collect_refs(node[2], acc)
    return acc


def _to3(off, dim):
    return (off[0], 0, off[1]) if dim == 2 else tuple(off)


def parse_source(path: str) -> Stencil:
    name = os.path.basename(path)[:-2]
    text = open(path).read()
    macros = {}
    for key in ("BENCH_DIM", "BENCH_FPP", "BENCH_RAD"):
        m = re.search(rf"#define\s+{key}\s+(\d+)", text)
        if not m:
            raise ParseError(f"{path}: missing {key}")
        macros[key] = int(m.group(1))
    dim = macros["BENCH_DIM"]
    start = text.index("#pragma scop")
    end = text.index("#pragma endscop")
    body = text[start:end]
    lhs = "A[(t+1)%2]" + "".join(f"[{v}]" for v in "ijk"[:dim])
    at = body.index(lhs)
    stmt = body[at + len(lhs):]
    stmt = stmt[stmt.index("=") + 1: stmt.index(";")]
    first_line = text[:start].count("\n") + 1
    last_line = text[:end].count("\n") + 1
    ast = _Parser(tokenize(stmt), "ijk"[:dim]).expr()
    st = Stencil(name=name, dim=dim, fpp=macros["BENCH_FPP"],
                 rad=macros["BENCH_RAD"], kind="linear",
                 src_lines=(first_line, last_line))

    root = ast
    if root[0] == "bin" and root[1] == "/" and root[3][0] == "num" \
            and re.fullmatch(r"\d+", root[3][1]):
        st.divisor = int(root[3][1])
        root = root[2]
    chain = _flatten_chain(root)
    prods = [_as_product(n) for _, n in chain]
    if all(p is not None for p in prods):
        for (sign, _), (coef, off) in zip(chain, prods):
            st.taps.append(Tap(sign=sign, coef=coef, off=_to3(off, dim)))
    else:
        st.kind = "nonlinear"
        st.divisor = 0
        st.points = [_to3(o, dim) for o in _collect_refs(ast, [])]
    return st


def _ident(name):
    return name.replace("-", "_")


def render_header(stencils, src_dir_label="reference sources") -> str:
    out = []
    w = out.append
    w("/* Stencil descriptors: tap tables in source term order.")
    w(" *")
    w(" * GENERATED by an5d-artifact_amd/frontend/scop2desc.py from the")
    w(" * `#pragma scop` statement of each AN5D benchmark source (<name>.c).")
    w(" * Do not edit by hand; regenerate instead.")
    w(" *")
    w(" * AN5D_STENCILS(X): X(id, ident, dim, rad, fpp, ntaps, divisor, linear)")
    w(" * AN5D_TAPS_<ident>(T): T(sign, d_outer, d_mid, d_fast, coef_literal)")
    w(" *   evaluation: acc = c0*A[p0]; acc = acc (sign) c_n*A[p_n]; [acc = acc/divisor]")
    w(" *   2D sources map A[i+a][j+b] to (a, 0, b).")
    w(" * AN5D_POINTS_<ident>(P): P(d_outer, d_mid, d_fast), nonlinear stencils only.")
    w(" */")
    w("#ifndef AN5D_STENCILS_GEN_H")
    w("#define AN5D_STENCILS_GEN_H")
    w("")
    w("#define AN5D_NUM_STENCILS %d" % len(stencils))
    w("")
    w("#define AN5D_STENCILS(X) \\")
    for sid, st in enumerate(stencils):
        n = len(st.taps) if st.kind == "linear" else len(st.points)
        tail = " \\" if sid + 1 < len(stencils) else ""
        w(f"  X({sid}, {_ident(st.name)}, {st.dim}, {st.rad}, {st.fpp}, {n}, "
          f"{st.divisor}, {1 if st.kind == 'linear' else 0}){tail}")
    w("")
    for st in stencils:
        w(f"/* {st.name}.c:{st.src_lines[0]}-{st.src_lines[1]} */")
        if st.kind == "linear":
            items = " ".join(f"T({tp.sign:+d}, {tp.off[0]}, {tp.off[1]}, {tp.off[2]}, {tp.coef})"
                             for tp in st.taps)
            w(f"#define AN5D_TAPS_{_ident(st.name)}(T) {items}")
        else:
            items = " ".join(f"P({p[0]}, {p[1]}, {p[2]})" for p in st.points)
            w(f"#define AN5D_POINTS_{_ident(st.name)}(P) {items}")
    w("")
    w("#endif /* AN5D_STENCILS_GEN_H */")
    return "\n".join(out) + "\n"


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--src", required=True, help="directory holding <stencil>.c")
    ap.add_argument("--out", required=True, help="header to write")
    args = ap.parse_args(argv)
    stencils = [parse_source(os.path.join(args.src, n + ".c")) for n in STENCILS]
    for st in stencils:
        n = len(st.taps) if st.kind == "linear" else len(st.points)
        print(f"{st.name:12s} dim={st.dim} r={st.rad} fpp={st.fpp} {st.kind:9s} "
              f"terms={n} div={st.divisor}", file=sys.stderr)
    with open(args.out, "w") as f:
        f.write(render_header(stencils))


if __name__ == "__main__":
    main()
