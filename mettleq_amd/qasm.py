"""Strict OpenQASM 2.0 importer: the unitary subset the reference accepts.

Restates the reference importer's contract (reference src/mettleq/qasm.py:
290-563): ``parse_qasm_file(path) -> (n_qubits, operations)`` with canonical
op dicts ``{"name", "wires", "parameters"}`` for Device.execute and the
serving loop. Accepted: ``OPENQASM 2.0;`` first, ``qelib1.inc`` /
``stdgates.inc`` includes, qreg/creg declarations (qregs concatenated in
declaration order), the primitive gates below, register-wide single-qubit
gates, non-recursive user gates with bound parameter arithmetic (numbers,
``pi``, bound symbols, parentheses, unary +/-, binary + - * /), barriers as
no-ops and terminal measurements (validated, omitted from the op list).
Everything else (reset, classical control, opaque gates, other includes,
quantum work after a measurement, malformed statements) raises
:class:`QASMParseError` as ``<path>:<line>: <message> [<source>]``, the
reference's error format (qasm.py:33-44), so callers can match on it.

The parser is this package's own: statements are read line by line (the
reference's statement-per-line rule) and expressions are evaluated by a
small recursive-descent parser over a token list.
"""
from __future__ import annotations

import math
import re
from pathlib import Path
from typing import Dict, List, Optional, Sequence, Tuple

__all__ = ["QASMParseError", "parse_qasm", "parse_qasm_file", "GATE_SIGNATURES"]

# canonical name -> (parameter count, wire count) (reference qasm.py:24-31)
GATE_SIGNATURES: Dict[str, Tuple[int, int]] = {
    "H": (0, 1), "X": (0, 1), "Y": (0, 1), "Z": (0, 1), "S": (0, 1), "SDG": (0, 1),
    "T": (0, 1), "TDG": (0, 1), "SX": (0, 1), "RX": (1, 1), "RY": (1, 1), "RZ": (1, 1),
    "U1": (1, 1), "U2": (2, 1), "U3": (3, 1),
    "CNOT": (0, 2), "CZ": (0, 2), "SWAP": (0, 2), "ISWAP": (0, 2), "CH": (0, 2),
    "CPHASE": (1, 2), "CRX": (1, 2), "CRY": (1, 2), "CRZ": (1, 2),
    "CCX": (0, 3), "CSWAP": (0, 3),
}
# QASM spellings -> canonical names (reference qasm.py:165-181)
_SPELLINGS = {
    "cx": "CNOT", "cnot": "CNOT", "cz": "CZ", "swap": "SWAP", "iswap": "ISWAP", "ch": "CH",
    "x": "X", "y": "Y", "z": "Z", "h": "H", "s": "S", "sdg": "SDG", "t": "T", "tdg": "TDG",
    "sx": "SX", "rx": "RX", "ry": "RY", "rz": "RZ", "p": "U1", "phase": "U1", "u1": "U1",
    "u2": "U2", "u3": "U3", "u": "U3", "crx": "CRX", "cry": "CRY", "crz": "CRZ",
    "cu1": "CPHASE", "cp": "CPHASE", "cphase": "CPHASE", "ccx": "CCX", "toffoli": "CCX",
    "cswap": "CSWAP", "fredkin": "CSWAP",
}
_INCLUDES = ("qelib1.inc", "stdgates.inc")
_IDENT = r"[A-Za-z_]\w*"


class QASMParseError(ValueError):
    """A failure at a source line of the supported OpenQASM subset."""

    def __init__(self, path, line: int, message: str, source: str = ""):
        text = f"{path}:{int(line)}: {message}"
        if source:
            text += f" [{source.strip()}]"
        super().__init__(text)
        self.path = path
        self.line = int(line)
        self.source = source


class _Where:
    __slots__ = ("path", "line", "source")

    def __init__(self, path, line: int, source: str):
        self.path, self.line, self.source = path, line, source

    def fail(self, message: str):
        raise QASMParseError(self.path, self.line, message, self.source)


def _canonical(name: str) -> str:
    return _SPELLINGS.get(name.lower(), name.upper())


# ---------------------------------------------------------------------------
# parameter expressions: + - * / unary +/-, parentheses, numbers, pi, symbols
# ---------------------------------------------------------------------------

_TOKEN = re.compile(r"\s*(?:(\d+\.\d*(?:[eE][+-]?\d+)?|\.\d+(?:[eE][+-]?\d+)?|\d+(?:[eE][+-]?\d+)?)"
                    r"|(" + _IDENT + r")|(\*\*|[-+*/()]))")


def _tokens(text: str, where: _Where) -> List[Tuple[str, str]]:
    out, pos = [], 0
    text = text.strip()
    if not text:
        where.fail("empty parameter expression")
    while pos < len(text):
        m = _TOKEN.match(text, pos)
        if not m or m.end() == pos:
            where.fail(f"malformed expression {text!r}")
        num, ident, op = m.groups()
        if num is not None:
            out.append(("num", num))
        elif ident is not None:
            out.append(("id", ident))
        else:
            if op == "**":
                where.fail(f"unsupported operator '**' in expression {text!r}")
            out.append(("op", op))
        pos = m.end()
    return out


class _Expr:
    """Expression tree kept unevaluated until its symbols are bound."""

    def __init__(self, text: str, symbols: Sequence[str], where: _Where):
        self.text = text
        self._toks = _tokens(text, where)
        self._i = 0
        self._where = where
        self._symbols = set(symbols)
        self.tree = self._sum()
        if self._i != len(self._toks):
            where.fail(f"malformed expression {text!r}")

    def _peek(self):
        return self._toks[self._i] if self._i < len(self._toks) else (None, None)

    def _take(self):
        tok = self._peek()
        self._i += 1
        return tok

    def _sum(self):
        node = self._product()
        while self._peek() in (("op", "+"), ("op", "-")):
            node = (self._take()[1], node, self._product())
        return node

    def _product(self):
        node = self._unary()
        while self._peek() in (("op", "*"), ("op", "/")):
            node = (self._take()[1], node, self._unary())
        return node

    def _unary(self):
        if self._peek() in (("op", "+"), ("op", "-")):
            sign = self._take()[1]
            return ("neg" if sign == "-" else "pos", self._unary())
        return self._atom()

    def _atom(self):
        kind, val = self._take()
        if kind == "num":
            return ("num", float(val))
        if kind == "id":
            if val == "pi" or val in self._symbols:
                return ("sym", val)
            self._where.fail(f"unknown symbol {val!r} in expression")
        if (kind, val) == ("op", "("):
            node = self._sum()
            if self._take() != ("op", ")"):
                self._where.fail(f"malformed expression {self.text!r}")
            return node
        self._where.fail(f"malformed expression {self.text!r}")

    def value(self, env: Dict[str, float], where: _Where) -> float:
        def ev(node):
            tag = node[0]
            if tag == "num":
                return node[1]
            if tag == "sym":
                if node[1] == "pi":
                    return math.pi
                if node[1] not in env:
                    where.fail(f"unbound symbol {node[1]!r} in expression")
                return env[node[1]]
            if tag == "neg":
                return -ev(node[1])
            if tag == "pos":
                return ev(node[1])
            a, b = ev(node[1]), ev(node[2])
            if tag == "+":
                return a + b
            if tag == "-":
                return a - b
            if tag == "*":
                return a * b
            if b == 0.0:
                where.fail("division by zero in expression")
            return a / b

        v = float(ev(self.tree))
        if not math.isfinite(v):
            where.fail("expression result must be finite")
        return v


def _split_top(text: str, where: _Where, what: str) -> List[str]:
    """Comma-split outside parentheses."""
    parts, depth, cur = [], 0, []
    for ch in text:
        if ch == "(":
            depth += 1
        elif ch == ")":
            depth -= 1
        if ch == "," and depth == 0:
            parts.append("".join(cur).strip())
            cur = []
        else:
            cur.append(ch)
    parts.append("".join(cur).strip())
    if any(not p for p in parts):
        where.fail(f"empty {what} between commas")
    return parts


# ---------------------------------------------------------------------------
# statements
# ---------------------------------------------------------------------------

_APPLY = re.compile(r"^(" + _IDENT + r")\s*(?:\((.*)\))?\s+([^;]+?)\s*;$")
_REG = re.compile(r"^(qreg|creg)\s+(" + _IDENT + r")\s*\[\s*(\d+)\s*\]\s*;$")
_INCLUDE = re.compile(r'^include\s+"([^"]+)"\s*;$')
_GATE_HEAD = re.compile(r"^gate\s+(" + _IDENT + r")\s*(?:\(([^)]*)\))?\s+([^{}]+?)\s*\{\s*$")
_INDEXED = re.compile(r"^(" + _IDENT + r")\s*\[\s*(\d+)\s*\]$")
_MEASURE = re.compile(r"^measure\s+(.+?)\s*->\s*(.+?)\s*;$")


def _application(source: str, where: _Where):
    m = _APPLY.match(source)
    if not m:
        where.fail("unsupported or malformed statement")
    name, params, targets = m.groups()
    plist = _split_top(params, where, "parameter") if params is not None else []
    tlist = _split_top(targets, where, "wire operand")
    return name, plist, tlist


def _check_primitive(name: str, params: Sequence[float], wires: Sequence[int], where: _Where) -> str:
    canon = _canonical(name)
    sig = GATE_SIGNATURES.get(canon)
    if sig is None:
        where.fail(f"unsupported gate {name!r}")
    npar, nwire = sig
    if len(params) != npar:
        where.fail(f"gate {canon} expects {npar} parameters, got {len(params)}")
    if len(wires) != nwire:
        where.fail(f"gate {canon} expects {nwire} wires, got {len(wires)}")
    if len(set(wires)) != len(wires):
        where.fail(f"gate {canon} requires distinct wires")
    return canon


class _GateDef:
    def __init__(self, name, params, args, body):
        self.name, self.params, self.args, self.body = name, params, args, body


def _resolve(operand: str, regs: Dict[str, Tuple[int, int]], kind: str, where: _Where) -> int:
    m = _INDEXED.match(operand.strip())
    if not m:
        where.fail(f"expected indexed {kind} operand, got {operand!r}")
    reg, idx = m.group(1), int(m.group(2))
    if reg not in regs:
        where.fail(f"unknown {kind} register {reg!r}")
    base, size = regs[reg]
    if idx >= size:
        where.fail(f"{kind} index {reg}[{idx}] is out of range for size {size}")
    return base + idx


def _expand(gd: _GateDef, wires: Sequence[int], values: Sequence[float], defs: Dict[str, _GateDef],
            where: _Where, stack: Tuple[str, ...] = ()) -> List[dict]:
    if len(wires) != len(gd.args):
        where.fail(f"gate {gd.name} expects {len(gd.args)} wires, got {len(wires)}")
    if len(values) != len(gd.params):
        where.fail(f"gate {gd.name} expects {len(gd.params)} parameters, got {len(values)}")
    if len(set(wires)) != len(wires):
        where.fail(f"gate {gd.name} requires distinct wires")
    if gd.name in stack:
        where.fail(f"recursive gate definition involving {gd.name}")
    env = dict(zip(gd.params, values))
    wire_of = dict(zip(gd.args, wires))
    out: List[dict] = []
    for line, source, name, exprs, targets in gd.body:
        here = _Where(where.path, line, source)
        vals = [e.value(env, here) for e in exprs]
        ws = [wire_of[t] for t in targets]
        canon = _canonical(name)
        if canon in defs:
            out.extend(_expand(defs[canon], ws, vals, defs, here, stack + (gd.name,)))
        else:
            out.append({"name": _check_primitive(name, vals, ws, here), "wires": ws, "parameters": vals})
    return out


def parse_qasm(text: str, path="<string>") -> Tuple[int, List[dict]]:
    """Parse QASM source text; `path` labels error messages."""
    lines = [(i, raw.split("//", 1)[0].strip()) for i, raw in enumerate(text.splitlines(), start=1)]
    qregs: Dict[str, Tuple[int, int]] = {}
    cregs: Dict[str, Tuple[int, int]] = {}
    defs: Dict[str, _GateDef] = {}
    ops: List[dict] = []
    nq = nc = 0
    version = False
    measured = False
    i = 0
    while i < len(lines):
        line, src = lines[i]
        i += 1
        if not src:
            continue
        here = _Where(path, line, src)
        if not version:
            if src != "OPENQASM 2.0;":
                here.fail("expected 'OPENQASM 2.0;' as the first statement")
            version = True
            continue
        if src.startswith("OPENQASM"):
            here.fail("duplicate or unsupported OPENQASM version declaration")
        if src.startswith("include"):
            m = _INCLUDE.match(src)
            if not m:
                here.fail("malformed include statement")
            if m.group(1) not in _INCLUDES:
                here.fail(f"include {m.group(1)!r} is not supported")
            continue
        if src.startswith(("qreg", "creg")):
            m = _REG.match(src)
            if not m:
                here.fail("malformed register declaration")
            kind, name, size = m.group(1), m.group(2), int(m.group(3))
            if size < 1:
                here.fail("register size must be positive")
            if name in qregs or name in cregs:
                here.fail(f"duplicate register name {name!r}")
            if kind == "qreg":
                qregs[name] = (nq, size)
                nq += size
            else:
                cregs[name] = (nc, size)
                nc += size
            continue
        if src.startswith("gate "):
            if measured:
                here.fail("measurements must be terminal")
            head = src
            while "{" not in head:
                if i >= len(lines):
                    here.fail("unterminated gate definition header")
                head += " " + lines[i][1]
                i += 1
            if "}" in head:
                here.fail("single-line gate definitions are not supported")
            m = _GATE_HEAD.match(head)
            if not m:
                here.fail("malformed gate definition header")
            gname = _canonical(m.group(1))
            params = [p.strip() for p in (m.group(2) or "").split(",") if p.strip()]
            args = [a.strip() for a in m.group(3).split(",") if a.strip()]
            if len(set(params)) != len(params):
                here.fail("duplicate gate parameter name")
            if not args or len(set(args)) != len(args):
                here.fail("gate wire arguments must be non-empty and unique")
            if gname in defs or gname in GATE_SIGNATURES:
                here.fail(f"duplicate or reserved gate name {gname!r}")
            body = []
            closed = False
            while i < len(lines):
                bline, bsrc = lines[i]
                i += 1
                if not bsrc:
                    continue
                if bsrc == "}":
                    closed = True
                    break
                bwhere = _Where(path, bline, bsrc)
                if "{" in bsrc or "}" in bsrc:
                    bwhere.fail("malformed gate body braces")
                if bsrc.lower().startswith("barrier"):
                    continue
                name, plist, tlist = _application(bsrc, bwhere)
                for t in tlist:
                    if t not in args:
                        bwhere.fail(f"unknown gate wire argument {t!r}")
                body.append((bline, bsrc, name, [_Expr(p, params, bwhere) for p in plist], tlist))
            if not closed:
                here.fail("unterminated gate definition")
            defs[gname] = _GateDef(gname, params, args, body)
            continue
        low = src.lower()
        if low.startswith("opaque "):
            here.fail("opaque gates are not supported")
        if low.startswith("reset"):
            here.fail("reset is not supported by the unitary QASM importer")
        if low.startswith("if"):
            here.fail("classical control is not supported by the unitary QASM importer")
        if low.startswith("barrier"):
            body = src[len("barrier"):].strip()
            if not body.endswith(";") or not body[:-1].strip():
                here.fail("malformed barrier statement")
            for operand in _split_top(body[:-1], here, "barrier operand"):
                if operand not in qregs:
                    _resolve(operand, qregs, "quantum", here)
            continue
        if low.startswith("measure"):
            m = _MEASURE.match(src)
            if not m:
                here.fail("malformed measurement statement")
            qs, cs = m.group(1).strip(), m.group(2).strip()
            if qs in qregs or cs in cregs:
                if not (qs in qregs and cs in cregs):
                    here.fail("whole-register measurement requires quantum and classical registers")
                if qregs[qs][1] != cregs[cs][1]:
                    here.fail("measurement register sizes must match")
            else:
                _resolve(qs, qregs, "quantum", here)
                _resolve(cs, cregs, "classical", here)
            measured = True
            continue
        if measured:
            here.fail("measurements must be terminal; quantum work follows measurement")
        name, plist, tlist = _application(src, here)
        vals = [_Expr(p, (), here).value({}, here) for p in plist]
        canon = _canonical(name)
        if len(tlist) == 1 and tlist[0] in qregs:
            if canon in defs:
                here.fail("user-defined gates require explicit indexed wires")
            sig = GATE_SIGNATURES.get(canon)
            if sig is None or sig[1] != 1:
                here.fail("register-wide application requires a single-qubit gate")
            base, size = qregs[tlist[0]]
            for w in range(base, base + size):
                ops.append({"name": _check_primitive(name, vals, [w], here), "wires": [w],
                            "parameters": list(vals)})
            continue
        wires = [_resolve(t, qregs, "quantum", here) for t in tlist]
        if canon in defs:
            ops.extend(_expand(defs[canon], wires, vals, defs, here))
            continue
        ops.append({"name": _check_primitive(name, vals, wires, here), "wires": wires, "parameters": vals})
    if not version:
        raise QASMParseError(path, 1, "missing OPENQASM 2.0 version declaration")
    return nq, ops


def parse_qasm_file(path) -> Tuple[int, List[dict]]:
    """Parse a QASM file (strict UTF-8) into (n_qubits, canonical ops)."""
    p = Path(path)
    return parse_qasm(p.read_text(encoding="utf-8", errors="strict"), p)
