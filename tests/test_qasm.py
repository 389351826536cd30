"""Strict OpenQASM 2.0 importer (mettleq_amd.qasm) and the reference's QASM
fixtures (tests/golden/qasm) through the oracle.

Mirrors the reference's tests: src/tests/mlxQQasmStrictnessTest.py (error
lines and messages) and the fixture checks in src/tests/mlxQCoreTest.py
(:125-145, :569-573, :743-760, :844-867), plus the closed-form answers the
fixture circuits are built for.
"""
import math
from pathlib import Path

import numpy as np
import pytest

from mettleq_amd.qasm import QASMParseError, parse_qasm, parse_qasm_file
from oracle import sv_oracle as O

GOLDEN = Path(__file__).resolve().parent / "golden" / "qasm"


def _write(tmp_path, text):
    p = tmp_path / "circuit.qasm"
    p.write_text(text, encoding="utf-8")
    return p


def _probs(n, ops):
    return np.abs(O.run_ops(n, ops)) ** 2


# --- strictness (mlxQQasmStrictnessTest.py) ---------------------------------


def test_bound_parameter_arithmetic_is_evaluated(tmp_path):
    p = _write(tmp_path, """OPENQASM 2.0;
include "qelib1.inc";
qreg q[1];
gate twist(theta, phi) a {
  rz(-(theta/2) + pi/8 + phi*2) a;
}
twist(pi/2, 0.125) q[0];
""")
    n, ops = parse_qasm_file(p)
    assert n == 1
    assert len(ops) == 1 and ops[0]["name"] == "RZ" and ops[0]["wires"] == [0]
    assert ops[0]["parameters"][0] == pytest.approx(-math.pi / 8 + 0.25, abs=1e-12)


@pytest.mark.parametrize("expr", ["secret", "pi**2"])
def test_unknown_or_unsupported_expressions_name_the_line(tmp_path, expr):
    p = _write(tmp_path, f"""OPENQASM 2.0;
include "qelib1.inc";
qreg q[1];
rx({expr}) q[0];
""")
    with pytest.raises(QASMParseError, match=r"circuit\.qasm:4:.*expression"):
        parse_qasm_file(p)


@pytest.mark.parametrize("source,line,feature", [
    ('OPENQASM 2.0;\ninclude "qelib1.inc";\nqreg q[1];\nreset q[0];\n', 4, "reset"),
    ('OPENQASM 2.0;\ninclude "qelib1.inc";\nqreg q[1];\ncreg c[1];\nif(c==1) x q[0];\n', 5,
     "classical control"),
])
def test_reset_and_classical_control_rejected(tmp_path, source, line, feature):
    p = _write(tmp_path, source)
    with pytest.raises(QASMParseError, match=rf"circuit\.qasm:{line}:.*{feature}.*not supported"):
        parse_qasm_file(p)


def test_terminal_measurements_only(tmp_path):
    ok = _write(tmp_path, 'OPENQASM 2.0;\ninclude "qelib1.inc";\nqreg q[1];\ncreg c[1];\n'
                          'h q[0];\nmeasure q[0] -> c[0];\n')
    assert parse_qasm_file(ok)[1] == [{"name": "H", "wires": [0], "parameters": []}]
    bad = _write(tmp_path, 'OPENQASM 2.0;\ninclude "qelib1.inc";\nqreg q[1];\ncreg c[1];\n'
                           'h q[0];\nmeasure q[0] -> c[0];\nz q[0];\n')
    with pytest.raises(QASMParseError, match=r"circuit\.qasm:7:.*terminal"):
        parse_qasm_file(bad)


@pytest.mark.parametrize("source,line,message", [
    ('OPENQASM 2.0;\ninclude "unknown.inc";\nqreg q[1];\n', 2, "include"),
    ('OPENQASM 2.0;\ninclude "qelib1.inc";\nqreg q[1];\nnonsense;\n', 4, "statement"),
    ('OPENQASM 2.0;\ninclude "qelib1.inc";\nqreg q[2];\ngate pair a,b {\n  cx a,b;\n}\npair q[0];\n',
     7, "expects 2 wires"),
])
def test_unknown_includes_statements_and_arity(tmp_path, source, line, message):
    p = _write(tmp_path, source)
    with pytest.raises(QASMParseError, match=rf"circuit\.qasm:{line}:.*{message}"):
        parse_qasm_file(p)


def test_more_rejections_and_register_forms():
    with pytest.raises(QASMParseError, match=r"<string>:1:.*OPENQASM 2.0"):
        parse_qasm("qreg q[1];")
    with pytest.raises(QASMParseError, match=r":3:.*out of range"):
        parse_qasm('OPENQASM 2.0;\nqreg q[2];\nx q[2];')
    with pytest.raises(QASMParseError, match=r":3:.*distinct wires"):
        parse_qasm('OPENQASM 2.0;\nqreg q[2];\ncx q[0],q[0];')
    with pytest.raises(QASMParseError, match=r":3:.*single-qubit gate"):
        parse_qasm('OPENQASM 2.0;\nqreg q[2];\ncx q;')
    with pytest.raises(QASMParseError, match=r":3:.*opaque"):
        parse_qasm('OPENQASM 2.0;\nqreg q[2];\nopaque g a;')
    with pytest.raises(QASMParseError, match=r":3:.*division by zero"):
        parse_qasm('OPENQASM 2.0;\nqreg q[1];\nrx(1/0) q[0];')
    # two quantum registers concatenate; register-wide gates broadcast;
    # barriers are no-ops; aliases map to the canonical names
    n, ops = parse_qasm('OPENQASM 2.0;\nqreg a[2];\nqreg b[1];\nh a;\nbarrier a,b[0];\n'
                        'cp(pi/4) a[1],b[0];\ncu1(0.5) b[0],a[0];\nu(1,2,3) b[0];\nccx a[0],a[1],b[0];')
    assert n == 3
    assert [o["name"] for o in ops] == ["H", "H", "CPHASE", "CPHASE", "U3", "CCX"]
    assert ops[2]["wires"] == [1, 2] and ops[2]["parameters"] == [pytest.approx(math.pi / 4)]


# --- the reference's fixtures ----------------------------------------------


def test_bell_fixture_counts_and_probabilities():
    n, ops = parse_qasm_file(GOLDEN / "bell.qasm")
    assert n == 2 and len(ops) >= 2
    p = _probs(n, ops)
    assert np.allclose(p, [0.5, 0, 0, 0.5], atol=1e-12)


def test_inverse_qft_fixture_rejected_at_line_14():
    with pytest.raises(QASMParseError, match=r"inverseqft_n4\.qasm:14:.*classical control"):
        parse_qasm_file(GOLDEN / "inverseqft_n4.qasm")


@pytest.mark.parametrize("name", ["qft_n4", "deutsch_n2", "grover_2qubit", "wstate_n3", "vqe_n4"])
def test_fixture_norms(name):
    n, ops = parse_qasm_file(GOLDEN / f"{name}.qasm")
    assert abs(_probs(n, ops).sum() - 1.0) < 1e-12


def test_fixture_closed_forms():
    # Grover on 2 qubits with the |11> oracle finds |11> with certainty
    n, ops = parse_qasm_file(GOLDEN / "grover_2qubit.qasm")
    assert np.allclose(_probs(n, ops), [0, 0, 0, 1], atol=1e-12)
    # Deutsch for the balanced f(x) = x: qubit 0 (the MSB) reads 1
    n, ops = parse_qasm_file(GOLDEN / "deutsch_n2.qasm")
    p = _probs(n, ops)
    assert abs(p[2] + p[3] - 1.0) < 1e-12
    # W state: 1/3 on each one-hot state (the fixture's angle 1.91063 is
    # 2 acos(1/sqrt 3) rounded to 6 digits)
    n, ops = parse_qasm_file(GOLDEN / "wstate_n3.qasm")
    p = _probs(n, ops)
    assert np.allclose(p[[1, 2, 4]], 1 / 3, atol=5e-6) and abs(p[[1, 2, 4]].sum() - 1) < 1e-10
    # QFT of a basis state has uniform magnitudes
    n, ops = parse_qasm_file(GOLDEN / "qft_n4.qasm")
    assert np.allclose(_probs(n, ops), 1 / 16, atol=1e-12)


def test_qasm_request_for_the_serving_loop():
    from mettleq_amd.serving import qasm_request
    r = qasm_request(GOLDEN / "bell.qasm", id=4, shots=10, seed=9)
    assert (r.id, r.n, r.shots, r.seed) == (4, 2, 10, 9)
    assert [o["name"] for o in r.ops] == ["H", "CNOT"]
    r2 = qasm_request('OPENQASM 2.0;\nqreg q[3];\nccx q[0],q[1],q[2];\n')
    assert r2.n == 3 and r2.ops[0]["name"] == "CCX"
    from mettleq_amd.circuit import encode_ops
    codes, wires, params = encode_ops(r2.ops)
    assert list(wires[:3]) == [0, 1, 2]
