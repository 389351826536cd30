"""CPU emulation of the device-side semantics (test infrastructure).

Applies the host's device-gate lists (msv_gate records) and fused layer ops
with the oracle, so host translation, gate decompositions and fusion can be
checked on the CPU against unfused dense application.
"""
import numpy as np

from mettleq_amd import _native as nat
from oracle import sv_oracle as O


def gate_matrix2(g):
    m = np.array([complex(g.m[2 * i], g.m[2 * i + 1]) for i in range(4)])
    return m


def apply_gate(state, n, g):
    kind = g.kind
    if kind == nat.MSV_G_U:
        t = g.q[0]
        cs = [g.q[1 + c] for c in range(g.ncontrols)]
        u = gate_matrix2(g).reshape(2, 2)
        if cs:
            return O.apply_controlled(state, n, u, cs, t)
        return O.apply_dense(state, n, u, [t])
    if kind == nat.MSV_G_DIAG1:
        d = gate_matrix2(g)[:2]
        return O.apply_dense(state, n, np.diag(d), [g.q[0]])
    if kind == nat.MSV_G_DIAG2:
        d = gate_matrix2(g)
        return O.apply_dense(state, n, np.diag(d), [g.q[0], g.q[1]])
    if kind == nat.MSV_G_SWAP:
        a, b = g.q[0], g.q[1]
        cs = [g.q[2 + c] for c in range(g.ncontrols)]
        if cs:
            m = np.eye(1 << (len(cs) + 2), dtype=complex)
            m[-4:, -4:] = O.SWAP()
            return O.apply_dense(state, n, m, cs + [a, b])
        return O.apply_dense(state, n, O.SWAP(), [a, b])
    raise ValueError(kind)


def apply_gates(state, n, gates):
    for g in gates:
        state = apply_gate(state, n, g)
    return state


def apply_fused(state, n, fused):
    """Semantics of each fused op (what the native entry point computes)."""
    from mettleq_amd import gates as G
    for op in fused:
        name = op["name"]
        if name == "_GATE":
            state = apply_gates(state, n, op["gates"])
        elif name == "_U2LISTLAYER":
            mats = np.asarray(op["mats"]).reshape(n, 2, 2)
            for q in op["active"]:
                state = O.apply_dense(state, n, mats[q], [q])
        elif name in ("_U2LAYER", "_RXLAYER"):
            u = (G.single_qubit_matrix("RX", [op["theta"]]) if name == "_RXLAYER"
                 else G.single_qubit_matrix(op["gate"], op.get("params", [])))
            for q in range(n):
                state = O.apply_dense(state, n, u, [q])
        elif name == "_XORLAYER":
            state = O.xor_affine_gather(state, n, op["rows"], op["c"])
        elif name == "_QFTSTAGE":
            j, m = op["j"], op["m"]
            if not op["inverse"]:
                state = O.apply_dense(state, n, O.H(), [j])
                for k in range(j + 1, m):
                    state = O.apply_dense(state, n, O.CPHASE(np.pi / 2 ** (k - j)), [k, j])
            else:
                for k in reversed(range(j + 1, m)):
                    state = O.apply_dense(state, n, O.CPHASE(-np.pi / 2 ** (k - j)), [k, j])
                state = O.apply_dense(state, n, O.H(), [j])
        elif name == "_QFTLAYER":
            state = O.qft(state, n, list(range(n)))
        elif name in ("_DIAGLAYER", "_DIAGWEIGHTED"):
            th = op["thetas"] if name == "_DIAGWEIGHTED" else [op["theta"]] * len(op["bonds"])
            state = state * O.cphase_layer_diag(n, op["bonds"], th)
        elif name in ("_ZZLAYER", "_ZZWEIGHTED"):
            th = op["thetas"] if name == "_ZZWEIGHTED" else [op["theta"]] * len(op["bonds"])
            state = state * O.zz_layer_diag(n, op["bonds"], th)
        elif name in ("_XXLAYER", "_YYLAYER"):
            pname = "XXPHASE" if name == "_XXLAYER" else "YYPHASE"
            for a, b in op["bonds"]:
                state = O.apply_dense(state, n, O.pauli_pair_phase(pname, op["theta"]), [a, b])
        elif name == "_CHAINPHASE_RXLAYER":
            th = [op["theta"]] * len(op["bonds"])
            diag = (O.zz_layer_diag(n, op["bonds"], th) if op["phase_kind"] == "zz"
                    else O.cphase_layer_diag(n, op["bonds"], th))
            state = state * diag
            u = G.single_qubit_matrix("RX", [op["rx_theta"]])
            for q in range(n):
                state = O.apply_dense(state, n, u, [q])
        elif name == "_CULISTLAYER":
            for group in op["groups"]:
                for g in group:
                    state = O.apply_controlled(state, n, np.asarray(g["matrix"]).reshape(2, 2),
                                               [g["control"]], g["target"])
        else:
            raise ValueError(name)
    return state
