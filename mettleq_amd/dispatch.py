"""Dispatch of fused operations onto a state's native entry points.

The routing of the reference's Device._apply / execute loop
(device.py:236-423): each structured layer goes to its layer function (the
reference's shaders.* seam, mettleq_amd.layers), and the single gates the
matcher leaves are queued and handed to the native planner in one
msv_apply_gates call. The same routine drives a real state (Device.execute)
and a virtual one (NativeState.planner), whose native planner counts the
sweeps and logs the kernels the real state would launch: the execution
plan's prediction and the fusion pass's cost model both come from it.
"""
from __future__ import annotations

from typing import Any, Dict, List, Optional

import numpy as np

from . import gates as G
from . import layers as L


def dispatch_op(st, op: Dict[str, Any], n: int) -> None:
    """One fused (structured) operation on `st`."""
    name = op.get("name")
    if name == "_GATE":
        st.apply_gates(op["gates"])
    elif name == "_U2LISTLAYER":
        st.apply_1q_layer(list(np.asarray(op["mats"]).reshape(n, 2, 2)), active=op.get("active"))
    elif name in ("_U2LAYER", "_RXLAYER"):
        if name == "_RXLAYER":
            u = G.single_qubit_matrix("RX", [op["theta"]])
        else:
            u = G.single_qubit_matrix(op["gate"], op.get("params", []))
        st.apply_1q_layer([u] * n)
    elif name == "_XORLAYER":
        st.apply_xor_affine(op["rows"], op["c"])
    elif name == "_QFTSTAGE":
        st.apply_qft_stage(op["j"], op["m"], inverse=op["inverse"])
    elif name == "_QFTLAYER":
        st.apply_qft(0, n, inverse=False)
    elif name == "_DIAGLAYER":
        L.diag_pair_layer(st, op["theta"], n, op["bonds"])
    elif name == "_DIAGWEIGHTED":
        L.diag_weighted_layer(st, n, op["bonds"], op["thetas"])
    elif name == "_ZZLAYER":
        L.zz_chain_layer(st, op["theta"], n, op["bonds"])
    elif name == "_ZZWEIGHTED":
        L.zz_weighted_layer(st, n, op["bonds"], op["thetas"])
    elif name == "_XXLAYER":
        L.xx_layer(st, op["theta"], n, op["bonds"])
    elif name == "_YYLAYER":
        L.yy_layer(st, op["theta"], n, op["bonds"])
    elif name == "_CULISTLAYER":
        L.controlled_u_groups(st, n, op["groups"])
    elif name == "_CHAINPHASE_RXLAYER":
        # the diagonal stays pending and the RX layer's first sweep reads it
        if op["phase_kind"] == "zz":
            L.zz_chain_layer(st, op["theta"], n, op["bonds"])
        else:
            L.diag_pair_layer(st, op["theta"], n, op["bonds"])
        st.apply_1q_layer([G.single_qubit_matrix("RX", [op["rx_theta"]])] * n)
    else:
        raise ValueError(f"unknown fused op {name}")


def dispatch(st, fused: List[Dict[str, Any]], n: int, per_op: Optional[list] = None) -> None:
    """All fused operations in order; consecutive `_GATE` entries are merged
    into one gate list. With `per_op` (a list), appends for each dispatched
    entry (optimized-operation indices, kernel kinds launched during it)."""
    pending: List = []
    pending_idx: List[int] = []

    def run(idx, fn):
        if per_op is None:
            fn()
            return
        st.dispatch_log(reset=True)
        fn()
        per_op.append((idx, st.dispatch_log()))

    def flush():
        if pending:
            gates = list(pending)
            run(list(pending_idx), lambda: st.apply_gates(gates))
            pending.clear()
            pending_idx.clear()

    for i, op in enumerate(fused):
        if op.get("name") == "_GATE":
            pending.extend(op["gates"])
            pending_idx.append(i)
            continue
        flush()
        run([i], lambda op=op: dispatch_op(st, op, n))
    flush()


def planned_run(fused: List[Dict[str, Any]], n: int, dtype: str = "complex64", basis: bool = True,
                per_op: Optional[list] = None) -> Dict[str, Any]:
    """The native planner's account of `fused` on a virtual n-qubit state
    (msv_state_create_virtual): starting from a deferred basis state (basis)
    or an already written one, then the final flush a read would trigger.
    Returns the sweeps and the kernel kinds in launch order."""
    from .native_state import NativeState
    with NativeState.planner(n, dtype) as v:
        if basis:
            v.reset()
        v.dispatch_log(reset=True)
        v.sweep_count(reset=True)
        dispatch(v, fused, n, per_op)
        layer_sweeps = v.sweep_count()
        if per_op is not None:
            v.dispatch_log(reset=True)
        v.sync()
        kernels = v.dispatch_log()
        total = v.sweep_count()
    if per_op is not None:
        per_op.append((None, kernels))
        kernels = [k for _, ks in per_op for k in ks]
    return {"sweeps": total, "sweeps_before_flush": layer_sweeps, "kernels": kernels}
