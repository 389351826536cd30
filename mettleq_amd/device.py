"""Device: runtime gate fusion and dispatch onto the HIP layer kernels.

The fusion pass (mettleq_amd.fusion) restates the reference's pattern
matcher (device.py:556-1087); the dispatcher here (device.py:236-423)
routes each structured layer to its native entry point. Everything the matcher leaves as single gates is queued and handed to the
native planner in one msv_apply_gates call, which packs the ordered gates
into tile sweeps (up to ten target qubits and any number of diagonal or
control bits per sweep). Dispatch is asynchronous on the state's stream;
``synchronize`` waits for it.
"""
from __future__ import annotations

import os
from typing import Any, Dict, List, Optional

import numpy as np

from . import gates as G
from . import layers as L
from .execution import build_execution_plan, mark_synchronized
from .fusion import fuse_operations
from .sim import StateVectorSimulator

class Device:
    def __init__(self, wires: int, shots: int = 1000, backend: Optional[str] = None, mps_opts=None, *,
                 allow_unsafe_statevector: Optional[bool] = None,
                 execution_device: Optional[str] = None, dtype: str = "complex64",
                 device_index: int = 0, preflight_mode: str = "out_of_place",
                 reserved_bytes: int = 0, **_ignored):
        self.wires = int(wires)
        self.shots = int(shots)
        backend = (backend or os.environ.get("METTLEQ_BACKEND", "sv")).strip().lower()
        if backend not in ("sv", "statevector"):
            raise ValueError("this engine implements the exact statevector backend only ('sv')")
        if execution_device is not None and str(execution_device).strip().lower() not in ("gpu", "auto"):
            raise ValueError("execution_device must be 'gpu' (the HIP engine) or None")
        self.backend = "sv"
        self.execution_device = "gpu"
        self._allow_unsafe_statevector = allow_unsafe_statevector
        self.dtype = dtype
        self.sim = StateVectorSimulator(self.wires, allow_unsafe_statevector=allow_unsafe_statevector,
                                        dtype=dtype, device=device_index,
                                        preflight_mode=preflight_mode, reserved_bytes=reserved_bytes)
        self.statevector_preflight = self.sim.preflight
        self.last_execution_plan: Optional[Dict[str, Any]] = None

    # -- lifecycle ---------------------------------------------------------
    def reset(self) -> None:
        self.sim.reset()

    def close(self) -> None:
        self.sim.close()

    @property
    def native(self):
        return self.sim.native

    def synchronize(self):
        self.native.sync()
        mark_synchronized(self.last_execution_plan, self.native.sweep_count())
        return self.native

    # -- execution ---------------------------------------------------------
    def execute(self, operations: List[Dict[str, Any]], *, report: Optional[bool] = None):
        ops = [dict(op) for op in operations]
        fused = self.fuse(ops)
        if report is None:
            report = os.environ.get("METTLEQ_EXECUTION_REPORT", "0") == "1"
        self.native.sweep_count(reset=True)
        self.last_execution_plan = (
            build_execution_plan(n_qubits=self.wires, operations=ops, optimized_operations=fused,
                                 statevector_preflight_data=self.statevector_preflight, dtype=self.dtype)
            if report else None)
        self.dispatch(fused)
        return self.native

    def explain(self, operations: List[Dict[str, Any]]) -> Dict[str, Any]:
        fused = self.fuse([dict(op) for op in operations])
        return build_execution_plan(n_qubits=self.wires, operations=list(operations),
                                    optimized_operations=fused,
                                    statevector_preflight_data=self.statevector_preflight,
                                    dtype=self.dtype)

    def dispatch(self, fused: List[Dict[str, Any]]) -> None:
        n = self.wires
        st = self.native
        pending: List = []

        def flush():
            if pending:
                st.apply_gates(pending)
                pending.clear()

        for op in fused:
            name = op.get("name")
            if name == "_GATE":
                pending.extend(op["gates"])
                continue
            flush()
            if name == "_U2LISTLAYER":
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
        flush()

    # -- fusion ------------------------------------------------------------
    def fuse(self, operations: List[Dict[str, Any]]) -> List[Dict[str, Any]]:
        return fuse_operations(operations, self.wires)

    # -- measurement passthroughs ---------------------------------------------
    def sample(self, shots: int = None, wires: Optional[List[int]] = None):
        return self.sim.sample(self.shots if shots is None else int(shots), wires)

    def counts(self, shots: int = None, wires: Optional[List[int]] = None):
        return self.sim.sample_counts(self.shots if shots is None else int(shots), wires)
