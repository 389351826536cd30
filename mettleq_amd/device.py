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

from .dispatch import dispatch as _dispatch
from .execution import build_execution_plan, mark_synchronized, record_observed_dispatches
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
        if backend not in ("sv", "statevector", "mps"):
            raise ValueError("backend must be 'sv' (exact statevector) or 'mps' (matrix product state)")
        if execution_device is not None and str(execution_device).strip().lower() not in ("gpu", "auto"):
            raise ValueError("execution_device must be 'gpu' (the HIP engine) or None")
        self.execution_device = "gpu"
        self._allow_unsafe_statevector = allow_unsafe_statevector
        self.dtype = dtype
        self.last_execution_plan: Optional[Dict[str, Any]] = None
        if backend == "mps":
            # device.py:91-128: the MPS method with MPSOptions(dmax, eps)
            from .mps import MPSOptions, MPSState
            self.backend = "mps"
            self.sim = MPSState(self.wires, mps_opts or MPSOptions(), device=device_index)
            self.statevector_preflight = None
            return
        self.backend = "sv"
        self.sim = StateVectorSimulator(self.wires, allow_unsafe_statevector=allow_unsafe_statevector,
                                        dtype=dtype, device=device_index,
                                        preflight_mode=preflight_mode, reserved_bytes=reserved_bytes)
        self.statevector_preflight = self.sim.preflight

    # -- lifecycle ---------------------------------------------------------
    def reset(self) -> None:
        self.sim.reset()

    def close(self) -> None:
        self.sim.close()

    @property
    def native(self):
        if self.backend == "mps":
            raise AttributeError("the MPS backend has no statevector; use Device.sim")
        return self.sim.native

    def synchronize(self):
        if self.backend == "mps":
            return self.sim
        if self.last_execution_plan is None:
            self.native.sync()
            return self.native
        self.native.dispatch_log(reset=True)
        self.native.sync()
        mark_synchronized(self.last_execution_plan, self.native.sweep_count(),
                          flush_kernels=self.native.dispatch_log())
        return self.native

    # -- execution ---------------------------------------------------------
    def execute(self, operations: List[Dict[str, Any]], *, report: Optional[bool] = None):
        ops = [dict(op) for op in operations]
        if self.backend == "mps":
            # op by op through the reference's routing to device gates
            # (device.py:1138-1290); each gate is one MPS update
            from .fusion import gates_for_op
            from .mps import apply_device_gate
            for op in ops:
                for g in gates_for_op(op, self.wires):
                    apply_device_gate(self.sim, g)
            return self.sim
        fused = self.fuse(ops)
        if report is None:
            report = os.environ.get("METTLEQ_EXECUTION_REPORT", "0") == "1"
        basis = True
        if report:
            from . import _native as nat
            pend = self.native.pending()
            if pend & ~nat.MSV_PENDING_BASIS:
                # work deferred by an earlier call runs first, so that this
                # call's observed launches are its own
                self.native.sync()
                pend = self.native.pending()
            basis = bool(pend & nat.MSV_PENDING_BASIS)
        self.native.sweep_count(reset=True)
        self.last_execution_plan = (
            build_execution_plan(n_qubits=self.wires, operations=ops, optimized_operations=fused,
                                 statevector_preflight_data=self.statevector_preflight, dtype=self.dtype,
                                 basis=basis)
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
        if self.last_execution_plan is not None:
            per_op: list = []
            _dispatch(self.native, fused, self.wires, per_op)
            record_observed_dispatches(self.last_execution_plan, per_op)
        else:
            _dispatch(self.native, fused, self.wires)

    # -- fusion ------------------------------------------------------------
    def fuse(self, operations: List[Dict[str, Any]]) -> List[Dict[str, Any]]:
        return fuse_operations(operations, self.wires)

    # -- measurement passthroughs ---------------------------------------------
    def sample(self, shots: int = None, wires: Optional[List[int]] = None):
        return self.sim.sample(self.shots if shots is None else int(shots), wires)

    def counts(self, shots: int = None, wires: Optional[List[int]] = None):
        return self.sim.sample_counts(self.shots if shots is None else int(shots), wires)
