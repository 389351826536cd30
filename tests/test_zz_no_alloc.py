"""msv_sample_noalloc (allocation-free sampling through the library's
static device scratch) against msv_sample: bit-identical shots on calls in
a row with shrinking inputs, chunked shots, the refusal above 31 qubits and
the refusal when a state's deferred work would need a new buffer."""
import ctypes
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(st, shots, wires, seed, stream_id=0):
    from mettleq_amd import _native as nat
    out = np.zeros(shots, dtype=np.uint64)
    w = None if wires is None else np.asarray(wires, dtype=np.int32)
    rc = nat.lib().msv_sample_noalloc(st.handle, nat.as_ptr(w, ctypes.c_int32) if w is not None else None,
                                      0 if w is None else len(w), shots, seed, stream_id,
                                      nat.as_ptr(out, ctypes.c_uint64))
    return rc, out


def _prepare(st, n, seed):
    """A circuit that leaves deferred work behind (the last phase layer, and
    above 13 qubits the last 1q layer's pending rotations): the
    allocation-free call flushes it."""
    from mettleq_amd import _native as nat
    from mettleq_amd import gates as G
    from mettleq_amd.layers import compose_inverse_affine
    from mettleq_amd.workloads import ring_bonds
    rng = np.random.default_rng(seed)
    st.reset()
    for _ in range(2):
        st.apply_1q_layer([G.single_qubit_matrix("U3", rng.uniform(0, 2 * math.pi, 3)) for _ in range(n)])
        bonds = ring_bonds(n)
        st.apply_phase_layer([nat.MSV_PHASE_ZZ] * len(bonds), bonds, list(rng.uniform(0, 1.5, len(bonds))))
        rows, c = compose_inverse_affine([{"name": "CNOT", "wires": [i, i + 1]} for i in range(n - 1)], n)
        st.apply_xor_affine(rows, c)
    st.apply_1q_layer([G.single_qubit_matrix("U3", rng.uniform(0, 2 * math.pi, 3)) for _ in range(n)])
    bonds = ring_bonds(n)
    st.apply_phase_layer([nat.MSV_PHASE_ZZ] * len(bonds), bonds, list(rng.uniform(0, 1.5, len(bonds))))


@pytest.mark.parametrize("calls", [[(22, 350_000, None, 11), (20, 3000, [19, 0, 7, 3], 12),
                                    (12, 500, [5, 2], 13)]])
def test_noalloc_sampling_matches_msv_sample(gpu, calls):
    from mettleq_amd import _native as nat
    from mettleq_amd.native_state import NativeState
    for n, shots, wires, seed in calls:
        print(f"noalloc sample: n={n} shots={shots} wires={wires} seed={seed}", flush=True)
        with NativeState(n) as st:
            _prepare(st, n, seed)
            assert st.pending() != 0  # deferred work for the allocation-free call to flush
            rc, got = _noalloc(st, shots, wires, seed, stream_id=3)
            assert rc == nat.MSV_OK, nat.last_error()
            assert st.pending() == 0
            want = st.sample(shots, wires, seed=seed, stream_id=3)
            assert np.array_equal(got, want)
            # and again on the flushed state (the scratch keeps the last call's contents)
            rc, again = _noalloc(st, shots, wires, seed, stream_id=3)
            assert rc == nat.MSV_OK and np.array_equal(again, want)


def test_noalloc_refuses_what_needs_allocation(gpu):
    from mettleq_amd import _native as nat
    from mettleq_amd.native_state import NativeState
    with NativeState(14) as st:
        # a fresh state has no staging ring yet: its deferred phase layer cannot
        # be launched without allocating one
        st.apply_phase_layer([nat.MSV_PHASE_ZZ], [(0, 1)], [0.3])
        rc, _ = _noalloc(st, 10, None, 1)
        assert rc == nat.MSV_E_REFUSED
        assert "msv_state_sync" in nat.last_error()
        assert st.pending() == nat.MSV_PENDING_DIAG  # nothing was lost
        st.sync()
        rc, got = _noalloc(st, 10, None, 1)
        assert rc == nat.MSV_OK
        assert np.array_equal(got, st.sample(10, None, seed=1))
    with NativeState(32) as big:
        rc, _ = _noalloc(big, 10, None, 1)
        assert rc == nat.MSV_E_REFUSED
        assert "31 qubits" in nat.last_error()
