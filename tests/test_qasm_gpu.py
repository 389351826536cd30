"""The reference's QASM fixtures (tests/golden/qasm) through the HIP engine:
Device.execute (fusion + dispatch) and the serving loop's native circuit
ingestion, against the complex128 oracle (5e-6 max-abs for complex64
amplitudes, the reference's kernel A/B budget)."""
from pathlib import Path

import numpy as np
import pytest

from mettleq_amd.qasm import parse_qasm_file
from oracle import sv_oracle as O

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden" / "qasm"
FIXTURES = ["bell", "qft_n4", "deutsch_n2", "grover_2qubit", "wstate_n3", "vqe_n4"]


@pytest.mark.parametrize("name", FIXTURES)
def test_fixture_amplitudes_through_device(gpu, name):
    from mettleq_amd.device import Device
    n, ops = parse_qasm_file(GOLDEN / f"{name}.qasm")
    want = O.run_ops(n, ops)
    dev = Device(n)
    try:
        dev.execute(ops)
        got = dev.native.to_numpy()
    finally:
        dev.close()
    assert np.max(np.abs(got - want)) < 5e-6


@pytest.mark.parametrize("name", FIXTURES)
def test_fixture_served_probabilities_and_shots(gpu, name):
    from mettleq_amd.native_state import NativeState
    from mettleq_amd.serving import qasm_request, run_request
    req = qasm_request(GOLDEN / f"{name}.qasm", id=3, shots=4000, seed=77)
    want = np.abs(O.run_ops(req.n, req.ops)) ** 2
    with NativeState(req.n) as st:
        outcomes, counts, probs = run_request(st, req, keep_probabilities=True)
        amps = st.to_numpy()
    assert np.max(np.abs(probs - want)) < 5e-6
    assert counts.sum() == 4000
    # shots only land on outcomes of nonzero probability, and match the
    # counter-based sampler of the CPU restatement on the same amplitudes
    assert np.all(want[outcomes] > 1e-9)
    ref = O.sample_indices(O.amp_probs_c64(amps), 4000, 77, 0)
    ro, rc = np.unique(ref, return_counts=True)
    assert np.array_equal(ro, outcomes) and np.array_equal(rc, counts)
