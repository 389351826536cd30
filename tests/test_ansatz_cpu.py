"""CPU checks of the layered-ansatz oracle (energies, parameter shift,
central differences) against closed forms; the GPU engine is compared with
this oracle in tests/test_ansatz_gpu.py."""
import math

import numpy as np
import pytest

from oracle import sv_oracle as O


@pytest.mark.parametrize("t", [0.3, 1.7, -2.2, math.pi])
def test_ry_expectation_and_parameter_shift(t):
    """<Z> after RY(t) is cos t and its parameter-shift derivative is -sin t
    (tests/mlxQSdkIntegrationsTest.py:182-192 of the reference)."""
    g = [("RY", 0, 0)]
    terms = [((0,), 1.0)]
    assert abs(O.ansatz_energy(1, g, [t], terms) - math.cos(t)) < 1e-12
    assert abs(O.ansatz_param_shift(1, g, [t], terms)[0] + math.sin(t)) < 1e-12
    assert abs(O.ansatz_central_diff(1, g, [t], terms)[0] + math.sin(t)) < 1e-7


def test_product_state_zz_closed_form():
    a, b = 0.7, -1.3
    g = [("RY", 0, 0), ("RY", 1, 1), ("CPHASE", 0, 1, math.pi)]  # CZ commutes with Z0 Z1
    terms = [((0, 1), 1.0)]
    assert abs(O.ansatz_energy(2, g, [a, b], terms) - math.cos(a) * math.cos(b)) < 1e-12
    ps = O.ansatz_param_shift(2, g, [a, b], terms)
    assert np.allclose(ps, [-math.sin(a) * math.cos(b), -math.cos(a) * math.sin(b)], atol=1e-12)


def test_rz_on_computational_state_has_zero_gradient():
    g = [("RZ", 0, 0), ("RY", 1, 1)]
    terms = [((0,), 1.0), ((1,), 0.5)]
    ps = O.ansatz_param_shift(2, g, [0.4, 0.9], terms)
    assert abs(ps[0]) < 1e-12
    assert abs(ps[1] + 0.5 * math.sin(0.9)) < 1e-12


@pytest.mark.parametrize("seed", range(4))
def test_parameter_shift_equals_central_differences(seed):
    rng = np.random.default_rng(seed)
    n = 4
    gates = []
    p = 0
    for layer in range(2):
        for q in range(n):
            gates.append((["RX", "RY", "RZ"][rng.integers(3)], q, p))
            p += 1
        gates.append(("CPHASE", 0, 2, math.pi))
        gates.append(("ZZPHASE", 1, 3, float(rng.uniform(0, 3))))
    gates.append(("RY", 2, 0))  # shared parameter
    params = rng.uniform(-math.pi, math.pi, p)
    terms = [((0, 1), 1.0), ((2,), -0.7), ((1, 3), 0.25)]
    ps = O.ansatz_param_shift(n, gates, params, terms)
    fd = O.ansatz_central_diff(n, gates, params, terms)
    assert np.max(np.abs(ps - fd)) < 1e-7


def test_observable_diag_msb_first():
    h = O.z_observable_diag(3, [((0,), 1.0)])
    # qubit 0 is the most significant bit
    assert list(h) == [1, 1, 1, 1, -1, -1, -1, -1]
