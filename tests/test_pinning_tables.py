"""Host pinning tables (nqs.pinning, no GPU): every ExtField the factories build evaluates, on
the host, to the oracle's restatement of src/nqs/pinning.py:20-119 at random points; GaussianPin
rows carry the centre, width, v0 and field in the order the kernel reads them."""
import numpy as np
import pytest
import torch

from oracle import pinning as opn
from nqs import pinning as npn
from nqs.local_energy import pin_table

A = np.array([[7.0, 3.5], [0.0, 7.0 * np.sqrt(3) / 2]])
B = 2 * np.pi * np.linalg.inv(A).T
Q1, Q2 = B[:, 0], B[:, 1]
Q = np.array([0.4, -0.3])
KEYS = [(0, 0), (1, 0), (0, 1), (-1, -1), (2, -1)]
VALS = [(0.1, 0.0, 0.5), (0.2j, -0.1, 0.3 - 0.1j), (0.05, 0.4j, 1.0), (-0.3j, 0.1 + 0.2j, 0.7), (0.01, 0.02, 0.03j)]

CASES = {
    "skx": (lambda m: m.make_pin("skx", Q1, Q2, 0.1)),
    "skx_chiral_m0": (lambda m: m.make_skx_pin(Q1, Q2, 0.07, m0=-0.4, chirality=-1.0)),
    "skx_bloch": (lambda m: m.make_pin("skx_bloch", Q1, Q2, 0.1)),
    "fm": (lambda m: m.make_pin("fm", None, None, 0.2, axis=(1.0, -2.0, 0.5))),
    "spiral": (lambda m: m.make_pin("spiral", Q, None, 0.2)),
    "spiral_xy": (lambda m: m.make_pin("spiral_xy", Q, None, 0.2)),
    "overhauser": (lambda m: m.make_pin("overhauser", Q, None, 0.2)),
    "hf_skx": (lambda m: m.make_hf_skx_pin(KEYS, VALS, Q1, Q2, 0.15)),
}


@pytest.mark.parametrize("name", list(CASES))
def test_texture_tables_match_oracle(name):
    r = np.random.default_rng(3).uniform(-5, 12, (40, 2))
    host = CASES[name](npn)(r)
    ref = CASES[name](opn)(torch.tensor(r)).numpy()
    assert np.max(np.abs(host - ref)) < 1e-14


def test_fluid_is_no_field():
    assert npn.make_pin("fluid", Q1, Q2, 0.1) is None


def test_gaussian_pin_rows():
    pins = npn.make_gaussian_pins([(1.0, 2.0), (3.0, 4.0)], 0.05, 5.0, kind="spin", axis=(0, 0, 2))
    t = pin_table(pins)
    assert t.shape == (2, 8)
    assert np.array_equal(t[1], [3.0, 4.0, 5.0, 0.0, 0.0, 0.0, 0.05, 0.0])
    s = pin_table(npn.make_gaussian_pins([(1.0, 2.0)], 0.05, 5.0, kind="scalar"))
    assert np.array_equal(s[0], [1.0, 2.0, 5.0, 0.05, 0.0, 0.0, 0.0, 0.0])
    with pytest.raises(TypeError):
        pin_table([(1.0, 2.0, 0.05, 5.0)])
    with pytest.raises(ValueError):
        pin_table([npn.GaussianPin((0.0, 0.0), 0.0)])
