"""The committed golden fixture (tests/golden/c1.npz, written by
tests/golden/make_golden.py from the oracle at the small plain-method configuration)
stays reproducible: the seeded input generator gives the same CSR, and the oracle
gives the same outputs."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_golden as MG  # noqa: E402


def _fixture():
    return np.load(os.path.join(HERE, "golden", "c1.npz"), allow_pickle=False)


def test_fixture_inputs_reproduce():
    z = _fixture()
    indptr, indices, data = MG.inputs()
    assert MG.csr_digest(indptr, indices, data) == str(z["csr_sha256"])


def test_fixture_oracle_reproduces(oracle):
    z = _fixture()
    out = MG.oracle_outputs(*MG.inputs())
    for k, v in out.items():
        np.testing.assert_array_equal(np.asarray(v), z[k], err_msg=k)
