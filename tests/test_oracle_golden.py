"""The committed golden vectors (tests/golden/sart_32x32x8_5tilts.npz, written by
tests/golden/make_golden.py from the FP64 oracle) are what the oracle computes today: a change of
the oracle's projector, backprojector or SART composition shows here."""
import sys
from pathlib import Path

import numpy as np
import pytest

sys.path.insert(0, str(Path(__file__).resolve().parent / "golden"))
import make_golden as MG  # noqa: E402

FIXTURE = Path(__file__).resolve().parent / "golden" / "sart_32x32x8_5tilts.npz"


@pytest.fixture(scope="module")
def golden():
    with np.load(FIXTURE, allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


def test_fixture_inputs_are_the_seeded_ones(golden):
    vol, projs, resid = MG.inputs()
    assert np.array_equal(golden["volume"], vol)
    assert np.array_equal(golden["projections"], projs)
    assert np.array_equal(golden["residual"], resid)
    assert np.array_equal(golden["tilts"], MG.TILTS)


def test_oracle_reproduces_golden_vectors(golden):
    now = MG.compute()
    for key in ("march_proj", "march_dist", "bp_os2", "sart_1iter"):
        ref = golden[key]
        assert np.abs(ref).max() > 0, key
        np.testing.assert_allclose(now[key], ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max(), err_msg=key)
