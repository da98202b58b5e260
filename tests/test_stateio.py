"""The saved-state JSON document (src/main.rs:3800-3827, 4290-4296)."""
import json
from pathlib import Path

import numpy as np
import pytest

from qoa_amd import NoiseConfig
from qoa_amd import stateio

GOLDEN = Path(__file__).resolve().parent / "golden" / "saved_state_2q.json"


def test_parse_serde_pretty_document():
    """A document in serde_json's pretty layout for (QuantumState, Vec<f64>)
    with a Bell-like state, Fixed(0.01) noise and two registers."""
    d = stateio.parse_document(GOLDEN.read_text())
    assert d["n"] == 2
    assert np.array_equal(d["amps"], np.array([0.7071067811865475, 0, 0, 0.7071067811865475j]))
    assert d["noise"].kind == "Fixed" and d["noise"].p == 0.01
    assert d["registers"] == [1.0, 2.5] and d["regs"].size == 10
    assert d["status"] == {"nan": False, "div_by_zero": False, "overflow": False}


def test_document_reproduces_golden_text():
    amps = np.array([0.7071067811865475, 0, 0, 0.7071067811865475j])
    doc = stateio.document(2, amps, NoiseConfig.Fixed(0.01), registers=[1.0, 2.5])
    assert json.dumps(doc, indent=2) == GOLDEN.read_text().rstrip("\n")


@pytest.mark.parametrize("noise", [None, "Random", "Ideal"])
def test_noise_variants_round_trip(noise):
    cfg = None if noise is None else NoiseConfig(noise)
    doc = stateio.document(1, np.array([1, 0]), cfg)
    back = stateio.parse_document(json.dumps(doc))
    assert (back["noise"] is None) if noise is None else back["noise"].kind == noise


def test_rejects_malformed():
    with pytest.raises(ValueError):
        stateio.parse_document("[{}, []]")
    doc = stateio.document(2, np.zeros(4))
    doc[0]["amps"] = doc[0]["amps"][:3]
    with pytest.raises(ValueError):
        stateio.parse_document(json.dumps(doc))
