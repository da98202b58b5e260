"""QOA's saved-state JSON format (``qoa run --save-state/--load-state``).

The reference writes ``serde_json::to_string_pretty(&(qs, registers))``
(src/main.rs:4290-4296) and reads it back with
``serde_json::from_str::<(QuantumState, Vec<f64>)>`` (src/main.rs:3800-3827).
The document is a two-element array: the QuantumState object
(src/runtime/quantum_state.rs:37-53, ``rng`` skipped) and the register list.
Complex numbers are ``[re, im]`` pairs (num-complex's serde form);
``noise_config`` is ``null``, ``"Random"``, ``"Ideal"`` or ``{"Fixed": p}``.
"""
from __future__ import annotations

import json
from typing import Optional

import numpy as np

from .state import NoiseConfig, QuantumState


def _noise_to_json(cfg: Optional[NoiseConfig]):
    if cfg is None:
        return None
    if cfg.kind == "Fixed":
        return {"Fixed": cfg.p}
    return cfg.kind


def _noise_from_json(v) -> Optional[NoiseConfig]:
    if v is None:
        return None
    if isinstance(v, str):
        if v not in ("Random", "Ideal"):
            raise ValueError(f"unknown noise_config variant {v!r}")
        return NoiseConfig(v)
    if isinstance(v, dict) and list(v) == ["Fixed"]:
        return NoiseConfig.Fixed(float(v["Fixed"]))
    raise ValueError(f"malformed noise_config {v!r}")


def document(n: int, amps: np.ndarray, noise: Optional[NoiseConfig] = None, registers=(),
             regs=None, status=None) -> list:
    """The (QuantumState, Vec<f64>) document for host-side amplitudes."""
    n_regs = max(n, 10)  # QuantumState::new allocates n.max(10) registers (:70)
    regs = np.zeros(n_regs, dtype=np.complex128) if regs is None else np.asarray(regs)
    status = status or {"nan": False, "div_by_zero": False, "overflow": False}
    obj = {
        "n": int(n),
        "amps": [[float(a.real), float(a.imag)] for a in np.asarray(amps)],
        "regs": [[float(r.real), float(r.imag)] for r in regs],
        "status": dict(status),
        "noise_config": _noise_to_json(noise),
    }
    return [obj, [float(x) for x in registers]]


def parse_document(text: str) -> dict:
    """Validates and decodes a saved-state document (no device needed)."""
    doc = json.loads(text)
    if not (isinstance(doc, list) and len(doc) == 2 and isinstance(doc[0], dict)):
        raise ValueError("expected [QuantumState, registers]")
    obj, registers = doc
    for key in ("n", "amps", "regs", "status", "noise_config"):
        if key not in obj:
            raise ValueError(f"missing field `{key}`")
    n = int(obj["n"])
    amps = np.array([complex(re, im) for re, im in obj["amps"]], dtype=np.complex128)
    if amps.size != (1 << n):
        raise ValueError(f"amps has {amps.size} entries, expected {1 << n}")
    regs = np.array([complex(re, im) for re, im in obj["regs"]], dtype=np.complex128)
    return {"n": n, "amps": amps, "regs": regs, "status": obj["status"],
            "noise": _noise_from_json(obj["noise_config"]),
            "registers": [float(x) for x in registers]}


def state_to_document(qs: QuantumState, registers=(), regs=None, status=None) -> list:
    return document(qs.n, qs.download(), qs.noise_config, registers, regs, status)


def dumps(qs: QuantumState, registers=(), **kw) -> str:
    """serde_json's pretty form: 2-space indentation, one value per line."""
    return json.dumps(state_to_document(qs, registers, **kw), indent=2)


def save_state(path: str, qs: QuantumState, registers=(), **kw) -> None:
    with open(path, "w") as f:
        f.write(dumps(qs, registers, **kw))


def loads(text: str) -> tuple[QuantumState, list, dict]:
    """Returns (state on the device, registers, extra fields: regs and status)."""
    d = parse_document(text)
    qs = QuantumState(d["n"], d["noise"])
    qs.upload(d["amps"])  # stored verbatim: the reference does not renormalise on load
    return qs, d["registers"], {"regs": d["regs"], "status": d["status"]}


def load_state(path: str):
    with open(path) as f:
        return loads(f.read())
