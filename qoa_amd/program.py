"""QOA assembly (``.qoa``) for the gate path: text -> bytecode -> execution.

The compiler follows the reference's line handling (src/main.rs:1677-1700,
src/instructions.rs:1849-1940): blank lines and lines starting with ``;`` or
``//`` are skipped, ``NAME:`` is a label (no bytes), inline ``;`` comments are
stripped, tokens split on whitespace outside ``"..."`` literals, the mnemonic
is matched case-insensitively through the short-name table, and arguments
parse as u8 / f64 / ``'c'`` / ``"text"`` exactly. Encodings are the
reference's bytes (src/instructions.rs:631-1406), so programs compiled here
and by the reference are the same bytes.

Only the gate path is covered: the quantum instructions, measurement, the
character/register instructions the example programs use, and the
instructions the reference VM accepts but ignores. A line outside that set
raises ``UnsupportedInstruction`` (``strict=True``) or is skipped with a
warning, as the reference skips lines it cannot parse.

The interpreter reproduces the reference VM (src/main.rs:2314-2760) for these
opcodes, quirks included:

* ``QINIT`` (0x04), ``RESET`` (0x0A), ``SETPHASE`` (0x1D), ``PHASE`` (0x24),
  ``ENTANGLE``/``ENTANGLEBELL`` (0x11/0x12) and the position/wavelength ops
  are read and ignored;
* ``CHARLOAD`` is encoded in 3 bytes but the VM advances 4 (the next
  instruction loses its first byte);
* ``LOOPSTART`` only advances; ``LOOPEND`` with no open loop is reported and
  skipped, so loops run once;
* ``CZ``/``CPHASE`` go through ``QuantumState::apply_cz`` /
  ``apply_controlled_phase``, i.e. with the integer ``c|t`` as the mask
  (``cz_ref_masks``);
* measurement errors are reported and execution continues.

The qubit count is the VM's first-pass maximum of operand bytes + 1.
"""
from __future__ import annotations

import re
import struct
import sys
from dataclasses import dataclass, field
from typing import Callable, Optional

# canonical mnemonic -> (opcode, argument kinds); q = u8, f = f64,
# c = 'x' char literal, s = "text" (NUL-terminated), L = "text" (u64 length prefix)
_TABLE = {
    "QINIT": (0x04, "q"), "INITQUBIT": (0x04, "q"), "QINITQ": (0x04, "q"),
    "QMEAS": (0x32, "q"), "MEASURE": (0x32, "q"),
    "APPLYHADAMARD": (0x05, "q"),
    "CONTROLLEDNOT": (0x17, "qq"),
    "APPLYPHASEFLIP": (0x06, "q"), "APPLYBITFLIP": (0x07, "q"),
    "APPLYTGATE": (0x0D, "q"), "APPLYSGATE": (0x0E, "q"),
    "PHASESHIFT": (0x08, "qf"),
    "RESET": (0x0A, "q"), "QRESET": (0x0A, "q"),
    "SWAP": (0x0B, "qq"), "CONTROLLEDSWAP": (0x0C, "qqq"),
    "ENTANGLE": (0x11, "qq"), "ENTANGLEBELL": (0x12, "qq"),
    "RESETALL": (0x1C, ""), "SETPHASE": (0x1D, "qf"),
    "RX": (0x22, "qf"), "RY": (0x23, "qf"), "RZ": (0x0F, "qf"), "PHASE": (0x24, "qf"),
    "CZ": (0x1E, "qq"), "CONTROLLEDPHASEROTATION": (0x35, "qqf"),
    "CHARLOAD": (0x31, "qc"), "CHAROUT": (0x18, "q"), "REGSET": (0x21, "qf"),
    "LOOPSTART": (0x01, "q"), "LOOPEND": (0x10, ""),
    "SETPOS": (0x74, "qff"), "SETWL": (0x75, "qf"), "WLSHIFT": (0x76, "qf"),
    "MOVE": (0x77, "qff"), "MARKOBSERVED": (0x79, "q"), "RELEASE": (0x7A, "q"),
    "HALT": (0xFF, ""), "VERBOSELOG": (0x87, "qs"), "PRINTLN": (0x9A, "L"),
}

# short names (src/instructions.rs:1919-2039, the entries for the subset)
_ALIASES = {
    "QI": "QINIT", "IQ": "INITQUBIT", "QM": "QMEAS", "MEAS": "MEASURE", "M": "MEASURE",
    "HAD": "APPLYHADAMARD", "H": "APPLYHADAMARD", "CNOT": "CONTROLLEDNOT",
    "CN": "CONTROLLEDNOT", "Z": "APPLYPHASEFLIP", "X": "APPLYBITFLIP", "T": "APPLYTGATE",
    "S": "APPLYSGATE", "P": "PHASESHIFT", "RST": "RESET", "RSTQ": "RESET", "QR": "QRESET",
    "CSWAP": "CONTROLLEDSWAP", "CS": "CONTROLLEDSWAP", "EBELL": "ENTANGLEBELL",
    "EB": "ENTANGLEBELL", "RSTALL": "RESETALL", "RSA": "RESETALL", "VLOG": "VERBOSELOG",
    "VL": "VERBOSELOG", "SETP": "SETPHASE", "SPH": "SETPHASE", "CLOAD": "CHARLOAD",
    "CLD": "CHARLOAD", "COUT": "CHAROUT", "CO": "CHAROUT", "RSET": "REGSET",
    "RGST": "REGSET", "LSTART": "LOOPSTART", "LS": "LOOPSTART", "LEND": "LOOPEND",
    "LE": "LOOPEND", "CPHASE": "CONTROLLEDPHASEROTATION",
    "APPLYCPHASE": "CONTROLLEDPHASEROTATION", "CPR": "CONTROLLEDPHASEROTATION",
    "SPOS": "SETPOS", "STP": "SETPOS", "SWL": "SETWL", "SW": "SETWL", "WLS": "WLSHIFT",
    "WLSH": "WLSHIFT", "MOV": "MOVE", "MV": "MOVE", "MOBS": "MARKOBSERVED",
    "MO": "MARKOBSERVED", "REL": "RELEASE", "RL": "RELEASE", "HLT": "HALT",
    "PLN": "PRINTLN", "PSE": "PHASE",
}


class UnsupportedInstruction(ValueError):
    """A line outside the gate-path subset (or not a valid instruction)."""


@dataclass
class Compiled:
    payload: bytes
    skipped: list = field(default_factory=list)  # (line number, text, reason)


def tokenize(part: str) -> list[str]:
    """Whitespace split that keeps "..." literals (with their quotes) whole."""
    toks, cur, in_str = [], "", False
    for ch in part:
        if ch == '"':
            if in_str:
                toks.append(cur + ch)
                cur, in_str = "", False
            else:
                if cur:
                    toks.append(cur)
                cur, in_str = ch, True
        elif ch.isspace() and not in_str:
            if cur:
                toks.append(cur)
                cur = ""
        else:
            cur += ch
    if cur:
        toks.append(cur)
    return toks


_U8 = re.compile(r"^\+?[0-9]+$")


def _u8(tok: str) -> int:
    if not _U8.match(tok) or int(tok) > 255:
        raise UnsupportedInstruction(f"invalid u8: {tok}")
    return int(tok)


def _f64(tok: str) -> float:
    if "_" in tok or tok != tok.strip():
        raise UnsupportedInstruction(f"invalid f64: {tok}")
    try:
        return float(tok)
    except ValueError:
        raise UnsupportedInstruction(f"invalid f64: {tok}") from None


def _string(tok: str) -> bytes:
    if len(tok) >= 2 and tok[0] == '"' and tok[-1] == '"':
        return tok[1:-1].encode()
    raise UnsupportedInstruction(f"invalid string literal: {tok}")


def encode_line(line: str) -> Optional[bytes]:
    """The bytes of one source line (None for a label, a comment or a blank
    line); raises UnsupportedInstruction for anything else it cannot encode."""
    t = line.strip()
    if not t or t.startswith(";") or t.startswith("//"):
        return None
    if t.endswith(":"):
        if t[:-1] == "":
            raise UnsupportedInstruction("empty label name")
        return None
    part = t.split(";", 1)[0].strip()
    if not part:
        raise UnsupportedInstruction("empty instruction after stripping comment")
    toks = tokenize(part)
    name = toks[0].upper()
    name = _ALIASES.get(name, name)
    if name == "LABEL" and len(toks) == 2:
        return None
    if name not in _TABLE:
        raise UnsupportedInstruction(f"{toks[0]} is not in the gate-path instruction subset")
    op, kinds = _TABLE[name]
    args = toks[1:]
    if len(args) != len(kinds):
        raise UnsupportedInstruction(f"{name} takes {len(kinds)} arguments, got {len(args)}")
    out = bytearray([op])
    for kind, tok in zip(kinds, args):
        if kind == "q":
            out.append(_u8(tok))
        elif kind == "f":
            out += struct.pack("<d", _f64(tok))
        elif kind == "c":
            if len(tok) == 3 and tok[0] == "'" and tok[2] == "'":
                out.append(ord(tok[1]) & 0xFF)
            else:
                raise UnsupportedInstruction("usage: charload <reg> '<char>'")
        elif kind == "s":
            out += _string(tok) + b"\0"
        else:  # L
            s = _string(tok)
            out += struct.pack("<Q", len(s)) + s
    return bytes(out)


def compile_text(text: str, strict: bool = True) -> Compiled:
    """Compiles a .qoa program; lenient mode skips (and records) the lines the
    subset cannot encode, as the reference compiler skips unparseable lines."""
    payload = bytearray()
    skipped = []
    for no, line in enumerate(text.splitlines(), 1):
        try:
            b = encode_line(line)
        except UnsupportedInstruction as e:
            if strict:
                raise UnsupportedInstruction(f"line {no}: {line.strip()!r}: {e}") from None
            skipped.append((no, line.strip(), str(e)))
            print(f"warning: skipping line {no} {line.strip()!r}: {e}", file=sys.stderr)
            continue
        if b:
            payload += b
    return Compiled(bytes(payload), skipped)


# ---- executable container (src/main.rs:1703-1745) -------------------------
MAGIC_QEXE = b"QEXE"
_MAGICS = (b"QEX ", b"QEXE", b"OEXE", b"QOX ", b"XEXE", b"QX\0\0")


def write_exe(payload: bytes, magic: bytes = MAGIC_QEXE) -> bytes:
    return magic + bytes([1]) + struct.pack("<I", len(payload)) + payload


def read_exe(data: bytes) -> bytes:
    if len(data) < 9 or data[:4] not in _MAGICS:
        raise ValueError("not a QOA executable (header)")
    n = struct.unpack("<I", data[5:9])[0]
    if len(data) < 9 + n:
        raise ValueError("truncated QOA executable")
    return data[9:9 + n]


# ---- interpreter ------------------------------------------------------------
# bytes per opcode as the VM advances (CHARLOAD: 4, one more than encoded)
_FIXED = {0x04: 2, 0x05: 2, 0x06: 2, 0x07: 2, 0x0D: 2, 0x0E: 2, 0x0A: 2, 0x32: 2, 0x18: 2,
          0x79: 2, 0x7A: 2, 0x01: 2, 0x08: 10, 0x22: 10, 0x23: 10, 0x0F: 10, 0x24: 10,
          0x1D: 10, 0x21: 10, 0x75: 10, 0x76: 10, 0x17: 3, 0x1E: 3, 0x0B: 3, 0x11: 3,
          0x12: 3, 0x0C: 4, 0x31: 4, 0x35: 11, 0x74: 18, 0x77: 18, 0x1C: 1, 0x10: 1,
          0xFF: 1, 0x00: 1, 0x8D: 1}
# opcodes whose operand bytes the VM's first pass counts as qubits
_SCAN_Q = {0x04: 1, 0x05: 1, 0x06: 1, 0x07: 1, 0x0D: 1, 0x0E: 1, 0x0A: 1, 0x32: 1, 0x18: 1,
           0x79: 1, 0x7A: 1, 0x08: 1, 0x22: 1, 0x23: 1, 0x0F: 1, 0x24: 1, 0x1D: 1, 0x21: 1,
           0x75: 1, 0x76: 1, 0x74: 1, 0x77: 1, 0x17: 2, 0x1E: 2, 0x0B: 2, 0x11: 2, 0x12: 2,
           0x0C: 1, 0x31: 1, 0x87: 1, 0x35: 2}


def _length(payload: bytes, i: int, skip_unknown: bool = False) -> int:
    op = payload[i]
    if op in _FIXED:
        return _FIXED[op]
    if op == 0x87:  # q, NUL-terminated text
        end = payload.find(b"\0", i + 2)
        return (len(payload) if end < 0 else end + 1) - i
    if op == 0x9A:
        return 9 + struct.unpack("<Q", payload[i + 1:i + 9])[0]
    if skip_unknown:  # the VM's handling of an opcode it does not know: one byte
        return 1
    raise UnsupportedInstruction(f"opcode 0x{op:02X} at byte {i} is outside the gate-path subset")


def scan_qubits(payload: bytes, skip_unknown: bool = False) -> int:
    """The VM's first pass (src/main.rs:1880-2285): max operand byte + 1.
    Opcodes outside the subset raise, or (skip_unknown) advance one byte as
    the VM does for opcodes it does not know."""
    i, mx = 0, 0
    while i < len(payload):
        op = payload[i]
        k = _SCAN_Q.get(op, 0) if (op in _FIXED or op in (0x87, 0x9A)) else 0
        for j in range(k):
            if i + 1 + j < len(payload):
                mx = max(mx, payload[i + 1 + j])
        i += _length(payload, i, skip_unknown)
    return 0 if (mx == 0 and not payload) else mx + 1


@dataclass
class RunResult:
    n_qubits: int
    measurements: list = field(default_factory=list)  # (byte offset, qubit, outcome)
    output: str = ""
    registers: list = field(default_factory=lambda: [0.0] * 24)
    log: list = field(default_factory=list)
    halted: bool = False


def execute(payload: bytes, backend, registers: Optional[list] = None,
            echo: Callable[[str], None] = lambda s: None,
            skip_unknown: bool = False) -> RunResult:
    """Runs `payload` on `backend` (methods apply(op_name, q0, q1, q2, theta),
    measure(q) -> int raising ValueError for a bad qubit, attribute n).
    An opcode outside the subset raises UnsupportedInstruction, or with
    skip_unknown is logged and skipped one byte, as the VM skips opcodes it
    does not know (the VM itself runs the ones it knows)."""
    res = RunResult(n_qubits=backend.n)
    if registers is not None:
        res.registers = list(registers)
    regs = res.registers
    loops: list = []
    i = 0

    def f64(at):
        return struct.unpack("<d", payload[at:at + 8])[0]

    while i < len(payload):
        op = payload[i]
        step = _length(payload, i, skip_unknown)
        if i + step > len(payload):
            res.log.append(f"incomplete instruction at byte {i}")
            break
        if op == 0x05:
            backend.apply("H", payload[i + 1])
        elif op == 0x06:
            backend.apply("Z", payload[i + 1])
        elif op == 0x07:
            backend.apply("X", payload[i + 1])
        elif op == 0x0D:
            backend.apply("T", payload[i + 1])
        elif op == 0x0E:
            backend.apply("S", payload[i + 1])
        elif op == 0x08:
            backend.apply("PHASESHIFT", payload[i + 1], theta=f64(i + 2))
        elif op == 0x0F:
            backend.apply("RZ", payload[i + 1], theta=f64(i + 2))
        elif op == 0x22:
            backend.apply("RX", payload[i + 1], theta=f64(i + 2))
        elif op == 0x23:
            backend.apply("RY", payload[i + 1], theta=f64(i + 2))
        elif op == 0x0B:
            backend.apply("SWAP", payload[i + 1], payload[i + 2])
        elif op == 0x0C:
            backend.apply("CSWAP", payload[i + 1], payload[i + 2], payload[i + 3])
        elif op == 0x17:
            backend.apply("CNOT", payload[i + 1], payload[i + 2])
        elif op == 0x1E:
            backend.apply("CZ", payload[i + 1], payload[i + 2])
        elif op == 0x35:
            backend.apply("CPHASE", payload[i + 1], payload[i + 2], theta=f64(i + 3))
        elif op == 0x1C:
            backend.apply("RESET_ALL")
        elif op in (0x32, 0x18):
            q = payload[i + 1]
            try:
                o = backend.measure(q)
            except ValueError as e:
                res.log.append(f"error measuring qubit {q}: {e}")
                o = 0
            if op == 0x32:
                res.measurements.append((i, q, o))
            else:
                res.output += chr(o & 0xFF)
                echo(chr(o & 0xFF))
        elif op == 0x21:
            r = payload[i + 1]
            if r < len(regs):
                regs[r] = f64(i + 2)
            else:
                res.log.append(f"invalid register index {r} for regset at byte {i}")
        elif op == 0x31:
            r = payload[i + 1]
            if r < len(regs):
                regs[r] = float(payload[i + 2])
            else:
                res.log.append(f"invalid register index {r} for charload at byte {i}")
        elif op == 0x10:
            if loops:
                start, left = loops.pop()
                if left > 1:
                    loops.append((start, left - 1))
                    i = start
                    continue
            else:
                res.log.append(f"loopend without matching loopstart at byte {i}")
        elif op == 0x87:
            res.log.append(payload[i + 2:i + step - 1].decode(errors="replace"))
        elif op == 0x9A:
            text = payload[i + 9:i + step].decode(errors="replace")
            res.output += text + "\n"
            echo(text + "\n")
        elif op == 0xFF:
            res.halted = True
            break
        elif op not in _FIXED:
            res.log.append(f"unknown opcode 0x{op:02X} at byte {i}, skipping")
        # 0x04 0x0A 0x11 0x12 0x1D 0x24 0x01 0x74-0x77 0x79 0x7A 0x00 0x8D: read, ignored
        i += step
    return res


class HipBackend:
    """Executes on the device through QuantumState (layer-2 semantics)."""

    def __init__(self, qs):
        self.qs = qs
        self.n = qs.n

    def apply(self, op: str, q0: int = 0, q1: int = 0, q2: int = 0, theta: float = 0.0) -> None:
        self.qs._apply(op, q0, q1, q2, theta)

    def measure(self, q: int) -> int:
        return self.qs.measure(q)


def run_text(text: str, seed: Optional[int] = None, noise_config=None, strict: bool = True,
             echo: Callable[[str], None] = lambda s: None, skip_unknown: bool = False):
    """Compiles and runs a program on the device; returns (QuantumState,
    RunResult). The state has the VM's qubit count and uses the reference's
    CZ/CPHASE mask reading."""
    from .state import QuantumState

    payload = compile_text(text, strict=strict).payload
    n = max(1, scan_qubits(payload, skip_unknown))
    qs = QuantumState(n, noise_config=noise_config, cz_ref_masks=True)
    if seed is not None:
        qs.seed_rng(seed)
    res = execute(payload, HipBackend(qs), echo=echo, skip_unknown=skip_unknown)
    return qs, res


def top_amplitudes(amps, top_n: int = 10, noise_strength: float = 0.0, rng=None):
    """print_amplitudes' listing (src/main.rs:1748-1817): (basis string with
    qubit n-1 first, re, im, prob) sorted by probability, ties in index
    order; top_n = 0 lists all. noise_strength adds the reference's display
    noise (normal, per component) from `rng`."""
    import numpy as np

    a = np.asarray(amps)
    n = max(1, int(a.size).bit_length() - 1)
    re, im = a.real.copy(), a.imag.copy()
    if noise_strength:
        rng = rng or np.random.default_rng()
        re += rng.standard_normal(a.size) * noise_strength
        im += rng.standard_normal(a.size) * noise_strength
    prob = re * re + im * im
    order = np.argsort(-prob, kind="stable")
    if top_n:
        order = order[:top_n]
    return [(format(int(i), f"0{n}b"), float(re[i]), float(im[i]), float(prob[i])) for i in order]
