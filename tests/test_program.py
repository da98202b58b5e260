"""QOA assembly for the gate path (host code): encodings equal the
reference's bytes, the interpreter reproduces the VM's handling (ignored
opcodes, the CHARLOAD advance, loops that run once, CZ masks), checked on
the CPU with the oracle's QuantumState restatement as the backend."""
import struct
from pathlib import Path

import numpy as np
import pytest

import pyoracle as po
from qoa_amd import program as P
from qoa_amd._native import QOA_CZ_REF_MASKS
from qoa_amd.state import gates_array

REF_PROGRAMS = Path("/root/reference/qoa_files")


class OracleBackend:
    """QuantumState semantics on the CPU (normalize after every gate)."""

    def __init__(self, n: int, seed: int = 0):
        self.qs = po.QS(n, seed=seed, flags=QOA_CZ_REF_MASKS)
        self.n = self.qs.n

    def apply(self, op, q0=0, q1=0, q2=0, theta=0.0):
        self.qs.apply(gates_array([(op, q0, q1, q2, theta, 0)]))

    def measure(self, q):
        return self.qs.measure(q)


def f64(x):
    return struct.pack("<d", x)


@pytest.mark.parametrize("line,want", [
    ("H 3", b"\x05\x03"), ("had 3", b"\x05\x03"), ("APPLYHADAMARD 3", b"\x05\x03"),
    ("CNOT 1 2", b"\x17\x01\x02"), ("CN 1 2", b"\x17\x01\x02"),
    ("X 4 ; flip it", b"\x07\x04"), ("APPLYBITFLIP 4", b"\x07\x04"),
    ("RZ 0 3", b"\x0f\x00" + f64(3.0)), ("P 2 -1.5e-1", b"\x08\x02" + f64(-0.15)),
    ("CPHASE 1 0 0.785", b"\x35\x01\x00" + f64(0.785)),
    ("CONTROLLEDPHASEROTATION 2 5 1", b"\x35\x02\x05" + f64(1.0)),
    ("CZ 3 4", b"\x1e\x03\x04"), ("SWAP 0 7", b"\x0b\x00\x07"), ("CS 1 2 3", b"\x0c\x01\x02\x03"),
    ("QMEAS 9", b"\x32\x09"), ("MEAS 9", b"\x32\x09"), ("QINIT 24", b"\x04\x18"),
    ("PHASE 1 0.5", b"\x24\x01" + f64(0.5)), ("SETPHASE 1 0.5", b"\x1d\x01" + f64(0.5)),
    ("CHARLOAD 2 'A'", b"\x31\x02\x41"), ("CHAROUT 6", b"\x18\x06"),
    ("REGSET 3 2.5", b"\x21\x03" + f64(2.5)), ("HALT", b"\xff"), ("RSA", b"\x1c"),
    ('VERBOSELOG 1 "hi there"', b"\x87\x01hi there\x00"),
    ('PRINTLN "ok"', b"\x9a" + struct.pack("<Q", 2) + b"ok"),
    ("SETPOS 2 1.0 2.0", b"\x74\x02" + f64(1.0) + f64(2.0)),
])
def test_encodings(line, want):
    assert P.encode_line(line) == want


@pytest.mark.parametrize("line", ["", "   ", "; a comment", "// another", "LOOP_A:"])
def test_lines_without_bytes(line):
    assert P.encode_line(line) is None


@pytest.mark.parametrize("line", ["QGATE 0 H", "H 256", "H -1", "RZ 0", "CHARLOAD 0 A",
                                  "H 1 2", ":", "X ; only a mnemonic", "RZ 0 1_0"])
def test_rejected_lines(line):
    with pytest.raises(P.UnsupportedInstruction):
        P.encode_line(line)


def test_compile_strict_and_lenient():
    text = "QINIT 0\nQGATE 0 H\nH 0\n"
    with pytest.raises(P.UnsupportedInstruction):
        P.compile_text(text)
    c = P.compile_text(text, strict=False)
    assert c.payload == b"\x04\x00\x05\x00" and [s[0] for s in c.skipped] == [2]


def test_exe_container_round_trip():
    payload = P.compile_text("H 0\nCNOT 0 1\nQMEAS 1\n").payload
    exe = P.write_exe(payload)
    assert exe[:4] == b"QEXE" and exe[4] == 1 and struct.unpack("<I", exe[5:9])[0] == len(payload)
    assert P.read_exe(exe) == payload
    with pytest.raises(ValueError):
        P.read_exe(b"ELF\0\1\0\0\0\0")


def test_qubit_count_is_the_vm_scan():
    assert P.scan_qubits(P.compile_text("H 0\nQMEAS 4\n").payload) == 5
    assert P.scan_qubits(P.compile_text("CPHASE 2 6 1.0\n").payload) == 7
    # the VM counts register operands too
    assert P.scan_qubits(P.compile_text("H 0\nREGSET 11 1.0\n").payload) == 12


def test_bell_pair_outcomes_agree():
    payload = P.compile_text("H 0\nCNOT 0 1\nQMEAS 0\nQMEAS 1\n").payload
    seen = set()
    for seed in range(12):
        res = P.execute(payload, OracleBackend(2, seed))
        (_, _, a), (_, _, b) = res.measurements
        assert a == b
        seen.add(a)
    assert seen == {0, 1}


def test_ignored_opcodes_and_cz_masks():
    """RESET, PHASE, SETPHASE, QINIT do nothing in the VM; CZ 1 2 acts with
    mask 1|2 = 3, i.e. on qubits 0 and 1."""
    b = OracleBackend(3, 1)
    P.execute(P.compile_text("X 0\nRESET 0\nPHASE 0 1.0\nSETPHASE 0 1.0\nQINIT 2\n").payload, b)
    want = np.zeros(8, complex)
    want[1] = 1
    assert np.array_equal(b.qs.a, want)
    b = OracleBackend(3, 1)
    P.execute(P.compile_text("X 0\nX 1\nCZ 1 2\n").payload, b)
    assert b.qs.a[3] == -1


def test_charload_advance_quirk():
    """CHARLOAD is 3 bytes but the VM advances 4: X 1 loses its opcode byte,
    its operand 0x01 reads as LOOPSTART and swallows H's opcode."""
    payload = P.compile_text("CHARLOAD 0 'A'\nX 1\nH 0\n").payload
    assert payload == b"\x31\x00\x41\x07\x01\x05\x00"
    b = OracleBackend(2, 0)
    res = P.execute(payload, b)
    assert res.registers[0] == 65.0
    assert b.qs.a[0] == 1 and np.count_nonzero(b.qs.a) == 1


def test_misaligned_stream_raises_or_skips_unknown_opcodes():
    """After CHARLOAD the VM reads CHAROUT's operand (0x01) as LOOPSTART,
    which takes X's opcode as its operand, and X's operand (0x02) lands as an
    opcode outside the subset: it raises by default and is skipped one byte
    with skip_unknown (the VM's rule for opcodes it does not know)."""
    payload = P.compile_text("CHARLOAD 0 'A'\nCHAROUT 1\nX 2\n").payload
    assert payload == b"\x31\x00\x41\x18\x01\x07\x02"
    with pytest.raises(P.UnsupportedInstruction):
        P.scan_qubits(payload)
    n = P.scan_qubits(payload, skip_unknown=True)
    res = P.execute(payload, OracleBackend(n, 0), skip_unknown=True)
    assert any("unknown opcode" in m for m in res.log)


def test_loops_run_once_and_halt_stops():
    b = OracleBackend(1, 0)
    res = P.execute(P.compile_text("LOOPSTART 0\nX 0\nLOOPEND\nHALT\nX 0\n").payload, b)
    assert b.qs.a[1] == 1 and res.halted
    assert any("loopend without matching loopstart" in m for m in res.log)


def test_charout_and_println_output():
    b = OracleBackend(1, 0)
    res = P.execute(P.compile_text('X 0\nCHAROUT 0\nPRINTLN "done"\n').payload, b)
    assert res.output == "\x01done\n"


def test_top_amplitudes_listing():
    a = np.zeros(8, complex)
    a[5], a[2], a[6] = 0.8, 0.36j, 0.48
    rows = P.top_amplitudes(a, 2)
    assert [r[0] for r in rows] == ["101", "110"] and abs(rows[0][3] - 0.64) < 1e-15
    assert len(P.top_amplitudes(a, 0)) == 8


@pytest.mark.skipif(not REF_PROGRAMS.exists(), reason="reference example programs not present")
def test_reference_example_programs_compile():
    """Every line of the reference's teleport example is in the subset; the
    other examples compile leniently (lines outside the subset recorded)."""
    full = P.compile_text((REF_PROGRAMS / "teleport.qoa").read_text())
    assert len(full.payload) > 0 and P.scan_qubits(full.payload) == 25
    for f in sorted(REF_PROGRAMS.glob("*.qoa")):
        c = P.compile_text(f.read_text(errors="replace"), strict=False)
        assert len(c.payload) > 0 or c.skipped, f.name
