"""Writes tests/golden/kats.json: the reference's gate known-answer tests as data.

Each case is transcribed from src/test.rs (planetryan/qoa): the initial
amplitudes, the call (function name and arguments exactly as the reference
test passes them) and the expected amplitudes, evaluated to doubles. The
reference asserts each component to 1e-9 (src/test.rs:33-59).

    python tests/golden/make_kats.py
"""
from __future__ import annotations

import cmath
import json
import math
from pathlib import Path

S2 = 1.0 / math.sqrt(2.0)
PI = math.pi


def basis(n: int, idx: int = 0, val: complex = 1.0):
    a = [0j] * (1 << n)
    a[idx] = val
    return a


def polar(th: float) -> complex:
    return cmath.exp(1j * th)


CASES = []


def case(name, src, fn, args, inp, exp, eps=1e-9, level="l1"):
    CASES.append(dict(name=name, src=src, fn=fn, args=args, level=level, eps=eps,
                      input=[[z.real, z.imag] for z in map(complex, inp)],
                      expected=[[z.real, z.imag] for z in map(complex, exp)]))


# ---- src/test.rs:63-471 (arch-independent gate KATs) ----
case("hadamard_on_0_qubit", "src/test.rs:63-74", "hadamard", [S2, 0.0, 0], basis(1), [S2, S2])
case("hadamard_on_1_qubit", "src/test.rs:76-96", "hadamard", [S2, 0.0, 1], basis(2, 0b01),
     [0, S2, 0, S2])
case("hadamard_on_superposition", "src/test.rs:98-109", "hadamard", [S2, 0.0, 0], [S2, S2], [1, 0])
case("x_on_0_qubit", "src/test.rs:113-119", "x", [0], basis(1), [0, 1])
case("x_on_1_qubit", "src/test.rs:121-136", "x", [1], basis(2, 0b01), [0, 0, 0, 1])
case("y_on_0_qubit", "src/test.rs:140-146", "y", [0], basis(1), [0, 1j])
case("y_on_1_qubit", "src/test.rs:148-155", "y", [0], [1j, 0], [0, -1])
case("z_on_0", "src/test.rs:159-164", "z", [0], basis(1), [1, 0])
case("z_on_1", "src/test.rs:166-169", "z", [0], [0, 1], [0, -1])
case("t_on_0", "src/test.rs:174-179", "t", [0], basis(1), [1, 0])
case("t_on_1", "src/test.rs:181-187", "t", [0], [0, 1], [0, complex(S2, S2)])
case("s_on_0", "src/test.rs:192-197", "s", [0], basis(1), [1, 0])
case("s_on_1", "src/test.rs:199-202", "s", [0], [0, 1], [0, 1j])
case("phaseshift_on_0", "src/test.rs:207-212", "phaseshift", [0, PI / 3], basis(1), [1, 0])
case("phaseshift_on_1", "src/test.rs:214-220", "phaseshift", [0, PI / 3], [0, 1],
     [0, complex(math.cos(PI / 3), math.sin(PI / 3))])
case("reset_on_0", "src/test.rs:225-230", "reset", [0], basis(1), [1, 0])
case("reset_on_1", "src/test.rs:232-236", "reset", [0], [0, 1], [1, 0])
case("reset_on_plus", "src/test.rs:238-245", "reset", [0], [S2, S2], [1, 0])
case("swap_2_qubits", "src/test.rs:250-264", "swap", [0, 1], basis(2, 0b01), [0, 0, 1, 0])
case("swap_3_qubits", "src/test.rs:268-288", "swap", [0, 1], basis(3, 0b001),
     [0, 0, 1, 0, 0, 0, 0, 0])
case("controlled_swap_3_qubits", "src/test.rs:293-313", "controlled_swap", [2, 1, 0],
     basis(3, 0b101), [0, 0, 0, 0, 0, 0, 1, 0])
c4, s4 = math.cos(PI / 4), math.sin(PI / 4)
case("rx_on_0", "src/test.rs:318-330", "rx", [0, PI / 2], basis(1), [complex(c4, 0), complex(0, -s4)])
case("ry_on_0", "src/test.rs:334-344", "ry", [0, PI / 2], basis(1), [complex(c4, 0), complex(s4, 0)])
case("rz_on_0", "src/test.rs:348-357", "rz", [0, PI / 2], basis(1), [cmath.exp(-1j * PI / 4), 0])
case("rz_on_1", "src/test.rs:359-362", "rz", [0, PI / 2], [0, 1], [0, cmath.exp(1j * PI / 4)])
case("cnot_2_qubits_00", "src/test.rs:367-378", "cnot", [1, 0], basis(2), [1, 0, 0, 0])
case("cnot_2_qubits_10", "src/test.rs:382-393", "cnot", [1, 0], basis(2, 0b10), [0, 0, 0, 1])
case("cz_2_qubits_00", "src/test.rs:398-409", "cz", [1, 0], basis(2), [1, 0, 0, 0])
case("cz_2_qubits_11", "src/test.rs:413-424", "cz", [1, 0], basis(2, 0b11), [0, 0, 0, -1])
case("cphase_2_qubits_00", "src/test.rs:429-441", "cphase", [1, 0, PI / 4], basis(2), [1, 0, 0, 0])
case("cphase_2_qubits_11", "src/test.rs:445-456", "cphase", [1, 0, PI / 4], basis(2, 0b11),
     [0, 0, 0, complex(math.cos(PI / 4), math.sin(PI / 4))])
case("reset_all", "src/test.rs:460-471", "reset_all", [],
     [complex(0.5, 0.5), 1j, 1, complex(0.2, 0.8)], [1, 0, 0, 0])

# ---- src/test.rs:815-1046 (x86_64 module) ----
case("x86_hadamard", "src/test.rs:828-843", "hadamard", [S2, 0.0, 0], basis(1), [S2, S2])
case("x86_x", "src/test.rs:845-851", "x", [0], basis(1), [0, 1])
case("x86_y_0", "src/test.rs:853-859", "y", [0], basis(1), [0, 1j])
case("x86_y_1", "src/test.rs:861-867", "y", [0], [0, 1], [-1j, 0])
case("x86_z", "src/test.rs:870-877", "z", [0], [0, 1], [0, -1])
case("x86_t", "src/test.rs:879-889", "phaseshift", [0, PI / 4], [0, 1], [0, cmath.exp(1j * PI / 4)])
case("x86_s", "src/test.rs:891-901", "phaseshift", [0, PI / 2], [0, 1], [0, cmath.exp(1j * PI / 2)])
case("x86_phaseshift", "src/test.rs:903-914", "phaseshift", [0, PI / 3], [0, 1],
     [0, cmath.exp(1j * PI / 3)])
nrm = math.sqrt((0.5 * 0.5 + 0.5 * 0.5) + 1.0)
case("x86_reset_unnormalised", "src/test.rs:916-931", "reset", [0],
     [complex(0.5, 0.5), 1j, 1, complex(0.2, 0.8)],
     [complex(0.5, 0.5) / nrm, 0, complex(1, 0) / nrm, 0])
case("x86_reset_single", "src/test.rs:933-937", "reset", [0], [1j, 1], [1j, 0])
case("x86_rx", "src/test.rs:940-951", "rx", [0, PI / 2], basis(1), [complex(S2, 0), complex(0, -S2)])
case("x86_ry", "src/test.rs:953-963", "ry", [0, PI / 2], basis(1), [complex(S2, 0), complex(S2, 0)])
case("x86_rz_0", "src/test.rs:965-976", "rz", [0, PI / 2], basis(1), [cmath.exp(-1j * PI / 4), 0])
case("x86_rz_1", "src/test.rs:978-982", "rz", [0, PI / 2], [0, 1], [0, cmath.exp(1j * PI / 4)])
case("x86_cnot", "src/test.rs:985-996", "cnot", [1, 0], basis(2), [1, 0, 0, 0])
case("x86_cz", "src/test.rs:998-1014", "cz", [1, 0], basis(2, 3), [0, 0, 0, -1])
case("x86_cphase", "src/test.rs:1016-1033", "cphase", [1, 0, PI / 4], basis(2, 3),
     [0, 0, 0, cmath.exp(1j * PI / 4)])
case("x86_reset_all", "src/test.rs:1035-1046", "reset_all", [],
     [complex(0.5, 0.5), 1j, 1, complex(0.2, 0.8)], [1, 0, 0, 0])

# ---- src/test.rs:2206-2530 (execute_arithmetic: wrapper + normalize) ----
case("qs_qinit", "src/test.rs:2206-2225", "QINIT", [0], [0, 1], [1, 0], level="qs")
case("qs_h", "src/test.rs:2271-2280", "H", [0], basis(1), [S2, S2], level="qs")
case("qs_cnot", "src/test.rs:2293-2306", "CNOT", [1, 0], basis(2, 0b10), [0, 0, 0, 1], level="qs")
case("qs_z", "src/test.rs:2308-2315", "Z", [0], [0, 1], [0, -1], level="qs")
case("qs_x", "src/test.rs:2317-2322", "X", [0], basis(1), [0, 1], level="qs")
case("qs_t", "src/test.rs:2324-2332", "T", [0], [0, 1], [0, cmath.exp(1j * PI / 4)], level="qs")
case("qs_s", "src/test.rs:2334-2342", "S", [0], [0, 1], [0, cmath.exp(1j * PI / 2)], level="qs")
case("qs_p", "src/test.rs:2344-2353", "PHASESHIFT", [0, PI / 2], [0, 1],
     [0, cmath.exp(1j * PI / 2)], level="qs")
case("qs_rst", "src/test.rs:2355-2362", "RESET", [0], [0, 1], [1, 0], level="qs")
case("qs_cswap", "src/test.rs:2364-2381", "CSWAP", [2, 1, 0], basis(3, 0b101),
     [0, 0, 0, 0, 0, 0, 1, 0], level="qs")
case("qs_ebell", "src/test.rs:2383-2390", "EBELL", [0, 1], basis(2), [S2, 0, 0, S2], level="qs")
case("qs_setp", "src/test.rs:2472-2481", "PHASESHIFT", [0, PI / 2], [0, 1],
     [0, cmath.exp(1j * PI / 2)], level="qs")
case("qs_phase_ionq", "src/test.rs:2519-2530", "PHASESHIFT", [0, PI / 4], [0, 1],
     [0, cmath.exp(1j * PI / 4)], level="qs")
case("qs_rstall", "src/test.rs:2455-2462", "RESET_ALL", [], basis(2, 0b11), [1, 0, 0, 0],
     level="qs")


def main() -> None:
    out = Path(__file__).with_name("kats.json")
    out.write_text(json.dumps({"source": "planetryan/qoa src/test.rs", "eps": 1e-9,
                               "cases": CASES}, indent=1))
    print(f"wrote {len(CASES)} cases to {out}")


if __name__ == "__main__":
    main()
