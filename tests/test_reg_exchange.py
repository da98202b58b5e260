"""Index model of the register kernels' LDS exchanges (CPU): the 2^q rounds
(q = 0, 1, 2 by tile size; at 12 qubits 2 in the fused kernel and 1 in the
batched kernels) of a layout change and of thread-bit X flips
through one buffer,
with the buffer indexing written as the device code computes it (squeeze,
the XOR swizzle, per-thread part XOR per-register part). Every write lands inside its buffer, no two
writes of a round collide, every read finds a value written in the same
round, and the result is the intended permutation of the tile."""
import itertools

import numpy as np
import pytest

from tile_emulator import plan_program

K_REG = 4
PLAN_FORCE_REG, PLAN_SWAPS = 1 << 31, 1 << 29  # internal planner flags (planner.hpp)


def swz(x):
    return x ^ ((x >> 4) & 15)


def squeeze(l, bits):
    for p in sorted(bits, reverse=True):
        l = ((l >> (p + 1)) << p) | (l & ((1 << p) - 1))
    return l


def split(k, batched):
    return 0 if k <= 11 else 1 if (k == 12 and batched) else 2


def stride(k, batched):
    return 1 << (k - split(k, batched))


def set_layout(tid, k, rpos):
    rmask = sum(1 << b for b in rpos)
    l, src = 0, 0
    for b in range(k):
        if rmask >> b & 1:
            continue
        l |= ((tid >> src) & 1) << b
        src += 1
    return l


def regl(rpos, r):
    return sum(1 << rpos[j] for j in range(K_REG) if r >> j & 1)


def xchg(vals, k, rold, rnew, tw_of, tr_of, A, B, batched):
    """vals[t][r]: register r of thread t. Returns the registers after the
    exchange; tw_of / tr_of give a thread's write / read thread part."""
    T = 1 << (k - K_REG)
    Q = split(k, batched)
    sl = [A, B][:Q]
    bits = [rold[j] for j in sl]
    assert all(rnew[j] == rold[j] for j in sl)

    def offs(rpos):
        return [0 if j in sl else swz(squeeze(1 << rpos[j], bits)) for j in range(K_REG)]

    ow, orr = offs(rold), offs(rnew)
    def off(o, r):
        x = 0
        for j in range(K_REG):
            if r >> j & 1:
                x ^= o[j]
        return x
    rnd = lambda r: sum(((r >> j) & 1) << i for i, j in enumerate(sl))
    S = stride(k, batched)
    assert S <= 2048  # 32 KiB
    out = [list(v) for v in vals]
    for c in range(1 << Q):
        lds = [None] * S  # a round's reads see only that round's writes
        base = 0
        written = set()
        for t in range(T):
            bw = swz(squeeze(tw_of(t), bits))
            for r in range(1 << K_REG):
                if rnd(r) != c:
                    continue
                i = bw ^ off(ow, r)
                assert 0 <= i < S and i not in written
                written.add(i)
                lds[base + i] = vals[t][r]
        for t in range(T):
            br = swz(squeeze(tr_of(t), bits))
            for r in range(1 << K_REG):
                if rnd(r) != c:
                    continue
                i = br ^ off(orr, r)
                assert i in written
                out[t][r] = lds[base + i]
    return out


def tile_regs(k, rpos):
    """vals[t][r] = the local index held there (a tile whose amplitude is its index)."""
    T = 1 << (k - K_REG)
    return [[set_layout(t, k, rpos) | regl(rpos, r) for r in range(1 << K_REG)] for t in range(T)]


def test_swizzle_is_a_linear_bijection():
    """The buffer index map must be a bijection of [0, 2^m) for every buffer
    size and linear over XOR (the kernels split an index into a per-thread
    and a per-register part and XOR their images)."""
    for m in range(1, 12):
        img = [swz(x) for x in range(1 << m)]
        assert sorted(img) == list(range(1 << m))
    rng = np.random.default_rng(3)
    for a, b in rng.integers(0, 1 << 11, size=(200, 2)):
        assert swz(int(a) ^ int(b)) == swz(int(a)) ^ swz(int(b))


@pytest.mark.parametrize("batched", [False, True])
@pytest.mark.parametrize("k", [5, 6, 8, 10, 11, 12, 13])
def test_layout_change_rounds(k, batched):
    rng = np.random.default_rng(k)
    for _ in range(40 if k < 11 else 8):
        rold = [int(x) for x in rng.choice(k, K_REG, replace=False)]
        A, B = sorted(int(x) for x in rng.choice(K_REG, 2, replace=False))
        free = [b for b in range(k) if b not in rold]
        rnew = list(rold)
        for j in range(K_REG):
            if j not in (A, B) and free and rng.random() < 0.7:
                rnew[j] = free.pop(int(rng.integers(len(free))))
        vals = tile_regs(k, rold)
        out = xchg(vals, k, rold, rnew, lambda t: set_layout(t, k, rold),
                   lambda t: set_layout(t, k, rnew), A, B, batched)
        assert out == tile_regs(k, rnew)


@pytest.mark.parametrize("k", [6, 9, 11, 12])
def test_thread_bit_flip_rounds(k):
    rng = np.random.default_rng(100 + k)
    for _ in range(10):
        rpos = [int(x) for x in rng.choice(k, K_REG, replace=False)]
        tbits = [b for b in range(k) if b not in rpos]
        xt = sum(1 << b for b in tbits if rng.random() < 0.4) or (1 << tbits[0])
        vals = tile_regs(k, rpos)
        out = xchg(vals, k, rpos, rpos, lambda t: set_layout(t, k, rpos),
                   lambda t: set_layout(t, k, rpos) ^ xt, 2, 3, True)  # noise: batched only
        want = [[v ^ xt for v in row] for row in vals]
        assert out == want


def swaps_of(op):
    sw = int(op["aux"]) & 0x7FFFFFFF
    out = []
    for i in range(sw >> 28):
        e = (sw >> (7 * i)) & 0x7F
        out.append((e & 3, (e >> 2) & 15, bool(e & 64)))
    return out


def run_swaps(vals, k, slots, tb, swaps):
    """The device's data movement for a swap list (registers of every thread
    hold local indices). Slot j <-> thread bit m runs through slot 3: slots j
    and 3 trade registers, then with s = 3 a thread whose bit m is 0 keeps
    x = v[r] (bit s of r clear) and receives its partner's x into y =
    v[r | 2^s]; a thread whose bit m is 1 keeps y and receives its partner's y
    into x. Slot <-> slot renames registers."""
    T = len(vals)
    for j, opd, rr in swaps:
        if not rr and j != 3:
            run_swaps(vals, k, slots, tb, [(j, 3, True)])
            j = 3
        if rr:
            i2 = opd
            for t in range(T):
                row = list(vals[t])
                for r in range(16):
                    if (r >> j) & 1 and not (r >> i2) & 1:
                        r2 = r ^ (1 << j) ^ (1 << i2)
                        row[r], row[r2] = row[r2], row[r]
                vals[t] = row
            slots[j], slots[i2] = slots[i2], slots[j]
            continue
        m = opd
        old = [list(row) for row in vals]
        for t in range(T):
            partner = t ^ (1 << m)
            for r in range(16):
                if (r >> j) & 1:
                    continue
                r1 = r | (1 << j)
                if (t >> m) & 1:
                    vals[t][r] = old[partner][r1]
                else:
                    vals[t][r1] = old[partner][r]
        slots[j], tb[m] = tb[m], slots[j]


def expect(k, slots, tb):
    T = 1 << (k - K_REG)
    return [[sum(((t >> m) & 1) << tb[m] for m in range(k - K_REG)) |
             sum(((r >> j) & 1) << slots[j] for j in range(K_REG)) for r in range(16)]
            for t in range(T)]


@pytest.mark.parametrize("tile", [8, 10, 11, 12, 13])
def test_planner_swap_lists_move_the_tile(tile):
    """Every layout op of a pass, applied with the device's data movement,
    leaves each thread's registers holding the amplitudes the planner's slots
    name; a pass ends with the lowest local bits (up to three) on the lowest
    thread index bits."""
    from qoa_amd import workloads as W
    g = W.layered_random(16, 12, seed=5, ladders=True)
    # the batched kernel's plans (swaps for tiles of up to 10 qubits)
    ops, spos, rng, k = plan_program(16, g, tile_qubits=tile, flags=PLAN_SWAPS | PLAN_FORCE_REG)
    seen = 0
    swaps = 0
    for b0, b1, reg in rng:
        if not reg:
            continue
        slots = list(range(k - K_REG, k))
        tb = list(range(k - K_REG))
        vals = expect(k, slots, tb)
        for op in ops[b0:b1]:
            if int(op["kind"]) != 9:
                continue
            seen += 1
            if not int(op["aux"]) >> 31:  # whole-tile exchange (modelled above)
                slots = [int(x) for x in op["ins"]]
                tb = [b for b in range(k) if b not in slots]
                vals = expect(k, slots, tb)
                continue
            run_swaps(vals, k, slots, tb, swaps_of(op))
            swaps += 1
            assert slots == [int(x) for x in op["ins"]]
            assert vals == expect(k, slots, tb)
        nlow = min(3, k - K_REG)
        assert sorted(tb[:nlow]) == list(range(nlow))
    assert seen > 0
    assert (swaps > 0) == (k <= 10)


def test_planner_mat4_slots():
    """A 4x4 op runs with its lower local bit in register slot 2 and its
    higher one in slot 3."""
    from test_planner import OP, mixed_circuit
    g, mats = mixed_circuit(14, 300, seed=11)
    ops, spos, rng, k = plan_program(14, g, tile_qubits=12, matrices=mats)
    seen = 0
    for b0, b1, reg in rng:
        if not reg:
            continue
        layout = list(range(k - K_REG, k))
        for op in ops[b0:b1]:
            kind = int(op["kind"])
            if kind == 9:
                layout = [int(x) for x in op["ins"]]
            elif kind == 8:
                seen += 1
                lo, hi = sorted((int(op["t"]), int(op["t2"])))
                assert (int(op["ins"][0]), int(op["ins"][1])) == (2, 3)
                assert layout[2] == lo and layout[3] == hi
    assert seen > 0
