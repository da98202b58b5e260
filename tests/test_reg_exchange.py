"""Index model of the register kernels' LDS exchanges (CPU): the 2^q rounds
(q = 0, 1, 2 by tile size) of a layout change and of thread-bit X flips,
with the buffer indexing written as the device code computes it (squeeze,
the XOR swizzle, per-thread part XOR per-register part). Every write lands inside its buffer, no two
writes of a round collide, every read finds a value written in the same
round, and the result is the intended permutation of the tile."""
import itertools

import numpy as np
import pytest

from tile_emulator import plan_program

K_REG = 4


def swz(x):
    return x ^ ((x >> 4) & 15)


def squeeze(l, bits):
    for p in sorted(bits, reverse=True):
        l = ((l >> (p + 1)) << p) | (l & ((1 << p) - 1))
    return l


def split(k):
    return 0 if k <= 10 else 1 if k == 11 else 2


def stride(k):
    return 1 << (k - split(k))


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


def xchg(vals, k, rold, rnew, tw_of, tr_of, A, B):
    """vals[t][r]: register r of thread t. Returns the registers after the
    exchange; tw_of / tr_of give a thread's write / read thread part."""
    T = 1 << (k - K_REG)
    Q = split(k)
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
    S = stride(k)
    lds = [None] * (2 * S)
    out = [list(v) for v in vals]
    for c in range(1 << Q):
        base = (c & 1) * S
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


@pytest.mark.parametrize("k", [5, 6, 8, 10, 11, 12])
def test_layout_change_rounds(k):
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
                   lambda t: set_layout(t, k, rnew), A, B)
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
                   lambda t: set_layout(t, k, rpos) ^ xt, 2, 3)
        want = [[v ^ xt for v in row] for row in vals]
        assert out == want


@pytest.mark.parametrize("tile", [8, 10, 11, 12])
def test_planner_layout_ops_keep_split_slots(tile):
    """A layout op keeps the bits of the slots its exchange is keyed by and
    changes at most 4 - q slots; every pass ends with the lowest local bits
    (up to three) outside the registers."""
    from qoa_amd import workloads as W
    g = W.layered_random(16, 12, seed=5, ladders=True)
    ops, spos, rng, k = plan_program(16, g, tile_qubits=tile)
    q = split(k)
    seen = 0
    for b0, b1, reg in rng:
        if not reg:
            continue
        layout = list(range(k - K_REG, k))
        for op in ops[b0:b1]:
            if int(op["kind"]) != 9:
                continue
            seen += 1
            new = [int(x) for x in op["ins"]]
            a, b = int(op["t"]), int(op["t2"])
            assert a < b and len(set(new)) == K_REG
            for s in [a, b][:q]:
                assert layout[s] == new[s]
            assert sum(x != y for x, y in zip(layout, new)) <= K_REG - q
            layout = new
        low = (1 << min(3, k - K_REG)) - 1
        assert not any(low >> b & 1 for b in layout)
    assert seen > 0


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
