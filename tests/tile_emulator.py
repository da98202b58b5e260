"""CPU emulator of the fused-pass op programs (test infrastructure).

Runs a plan produced by qoa_plan_program exactly as the device kernels define
it: per pass, per tile (base = tile index spread over the bits outside S),
every op on the tile's local indices with the per-gate arithmetic (products
written out so each multiply/add rounds on its own, as the kernels built
without FMA do). Register-mode programs are checked for consistency: a 2x2,
SWAP or MAT4 op's register slot must hold its target bit under the current
layout.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from qoa_amd import _native as N
from qoa_amd.state import gates_array

TILE_OP = np.dtype([("kind", "u1"), ("t", "u1"), ("t2", "u1"), ("nins", "u1"), ("ins", "u1", 4),
                    ("rsel", "u1"), ("rt", "u1"), ("outside", "u1"), ("nq", "u1"), ("tl", "<u4"),
                    ("gmask", "<u8"), ("gval", "<u8"), ("aux", "<u4"), ("lmask", "<u4"),
                    ("gi", "<u4"), ("psel", "u1"), ("qphys", "u1", 3), ("c", "<f8", 8)])
TK_H, TK_M, TK_X, TK_Y, TK_DNEG, TK_DPHASE, TK_RZ, TK_SWAP, TK_MAT4, TK_LAYOUT, TK_DIAGN = range(11)
# a merged diagonal term: register slots, thread bits, tile condition, factor
DIAG_TERM = np.dtype([("rs", "<u4"), ("tl", "<u4"), ("gmask", "<u8"), ("gval", "<u8"),
                      ("c", "<f8", 2), ("pad", "<u4", 2)])


def plan_program(n, gates, tile_qubits=0, low_qubits=0, flags=0, matrices=None):
    lib = N.load()
    f = lib.qoa_plan_program
    f.restype = C.c_int
    assert lib.qoa_tile_op_size() == TILE_OP.itemsize
    g = gates_array(gates)
    m = None if matrices is None else np.ascontiguousarray(matrices, dtype=np.complex128)
    opts = N.QoaFuseOpts(tile_qubits, low_qubits, flags, 0)
    nops, npass, k = C.c_uint32(), C.c_uint32(), C.c_uint32()
    mp = None if m is None else N.ptr(m)
    N.check(f(C.c_uint32(n), N.ptr(g), C.c_uint32(len(g)), mp, C.byref(opts), None,
              C.c_uint32(0), None, None, C.c_uint32(0), C.byref(nops), C.byref(npass),
              C.byref(k)))
    ops = np.zeros(nops.value, dtype=TILE_OP)
    spos = np.zeros((npass.value, 16), dtype=np.uint8)
    rng = np.zeros((npass.value, 3), dtype=np.uint32)
    N.check(f(C.c_uint32(n), N.ptr(g), C.c_uint32(len(g)), mp, C.byref(opts), N.ptr(ops),
              C.c_uint32(len(ops)), N.ptr(spos), N.ptr(rng), C.c_uint32(npass.value),
              C.byref(nops), C.byref(npass), C.byref(k)))
    return ops, spos, rng, k.value


def plan_store_positions(n, gates, tile_qubits=0, low_qubits=0, flags=0,
                         matrices=None) -> np.ndarray:
    """Per pass, the physical position each local bit is stored to."""
    lib = N.load()
    f = lib.qoa_plan_store_positions
    f.argtypes, f.restype = N._SIGS["qoa_plan_store_positions"]
    g = gates_array(gates)
    m = None if matrices is None else np.ascontiguousarray(matrices, dtype=np.complex128)
    opts = N.QoaFuseOpts(tile_qubits, low_qubits, flags, 0)
    cnt = C.c_uint32()
    mp = None if m is None else N.ptr(m)
    N.check(f(n, N.ptr(g), len(g), mp, C.byref(opts), None, 0, C.byref(cnt)))
    out = np.zeros((cnt.value, 16), dtype=np.uint8)
    N.check(f(n, N.ptr(g), len(g), mp, C.byref(opts), N.ptr(out), cnt.value, C.byref(cnt)))
    return out


def plan_pool(n, gates, tile_qubits=0, low_qubits=0, flags=0, matrices=None) -> np.ndarray:
    """The plan's coefficient pool (MAT4 matrices, merged diagonal terms)."""
    lib = N.load()
    f = lib.qoa_plan_pool
    f.argtypes, f.restype = N._SIGS["qoa_plan_pool"]
    g = gates_array(gates)
    m = None if matrices is None else np.ascontiguousarray(matrices, dtype=np.complex128)
    opts = N.QoaFuseOpts(tile_qubits, low_qubits, flags, 0)
    cnt = C.c_uint32()
    mp = None if m is None else N.ptr(m)
    N.check(f(n, N.ptr(g), len(g), mp, C.byref(opts), None, 0, C.byref(cnt)))
    pool = np.zeros(max(cnt.value, 1), dtype=np.float64)
    N.check(f(n, N.ptr(g), len(g), mp, C.byref(opts), N.ptr(pool), pool.size, C.byref(cnt)))
    return pool[: cnt.value]


def _cmul(xr, xi, yr, yi):
    return xr * yr - xi * yi, xr * yi + xi * yr


def _deposit(x: int, positions) -> int:
    for b in positions:
        lo = x & ((1 << b) - 1)
        x = ((x >> b) << (b + 1)) | lo
    return x


def _check_reg_fields(op, kind: int, lm: int, layout) -> None:
    """Register mode splits the controls by the current layout: rsel has slot
    j set iff local bit layout[j] is a control, tl holds the other controls,
    rt is the register slot of an RZ target (0xFF for a thread bit)."""
    R = sum(1 << b for b in layout)
    rsel = sum(1 << j for j, b in enumerate(layout) if lm >> b & 1)
    assert int(op["rsel"]) == rsel and int(op["tl"]) == lm & ~R, "control split"
    if kind == TK_RZ:
        t = int(op["t"])
        want = layout.index(t) if t in layout else 0xFF
        assert int(op["rt"]) == want, "RZ slot"


def run_program(a: np.ndarray, n: int, ops, spos, rng, k: int, pool=None,
                batched: bool = False, spos_st=None) -> None:
    """Applies the program to the complex128 state `a` in place (`pool`: the
    plan's coefficient pool, needed for merged diagonal ops; `batched`: a
    plan for the batched kernels, whose 12-qubit exchanges keep one slot
    instead of two; `spos_st`: the store positions of relabelling plans, whose
    register passes must then store the three low positions from thread
    index bits 0..2)."""
    re, im = a.real.copy(), a.imag.copy()
    T = 1 << k
    loc = np.arange(T)
    for p in range(len(rng)):
        pos = [int(x) for x in spos[p][:k]]
        pos_st = pos if spos_st is None else [int(x) for x in spos_st[p][:k]]
        assert sorted(pos_st) == sorted(pos), "a tile stores to its own positions"
        off = np.zeros(T, dtype=np.int64)
        off_st = np.zeros(T, dtype=np.int64)
        for j, (b, bs) in enumerate(zip(pos, pos_st)):
            off |= ((loc >> j) & 1) << b
            off_st |= ((loc >> j) & 1) << bs
        b0, b1, reg_mode = (int(x) for x in rng[p])
        if spos_st is not None and reg_mode:  # the loads: lanes 0..2 on positions 0..2
            assert sorted(pos[:3]) == [0, 1, 2], "coalesced loads"
        for tile in range(1 << (n - k)):
            layout = list(range(k - 4, k))  # every tile starts in the initial layout
            base = _deposit(tile, sorted(pos))
            idx = base | off
            idx_st = base | off_st
            vr, vi = re[idx].copy(), im[idx].copy()
            for op in ops[b0:b1]:
                kind = int(op["kind"])
                if kind == TK_LAYOUT:
                    new = [int(x) for x in op["ins"]]
                    sw = int(op["aux"])
                    if not sw >> 31:
                        # whole-tile exchange: the first q of slots (t, t2)
                        # keep their bit (the rounds are keyed by those bits)
                        q = 0 if k <= 11 else 1 if (k == 12 and batched) else 2
                        for ks in [int(op["t"]), int(op["t2"])][:q]:
                            assert layout[ks] == new[ks], "kept slots"
                        assert sorted(new) == sorted(set(new)), "slots hold distinct bits"
                        layout = new
                        continue
                    # a list of swaps: slot <-> thread index bit (a bit from
                    # outside the registers enters the slot) or slot <-> slot
                    sw &= 0x7FFFFFFF
                    cur = list(layout)
                    for i in range(sw >> 28):
                        e = (sw >> (7 * i)) & 0x7F
                        j, opd = e & 3, (e >> 2) & 15
                        if e & 64:
                            cur[j], cur[opd] = cur[opd], cur[j]
                        else:  # through slot 3: slot j takes slot 3's bit
                            cur[j], cur[3] = cur[3], None
                    for j in range(4):
                        if cur[j] is None:
                            assert new[j] not in layout, "a swapped-in bit was a register bit"
                        else:
                            assert new[j] == cur[j], "slot changed without a swap"
                    assert sorted(new) == sorted(set(new)), "slots hold distinct bits"
                    layout = new
                    continue
                if kind == TK_DIAGN:
                    # merged diagonal run: per group of terms with the same
                    # register slots rs, the product of the factors whose
                    # tile and thread conditions hold, into the amplitudes
                    # whose register bits hold rs
                    assert reg_mode and pool is not None
                    cnt, at = int(op["lmask"]), int(op["aux"])
                    terms = pool[at:at + 6 * cnt].view(DIAG_TERM)
                    assert len(terms) == cnt
                    slots = terms["rs"] & 0xF
                    assert np.all(np.diff(slots.astype(np.int64)) >= 0), "terms grouped"
                    # the group-end flag (bit 31) marks exactly each group's last term
                    ends = (terms["rs"] >> 31).astype(bool)
                    assert np.array_equal(ends, np.append(np.diff(slots) != 0, True)), "group ends"
                    spare = pool[at + 6 * cnt:at + 6 * cnt + 6]
                    assert len(spare) == 6 and not spare.any(), "spare term after the terms"
                    R = sum(1 << b for b in layout)
                    for rs in np.unique(slots):
                        ar, ai = np.ones(T), np.zeros(T)
                        for t in terms[slots == rs]:
                            if (base & int(t["gmask"])) != int(t["gval"]):
                                continue
                            tl = int(t["tl"])
                            assert tl & R == 0, "thread condition on a register bit"
                            on = (loc & tl) == tl
                            cr = np.where(on, t["c"][0], 1.0)
                            ci = np.where(on, t["c"][1], 0.0)
                            ar, ai = _cmul(ar, ai, cr, ci)
                        need = sum(1 << layout[j] for j in range(4) if int(rs) >> j & 1)
                        sel = (loc & need) == need
                        r2, i2 = _cmul(vr, vi, ar, ai)
                        vr, vi = np.where(sel, r2, vr), np.where(sel, i2, vi)
                    continue
                lm = int(op["lmask"])
                if reg_mode:
                    _check_reg_fields(op, kind, lm, layout)
                if (base & int(op["gmask"])) != int(op["gval"]):
                    continue
                c = op["c"]
                if kind in (TK_H, TK_M, TK_X, TK_Y):
                    t = int(op["t"])
                    if reg_mode:
                        assert layout[int(op["ins"][0])] == t, "register slot mismatch"
                    l0 = loc[((loc >> t) & 1 == 0) & ((loc & lm) == lm)]
                    l1 = l0 | (1 << t)
                    xr, xi, yr, yi = vr[l0], vi[l0], vr[l1], vi[l1]
                    if kind == TK_H:
                        o0 = _cmul(c[0], c[1], xr + yr, xi + yi)
                        o1 = _cmul(c[0], c[1], xr - yr, xi - yi)
                    elif kind == TK_M:
                        a0 = _cmul(c[0], c[1], xr, xi)
                        a1 = _cmul(c[2], c[3], yr, yi)
                        b0_ = _cmul(c[4], c[5], xr, xi)
                        b1_ = _cmul(c[6], c[7], yr, yi)
                        o0 = (a0[0] + a1[0], a0[1] + a1[1])
                        o1 = (b0_[0] + b1_[0], b0_[1] + b1_[1])
                    elif kind == TK_X:
                        o0, o1 = (yr, yi), (xr, xi)
                    else:
                        o0 = _cmul(yr, yi, 0.0, -1.0)
                        o1 = _cmul(xr, xi, 0.0, 1.0)
                    vr[l0], vi[l0] = o0
                    vr[l1], vi[l1] = o1
                elif kind in (TK_DNEG, TK_DPHASE):
                    sel = loc[(loc & lm) == lm]
                    if kind == TK_DNEG:
                        vr[sel], vi[sel] = vr[sel] * -1.0, vi[sel] * -1.0
                    else:
                        vr[sel], vi[sel] = _cmul(vr[sel], vi[sel], c[0], c[1])
                elif kind == TK_RZ:
                    t = int(op["t"])
                    one = (loc >> t) & 1 == 1
                    r1, i1 = _cmul(vr, vi, c[2], c[3])
                    r0, i0 = _cmul(vr, vi, c[0], c[1])
                    vr, vi = np.where(one, r1, r0), np.where(one, i1, i0)
                elif kind == TK_SWAP:
                    t, t2 = int(op["t"]), int(op["t2"])
                    if reg_mode:
                        assert {layout[int(op["ins"][0])], layout[int(op["ins"][1])]} == {t, t2}
                    m1, m2 = 1 << t, 1 << t2
                    l = loc[((loc & (m1 | m2)) == 0) & ((loc & lm) == lm)]
                    ur, ui = vr[l | m2].copy(), vi[l | m2].copy()
                    vr[l | m2], vi[l | m2] = vr[l | m1], vi[l | m1]
                    vr[l | m1], vi[l | m1] = ur, ui
                else:
                    raise NotImplementedError("MAT4 is covered by the GPU tests")
            if spos_st is not None and reg_mode:  # the stores, from the final layout
                lanes = [b for b in range(k) if b not in layout][:3]
                assert sorted(pos_st[b] for b in lanes) == [0, 1, 2], "coalesced stores"
            re[idx_st], im[idx_st] = vr, vi
    a.real[:] = re
    a.imag[:] = im
