"""The persistent pobtaf launch's task order (no GPU): every readiness flag a task waits on is raised
by a task listed before it, so the oldest unfinished task can always proceed and the grid drains.

The table is the library's own (`dh_dag_table`); the flags each task kind waits on and raises are
restated from `dag_task` in dalia_amd/csrc/dalia_hip.hip (OLD_*: previous block's terms and the own
terms below the lookahead; PRE: the chain helpers; CHAIN: the diagonal tile, which also applies the
term k = c-2 itself; DIAG/LOWER/ARROW: the finishing tasks, which also wait for the factor of their
column's diagonal tile).
"""

import ctypes

import pytest

from dalia_amd import _lib

PRE_D, PRE_L, CHAIN, DIAG, LOWER, ARROW, TIP, OLD_D, OLD_L, OLD_X, OLD_LW, OLD_A, QX = range(13)


def block_table(lib, nbt, last, arrow):
    look, slook = ctypes.c_int(), ctypes.c_int()
    cnt = lib.dh_dag_table(nbt, int(last), int(arrow), None, 0, ctypes.byref(look), ctypes.byref(slook))
    assert cnt > 0
    buf = (ctypes.c_int * cnt)()
    assert lib.dh_dag_table(nbt, int(last), int(arrow), buf, cnt, None, None) == cnt
    return [((v >> 24) & 0xff, (v >> 12) & 0xfff, v & 0xfff) for v in buf], look.value, slook.value


def waits_and_raises(i, kind, r, c, nbt, arrow, look, slook):
    """(flags waited on, flags raised) of task (kind, r, c) of block i."""
    D = lambda rr, kk, b=i: ("D", b, rr, kk)
    L = lambda rr, kk, b=i: ("L", b, rr, kk)
    A = lambda kk, b=i: ("A", b, kk)
    kold = max(0, r - look)
    kpre = r - 2 if r >= 2 else r - 1
    ks = max(0, c - slook)
    w = []
    if kind == TIP:
        w += [A(k) for k in range(nbt)] + ([("T", i - 1)] if i > 0 else [])
        return w, [("T", i)]
    if kind == OLD_D:
        if i > 0:
            w += [L(r, k, i - 1) for k in range(nbt - 1 if r == 0 else nbt)]
        w += [D(r, k) for k in range(kold)]
        return w, [("OD", i, r)]
    if kind == OLD_L:
        if i > 0:
            w += [L(r, k, i - 1) for k in range(nbt)] + [L(r - 1, k, i - 1) for k in range(nbt)]
        w += [D(r, k) for k in range(kold)] + [D(r - 1, k) for k in range(kold)]
        return w, [("OL", i, r)]
    if kind == PRE_D:
        w += [("OD", i, r)] + [D(r, k) for k in range(kold, kpre)]
        return w, [("PD", i, r)]
    if kind == PRE_L:
        w += [("OL", i, r)] + [D(r, k) for k in range(kold, kpre)] + [D(r - 1, k) for k in range(kold, kpre)]
        return w, [("PL", i, r)]
    if kind == QX:
        return [("PD", i, r), ("PL", i, r), D(r, r - 2), D(r - 1, r - 2)], [("Q", i, r)]
    if kind == CHAIN:
        if i >= 2 and arrow:
            w.append(A(c, i - 2))  # the tile-inverse slot's last reader in block i-2
        w.append(("OD", i, 0) if c == 0 else ("PD", i, c))
        if c >= 1:
            w.append(("PL", i, c))
            if c >= 2:  # the term k = c-2, applied by the chain task itself
                w += [D(c, c - 2), D(c - 1, c - 2)]
            w.append(D(c - 1, c - 1))
            return w, [D(c, c - 1), D(c, c)]
        if i > 0:
            w.append(L(0, nbt - 1, i - 1))
        return w, [D(c, c)]
    if kind == OLD_X:
        if i > 0:
            w += [L(r, k, i - 1) for k in range(nbt)] + [L(c, k, i - 1) for k in range(nbt)]
        w += [D(r, k) for k in range(ks)] + [D(c, k) for k in range(ks)]
        return w, [("X", i, r, c)]
    if kind == OLD_LW:
        w += [L(r, k) for k in range(ks)] + [D(c, k) for k in range(ks)]
        return w, [("LW", i, r, c)]
    if kind == OLD_A:
        if i > 0:
            w += [A(k, i - 1) for k in range(nbt)] + [L(c, k, i - 1) for k in range(nbt)]
        w += [A(k) for k in range(ks)] + [D(c, k) for k in range(ks)]
        return w, [("XA", i, c)]
    if kind == DIAG:
        w += [("X", i, r, c)] + [D(r, k) for k in range(ks, c)] + [D(c, k) for k in range(ks, c)]
        return w + [D(c, c)], [D(r, c)]
    if kind == LOWER:
        k0 = ks if c > slook else 0
        if c > slook:
            w.append(("LW", i, r, c))
        w += [L(r, k) for k in range(k0, c)] + [D(c, k) for k in range(k0, c)]
        return w + [D(c, c)], [L(r, c)]
    if kind == ARROW:
        w += [("XA", i, c)] + [A(k) for k in range(ks, c)] + [D(c, k) for k in range(ks, c)]
        return w + [D(c, c)], [A(c)]
    raise AssertionError(f"unknown task kind {kind}")


@pytest.mark.parametrize("arrow", [0, 1])
@pytest.mark.parametrize("nbt", [1, 2, 3, 4, 5, 8, 11, 13, 17])
def test_every_wait_is_on_an_earlier_task(nbt, arrow):
    lib = _lib.load()
    full, look, slook = block_table(lib, nbt, False, arrow)
    last, _, _ = block_table(lib, nbt, True, arrow)
    for n in (1, 2, 3, 5):
        raised = set()
        tasks = 0
        for i in range(n):
            for kind, r, c in (full if i < n - 1 else last):
                waits, sets = waits_and_raises(i, kind, r, c, nbt, arrow, look, slook)
                missing = [f for f in waits if f not in raised]
                assert not missing, (n, i, kind, r, c, missing[:3])
                assert not (set(sets) & raised), (n, i, kind, r, c)  # each flag raised once
                raised.update(sets)
                tasks += 1
        # every tile of the factor is final at the end
        for i in range(n):
            for rr in range(nbt):
                for cc in range(rr + 1):
                    assert ("D", i, rr, cc) in raised
                if i < n - 1:
                    assert all(("L", i, rr, cc) in raised for cc in range(nbt))
            if arrow:
                assert all(("A", i, cc) in raised for cc in range(nbt)) and ("T", i) in raised


def test_table_rejects_bad_arguments():
    lib = _lib.load()
    assert lib.dh_dag_table(0, 0, 0, None, 0, None, None) == -1
    assert lib.dh_dag_table(4, 0, 0, None, 5, None, None) == -1
