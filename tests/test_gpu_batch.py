"""Batched steps (f3d_batch): scenarios of different sizes, class counts and
member counts stepped together by one persistent launch per step, replayed
from HIP graphs.  Every scenario's result in the batch is bit-identical to
the same scenario stepped alone with the same tile shape, graph replays match
direct launches step by step, and the autotuner leaves the state untouched."""
import numpy as np
import pytest

from fall3d_amd import lib as F
from fall3d_amd import scenario as S
from oracle.oracle import OracleRun

pytestmark = pytest.mark.gpu


def _mixture(n, seed=7):
    # sizes chosen so both interior and edge tiles occur, nz small to fast
    return [S.config_uneven(seed + k, nmin=24, nmax=130, zmin=3, zmax=12, cmin=2, cmax=5)
            for k in range(n)]


def _runs(scs, tile):
    runs = []
    for sc in scs:
        r = F.Run(sc, precision=F.FP32, tile=tile)
        r.meteo_update()
        runs.append(r)
    return runs


def _close(runs):
    for r in runs:
        r.close()


@pytest.mark.parametrize("tile", [(58, 18, 0, 2), (58, 10, 0, 1), (58, 26, 0, 2)],
                         ids=lambda t: f"{t[0]}x{t[1]}m{t[3]}")
def test_batch_equals_each_scenario_alone(tile):
    scs = _mixture(5)
    nsteps = 9
    mix = _runs(scs, tile)
    try:
        b = F.Batch([r.ctx for r in mix], tile=(tile[0], tile[1], tile[3]))
        b.step(nsteps)
        b.sync()
        got = [(r.state(), r.budget(), r.ctx.clock()) for r in mix]
        b.close()
    finally:
        _close(mix)
    for k in (0, 2, 4):
        alone = _runs([scs[k]], tile)[0]
        try:
            alone.step(nsteps)
            (c, cdry), bud = alone.state(), alone.budget()
            clk = alone.ctx.clock()
        finally:
            alone.close()
        (cm, cdrym), budm, clkm = got[k]
        assert np.array_equal(c, cm), k
        assert np.array_equal(cdry[:, :, 1:], cdrym[:, :, 1:]), k
        for key in ("massvol", "massdep", "outmass", "erumass"):
            assert bud[key] == budm[key], (k, key)
        assert clk["iiter"] == clkm["iiter"] and clk["time"] == clkm["time"]


def test_graph_replays_match_direct_launches():
    scs = _mixture(3, seed=31)
    a, b = _runs(scs, None), _runs(scs, None)
    try:
        ba = F.Batch([r.ctx for r in a], graph=True)
        bb = F.Batch([r.ctx for r in b], graph=False)
        for step in range(6):
            ba.step(1)
            bb.step(1)
            ba.sync()
            bb.sync()
            for k in range(len(scs)):
                ca, _ = a[k].state()
                cb, _ = b[k].state()
                assert np.array_equal(ca, cb), (step, k)
        ba.close()
        bb.close()
    finally:
        _close(a)
        _close(b)


def test_batch_scenario_vs_oracle():
    scs = _mixture(3, seed=55)
    runs = _runs(scs, None)
    try:
        b = F.Batch([r.ctx for r in runs])
        b.step(10)
        b.sync()
        c, cdry = runs[1].state()
        b.close()
    finally:
        _close(runs)
    o = OracleRun(scs[1])
    o.meteo_update()
    for _ in range(10):
        o.step()
    ci, oi = c[1:-1, 1:-1, 1:-1], o.c[1:-1, 1:-1, 1:-1]
    e_c = float(np.abs(ci - oi).sum() / np.abs(oi).sum())
    load = cdry[:, :, 1:].sum(-1)
    ref = o.cdry[:, :, 1:].sum(-1)
    e_load = float(np.abs(load - ref).sum() / np.abs(ref).sum())
    assert e_c <= 1e-4, e_c
    assert e_load <= 1e-4, e_load


def test_ensemble_context_in_a_batch():
    base = S.config_small(nx=40, ny=33, nz=7, nc=3, seed=12, mean_u=4.0, mean_v=2.0)
    mem = [S.config_ensemble_member(base, 300 + k) for k in range(3)]
    other = _mixture(2, seed=77)
    ens_a, ens_b = F.Ensemble(base, mem), F.Ensemble(base, mem)
    runs = _runs(other, None)
    try:
        ens_a.meteo_update()
        ens_b.meteo_update()
        b = F.Batch([ens_a.ctx] + [r.ctx for r in runs])
        b.step(7)
        b.sync()
        ens_b.ctx.set_tile(*b.tile()[:2], 0, b.tile()[2])
        ens_b.step(7)
        for m in range(len(mem)):
            assert np.array_equal(ens_a.state(m)[0], ens_b.state(m)[0]), m
        b.close()
    finally:
        ens_a.close()
        ens_b.close()
        _close(runs)


def test_autotune_keeps_state_and_picks_a_candidate():
    scs = _mixture(3, seed=91)
    runs = _runs(scs, None)
    try:
        before = [r.state()[0] for r in runs]
        b = F.Batch([r.ctx for r in runs])
        (tx, ty, sl), ms = b.autotune(trials=2)
        assert tx == 58 and ty in (10, 18, 26) and sl in (1, 2) and ms > 0
        after = [r.state()[0] for r in runs]
        for x, y in zip(before, after):
            assert np.array_equal(x, y)
        b.step(4)
        b.sync()
        got = runs[0].state()[0]
        b.close()
        (t1, t2, t3), ms1 = runs[2].ctx.autotune(trials=2)
        assert ms1 > 0
    finally:
        _close(runs)
    alone = _runs([scs[0]], (tx, ty, 0, sl))[0]
    try:
        alone.step(4)
        assert np.array_equal(alone.state()[0], got)
    finally:
        alone.close()


def test_batch_mixed_parities_equal_alone():
    # contexts that joined at different times sit on different iiter
    # parities; each keeps its own sweep order in the batch
    scs = _mixture(3, seed=5)
    runs = _runs(scs, None)
    try:
        runs[0].step(1)
        runs[2].step(3)
        b = F.Batch([r.ctx for r in runs])
        b.step(5)
        b.sync()
        got = [r.state()[0] for r in runs]
        b.close()
    finally:
        _close(runs)
    for k, n in ((0, 6), (1, 5), (2, 8)):
        alone = _runs([scs[k]], None)[0]
        try:
            alone.step(n)
            assert np.array_equal(alone.state()[0], got[k]), k
        finally:
            alone.close()


def test_batch_refuses_a_context_that_turns_wet():
    # precipitation set on a context after it joined the batch: the next
    # batched step is refused (the batched kernel has no wet deposition)
    scs = _mixture(2, seed=41)
    runs = _runs(scs, None)
    try:
        b = F.Batch([r.ctx for r in runs])
        b.step(1)
        runs[1].ctx.set_precipitation(np.full((scs[1].nx, scs[1].ny), 2.0, order="F"))
        with pytest.raises(F.F3DError, match="wet deposition"):
            b.step(1)
        b.close()
    finally:
        _close(runs)


def test_context_calls_are_ordered_after_batched_steps():
    # budgets, state reads and f3d_step issued right after f3d_batch_step,
    # with no f3d_batch_sync, act on the state the batched steps left
    scs = _mixture(3, seed=63)
    a = _runs(scs, None)
    try:
        b = F.Batch([r.ctx for r in a])
        b.step(3)
        bud = [r.budget() for r in a]          # no batch sync
        a[1].step(1)                            # f3d_step on a member, no sync
        b.step(2)
        st = [r.state()[0] for r in a]
        bud2 = [r.budget() for r in a]
        b.close()
    finally:
        _close(a)
    for k, n in ((0, 5), (1, 6), (2, 5)):
        alone = _runs([scs[k]], None)[0]
        try:
            alone.step(3)
            b3 = alone.budget()
            alone.step(n - 3)
            assert np.array_equal(alone.state()[0], st[k]), k
            b5 = alone.budget()
        finally:
            alone.close()
        for key in ("massvol", "massdep", "outmass", "erumass"):
            assert b3[key] == bud[k][key], (k, key)
            assert b5[key] == bud2[k][key], (k, key)


def test_batches_sharing_a_context_stay_ordered():
    # back-to-back f3d_batch_step calls with no call on the contexts in
    # between skip re-ordering a batch after its members' streams; a context
    # that another batch stepped in the meantime is still waited for
    scs = _mixture(3, seed=71)
    a = _runs(scs, None)
    try:
        b1 = F.Batch([a[0].ctx, a[1].ctx])
        b2 = F.Batch([a[1].ctx, a[2].ctx])
        b1.step(2)
        b1.step(1)
        b2.step(3)
        b1.step(1)
        b1.step(2)
        b1.sync()
        b2.sync()
        st = [r.state() for r in a]
        b1.close()
        b2.close()
    finally:
        _close(a)
    for k, n in ((0, 6), (1, 9), (2, 3)):
        alone = _runs([scs[k]], None)[0]
        try:
            alone.step(n)
            c, cdry = alone.state()
        finally:
            alone.close()
        assert np.array_equal(c, st[k][0]), k
        assert np.array_equal(cdry[:, :, 1:], st[k][1][:, :, 1:]), k
