"""GPU checks of the batching form: degree buckets launched with per-bucket
shapes from a tile table (register budget, persistent grid), the
autotuner, and per-half-sweep graph replay.  Each configuration must give
the factors of the single-launch path (same arithmetic: rows are solved
by the same code whatever launch carries them) to 1e-6 relative, and the
bucketed, persistent, graph-replayed half-sweeps are checked against the
FP64 oracle solving from the same factors."""
import numpy as np
import pytest

import cumf
from cumf import als as G
from oracle import oracle as O
from tests import _data
from tests._checks import bf16_round, degree_sample, rel_rows, row_errors

pytestmark = pytest.mark.gpu


def _run(ds, f, lam, T0, iters=2, table=None, tune=False, **opt):
    h = G.ALS(f, lam, **{"timing": 1, **opt})
    h.set_dataset(ds)
    h.set_factors(theta=T0)
    if table is not None:
        h.load_tile_table(table)
    if tune:
        h.autotune(reps=1, path=tune)
    h.iterate(iters)
    X, T = h.get_factors()
    st = h.stats()
    h.close()
    return X, T, st


def _rel(a, b):
    return float(np.max(np.linalg.norm(a - b, axis=1) / np.maximum(np.linalg.norm(b, axis=1), 1e-30)))


def _table(tmp_path, split, variant_of, persistent_of, occupancy_of=None):
    lines = ["# test table"]
    for side in (0, 1):
        lines.append(f"side {side} split {split} finish {1 + side} fallback {2 - side}")
        for b in range(8):
            line = f"bucket {side} {b} 0 0 variant {variant_of(side, b)} persistent {persistent_of(side, b)}"
            if occupancy_of is not None:
                line += f" solver -1 occupancy {occupancy_of(side, b)}"
            lines.append(line)
    p = tmp_path / "tiles.txt"
    p.write_text("\n".join(lines) + "\n")
    return p


@pytest.mark.parametrize("solver,precision", [(G.CUMF_SOLVER_CHOLESKY, G.CUMF_PREC_FP32),
                                              (G.CUMF_SOLVER_CG, G.CUMF_PREC_BF16_GRAM)])
def test_bucket_shapes_match_single_launch(tmp_path, solver, precision):
    ds = _data.netflix_like_small()
    f, lam = 100, 0.058
    T0 = cumf.init_factors(ds.n, f, 41)
    Xr, Tr, _ = _run(ds, f, lam, T0, split=256, solver=solver, precision=precision)
    tab = _table(tmp_path, 256, lambda s, b: 1 + (b + s) % 2, lambda s, b: 0 if b % 3 else 96)
    Xb, Tb, st = _run(ds, f, lam, T0, table=tab, split=2048, solver=solver, precision=precision)
    assert st["failed_rows"] == 0
    assert _rel(Xb, Xr) <= 1e-6 and _rel(Tb, Tr) <= 1e-6, (_rel(Xb, Xr), _rel(Tb, Tr))
    # one launch per non-empty bucket and half-sweep
    assert st["task_launches"] > 4


def test_occupancy_caps_match_single_launch(tmp_path):
    """Buckets launched one task per group with at most 3, 6 or 7 wavefronts
    per CU (the LDS each group requests) give the factors of one uncapped
    launch."""
    ds = _data.netflix_like_small()
    f, lam = 100, 0.058
    T0 = cumf.init_factors(ds.n, f, 45)
    Xr, Tr, _ = _run(ds, f, lam, T0, split=512, solver=G.CUMF_SOLVER_CG, precision=G.CUMF_PREC_BF16_GRAM)
    tab = _table(tmp_path, 512, lambda s, b: 1 if b % 4 == 3 else 2, lambda s, b: 0,
                 lambda s, b: 3 if b % 4 == 3 else (0, 6, 7)[b % 3])
    Xc, Tc, st = _run(ds, f, lam, T0, table=tab, solver=G.CUMF_SOLVER_CG, precision=G.CUMF_PREC_BF16_GRAM)
    assert st["failed_rows"] == 0
    assert _rel(Xc, Xr) <= 1e-6 and _rel(Tc, Tr) <= 1e-6, (_rel(Xc, Xr), _rel(Tc, Tr))


def _dual_table(tmp_path, split, dual_variant=2):
    lines = []
    for side in (0, 1):
        lines.append(f"side {side} split {split} finish 2 fallback 2")
        for b in range(8):
            sol = 2 if b < 3 else -1
            lines.append(f"bucket {side} {b} 0 0 variant {dual_variant if b < 3 else 2} persistent 0 solver {sol}")
    p = tmp_path / "dual.txt"
    p.write_text("\n".join(lines) + "\n")
    return p


@pytest.mark.parametrize("f", [10, 32, 100, 128, 140])
def test_dual_form_matches_primal_and_oracle(tmp_path, f):
    """Rows of at most 95 ratings (degree buckets 0-2: two staged chunks up
    to 63 ratings, three up to 95) solved in the dual form,
    (Y Y^T + mu I) z = r and x = Y^T z, give the primal solution: within
    1e-4 of the BF16 primal path, residual <= 1e-5 against the FP64 system
    on the BF16 factors.  With split 48 every chunk of a longer row also
    lands in buckets 0-1: the dual kernel forms those partial Grams, and
    those rows match the primal path bit for bit."""
    ds = _data.netflix_like_small()
    lam = 0.058
    T0 = cumf.init_factors(ds.n, f, 46)
    rp = ds.csr_row_ptr
    deg = np.diff(rp)
    for split in (16384, 48):
        opt = dict(precision=G.CUMF_PREC_BF16_GRAM, solver=G.CUMF_SOLVER_CHOLESKY, split=split)
        Xp, Tp, _ = _run(ds, f, lam, T0, iters=1, **opt)
        Xd, Td, st = _run(ds, f, lam, T0, iters=1, table=_dual_table(tmp_path, split, 1 if f == 32 else 2), **opt)
        assert st["failed_rows"] == 0
        light = deg <= 95
        assert light.sum() > 100
        assert _rel(Xd[light], Xp[light]) <= 1e-4, _rel(Xd[light], Xp[light])
        if split == 48:
            heavy = ~light
            np.testing.assert_array_equal(Xd[heavy], Xp[heavy])
    # one X half-sweep from T0 checked against the FP64 oracle on light rows
    with G.ALS(f, lam, precision=G.CUMF_PREC_BF16_GRAM, solver=G.CUMF_SOLVER_CHOLESKY) as h:
        h.set_dataset(ds)
        h.load_tile_table(str(_dual_table(tmp_path, 16384)))
        h.set_factors(theta=T0)
        h.update_x()
        X1, _ = h.get_factors()
    from tests._checks import bf16_round, row_errors
    rows = np.flatnonzero((deg > 0) & (deg <= 95))[::7][:300]
    res, err = row_errors(rp, ds.csr_col, ds.csr_val, bf16_round(T0), X1, f, lam, rows)
    assert res.max() <= 1e-5 and err.max() <= 1e-3, (res.max(), err.max())


@pytest.mark.parametrize("f", [32, 100])
def test_fp32_dual_form_matches_primal_and_oracle(tmp_path, f):
    """On an FP32 handle the dual form of buckets 0-2 (K = Y Y^T on FP32 MFMA
    from the factors, one wavefront per row) gives the primal solution:
    within 1e-4 of the FP32 primal path, residual <= 1e-5 against the FP64
    system.  With split 48 the launches that hold split rows' chunks take
    the Cholesky kernel: those rows match the primal path bit for bit."""
    ds = _data.netflix_like_small()
    lam = 0.058
    T0 = cumf.init_factors(ds.n, f, 47)
    rp = ds.csr_row_ptr
    deg = np.diff(rp)
    light = deg <= 95
    for split in (16384, 48):
        opt = dict(solver=G.CUMF_SOLVER_CHOLESKY, split=split)
        Xp, _, _ = _run(ds, f, lam, T0, iters=1, **opt)
        Xd, _, st = _run(ds, f, lam, T0, iters=1, table=_dual_table(tmp_path, split), **opt)
        assert st["failed_rows"] == 0
        assert _rel(Xd[light], Xp[light]) <= 1e-4, _rel(Xd[light], Xp[light])
        if split == 48:
            np.testing.assert_array_equal(Xd[~light], Xp[~light])
    with G.ALS(f, lam, solver=G.CUMF_SOLVER_CHOLESKY) as h:
        h.set_dataset(ds)
        h.load_tile_table(str(_dual_table(tmp_path, 16384)))
        h.set_factors(theta=T0)
        h.update_x()
        X1, _ = h.get_factors()
    from tests._checks import row_errors
    rows = np.flatnonzero((deg > 0) & light)[::7][:300]
    res, err = row_errors(rp, ds.csr_col, ds.csr_val, T0, X1, f, lam, rows)
    assert res.max() <= 1e-5 and err.max() <= 1e-3, (res.max(), err.max())


def test_all_persistent_buckets_and_graph_replay(tmp_path):
    ds = _data.netflix_like_small()
    f, lam = 32, 0.058
    T0 = cumf.init_factors(ds.n, f, 42)
    Xr, Tr, _ = _run(ds, f, lam, T0, iters=3, split=512, solver=G.CUMF_SOLVER_CG)
    tab = _table(tmp_path, 512, lambda s, b: 2 - b % 2, lambda s, b: 64 + 32 * b)
    Xe, Te, _ = _run(ds, f, lam, T0, iters=3, table=tab, solver=G.CUMF_SOLVER_CG)
    Xg, Tg, st = _run(ds, f, lam, T0, iters=3, table=tab, solver=G.CUMF_SOLVER_CG, use_graph=1)
    assert _rel(Xe, Xr) <= 1e-6 and _rel(Te, Tr) <= 1e-6
    np.testing.assert_array_equal(Xg, Xe)
    np.testing.assert_array_equal(Tg, Te)
    assert st["iters"] == 6  # one graph replay per half-sweep


def test_autotune_writes_table_and_keeps_factors(tmp_path):
    ds = _data.netflix_like_small()
    f, lam = 32, 0.058
    T0 = cumf.init_factors(ds.n, f, 43)
    path = tmp_path / "tuned.txt"
    h = G.ALS(f, lam, solver=G.CUMF_SOLVER_CG, precision=G.CUMF_PREC_BF16_GRAM)
    h.set_dataset(ds)
    h.set_factors(theta=T0)
    h.iterate(1)
    X1, T1 = h.get_factors()
    h.autotune(reps=1, path=str(path))
    X2, T2 = h.get_factors()
    np.testing.assert_array_equal(X1, X2)
    np.testing.assert_array_equal(T1, T2)
    h.iterate(1)
    Xt, Tt = h.get_factors()
    h.close()
    text = path.read_text().splitlines()
    assert sum(l.startswith("side ") for l in text) == 2
    assert sum(l.startswith("bucket ") for l in text) == 16
    # the tuned shapes give the factors of the single-launch path
    h = G.ALS(f, lam, solver=G.CUMF_SOLVER_CG, precision=G.CUMF_PREC_BF16_GRAM)
    h.set_dataset(ds)
    h.set_factors(X=X1, theta=T1)
    h.load_tile_table(str(path))
    h.save_tile_table(str(tmp_path / "again.txt"))
    assert (tmp_path / "again.txt").read_text() == path.read_text()
    h.close()
    splits = {int(l.split()[3]) for l in text if l.startswith("side ")}
    h = G.ALS(f, lam, solver=G.CUMF_SOLVER_CG, precision=G.CUMF_PREC_BF16_GRAM, split=min(splits))
    h.set_dataset(ds)
    h.set_factors(X=X1, theta=T1)
    h.iterate(1)
    Xs, Ts = h.get_factors()
    h.close()
    if len(splits) == 1:
        solvers = {int(l.split()[-1]) for l in text if l.startswith("bucket ")}
        if solvers <= {-1, G.CUMF_SOLVER_CG}:
            assert _rel(Xt, Xs) <= 1e-6 and _rel(Tt, Ts) <= 1e-6
        else:
            # buckets the tuner moved to the Cholesky solver: both solutions
            # meet the 1e-5 relative residual on the same systems
            assert _rel(Xt, Xs) <= 1e-2 and _rel(Tt, Ts) <= 1e-2, (_rel(Xt, Xs), _rel(Tt, Ts))


def test_bad_tile_table_rejected(tmp_path):
    p = tmp_path / "bad.txt"
    p.write_text("side 0 split 0 finish 2 fallback 2\n")
    with G.ALS(16, 0.05) as h:
        with pytest.raises(G.CumfError):
            h.load_tile_table(str(p))
        p.write_text("bucket 0 9 0 0 variant 2 persistent 0\n")
        with pytest.raises(G.CumfError):
            h.load_tile_table(str(p))
        for occ in (2, 17, -1):  # occupancy caps outside 3..16 (0 = none)
            p.write_text(f"bucket 0 1 0 0 variant 2 persistent 0 solver -1 occupancy {occ}\n")
            with pytest.raises(G.CumfError):
                h.load_tile_table(str(p))
        # the dual form: a bucket that may hold 96+ ratings, a persistent grid
        for line in ("bucket 0 3 0 0 variant 2 persistent 0 solver 2", "bucket 0 1 0 0 variant 2 persistent 64 solver 2"):
            p.write_text(line + "\n")
            with pytest.raises(G.CumfError):
                h.load_tile_table(str(p))
        p.write_text("bucket 0 1 0 0 variant 2 persistent 0 solver 2\n")  # FP32 factors: the FP32 dual form
        h.load_tile_table(str(p))
    with G.ALS(16, 0.05, precision=G.CUMF_PREC_BF16_GRAM) as h:
        for line in ("bucket 0 3 0 0 variant 2 persistent 0 solver 2", "bucket 1 0 0 0 variant 2 persistent 64 solver 2",
                     "bucket 0 0 0 0 variant 2 persistent 0 solver 3"):
            p.write_text(line + "\n")
            with pytest.raises(G.CumfError):
                h.load_tile_table(str(p))
        p.write_text("bucket 0 2 0 0 variant 2 persistent 0 solver 2\n")
        h.load_tile_table(str(p))


def test_uniform_table_is_one_launch(tmp_path):
    """A side whose buckets share one shape and no persistent grid runs as one
    launch of all its tasks, with the factors of the table-free path."""
    ds = _data.netflix_like_small()
    f, lam = 32, 0.058
    T0 = cumf.init_factors(ds.n, f, 44)
    Xr, Tr, sr = _run(ds, f, lam, T0, iters=2, split=512, solver=G.CUMF_SOLVER_CG)
    tab = _table(tmp_path, 512, lambda s, b: 2, lambda s, b: 0)
    Xu, Tu, su = _run(ds, f, lam, T0, iters=2, table=tab, solver=G.CUMF_SOLVER_CG)
    assert _rel(Xu, Xr) <= 1e-6 and _rel(Tu, Tr) <= 1e-6
    assert sr["task_launches"] == 4 and su["task_launches"] == 4  # one per half-sweep


@pytest.mark.parametrize("solver,precision", [(G.CUMF_SOLVER_CHOLESKY, G.CUMF_PREC_FP32),
                                              (G.CUMF_SOLVER_CHOLESKY, G.CUMF_PREC_BF16_GRAM),
                                              (G.CUMF_SOLVER_CG, G.CUMF_PREC_BF16_GRAM)])
def test_batched_graph_half_sweeps_match_oracle(tmp_path, solver, precision):
    """Per-bucket shapes (both register budgets, persistent and one-group-per-
    task buckets, split rows) replayed from captured graphs: the X and Theta
    half-sweeps of the third iteration against the FP64 oracle's half-sweep
    from the same factors (BF16-rounded where the Gram reads BF16)."""
    ds = _data.netflix_like_small()
    f, lam = 100, 0.058
    bf = precision == G.CUMF_PREC_BF16_GRAM
    T0 = cumf.init_factors(ds.n, f, 45)
    tab = _table(tmp_path, 384, lambda s, b: 1 + (b + s) % 2, lambda s, b: 128 if b % 2 else 0)
    with G.ALS(f, lam, solver=solver, precision=precision, use_graph=1, timing=1) as h:
        h.set_dataset(ds)
        h.load_tile_table(str(tab))
        h.set_factors(theta=T0)
        h.iterate(2)
        _, T2 = h.get_factors()
        h.update_x()
        X3, _ = h.get_factors()
        h.update_theta()
        _, T3 = h.get_factors()
        st = h.stats()
    assert st["failed_rows"] == 0 and st["iters"] == 6
    src_t = bf16_round(T2) if bf else T2
    src_x = bf16_round(X3) if bf else X3
    Xo, _ = O.half_sweep(ds.csr_row_ptr, ds.csr_col, ds.csr_val, src_t, lam)
    To, _ = O.half_sweep(ds.csc_col_ptr, ds.csc_row, ds.csc_val, src_x, lam)
    ex, et = rel_rows(X3, Xo), rel_rows(T3, To)
    print(f"factor error vs oracle: X median {np.median(ex):.2e} max {ex.max():.2e}; "
          f"Theta median {np.median(et):.2e} max {et.max():.2e}; CG fallback rows {st['cg_fallback_rows']}")
    bar = 1e-5 if solver == G.CUMF_SOLVER_CHOLESKY else 2e-5
    users = degree_sample(ds.csr_row_ptr, 200, 8)
    items = degree_sample(ds.csc_col_ptr, 80, 9)
    rx, _ = row_errors(ds.csr_row_ptr, ds.csr_col, ds.csr_val, src_t, X3, f, lam, users)
    rt, _ = row_errors(ds.csc_col_ptr, ds.csc_row, ds.csc_val, src_x, T3, f, lam, items)
    assert rx.max() <= bar and rt.max() <= bar, (rx.max(), rt.max())
    # the factor error is the residual times the system's conditioning
    ebar = 1e-3 if solver == G.CUMF_SOLVER_CHOLESKY else 1e-2
    assert ex.max() <= ebar and et.max() <= ebar, (ex.max(), et.max())
