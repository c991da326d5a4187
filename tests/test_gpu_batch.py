"""Batched Jacobi-PCG: many uneven pressure systems in one persistent launch,
each converging to its own tolerance, in FP64 and mixed precision; each
system's solution matches the single-system solve and the oracle."""
import numpy as np
import pytest

from fortranfvcell_amd import problems as P

pytestmark = pytest.mark.gpu


def _systems(count, lo, hi, seed):
    sides = P.config_batch_sides(count, lo, hi, seed)
    out = []
    for i, n in enumerate(sides):
        sd = seed * 1000 + i
        mesh = P.quad_mesh(int(n))
        out.append(P.pressure_system(mesh, k=P.lognormal_k(sd, mesh.neq), seed=sd))
    return out


@pytest.mark.parametrize("precision", [0, 1])
def test_batch_matches_single_solves(fvc, orc, precision):
    systems = _systems(24, 8, 48, 7)
    fmts = [P.to_ell(s) for s in systems]
    batch = fvc.concat_ell(fmts)
    print(f"batch of {len(fmts)} systems, {batch['row_offsets'][-1]} rows, precision={precision}", flush=True)
    x, sts = fvc.solve_batch(batch, tol=1e-10, precision=precision, stop_rule=fvc.STOP_RELRES)
    ro = batch["row_offsets"]
    for i, (s, f) in enumerate(zip(systems, fmts)):
        assert sts[i]["status"] == 0, (i, sts[i])
        xi = x[ro[i]:ro[i + 1]]
        r = np.linalg.norm(s.b - P.scipy_csr(s) @ xi) / np.linalg.norm(s.b)
        assert r <= 1e-10 * 1.01, (i, r)
        xs, ss = fvc.solve(f, fvc.PCG, tol=1e-10, precision=precision, stop_rule=fvc.STOP_RELRES,
                           x0_rule=fvc.X0_ZERO)
        if precision == 0:
            # the same phases on the same rows with a smaller block group:
            # same algorithm, different reduction tree
            assert np.linalg.norm(xi - xs) <= 1e-8 * np.linalg.norm(xs)
        xo, _, so = orc.krylov("pcg", f, tol=1e-12)
        assert np.linalg.norm(xi - xo) <= 1e-6 * np.linalg.norm(xo)


def _renumbered(s, seed):
    """The ELLPACK arrays of s with its cells numbered in random order, and the
    order (new row r is old row p[r])."""
    n, w = s.neq, s.w
    p = np.random.default_rng(seed).permutation(n)
    q = np.empty(n, np.int64)
    q[p] = np.arange(n)
    nb = s.nbr[p]
    inner = nb >= 0
    ja = np.where(inner, q[np.where(inner, nb, 0)] + 1, 1).astype(np.int32)
    al = np.where(inner, s.off[p], 0.0)
    f = dict(matriz=4, neq=n, nad=int(inner.sum()), iax=np.array([w, n + 1], np.int32),
             ja=np.ascontiguousarray(ja.ravel()), ad=s.ad[p].copy(),
             al=np.ascontiguousarray(al.ravel()), au=None, b=s.b[p].copy(), w=w)
    return f, p


def test_batch_mixed_column_offsets_int16_and_int32(fvc):
    """Mixed batches read the FP32 inner iterations' columns as int16 offsets
    j - i when every slot of every system fits, else the int32 columns.  A
    batch in mesh numbering takes the first path; the same batch with a
    40,000-cell system renumbered at random (offsets above 32,767) the second."""
    small = _systems(6, 16, 48, 3)
    big = P.config_pressure_lognormal(200)
    renum, p = _renumbered(big, 11)
    for with_big in (False, True):
        systems = small + ([big] if with_big else [])
        fmts = [P.to_ell(s) for s in small] + ([renum] if with_big else [])
        batch = fvc.concat_ell(fmts)
        print(f"mixed batch of {len(fmts)} systems, renumbered member: {with_big}", flush=True)
        x, sts = fvc.solve_batch(batch, tol=1e-10, precision=1, stop_rule=fvc.STOP_RELRES)
        ro = batch["row_offsets"]
        for i, s in enumerate(systems):
            assert sts[i]["status"] == 0, (i, sts[i])
            # one system whose offsets do not fit sends the whole batch to int32
            assert bool(sts[i]["flags"] & fvc.FLAG_COL16) == (not with_big), (i, sts[i])
            xi = x[ro[i]:ro[i + 1]]
            if s is big:
                xo = np.empty(s.neq)
                xo[p] = xi
                xi = xo
            r = np.linalg.norm(s.b - P.scipy_csr(s) @ xi) / np.linalg.norm(s.b)
            assert r <= 1e-10 * 1.01, (i, r)


def test_batch_uneven_sizes_and_waves(fvc):
    # more systems than half the co-resident blocks: solved in waves
    systems = _systems(300, 4, 12, 11)
    batch = fvc.concat_ell([P.to_ell(s) for s in systems])
    x, sts = fvc.solve_batch(batch, tol=1e-10, stop_rule=fvc.STOP_RELRES)
    assert all(st["status"] == 0 for st in sts)
    ro = batch["row_offsets"]
    for i in (0, 150, 299):
        s = systems[i]
        xi = x[ro[i]:ro[i + 1]]
        assert np.linalg.norm(s.b - P.scipy_csr(s) @ xi) <= 1.01e-10 * np.linalg.norm(s.b)


@pytest.mark.parametrize("barrier", [0, 1], ids=["flags", "counters"])
@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_batch_shared_blocks_match_solo(fvc, precision, barrier):
    """Systems too small to keep a block busy share blocks, solved one after
    another on each.  A system of at most 64 rows runs as a group of one
    block whether it shares or is alone, so its solution and iteration count
    must equal those of a batch of its own, bit for bit.  A barrier epoch
    carried over wrongly from the previous system on the block, or state
    left behind, would show here."""
    tiny = _systems(120, 4, 8, 21)                 # 16 to 64 rows
    big = P.config_pressure_lognormal(160)         # 25,600 rows
    systems = tiny + [big]
    fmts = [P.to_ell(s) for s in systems]
    batch = fvc.concat_ell(fmts)
    print(f"shared blocks: {len(fmts)} systems, precision={precision} barrier={barrier}", flush=True)
    kw = dict(tol=1e-10, stop_rule=fvc.STOP_RELRES, precision=precision, barrier=barrier)
    x, sts = fvc.solve_batch(batch, **kw)
    ro = batch["row_offsets"]
    assert all(st["status"] == 0 for st in sts)
    r = np.linalg.norm(big.b - P.scipy_csr(big) @ x[ro[-2]:ro[-1]]) / np.linalg.norm(big.b)
    assert r <= 1.01e-10
    for i in range(0, len(tiny), 7):
        xs, ss = fvc.solve_batch(fvc.concat_ell([fmts[i]]), **kw)
        assert ss[0]["iterations"] == sts[i]["iterations"], i
        assert np.array_equal(xs, x[ro[i]:ro[i + 1]]), i


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_batch_launch_systems(fvc, precision):
    """batch_launch_systems caps the systems of one launch: the batch then
    runs in several launches one after another, each on all blocks.  Every
    system still converges; a system of at most 64 rows runs as a group of
    one block in any launch, so its solution is the same bits; a negative
    cap is refused."""
    systems = _systems(40, 6, 40, 17)
    fmts = [P.to_ell(s) for s in systems]
    batch = fvc.concat_ell(fmts)
    ro = batch["row_offsets"]
    kw = dict(tol=1e-10, stop_rule=fvc.STOP_RELRES, precision=precision)
    x0, s0 = fvc.solve_batch(batch, **kw)
    for per in (1, 7, 40):
        print(f"batch_launch_systems={per} precision={precision}", flush=True)
        x, sts = fvc.solve_batch(batch, batch_launch_systems=per, **kw)
        for i, s in enumerate(systems):
            assert sts[i]["status"] == 0, (per, i)
            xi = x[ro[i]:ro[i + 1]]
            assert np.linalg.norm(s.b - P.scipy_csr(s) @ xi) <= 1.01e-10 * np.linalg.norm(s.b), (per, i)
            if s.neq <= 64:
                assert np.array_equal(xi, x0[ro[i]:ro[i + 1]]), (per, i)
    with pytest.raises(fvc.FvcError):
        fvc.solve_batch(batch, batch_launch_systems=-1, **kw)


def test_batch_device_resident(fvc):
    systems = _systems(16, 16, 64, 5)
    batch = fvc.concat_ell([P.to_ell(s) for s in systems])
    db = fvc.DeviceBatch(batch)
    sts = db.solve(tol=1e-10, stop_rule=fvc.STOP_RELRES)
    xh, sth = fvc.solve_batch(batch, tol=1e-10, stop_rule=fvc.STOP_RELRES)
    assert np.array_equal(db.x.get(), xh)
    assert [s["iterations"] for s in sts] == [s["iterations"] for s in sth]


@pytest.fixture(scope="module")
def full_batch():
    """The batching config: 256 pressure systems, square quad meshes of side
    log-uniform in [32, 512], lognormal(0.5) conductivity, seeded zero-mean
    right-hand sides (10.2M rows in all)."""
    systems = P.config_batch()
    from fortranfvcell_amd import concat_ell
    return systems, concat_ell([P.to_ell(s) for s in systems])


@pytest.mark.parametrize("barrier", [0, 1], ids=["flags", "counters"])
@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_full_batch_config_every_system_converges(fvc, full_batch, precision, barrier):
    # Regression: a phase that reduces nothing (the mixed solve's x += S y)
    # must still order every wave's writes before the block's arrival, or a
    # neighbouring block's next SpMV reads stale x.
    systems, batch = full_batch
    print(f"full batch precision={precision} barrier={barrier} rows={batch['row_offsets'][-1]}", flush=True)
    db = fvc.DeviceBatch(batch)
    sts = db.solve(tol=1e-10, precision=precision, stop_rule=fvc.STOP_RELRES, barrier=barrier)
    print(f"  {sts[0]['seconds'] * 1e3:.1f} ms, {len(sts) / sts[0]['seconds']:.1f} solves/s, "
          f"max iterations {max(s['iterations'] for s in sts)}", flush=True)
    x = db.x.get()
    ro = batch["row_offsets"]
    worst = 0.0
    for i, s in enumerate(systems):
        assert sts[i]["status"] == 0, (i, sts[i])
        xi = x[ro[i]:ro[i + 1]]
        worst = max(worst, np.linalg.norm(s.b - P.scipy_csr(s) @ xi) / np.linalg.norm(s.b))
    print(f"  worst true relative residual {worst:.2e}", flush=True)
    assert worst <= 1.01e-10


@pytest.mark.parametrize("precision", [0, 1])
def test_batch_csrd_matches_ellpack(fvc, precision):
    # the same uneven batch in CSRD (per-system row pointers, local columns)
    # and ELLPACK: every system to true 1e-10, the two solutions within the
    # bar (the two layouts sum a row in different orders: the reference's
    # CSRD from the diagonal, its ELLPACK to it)
    systems = _systems(20, 8, 64, 11)
    ell = fvc.concat_ell([P.to_ell(s) for s in systems])
    csr = fvc.concat_csrd([P.to_csrd(s) for s in systems])
    print(f"CSRD batch of {len(systems)} systems, {csr['row_offsets'][-1]} rows, {csr['val_offsets'][-1]} values, "
          f"precision={precision}", flush=True)
    xe, se = fvc.solve_batch(ell, tol=1e-10, precision=precision, stop_rule=fvc.STOP_RELRES)
    xc, sc = fvc.solve_batch(csr, tol=1e-10, precision=precision, stop_rule=fvc.STOP_RELRES)
    ro = csr["row_offsets"]
    for i, s in enumerate(systems):
        assert sc[i]["status"] == 0 and se[i]["status"] == 0, (i, sc[i], se[i])
        xi = xc[ro[i]:ro[i + 1]]
        r = np.linalg.norm(s.b - P.scipy_csr(s) @ xi) / np.linalg.norm(s.b)
        assert r <= 1e-10 * 1.01, (i, r)
        ye = xe[ro[i]:ro[i + 1]]
        assert np.linalg.norm(xi - ye) <= 1e-8 * np.linalg.norm(ye), i
    # device-resident CSRD batch
    db = fvc.DeviceBatch(csr)
    sts = db.solve(tol=1e-10, precision=precision, stop_rule=fvc.STOP_RELRES)
    assert all(q["status"] == 0 for q in sts)
    assert np.linalg.norm(db.x.get() - xc) <= 1e-8 * np.linalg.norm(xc)


@pytest.mark.parametrize("rpw", [16, 256])
def test_batch_rows_per_wave(fvc, rpw):
    # fewer blocks for the small systems (at most neq / (rows x waves per
    # block) each): every system still converges to true 1e-10
    systems = _systems(30, 8, 96, 13)
    batch = fvc.concat_ell([P.to_ell(s) for s in systems])
    x, sts = fvc.solve_batch(batch, tol=1e-10, stop_rule=fvc.STOP_RELRES, batch_rows_per_wave=rpw, block_size=256)
    ro = batch["row_offsets"]
    for i, s in enumerate(systems):
        assert sts[i]["status"] == 0
        xi = x[ro[i]:ro[i + 1]]
        assert np.linalg.norm(s.b - P.scipy_csr(s) @ xi) <= 1.01e-10 * np.linalg.norm(s.b)
    with pytest.raises(fvc.FvcError):
        fvc.solve_batch(batch, tol=1e-10, batch_rows_per_wave=-1)


def test_batch_csrd_refusals(fvc):
    systems = _systems(3, 8, 16, 17)
    csr = fvc.concat_csrd([P.to_csrd(s) for s in systems])
    bad = dict(csr, ia=csr["ia"].copy())
    bad["ia"][int(csr["row_offsets"][1]) + 1 + 2] = 10_000     # a row pointer past the system's values
    with pytest.raises(fvc.FvcError):
        fvc.solve_batch(bad, tol=1e-10)
    bad = dict(csr, val_offsets=csr["val_offsets"][::-1].copy())
    with pytest.raises(fvc.FvcError):
        fvc.solve_batch(bad, tol=1e-10)
