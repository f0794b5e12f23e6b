"""Reconstruction server (lc_server_*): continuous batching with HBM-budget
admission.  Every result must be bitwise the result of solving that request
alone through the single-problem entry points, whatever else ran with it."""
import threading
import time

import numpy as np
import pytest

from conftest import announce

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lc():
    import laminocam_amd
    return laminocam_amd


@pytest.fixture(scope="module")
def problems(lc):
    rng = np.random.default_rng(17)
    g64 = lc.PolarGrid2D(np.arange(90) * np.pi / 90, 64)
    g96 = lc.PolarGrid2D(np.arange(60) * np.pi / 60, 96)
    p64 = lc.PointSpreadFunction2D(g64, 64)
    p96 = lc.PointSpreadFunction2D(g96, 96)
    th = np.arange(13) * 2 * np.pi / 13
    g3 = lc.PolarGrid(th, np.full(13, -np.pi / 6), np.zeros(13), 10, 32)
    p3 = lc.PointSpreadFunction(g3, (6, 32, 32))
    reqs = []
    for i in range(10):
        if i % 4 == 3:
            y = p3.convolve(rng.random((6, 32, 32)).astype(np.float32))
            reqs.append(dict(psf=p3, rhs=y, solver="cg", max_iter=8))
        elif i % 4 == 2:
            y = p96.convolve(rng.random((2, 96, 96)).astype(np.float32))
            reqs.append(dict(psf=p96, rhs=y, solver="sb", lam=0.05 + 0.01 * i, mu=10.0, outer=4,
                             inner=1))
        else:
            g, p, N = (g64, p64, 64) if i % 2 == 0 else (g96, p96, 96)
            sino = lc.forward2d(rng.random((1 + i % 3, N, N)), g)
            reqs.append(dict(psf=p, grid=g, sino=sino, solver="cg", max_iter=10))
    return reqs


def solve_alone(lc, r):
    p = r["psf"]
    if "sino" in r:
        N = p.image_size
        rhs = lc.backproj2d(r["sino"], r["grid"], (r["sino"].shape[0], N, N))
    else:
        rhs = r["rhs"]
    x0 = r["x0"] if "x0" in r else np.zeros(rhs.shape, np.float32)
    if r["solver"] == "cg":
        return lc.cgsolver(p, rhs, x0, r["max_iter"], 1e-5, 1e-5)[0]
    return lc.split_bregman(p, rhs, x0, r["lam"], r["mu"], r["outer"], r["inner"], 1e-5, 1e-5)[0]


def submit(srv, r):
    kw = {k: v for k, v in r.items() if k not in ("psf", "grid")}
    return srv.submit(r["psf"], **kw)


def test_server_same_alone_and_in_mix(lc, problems):
    announce("server mix vs alone")
    ref = [solve_alone(lc, r) for r in problems]
    with lc.Server(hbm_budget=8 << 30, lanes=3) as srv:
        tickets = []
        for i, r in enumerate(problems):
            tickets.append(submit(srv, r))
            if i % 3 == 2:
                time.sleep(0.002)   # later requests arrive while earlier ones run
        out = [srv.result(t, timeout=120.0) for t in tickets]
        st = srv.stats()
    for i, ((x, rep, times), xr) in enumerate(zip(out, ref)):
        assert x.shape == xr.shape, i
        assert np.array_equal(x, xr), f"request {i} differs from its solo solve"
        assert times[0] <= times[1] <= times[2]
    assert st["finished"] == len(problems) and st["inflight_bytes"] == 0


def test_server_budget_admission(lc, problems):
    # a budget that holds only one or two of these requests at a time still
    # completes every request, over several batches, with the same results
    announce("server budget admission")
    reqs = [r for r in problems if r["solver"] == "cg"][:4]
    ref = [solve_alone(lc, r) for r in reqs]
    with lc.Server(hbm_budget=5 << 20, lanes=2) as srv:
        tickets = [submit(srv, r) for r in reqs]
        out = [srv.result(t, timeout=120.0)[0] for t in tickets]
        st = srv.stats()
    for x, xr in zip(out, ref):
        assert np.array_equal(x, xr)
    assert st["batches"] >= 2


def test_server_rejects_request_above_budget(lc, problems):
    announce("server oversized request")
    with lc.Server(hbm_budget=1 << 20, lanes=1) as srv:
        with pytest.raises(lc.LcError):
            submit(srv, problems[0])


def test_server_threads_submit_concurrently(lc, problems):
    announce("server concurrent submitters")
    ref = [solve_alone(lc, r) for r in problems[:6]]
    res = [None] * 6
    with lc.Server(hbm_budget=8 << 30, lanes=2) as srv:
        def worker(i):
            t = submit(srv, problems[i])
            res[i] = srv.result(t, timeout=120.0)[0]
        ths = [threading.Thread(target=worker, args=(i,)) for i in range(6)]
        for t in ths:
            t.start()
        for t in ths:
            t.join(timeout=180.0)
    for x, xr in zip(res, ref):
        assert x is not None and np.array_equal(x, xr)


def test_server_borrowed_inputs_and_copy_out(lc, problems):
    # three ways in and out, one result: inputs copied at submit (the caller's
    # arrays are overwritten right after), inputs borrowed (read in place when
    # the request runs), and the C-level copy-out path (x_out NULL: the server
    # keeps the solution until lc_server_result copies it); with initial guesses
    announce("server borrowed inputs and copy-out")
    import ctypes as C
    from laminocam_amd import _lib
    rng = np.random.default_rng(23)
    reqs = []
    for r in problems[:4]:
        r = dict(r)
        p = r["psf"]
        shape = tuple(p.recon_dims) if isinstance(p, lc.PointSpreadFunction) else \
            ((r["sino"].shape[0] if "sino" in r else r["rhs"].shape[0]), p.image_size, p.image_size)
        r["x0"] = (0.1 * rng.random(shape)).astype(np.float32)
        reqs.append(r)
    ref = [solve_alone(lc, r) for r in reqs]
    with lc.Server(hbm_budget=8 << 30, lanes=2) as srv:
        copied = []
        for r in reqs:
            rc = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in r.items()}
            copied.append(submit(srv, rc))
            for v in rc.values():
                if isinstance(v, np.ndarray):
                    v[...] = 7.0    # the server holds its own copy
        borrowed = [srv.submit(r["psf"], borrow=True,
                               **{k: v for k, v in r.items() if k not in ("psf", "grid")})
                    for r in reqs]
        # C level, no x_out: a 2-D CG request from its right-hand side
        r0 = reqs[0]
        N = r0["psf"].image_size
        rhs0 = lc.backproj2d(r0["sino"], r0["grid"], (r0["sino"].shape[0], N, N))
        q = _lib.LcRequest()
        q.g2, q.nslices, q.solver, q.max_iter = r0["psf"].handle.value, rhs0.shape[0], 0, r0["max_iter"]
        q.outer, q.inner, q.tol, q.xtol = 1, 1, 1e-5, 1e-5
        q.rhs, q.x0 = rhs0.ctypes.data, r0["x0"].ctypes.data
        t = C.c_uint64()
        _lib.check(srv.lib.lc_server_submit(srv.handle, C.byref(q), C.byref(t)))
        d = C.c_int32()
        _lib.check(srv.lib.lc_server_wait(srv.handle, t, 120.0, C.byref(d)))
        assert d.value == 1
        xc = np.empty(rhs0.shape, np.float32)
        _lib.check(srv.lib.lc_server_result(srv.handle, t, xc.ctypes.data, None, None))
        out_c = [srv.result(k, timeout=120.0)[0] for k in copied]
        out_b = [srv.result(k, timeout=120.0)[0] for k in borrowed]
    for i, xr in enumerate(ref):
        assert np.array_equal(out_c[i], xr), f"copied request {i} differs from its solo solve"
        assert np.array_equal(out_b[i], xr), f"borrowed request {i} differs from its solo solve"
    assert np.array_equal(xc, ref[0])


def test_server_isolates_a_non_finite_request(lc, problems):
    # a request whose sinogram holds a NaN fails with a message; the requests
    # batched with it are bitwise their solo solves (nothing leaks across
    # problems of one batch)
    announce("server non-finite request isolation")
    reqs = [dict(r) for r in problems if r["solver"] == "cg" and "sino" in r][:3]
    bad = dict(reqs[0])
    bad["sino"] = reqs[0]["sino"].copy()
    bad["sino"][0, 3, 5] = np.nan
    ref = [solve_alone(lc, r) for r in reqs]
    with lc.Server(hbm_budget=8 << 30, lanes=1) as srv:
        tk = [submit(srv, r) for r in reqs[:1]] + [submit(srv, bad)] + [submit(srv, r) for r in reqs[1:]]
        good = [srv.result(tk[0], timeout=120.0)[0]] + \
            [srv.result(t, timeout=120.0)[0] for t in tk[2:]]
        with pytest.raises(lc.LcError, match="non-finite"):
            srv.result(tk[1], timeout=120.0)
        st = srv.stats()
    for x, xr in zip(good, ref):
        assert np.array_equal(x, xr)
    assert st["finished"] == 4 and st["inflight_bytes"] == 0


def test_server_isolates_a_failing_request(lc, problems):
    # a request that makes its batch raise (LC_REQ_TEST_FAIL, raised on the
    # host before any device work) fails alone: the batch is rerun one request
    # at a time and the others return their solo results
    announce("server failing request isolation")
    reqs = [r for r in problems if r["solver"] == "cg"][:4]
    ref = [solve_alone(lc, r) for r in reqs]
    with lc.Server(hbm_budget=8 << 30, lanes=1) as srv:
        # hold the lane on a first request so the next ones queue into one batch
        t_hold = submit(srv, reqs[0])
        tk = [submit(srv, r) for r in reqs[1:3]]
        t_bad = srv.submit(reqs[3]["psf"], test_fail=True,
                           **{k: v for k, v in reqs[3].items() if k not in ("psf", "grid")})
        tk.append(submit(srv, reqs[3]))
        x_hold = srv.result(t_hold, timeout=120.0)[0]
        out = [srv.result(t, timeout=120.0)[0] for t in tk]
        with pytest.raises(lc.LcError, match="LC_REQ_TEST_FAIL"):
            srv.result(t_bad, timeout=120.0)
        st = srv.stats()
    assert np.array_equal(x_hold, ref[0])
    for x, xr in zip(out, ref[1:]):
        assert np.array_equal(x, xr)
    assert st["finished"] == 5 and st["inflight_bytes"] == 0


def _stack_problem(lc, N, ns, nang, seed):
    rng = np.random.default_rng(seed)
    g = lc.PolarGrid2D(np.arange(nang) * np.pi / nang, N)
    p = lc.PointSpreadFunction2D(g, N)
    return p, p.convolve(rng.random((ns, N, N)).astype(np.float32))


def test_server_lanes_overlap_on_the_device(lc):
    # two groups that cannot share a batch (CG, Split-Bregman) go to two lanes;
    # the HIP-event intervals of their solves overlap on the device
    announce("server lanes overlap")
    p, y = _stack_problem(lc, 256, 24, 180, 5)
    with lc.Server(hbm_budget=16 << 30, lanes=2) as srv:
        t_cg = srv.submit(p, rhs=y, solver="cg", max_iter=150, tol=0.0, xtol=0.0)
        t_sb = srv.submit(p, rhs=y, solver="sb", lam=0.05, mu=10.0, outer=40, inner=2, tol=0.0,
                          xtol=0.0)
        srv.result(t_cg, timeout=120.0)
        srv.result(t_sb, timeout=120.0)
        tl = srv.timeline()
    assert len(tl) == 2 and tl[0][0] != tl[1][0], tl
    (a0, a1), (b0, b1) = tl[0][4:6], tl[1][4:6]
    print(f"device intervals (ms): lane {int(tl[0][0])} [{a0:.2f}, {a1:.2f}], "
          f"lane {int(tl[1][0])} [{b0:.2f}, {b1:.2f}]")
    assert a0 < a1 and b0 < b1
    assert max(a0, b0) < min(a1, b1), "the two lanes' solves did not overlap"


def test_server_admits_past_a_head_that_does_not_fit(lc):
    # budget for one large request and a little more: while the first large
    # request runs, the second does not fit, and small requests behind it
    # start before it instead of waiting
    announce("server admission past the head")
    # the large requests run long (16 x 512^2, 1000 Split-Bregman outer
    # iterations with the stopping tests off, so no early exit) and the small
    # ones are submitted well inside the first
    pb, yb = _stack_problem(lc, 512, 16, 180, 6)
    ps, ys = _stack_problem(lc, 64, 1, 90, 7)
    with lc.Server(hbm_budget=8 << 30, lanes=1) as probe:
        big_bytes = probe.request_bytes(pb, rhs=yb, solver="sb")
    with lc.Server(hbm_budget=int(1.5 * big_bytes), lanes=2) as srv:
        # warm both lanes (context, twiddles, graph capture) with two small
        # requests submitted together that cannot share a batch
        warm = [srv.submit(ps, rhs=ys, solver="cg", max_iter=2),
                srv.submit(ps, rhs=ys, solver="sb", lam=0.05, mu=10.0, outer=1, inner=1)]
        for t in warm:
            srv.result(t, timeout=60.0)
        t1 = srv.submit(pb, rhs=yb, solver="sb", lam=0.05, mu=10.0, outer=1000, inner=1,
                        tol=0.0, xtol=0.0)
        time.sleep(0.01)
        t2 = srv.submit(pb, rhs=yb, solver="sb", lam=0.05, mu=10.0, outer=1000, inner=1,
                        tol=0.0, xtol=0.0)
        ts = [srv.submit(ps, rhs=ys, solver="sb", lam=0.05, mu=10.0, outer=3, inner=1)
              for _ in range(4)]
        r1 = srv.result(t1, timeout=120.0)
        r2 = srv.result(t2, timeout=120.0)
        rs = [srv.result(t, timeout=120.0) for t in ts]
    start2 = r2[2][1]
    print(f"big request 2 starts {start2 - r1[2][1]:.3f} s after big 1; smalls start "
          f"{[round(r[2][1] - r1[2][1], 3) for r in rs]}")
    assert all(r[2][1] < start2 for r in rs)
    assert r1[1]["outer"] == r2[1]["outer"] == 1000 and np.array_equal(r1[0], r2[0])


def test_request_bytes_is_what_admission_charges(lc, problems):
    # lc_request_bytes: a budget of exactly that many bytes admits the
    # request, one byte less refuses it; the reference's fit rule on the
    # device's free memory (gpu::volume_fits_one_gpu)
    announce("server request bytes")
    from laminocam_amd.drivers import ReconParams, Regularizer
    for r in (problems[0], problems[2], problems[3]):
        kw = {k: v for k, v in r.items() if k in ("sino", "rhs")}
        with lc.Server(hbm_budget=8 << 30, lanes=1) as srv:
            b = srv.request_bytes(r["psf"], solver=r["solver"], **kw)
        with lc.Server(hbm_budget=b, lanes=1) as srv:
            x = srv.result(submit(srv, r), timeout=120.0)[0]
            assert np.all(np.isfinite(x))
        with lc.Server(hbm_budget=b - 1, lanes=1) as srv:
            with pytest.raises(lc.LcError, match="above the HBM budget"):
                submit(srv, r)
    small = ReconParams(regularizer=Regularizer.SPLIT_BREGMAN, recon_dims=(21, 511, 511))
    huge = ReconParams(regularizer=Regularizer.SPLIT_BREGMAN, recon_dims=(4096, 4096, 4096))
    assert lc.volume_fits_one_gpu(small) and not lc.volume_fits_one_gpu(huge)
