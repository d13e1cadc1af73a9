"""Serving loop (intp_server_* / intp_serve_*): a request's result is the
same whether it runs alone or in a concurrent mix, and equals the direct
build/eval calls; least-recently-used eviction, admission against the HBM
budget, and the counters."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SPECS = [
    dict(order=3, shape=(60, 50), periodic=[False, False], ranges=[(0.0, 1.0), (0.0, 2.0)], precision="f64"),
    dict(order=4, shape=(24, 20, 18), periodic=[True, False, True], ranges=[(0.0, 1.0)] * 3, precision="f64"),
    dict(order=5, shape=(40, 44), periodic=[False, True], ranges=[(-1.0, 1.0), (0.0, 1.0)],
         precision="f32_refined"),
    dict(order=1, shape=(300,), periodic=[False], ranges=[(0.0, 3.0)], precision="f64"),
]


@pytest.fixture(scope="module")
def amd():
    import bsplineinterpolation_amd as amd
    return amd


def _inputs(seed=11, npts=3000):
    import torch
    rng = np.random.default_rng(seed)
    vals, pts, orders = [], [], []
    for sp in SPECS:
        vals.append(torch.from_numpy(rng.standard_normal(sp["shape"])).cuda())
        lo = np.array([r[0] for r in sp["ranges"]])
        hi = np.array([r[1] for r in sp["ranges"]])
        D = len(sp["shape"])
        p = np.zeros((npts, 3))
        p[:, :D] = lo + (hi - lo) * rng.random((npts, D))
        pts.append(torch.from_numpy(p).cuda())
        orders.append([tuple([0] * D)] + [tuple(int(i == d) for i in range(D)) for d in range(D)])
    return vals, pts, orders


def _build(srv, i, vals):
    sp = SPECS[i]
    return srv.build(sp["order"], vals[i], periodic=sp["periodic"], ranges=sp["ranges"], precision=sp["precision"],
                     sweeps=2)


def test_gpu_serve_alone_equals_in_mix_and_direct_calls(amd):
    import torch
    vals, pts, orders = _inputs()
    # alone: one request at a time
    srv = amd.Server(streams=4)
    alone_ctrl, alone_out = [], []
    for i in range(len(SPECS)):
        sid = _build(srv, i, vals)
        srv.wait(sid)
        alone_ctrl.append(srv.spline(sid).control_points())
        rid = srv.eval(sid, pts[i], orders[i])
        alone_out.append(srv.wait(rid).cpu().numpy())
    srv.close()
    # in a mix: everything submitted at once, each eval right behind its
    # build, with extra builds and evaluations of other data in between
    srv = amd.Server(streams=4)
    rng = np.random.default_rng(5)
    ids, extra = [], []
    for rep in range(3):
        for i in range(len(SPECS)):
            other = [v + float(rng.standard_normal()) for v in vals]
            xid = _build(srv, i, other)
            extra.append(srv.eval(xid, pts[i], orders[i]))
            if rep == 1:
                sid = _build(srv, i, vals)
                ids.append((i, sid, srv.eval(sid, pts[i], orders[i])))
    outs = {rid: srv._hold[rid][1] for _, _, rid in ids}
    srv.drain()
    for i, sid, rid in ids:
        assert srv.poll(rid) == (amd._lib.INTP_REQ_DONE, 0)
        assert np.array_equal(srv.spline(sid).control_points(), alone_ctrl[i]), i
        assert np.array_equal(outs[rid].cpu().numpy(), alone_out[i]), i
    for rid in extra:
        assert srv.poll(rid)[0] == amd._lib.INTP_REQ_DONE
    srv.close()
    # the direct calls give the same numbers
    for i, sp in enumerate(SPECS):
        g = amd.InterpolationFunction(sp["order"], vals[i], periodic=sp["periodic"], ranges=sp["ranges"],
                                      precision=sp["precision"], sweeps=2)
        assert np.array_equal(g.control_points(), alone_ctrl[i]), i
        D = len(sp["shape"])
        direct = g.eval(pts[i][:, :D].cpu().numpy(), orders=orders[i])
        got = alone_out[i]
        assert np.array_equal(np.asarray(direct, dtype=got.dtype).reshape(got.shape), got), i
    torch.cuda.synchronize()


def test_gpu_serve_lru_eviction_and_admission(amd):
    import torch
    rng = np.random.default_rng(3)
    shape = (64, 64)
    one = 64 * 64 * 8
    srv = amd.Server(hbm_budget=3 * one + 100, streams=2)
    f = [torch.from_numpy(rng.standard_normal(shape)).cuda() for _ in range(4)]
    pts = torch.from_numpy(rng.random((100, 2))).cuda()
    ids = []
    for k in range(3):
        ids.append(srv.build(3, f[k], ranges=[(0, 1), (0, 1)]))
        srv.wait(ids[-1])
    srv.wait(srv.eval(ids[0], pts, [(0, 0)]))  # spline 0 becomes the most recently used
    ids.append(srv.build(3, f[3], ranges=[(0, 1), (0, 1)]))
    srv.wait(ids[-1])
    st = srv.stats()
    assert st["evictions"] == 1 and st["resident_splines"] == 3 and st["resident_bytes"] == 3 * one
    with pytest.raises(amd.IntpError) as e:  # spline 1 was the least recently used
        srv.wait(srv.eval(ids[1], pts, [(0, 0)]))
    assert e.value.code == amd._lib.INTP_E_EVICTED
    for k in (0, 2, 3):
        srv.wait(srv.eval(ids[k], pts, [(0, 0)]))
    # a build that can never fit fails; a released spline is gone
    big = torch.zeros((128, 128), dtype=torch.float64, device="cuda")
    with pytest.raises(MemoryError):
        srv.wait(srv.build(3, big, ranges=[(0, 1), (0, 1)]))
    srv.release(ids[2])
    with pytest.raises(amd.IntpError) as e:
        srv.wait(srv.eval(ids[2], pts, [(0, 0)]))
    assert e.value.code == amd._lib.INTP_E_EVICTED
    srv.close()


def test_gpu_serve_counters(amd):
    import torch
    vals, pts, orders = _inputs(seed=2, npts=500)
    srv = amd.Server(streams=3)
    sid = _build(srv, 0, vals)
    for _ in range(20):
        srv.eval(sid, pts[0], orders[0])
    srv.drain()
    st = srv.stats(reset=True)
    n = 60 * 50
    assert st["build"]["completed"] == 1 and st["eval"]["completed"] == 20
    assert st["build"]["bytes"] == 2 * 2 * n * 8
    assert st["eval"]["bytes"] == 20 * (500 * (2 * 8 + 3 * 8) + n * 8)
    for k in ("build", "eval"):
        assert 0 < st[k]["p50_ms"] <= st[k]["p99_ms"] <= st[k]["max_ms"]
    assert st["last_done_s"] >= st["first_submit_s"] >= 0
    again = srv.stats()
    assert again["eval"]["completed"] == 0 and again["resident_splines"] == 1
    srv.close()
    torch.cuda.synchronize()


def test_gpu_serve_concurrent_clients_and_nonuniform_grids(amd):
    """Several client threads submit builds (uniform and non-uniform grids,
    both precisions) and evaluations at once; each result equals the
    direct calls."""
    import threading
    import torch
    srv = amd.Server(streams=4)
    rng = np.random.default_rng(9)
    jobs = []
    for t in range(4):
        n0, n1 = 30 + 7 * t, 25 + 5 * t
        xs = np.sort(rng.random(n0)) * 2.0
        xs[0], xs[-1] = 0.0, 2.0
        f = torch.from_numpy(rng.standard_normal((n0, n1))).cuda()
        pts = np.column_stack([rng.random(2000) * 2.0, rng.random(2000), np.zeros(2000)])
        jobs.append(dict(order=3 + t % 3, f=f, coords=[xs, None], ranges=[None, (0.0, 1.0)],
                         precision="f64" if t % 2 == 0 else "f32_refined", pts=torch.from_numpy(pts).cuda()))
    results = [None] * len(jobs)
    errors = []

    def client(i):
        try:
            j = jobs[i]
            sid = srv.build(j["order"], j["f"], ranges=j["ranges"], coords=j["coords"], precision=j["precision"],
                            sweeps=2)
            outs = [srv.eval(sid, j["pts"], [(0, 0), (1, 0), (0, 1)]) for _ in range(3)]
            res = [srv.wait(r).cpu().numpy() for r in outs]
            srv.wait(sid)
            results[i] = (srv.spline(sid).control_points(), res)
        except Exception as e:  # reported below
            errors.append(e)

    threads = [threading.Thread(target=client, args=(i,)) for i in range(len(jobs))]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    assert not errors, errors
    for i, j in enumerate(jobs):
        g = amd.InterpolationFunction(j["order"], j["f"], ranges=j["ranges"], coords=j["coords"],
                                      precision=j["precision"], sweeps=2)
        ctrl, res = results[i]
        assert np.array_equal(ctrl, g.control_points()), i
        direct = g.eval(j["pts"][:, :2].contiguous(), orders=[(0, 0), (1, 0), (0, 1)]).cpu().numpy()
        for r in res:
            assert np.array_equal(r, direct), i
    srv.close()


def test_gpu_serve_memory_pressure(amd):
    """A budget of three splines and a burst of builds with evaluations
    queued behind them (no waits in between): builds wait for memory while
    pinned splines hold it, idle ones are evicted, and every request ends
    either done (with the direct call's result) or failed as evicted."""
    import torch
    rng = np.random.default_rng(17)
    shape = (48, 40)
    one = 48 * 40 * 8
    srv = amd.Server(hbm_budget=3 * one, streams=3, max_inflight=8)
    fs = [torch.from_numpy(rng.standard_normal(shape)).cuda() for _ in range(12)]
    pts = torch.from_numpy(np.column_stack([rng.random(4000), rng.random(4000)])).cuda()
    builds, evals = [], []
    for k, f in enumerate(fs):
        sid = srv.build(3, f, ranges=[(0, 1), (0, 1)])
        builds.append(sid)
        for back in range(3):  # this spline and the two before it
            if k - back >= 0:
                rid = srv.eval(builds[k - back], pts, [(0, 0), (1, 0)])
                evals.append((k - back, rid, srv._hold[rid][1]))
    srv.drain()
    st = srv.stats()
    assert st["build"]["completed"] == 12 and st["build"]["failed"] == 0
    assert st["resident_bytes"] <= 3 * one
    done = 0
    for k, rid, out in evals:
        state, status = srv.poll(rid)
        assert state in (amd._lib.INTP_REQ_DONE, amd._lib.INTP_REQ_FAILED)
        if state == amd._lib.INTP_REQ_DONE:
            g = amd.InterpolationFunction(3, fs[k], ranges=[(0, 1), (0, 1)])
            assert torch.equal(out, g.eval(pts, orders=[(0, 0), (1, 0)])), k
            done += 1
        else:
            assert status == amd._lib.INTP_E_EVICTED
    # each build's own evaluation is queued right behind it, so it runs
    assert done >= 12
    srv.close()
