"""Batched form (intp_batch_*): every spline of a mixed batch must come out
exactly as the single-spline path computes it (control points and values
bit for bit), graph replays must follow new buffer contents, and the
autotuner must not change results."""
import numpy as np
import pytest

from cases import rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def amd():
    import bsplineinterpolation_amd as amd
    return amd


@pytest.fixture(scope="module")
def work():
    from bsplineinterpolation_amd import datagen
    return datagen.workload_batch_mixed(seed=7, count=48, queries_per_spline=300)


def _setup(amd, work):
    import torch
    b = amd.Batch(work["specs"])
    assert b.total == work["values"].size
    vals = torch.from_numpy(work["values"]).cuda()
    pts = torch.from_numpy(np.ascontiguousarray(work["points"])).cuda()
    return b, vals, pts


def test_gpu_batch_equals_single_splines(amd, work, orc):
    import torch
    b, vals, pts = _setup(amd, work)
    out = b.run(vals, pts, work["point_offsets"])
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    po = work["point_offsets"]
    for s, sp in enumerate(work["specs"]):
        f = work["values"][b.offsets[s]:b.offsets[s + 1]].reshape(sp["shape"])
        g = amd.InterpolationFunction(sp["order"], f, periodic=sp["periodic"], ranges=sp["ranges"])
        assert np.array_equal(b.control_points(s), g.control_points().ravel()), s
        D = len(sp["shape"])
        q = work["points"][po[s]:po[s + 1], :D]
        assert np.array_equal(got[po[s]:po[s + 1]], g.eval(q)[0]), s
    # and the oracle for a few splines of each dimension
    for s in (0, 30, 45):
        sp = work["specs"][s]
        f = work["values"][b.offsets[s]:b.offsets[s + 1]].reshape(sp["shape"])
        o = orc.OracleSpline(sp["order"], f, periodic=sp["periodic"], ranges=sp["ranges"])
        D = len(sp["shape"])
        want = o.eval(np.ascontiguousarray(work["points"][po[s]:po[s + 1], :D]))[0]
        assert rel_err(got[po[s]:po[s + 1]], want) < 1e-12


def test_gpu_batch_graph_replay_and_autotune(amd, work):
    import torch
    b, vals, pts = _setup(amd, work)
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        plain = b.run(vals, pts, work["point_offsets"]).clone()
        out = torch.empty_like(plain)
        b.run(vals, pts, work["point_offsets"], out=out, graph=True)
        first = out.clone()
        b.run(vals, pts, work["point_offsets"], out=out, graph=True)  # replay
        second = out.clone()
        vals.mul_(2.0)  # exact scaling: a linear spline of 2 f is exactly 2 s(f)
        b.run(vals, pts, work["point_offsets"], out=out, graph=True)  # replay sees new values
        doubled = out.clone()
        vals.mul_(0.5)
        tuned = b.run(vals, pts, work["point_offsets"], autotune=True).clone()
    stream.synchronize()
    assert torch.equal(first, plain) and torch.equal(second, plain)
    assert torch.equal(doubled, 2.0 * plain)
    assert torch.equal(tuned, plain)
    rep = b.tuning()
    assert "tile_cap=" in rep and "->" in rep, rep
    print(rep)


def test_gpu_batch_graph_after_layout_change(amd, work):
    """A captured graph bakes in the eval work items.  An evaluation with
    a larger point layout reallocates them; a graph run with the first
    layout afterwards must not replay the stale graph, and equals the plain
    run (twice, so the recaptured graph is replayed too)."""
    import torch
    b, vals, pts = _setup(amd, work)
    po = work["point_offsets"]
    stream = torch.cuda.Stream()
    big_po = po * 4
    big_pts = pts.repeat(4, 1).contiguous()
    with torch.cuda.stream(stream):
        plain = b.run(vals, pts, po).clone()
        out = torch.empty_like(plain)
        b.run(vals, pts, po, out=out, graph=True)
        big = b.eval(big_pts, big_po)
        out.zero_()
        b.run(vals, pts, po, out=out, graph=True)
        again = out.clone()
        out.zero_()
        b.run(vals, pts, po, out=out, graph=True)
    stream.synchronize()
    assert torch.equal(again, plain) and torch.equal(out, plain)
    assert big.numel() == int(big_po[-1])


def test_gpu_batch_rejects_bad_buffers(amd, work):
    """Buffers that do not match the batch raise before any launch."""
    import torch
    b, vals, pts = _setup(amd, work)
    po = work["point_offsets"]
    with pytest.raises(ValueError):
        b.build(vals[:-1].contiguous())
    with pytest.raises(ValueError):
        b.build(vals.float())
    with pytest.raises(ValueError):
        b.eval(pts[:10].contiguous(), po)
    with pytest.raises(ValueError):
        b.eval(pts[:, :2].contiguous(), po)
    with pytest.raises(ValueError):
        b.eval(pts, po[:-1])
    bad = po.copy()
    bad[0] = -1
    with pytest.raises(ValueError):
        b.eval(pts, bad)
    with pytest.raises(ValueError):
        b.run(vals, pts, po, out=torch.empty(3, dtype=torch.float64, device="cuda"))


def test_gpu_batch_separate_build_and_eval(amd, work):
    import torch
    b, vals, pts = _setup(amd, work)
    plain = b.run(vals, pts, work["point_offsets"]).clone()
    b.build(vals)
    again = b.eval(pts, work["point_offsets"])
    # a different point layout: every other spline gets no points
    po = work["point_offsets"]
    counts = np.diff(po).copy()
    counts[1::2] = 0
    po2 = np.concatenate([[0], np.cumsum(counts)])
    sel = np.concatenate([np.arange(po[s], po[s + 1]) for s in range(0, len(counts), 2)])
    sub = b.eval(pts[torch.from_numpy(sel).cuda()].contiguous(), po2)
    torch.cuda.synchronize()
    assert torch.equal(again, plain)
    assert torch.equal(sub, plain[torch.from_numpy(sel).cuda()])


def test_gpu_batch_rejects_bad_specs(amd):
    with pytest.raises(amd.IntpError):
        amd.Batch([dict(order=3, shape=(10,), ranges=[(0, 1)]), dict(order=9, shape=(10,), ranges=[(0, 1)])])
    with pytest.raises(amd.IntpError):
        amd.Batch([dict(order=4, shape=(4,), periodic=[True], ranges=[(0, 1)])])


def test_gpu_batch_long_lines_and_singletons(amd):
    """1D splines longer than an LDS tile (solved in place by the line
    kernel), periodic and clamped, beside small 2D/3D ones; equal to the
    single-spline path."""
    import torch
    rng = np.random.default_rng(21)
    specs = [dict(order=5, shape=(20000,), periodic=[True], ranges=[(0.0, 1.0)]),
             dict(order=4, shape=(15000,), periodic=[False], ranges=[(0.0, 2.0)]),
             dict(order=1, shape=(2,), periodic=[False], ranges=[(0.0, 1.0)]),
             dict(order=3, shape=(5, 70), periodic=[False, True], ranges=[(0.0, 1.0), (0.0, 1.0)]),
             dict(order=2, shape=(9, 8, 66), periodic=[True, False, False], ranges=[(0.0, 1.0)] * 3)]
    b = amd.Batch(specs)
    vals = rng.standard_normal(b.total)
    q = 300
    pts = np.zeros((q * len(specs), 3))
    pts[:, :] = rng.random((q * len(specs), 3)) * 0.999
    pts[q:2 * q, 0] *= 2.0
    po = np.arange(len(specs) + 1) * q
    out = b.run(torch.from_numpy(vals).cuda(), torch.from_numpy(pts).cuda(), po).cpu().numpy()
    for s, sp in enumerate(specs):
        f = vals[b.offsets[s]:b.offsets[s + 1]].reshape(sp["shape"])
        g = amd.InterpolationFunction(sp["order"], f, periodic=sp["periodic"], ranges=sp["ranges"])
        assert np.array_equal(b.control_points(s), g.control_points().ravel()), s
        D = len(sp["shape"])
        assert np.array_equal(out[po[s]:po[s + 1]], g.eval(pts[po[s]:po[s + 1], :D])[0]), s


@pytest.mark.parametrize("order,periodic", [(2, False), (3, False), (5, False), (6, False), (7, False),
                                            (2, True), (3, True), (5, True), (6, True)])
def test_gpu_batch_split_lines(amd, orc, order, periodic):
    """INTP_BATCH_SPLIT_LINES: long 1D lines solved in parallel segments
    chained by precomputed responses.  Same linear system as the
    sequential solve, rounded differently: control points and values within
    1e-12 (norm-wise relative) of the sequential batch and of the oracle;
    short lines and multi-dimensional splines stay bit-identical."""
    import torch
    rng = np.random.default_rng(order * 10 + periodic)
    specs = [dict(order=order, shape=(4099,), periodic=[periodic], ranges=[(0.0, 1.0)]),
             dict(order=order, shape=(1500,), periodic=[periodic], ranges=[(-1.0, 2.0)]),
             dict(order=order, shape=(200,), periodic=[periodic], ranges=[(0.0, 1.0)]),
             dict(order=3, shape=(40, 50), periodic=[periodic, False], ranges=[(0.0, 1.0), (0.0, 1.0)])]
    b = amd.Batch(specs)
    x = np.linspace(0.0, 1.0, b.total)
    vals = torch.from_numpy(np.sin(40 * x) + 0.01 * rng.standard_normal(b.total)).cuda()
    q = 500
    pts = rng.random((q * len(specs), 3)) * 0.999
    pts[q:2 * q, 0] = -1.0 + 3.0 * pts[q:2 * q, 0]
    po = np.arange(len(specs) + 1) * q
    pd = torch.from_numpy(pts).cuda()
    seq = b.run(vals, pd, po).clone()
    ctrl_seq = [b.control_points(s) for s in range(len(specs))]
    split = b.run(vals, pd, po, split_lines=True, graph=True, stream=torch.cuda.Stream())
    torch.cuda.synchronize()
    for s in range(len(specs)):
        c = b.control_points(s)
        if s < 2:
            assert rel_err(c, ctrl_seq[s]) < 1e-12, s
            assert rel_err(split[po[s]:po[s + 1]].cpu().numpy(), seq[po[s]:po[s + 1]].cpu().numpy()) < 1e-12, s
        else:  # below the split threshold, or not 1D: the sequential solve
            assert np.array_equal(c, ctrl_seq[s]), s
            assert torch.equal(split[po[s]:po[s + 1]], seq[po[s]:po[s + 1]]), s
    f0 = vals[b.offsets[0]:b.offsets[1]].cpu().numpy()
    o = orc.OracleSpline(order, f0, periodic=[periodic], ranges=[(0.0, 1.0)])
    want = o.eval(np.ascontiguousarray(pts[:q, :1]))[0]
    assert rel_err(split[:q].cpu().numpy(), want) < 1e-12


def test_gpu_batch_tile_size_classes(amd):
    """The tiles of a pass are launched in three workgroup shapes by their
    LDS size (at most 20 KB: one wave; at most 39 KB: two waves; larger:
    256 threads, with the in-place runs).  Splines whose lines land in
    each class on every pass, periodic and clamped, band widths 1 to 4,
    give control points and values bit-identical to the single-spline
    path."""
    import torch
    rng = np.random.default_rng(33)
    # 64 lines per tile unless a line is long: 64 * (n | 1) * 8 bytes
    specs = [dict(order=3, shape=(40, 16), periodic=[False, False], ranges=[(0.0, 1.0)] * 2),    # small
             dict(order=5, shape=(30, 50), periodic=[True, False], ranges=[(0.0, 1.0)] * 2),     # medium
             dict(order=2, shape=(70, 150), periodic=[False, True], ranges=[(0.0, 1.0)] * 2),    # large
             dict(order=4, shape=(20, 255), periodic=[False, False], ranges=[(0.0, 1.0)] * 2),   # large, 32 lines
             dict(order=3, shape=(50, 30, 60), periodic=[False, True, False], ranges=[(0.0, 1.0)] * 3),
             dict(order=1, shape=(100, 9, 12), periodic=[True, False, False], ranges=[(0.0, 1.0)] * 3)]
    b = amd.Batch(specs)
    vals = rng.standard_normal(b.total)
    q = 200
    pts = rng.random((q * len(specs), 3)) * 0.999
    po = np.arange(len(specs) + 1) * q
    out = b.run(torch.from_numpy(vals).cuda(), torch.from_numpy(pts).cuda(), po).cpu().numpy()
    for s, sp in enumerate(specs):
        f = vals[b.offsets[s]:b.offsets[s + 1]].reshape(sp["shape"])
        g = amd.InterpolationFunction(sp["order"], f, periodic=sp["periodic"], ranges=sp["ranges"])
        assert np.array_equal(b.control_points(s), g.control_points().ravel()), s
        D = len(sp["shape"])
        assert np.array_equal(out[po[s]:po[s + 1]], g.eval(pts[po[s]:po[s + 1], :D])[0]), s
