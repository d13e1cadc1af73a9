"""A table of more than 2^32 control points: 524,289 x 8,192 FP64 (34 GB),
built from device values and evaluated with 64-bit offsets (the 32-bit
fast path stops at 2^32 elements).  The mesh is separable, f = g0 (x) g1,
so the spline is the product of the two 1D splines (the tensor-product
solve of a rank-one right-hand side is rank one); the 1D splines come
from the oracle.  Values and both first partials within 1e-12."""
import numpy as np
import pytest

from cases import rel_err

pytestmark = pytest.mark.gpu


def test_gpu_table_above_2_32_elements(orc):
    import torch
    import bsplineinterpolation_amd as amd
    if torch.cuda.get_device_properties(0).total_memory < (120 << 30):
        pytest.skip("needs about 100 GB of device memory")
    n0, n1 = 524289, 8192
    assert n0 * n1 > 1 << 32
    rng = np.random.default_rng(99)
    x0 = np.arange(n0) / (n0 - 1)
    g0 = np.cos(7.0 * x0) + 0.01 * rng.standard_normal(n0)
    g1 = np.sin(3.0 * np.arange(n1) / n1 * 2 * np.pi) + 1.5 + 0.01 * rng.standard_normal(n1)
    per, ranges = [False, True], [(0.0, 1.0), (0.0, 2.0)]
    print("building a %d x %d FP64 spline (%.1f GB)" % (n0, n1, n0 * n1 * 8 / 2 ** 30), flush=True)
    f = torch.outer(torch.from_numpy(g0).cuda(), torch.from_numpy(g1).cuda())
    g = amd.InterpolationFunction(3, f, periodic=per, ranges=ranges)
    del f
    torch.cuda.synchronize()
    q = np.stack([rng.uniform(0.0, 1.0, 20000), rng.uniform(0.0, 2.0, 20000)], axis=1)
    q[:4] = [[0.0, 0.0], [1.0, 1.999], [0.5, 1.0], [1e-9, 2.0 - 1e-9]]
    orders = [[0, 0], [1, 0], [0, 1]]
    got = g.eval(torch.from_numpy(q).cuda(), orders).cpu().numpy()
    del g
    s0 = orc.OracleSpline(3, g0, periodic=[per[0]], ranges=[ranges[0]])
    s1 = orc.OracleSpline(3, g1, periodic=[per[1]], ranges=[ranges[1]])
    v0, d0 = s0.eval(q[:, [0]], [[0], [1]])
    v1, d1 = s1.eval(q[:, [1]], [[0], [1]])
    want = [v0 * v1, d0 * v1, v0 * d1]
    for i, r in enumerate(orders):
        e = rel_err(got[i], want[i])
        print("order %s: rel_err %.3g" % (r, e), flush=True)
        assert e < 1e-12, (r, e)
