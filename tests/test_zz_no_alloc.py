"""The no-allocation mixed pressure solve (wf_mg_solve_mixed_static, its
workspace in the library's static device scratch) against the workspace form
(wf_mg_solve_mixed), both called through the C ABI on the same inputs: two
calls in a row, the second on a smaller grid, must give the same pp bit for
bit, the same cycle count and the same residual.  The size and flag checks
return their error codes on the host, before anything is launched."""
import ctypes as C

import numpy as np
import pytest

from test_gpu_mg import _Cfg
from test_gpu_mixed import topo_for

pytestmark = pytest.mark.gpu


def _both_forms(n3, levels, max_cycles, rtol, flags, seed):
    from warpflow import _hip
    from warpflow.device import array, current_stream, grid_of, zeros
    from warpflow.mixed import MixedMultigrid
    print("no-alloc solve: n3=%s levels=%d cycles=%d rtol=%g flags=%d seed=%d"
          % (n3, levels, max_cycles, rtol, flags, seed), flush=True)
    h = tuple(1.0 / n for n in n3)
    shape = tuple(n + 6 for n in n3)
    su = np.random.default_rng(seed).standard_normal(shape)
    # zero mean on the real cells: the singular Neumann system is then
    # consistent and the tolerance mode can reach rtol
    su[3:-3, 3:-3, 3:-3] -= su[3:-3, 3:-3, 3:-3].mean()
    su3 = array(su)
    mg = MixedMultigrid(_Cfg(levels), topo_for(n3), None, *h)
    g = grid_of(su3, 3)
    need = C.c_size_t(0)
    _hip.call("wf_mg_mixed_workspace_bytes", g, C.byref(need))
    assert need.value <= _hip.WF_STATIC_SCRATCH_BYTES
    work = zeros((need.value + 7) // 8)
    out = {}
    for form in ("workspace", "static"):
        pp = zeros(shape)  # ghost layers beyond the first stay zero, as the solve expects
        cyc, res = C.c_int(-1), C.c_double(-1.0)
        head = (mg._descs(), len(mg.levels), su3.ptr, pp.ptr, g, mg.n_pre, mg.n_post,
                mg.n_coarse, max_cycles, rtol, flags)
        if form == "static":
            _hip.call("wf_mg_solve_mixed_static", *head, C.byref(cyc), C.byref(res),
                      current_stream())
        else:
            _hip.call("wf_mg_solve_mixed", *head, work.ptr, work.nbytes, C.byref(cyc),
                      C.byref(res), current_stream())
        out[form] = (pp.numpy(), cyc.value, res.value)
    return out


def test_static_scratch_solve_matches_workspace_solve(gpu):
    # call 1: 64^3, fixed two cycles, gauge mean of pp (block sums in the scratch)
    a = _both_forms((64, 64, 64), 4, 2, 0.0, 2, seed=11)
    # call 2, smaller: tolerance mode with the defect mean (per-cycle residual
    # maxima and sums in the scratch, fewer partials than the first call left)
    b = _both_forms((40, 48, 56), 3, 30, 1e-8, 1 | 2, seed=12)
    for res in (a, b):
        pw, cw, rw = res["workspace"]
        ps, cs, rs = res["static"]
        assert np.isfinite(pw).all()
        assert np.array_equal(pw, ps)
        assert cw == cs and rw == rs
    assert a["static"][1] == 2
    assert 1 <= b["static"][1] < 30 and b["static"][2] <= 1e-8


def test_static_scratch_limits(gpu):
    from warpflow import _hip
    # 512^3 needs 4 MiB of partial sums: refused before any launch (levels unread)
    n = 512
    big = _hip.WfGrid(n, n, n, 3, n + 6, (n + 6) * (n + 6))
    need = C.c_size_t(0)
    _hip.call("wf_mg_mixed_workspace_bytes", big, C.byref(need))
    assert need.value > _hip.WF_STATIC_SCRATCH_BYTES
    code = _hip.lib.wf_mg_solve_mixed_static(None, 4, None, None, big, 6, 6, 300, 2, 0.0, 0,
                                             None, None, None)
    assert code == _hip.WF_ESCRATCH
    assert "static scratch" in _hip.lib.wf_error_string(code).decode()
    # the deferred gauge mean (flags & 4) would stay in the scratch: refused
    m = 32
    small = _hip.WfGrid(m, m, m, 3, m + 6, (m + 6) * (m + 6))
    code = _hip.lib.wf_mg_solve_mixed_static(None, 3, None, None, small, 6, 6, 300, 2, 0.0,
                                             2 | 4, None, None, None)
    assert code == _hip.WF_EINVAL
