"""ftk2h_extract_regular_noalloc (include/ftk2h.h) through the C ABI: the result it
writes into caller-allocated device arrays, with its work arrays in the library's
fixed scratch array, equals ftk2h_extract_regular's bit for bit, on two calls in a
row (the second on a smaller mesh, so stale scratch contents of the first must not
leak), and its capacity and argument errors are reported, not faulted."""
import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from ftk2_amd import _lib, engine, synth  # noqa: E402


def call_noalloc(mesh, kind, m, comps, node_cap, cell_cap, nv, dtype=torch.float64):
    views = (_lib.FieldView * len(comps))(*[engine.field_view(c) for c in comps])
    pd = _lib.PredicateDesc()
    pd.kind, pd.codimension, pd.dtype, pd.sos_q = kind, m, engine._dtype_of(comps[0]), 1e6
    keys = torch.zeros(max(node_cap, 1), dtype=torch.int64, device="cuda")
    lam = torch.zeros((max(node_cap, 1), m + 1), dtype=dtype, device="cuda")
    track = torch.zeros(max(node_cap, 1), dtype=torch.int32, device="cuda")
    cells = torch.zeros((max(cell_cap, 1), max(nv, 1)), dtype=torch.int32, device="cuda")
    ob = _lib.OutputBuffers(keys.data_ptr(), lam.data_ptr(), track.data_ptr(), None, cells.data_ptr(),
                            node_cap, cell_cap)
    nn, nc = C.c_uint64(), C.c_uint64()
    md = mesh.desc()
    torch.cuda.synchronize()
    rc = _lib.lib().ftk2h_extract_regular_noalloc(engine.context(0), C.byref(md), C.byref(pd), views, len(comps),
                                                  None, C.byref(ob), C.byref(nn), C.byref(nc), None)
    return rc, nn.value, nc.value, keys, lam, track, cells


def check_equal(mesh, kind, m, comps, nv, dtype=torch.float64):
    ref = engine.extract(mesh, kind, m, comps)
    dv = ref.device_view()
    n = dv.n_nodes
    ncell = dv.conn_cells[0] if dv.n_conn else 0
    rc, nn, nc, keys, lam, track, cells = call_noalloc(mesh, kind, m, comps, n + 7, ncell + 5, nv, dtype)
    assert rc == 0, _lib.lib().ftk2h_last_error()
    assert (nn, nc) == (n, ncell) and n > 0
    assert np.array_equal(keys[:n].cpu().numpy().view(np.uint64), ref.keys())
    assert np.array_equal(lam[:n].cpu().numpy(), ref.to_complex().lam)
    cx = ref.to_complex()
    assert np.array_equal(track[:n].cpu().numpy().astype(np.uint64), cx.track_id)
    if ncell:
        assert np.array_equal(cells[:ncell].cpu().numpy().view(np.uint32), cx.connectivity[nv - 1])
    return n, ncell


def dev(arrs):
    # device tensors indexed [x, y, (z), t] (x fastest in memory)
    return [torch.from_numpy(np.ascontiguousarray(a.transpose())).cuda().permute(*reversed(range(a.ndim)))
            for a in arrs]


def test_noalloc_equals_extract_two_calls():
    big = dev(synth.cp3d_config(n=(40, 36, 34), nt=4, seed=1, noise=0.02))
    small = dev(synth.cp3d_config(n=(18, 20, 16), nt=3, seed=4, noise=0.05))
    n1, c1 = check_equal(engine.RegularSimplicialMesh([40, 36, 34, 4]), 0, 3, big, 2)
    n2, c2 = check_equal(engine.RegularSimplicialMesh([18, 20, 16, 3]), 0, 3, small, 2)
    assert n2 < n1 and c2 < c1


def test_noalloc_cp2d_and_contour():
    uv = dev(synth.cp2d_config(nx=200, ny=150, nt=6, seed=4))
    check_equal(engine.RegularSimplicialMesh([200, 150, 6]), 0, 2, uv, 2)
    f = dev([synth.gaussians((48, 40, 36), 4, 16, seed=3, gradient=False, subtract_mean=True,
                             dtype=np.float32)[0]])
    check_equal(engine.RegularSimplicialMesh([48, 40, 36, 4]), 1, 1, f, 4, torch.float32)


def test_noalloc_errors():
    uv = dev(synth.cp2d_config(nx=64, ny=64, nt=4, seed=0))
    mesh = engine.RegularSimplicialMesh([64, 64, 4])
    ref = engine.extract(mesh, 0, 2, uv)
    n = ref.device_view().n_nodes
    assert n > 1
    rc, nn, nc, *_ = call_noalloc(mesh, 0, 2, uv, n - 1, 10 ** 6, 2)
    assert rc == 3  # FTK2H_E_OVERFLOW
    assert nn == n  # the required count is reported
    # star-fan cells are not built without allocation
    f = dev(synth.gaussians((12, 12, 12), 3, 4, seed=1, gradient=False))
    rc, *_ = call_noalloc(engine.RegularSimplicialMesh([12, 12, 12, 3]), 2, 2, [f[0], f[0]], 100, 100, 3,
                          torch.float64)
    assert rc == 2
    # a mesh whose work arrays exceed the scratch array
    rc, *_ = call_noalloc(engine.RegularSimplicialMesh([4096, 4096, 2]), 0, 2,
                          [torch.zeros((4096, 4096, 2), dtype=torch.float64, device="cuda")] * 2, 10, 10, 2)
    assert rc == 3
