"""The HIP adjoints of the transforms (no reference counterpart) are pinned by
mathematics: the transpose identity <T c, y> = <c, T^T y> on random vectors in
float64 (a wrong sign, index, boundary weight or pass order breaks it), central
finite differences of an MSE loss through autograd, and linearity."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

# (shape z,y,x), levels_xy, levels_z, dyadic
CASES = [
    ((24, 20, 28), 2, 2, True),
    ((17, 19, 23), 1, 1, True),
    ((33, 40, 36), 3, 3, True),
    ((12, 30, 26), 2, 1, False),
    ((9, 70, 65), 3, 0, False),
    ((100, 10, 9), 1, 4, False),
    ((64, 64, 64), 3, 3, True),
    ((10, 130, 140), 4, 1, False),
    ((6, 12, 12), 1, 1, True),      # short z lines (n < 9)
    ((12, 5, 7), 2, 1, False),
    ((3, 2, 10), 1, 1, True),       # lines of 2 and 3 samples
]


@pytest.fixture(scope="module")
def wz():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import waltz_amd
    from waltz_amd import autograd, stages
    waltz_amd.lib()
    return stages, autograd


def _dot(a, b):
    return float((a.double() * b.double()).sum())


@pytest.mark.parametrize("case", CASES, ids=lambda c: "x".join(map(str, c[0])) + f"_l{c[1]}{c[2]}")
@pytest.mark.parametrize("inverse", [False, True], ids=["dwt", "idwt"])
def test_transpose_identity_f64(wz, case, inverse):
    stages, _ = wz
    shape, lxy, lz, dyadic = case
    g = torch.Generator(device="cuda").manual_seed(5)
    c = torch.randn(shape, dtype=torch.float64, device="cuda", generator=g)
    y = torch.randn(shape, dtype=torch.float64, device="cuda", generator=g)
    tc = stages.transform(c, lxy, lz, dyadic, inverse=inverse)
    ty = stages.transform_adjoint(y, lxy, lz, dyadic, of_inverse=inverse)
    lhs, rhs = _dot(tc, y), _dot(c, ty)
    scale = float(tc.norm() * y.norm())
    assert abs(lhs - rhs) <= 1e-12 * scale, (lhs, rhs)


@pytest.mark.parametrize("case", CASES[:4], ids=lambda c: "x".join(map(str, c[0])))
def test_transpose_identity_f32(wz, case):
    stages, _ = wz
    shape, lxy, lz, dyadic = case
    g = torch.Generator(device="cuda").manual_seed(6)
    for inverse in (False, True):
        c = torch.randn(shape, dtype=torch.float32, device="cuda", generator=g)
        y = torch.randn(shape, dtype=torch.float32, device="cuda", generator=g)
        tc = stages.transform(c, lxy, lz, dyadic, inverse=inverse)
        ty = stages.transform_adjoint(y, lxy, lz, dyadic, of_inverse=inverse)
        lhs, rhs = _dot(tc, y), _dot(c, ty)
        assert abs(lhs - rhs) <= 1e-5 * float(tc.norm() * y.norm())


def test_adjoint_is_linear_and_in_place(wz):
    stages, _ = wz
    shape, lxy, lz, dyadic = (20, 18, 22), 1, 1, True
    a = torch.randn(shape, dtype=torch.float64, device="cuda")
    b = torch.randn(shape, dtype=torch.float64, device="cuda")
    ta = stages.transform_adjoint(a, lxy, lz, dyadic, of_inverse=True)
    tb = stages.transform_adjoint(b, lxy, lz, dyadic, of_inverse=True)
    tab = stages.transform_adjoint(2 * a - 3 * b, lxy, lz, dyadic, of_inverse=True)
    assert torch.allclose(tab, 2 * ta - 3 * tb, rtol=0, atol=1e-12)
    inplace = a.clone()
    stages.transform_adjoint(inplace, lxy, lz, dyadic, of_inverse=True, out=inplace)
    assert torch.equal(inplace, ta)


def test_autograd_matches_finite_differences(wz):
    """FP64 central differences of an MSE loss through idwt, on a 16^3 crop
    with one dyadic level (and through dwt)."""
    _, ag = wz
    rng = np.random.default_rng(3)
    shape = (16, 16, 16)
    levels = (1, 1, True)
    target = torch.from_numpy(rng.standard_normal(shape)).cuda()
    c0 = torch.from_numpy(rng.standard_normal(shape)).cuda()

    def loss(c):
        return ((ag.idwt(c, levels) - target) ** 2).mean()

    c = c0.clone().requires_grad_()
    loss(c).backward()
    h = 1e-6
    for flat in rng.choice(c0.numel(), 24, replace=False):
        e = torch.zeros_like(c0).view(-1)
        e[flat] = h
        e = e.view(shape)
        fd = (float(loss(c0 + e)) - float(loss(c0 - e))) / (2 * h)
        assert abs(float(c.grad.view(-1)[flat]) - fd) <= 1e-6 * max(1.0, abs(fd)), flat

    x = c0.clone().requires_grad_()
    (ag.dwt(x, levels) * target).sum().backward()
    for flat in rng.choice(c0.numel(), 8, replace=False):
        e = torch.zeros_like(c0).view(-1)
        e[flat] = h
        e = e.view(shape)
        fd = float(((ag.dwt(c0 + e, levels) - ag.dwt(c0 - e, levels)) * target).sum()) / (2 * h)
        assert abs(float(x.grad.view(-1)[flat]) - fd) <= 1e-6 * max(1.0, abs(fd)), flat


def test_idwt_of_dwt_roundtrip_and_default_levels(wz):
    _, ag = wz
    x = torch.randn((40, 36, 44), dtype=torch.float64, device="cuda")
    assert ag.default_levels(x.shape) == (3, 3, True)
    y = ag.idwt(ag.dwt(x))
    assert float((y - x).abs().max()) < 1e-12


@pytest.mark.parametrize("inverse", [False, True], ids=["dwt", "idwt"])
def test_dense_transpose_small(wz, inverse):
    """Dense matrices at 9x8x7 (one dyadic level; lines of 7, 8 and 9 samples):
    the adjoint's matrix equals the transform's matrix transposed."""
    stages, _ = wz
    shape, lxy, lz, dyadic = (9, 8, 7), 1, 1, True
    n = int(np.prod(shape))
    eye = torch.eye(n, dtype=torch.float64, device="cuda")
    fwd = torch.stack([stages.transform(eye[i].view(shape), lxy, lz, dyadic, inverse=inverse).reshape(-1)
                       for i in range(n)], dim=1)
    adj = torch.stack([stages.transform_adjoint(eye[i].view(shape), lxy, lz, dyadic, of_inverse=inverse).reshape(-1)
                       for i in range(n)], dim=1)
    assert float((adj - fwd.T).abs().max()) <= 1e-13 * float(fwd.abs().max())


def test_async_and_sync_entry_points_agree(wz):
    """waltz_amd_transform / _adjoint (synchronous, library scratch) and their
    _async forms (caller scratch, no synchronisation) give the same bits."""
    import ctypes as C
    from waltz_amd._lib import lib
    stages, _ = wz
    shape, lxy, lz, dyadic = CASES[3]
    nz, ny, nx = shape
    g = torch.Generator(device="cuda").manual_seed(11)
    x = torch.randn(shape, dtype=torch.float32, device="cuda", generator=g)
    for inverse in (0, 1):
        out = torch.empty_like(x)
        assert lib().waltz_amd_transform(4, inverse, x.data_ptr(), out.data_ptr(), nx, ny, nz, lxy, lz, int(dyadic),
                                         None) == 0
        assert torch.equal(out, stages.transform(x, lxy, lz, dyadic, inverse=bool(inverse)))
        assert lib().waltz_amd_transform_adjoint(4, inverse, x.data_ptr(), out.data_ptr(), nx, ny, nz, lxy, lz,
                                                 int(dyadic), None) == 0
        assert torch.equal(out, stages.transform_adjoint(x, lxy, lz, dyadic, of_inverse=bool(inverse)))


def test_autograd_step_replays_from_a_captured_graph(wz):
    """IDWT forward and its backward through the adjoint (for a fixed output
    gradient, so no reduction is captured) recorded in a graph: every replay,
    with eager work between, equals the eager step bit for bit."""
    _, autograd = wz
    shape = (32, 48, 40)
    g = torch.Generator(device="cuda").manual_seed(12)
    c0 = torch.randn(shape, dtype=torch.float32, device="cuda", generator=g)
    gout = torch.randn(shape, dtype=torch.float32, device="cuda", generator=g)
    c = c0.clone().requires_grad_()
    x = autograd.idwt(c)
    x.backward(gout)
    want_x, want_grad = x.detach().clone(), c.grad.clone()
    cs = c0.clone().requires_grad_()
    stream = torch.cuda.Stream()
    stream.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(stream):  # warm-up on the capture stream
        autograd.idwt(cs).backward(gout)
    torch.cuda.current_stream().wait_stream(stream)
    graph = torch.cuda.CUDAGraph()
    cs.grad = None
    with torch.cuda.graph(graph, stream=stream):
        gx = autograd.idwt(cs)
        gx.backward(gout)
    for _ in range(3):
        cs.grad.fill_(float("nan"))
        gx.detach().fill_(float("nan"))
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(gx.detach(), want_x)
        assert torch.equal(cs.grad, want_grad)
        _ = autograd.idwt(c0) * 2.0  # eager work between the replays
