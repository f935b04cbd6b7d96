"""Differentiable measurement operator on the GPU: adjoint dot-product test,
central finite differences, and the chi-squared gradient against the FP64
oracle (scaled-down version of the 1024^2 / 1e7-point configuration)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

C = 299792458.0


def setup(gpu, n=128, n_time=60, seed=3, dtype=torch.complex64):
    from astroviper_amd import synth
    from astroviper_amd.measurement import MeasurementOperator
    arr = synth.ArrayConfig(n_time=n_time, seed=seed)
    obs = synth.make_observation(arr, synth.SkyConfig())
    maxuv = np.max(np.hypot(obs.uvw[..., 0], obs.uvw[..., 1]))
    c = 0.45 * C / obs.freq[-1] / maxuv
    cell = (-c, c)
    rows = obs.uvw.shape[0] * obs.uvw.shape[1]
    op = MeasurementOperator(torch.from_numpy(obs.uvw.reshape(rows, 3)).to(gpu),
                             torch.from_numpy(obs.freq).to(gpu), (n, n), cell, n_pol=1,
                             dtype=dtype, device=gpu)
    return obs, op, cell


def blob_sky(n, nblob, rng):
    yy, xx = np.mgrid[0:n, 0:n]
    img = np.zeros((n, n))
    for _ in range(nblob):
        cx, cy = rng.uniform(0.2 * n, 0.8 * n, size=2)
        s = rng.uniform(2, 10)
        img += rng.uniform(0.1, 1.0) * np.exp(-((xx - cx) ** 2 + (yy - cy) ** 2) / (2 * s * s))
    return img


@pytest.mark.parametrize("dtype,tol", [(torch.complex64, 1e-6), (torch.complex128, 1e-12)])
def test_adjoint_dot_product(gpu, dtype, tol):
    obs, op, _ = setup(gpu, dtype=dtype)
    g = torch.Generator(device=gpu)
    g.manual_seed(31)
    x = torch.randn(op.shape, dtype=dtype, device=gpu, generator=g)
    y = torch.randn((op.n_samples, 1), dtype=dtype, device=gpu, generator=g)
    ax = op.forward(x)
    ahy = op.adjoint(y)
    lhs = torch.vdot(y.flatten().to(torch.complex128), ax.flatten().to(torch.complex128))
    rhs = torch.vdot(ahy.flatten().to(torch.complex128), x.flatten().to(torch.complex128))
    norm = torch.linalg.norm(ax.to(torch.complex128)) * torch.linalg.norm(y.to(torch.complex128))
    mismatch = (lhs - rhs).abs().item() / norm.item()
    print(f"adjoint mismatch {dtype}: {mismatch:.3e}")
    assert mismatch < tol


def test_forward_matches_oracle(gpu, oracle):
    obs, op, cell = setup(gpu, dtype=torch.complex128)
    rng = np.random.default_rng(4)
    sky = blob_sky(128, 10, rng)
    y = op.forward(torch.from_numpy(sky).to(gpu)).cpu().numpy()
    ref = oracle.measure(sky[None, None], obs.uvw, obs.freq, cell)
    ref = ref.reshape(y.shape)
    assert np.max(np.abs(y - ref)) <= 1e-6 * np.max(np.abs(ref))
    _, op32, _ = setup(gpu, dtype=torch.complex64)
    y32 = op32.forward(torch.from_numpy(sky).to(gpu)).cpu().numpy()
    assert np.max(np.abs(y32 - ref)) <= 1e-4 * np.max(np.abs(ref))


def test_finite_differences_fp64(gpu):
    obs, op, _ = setup(gpu, n=64, n_time=30, dtype=torch.complex128)
    rng = np.random.default_rng(5)
    sky = torch.from_numpy(blob_sky(64, 6, rng)).to(gpu)
    data = op.forward(sky * 1.1 + 0.01)
    x = sky.clone().requires_grad_(True)
    chi2 = (op.apply(x) - data).abs().pow(2).sum()
    chi2.backward()
    grad = x.grad.cpu().numpy()
    h = 1e-3
    pix = rng.integers(8, 56, size=(32, 2))
    for (i, j) in pix:
        e = torch.zeros_like(sky)
        e[i, j] = h
        fp = (op.forward(sky + e) - data).abs().pow(2).sum().item()
        fm = (op.forward(sky - e) - data).abs().pow(2).sum().item()
        fd = (fp - fm) / (2 * h)
        assert abs(fd - grad[i, j]) <= 1e-6 * max(1.0, abs(grad).max()), (i, j, fd, grad[i, j])


def test_chi2_gradient_against_fp64_oracle(gpu, oracle):
    obs, op, cell = setup(gpu, dtype=torch.complex64)
    rng = np.random.default_rng(6)
    sky = blob_sky(128, 12, rng)
    truth = blob_sky(128, 12, rng)
    d = oracle.measure(truth[None, None], obs.uvw, obs.freq, cell)
    x = torch.from_numpy(sky).to(gpu, torch.float32).requires_grad_(True)
    dt = torch.from_numpy(d.reshape(-1, 1)).to(gpu, torch.complex64)
    chi2 = (op.apply(x) - dt).abs().pow(2).sum()
    chi2.backward()
    g = x.grad.double().cpu().numpy()
    r = oracle.measure(sky[None, None], obs.uvw, obs.freq, cell) - d
    gref = 2 * np.real(oracle.measure_adjoint(r, obs.uvw, obs.freq, (128, 128), cell))[0, 0]
    # the adjoint divides by the taper (down to 1.6e-5 at the corners), so
    # the FP32 gradient is stated over the full field and where the taper
    # correction is at most 100x
    _, corr = oracle.create_prolate_spheroidal_kernel(100, 7, np.array([128, 128]))
    inner = corr >= 1e-2
    den = np.max(np.abs(gref))
    err = np.max(np.abs(g - gref)) / den
    err_in = np.max(np.abs(g - gref)[inner]) / den
    print(f"chi2 gradient FP32 vs FP64 oracle: full field {err:.3e}, taper >= 1e-2 {err_in:.3e}")
    assert err_in < 1e-4 and err < 1e-2


def test_full_size_operator(gpu, oracle):
    """The 1024 x 1024 sky (50 Gaussian blobs, sigma 2-10 px) and 1e7 uvw
    samples of the 27-antenna generator: adjoint dot products (FP32 and
    FP64), central finite differences on 32 pixels (FP64, step 1e-3) and the
    FP32 chi-squared gradient against the FP64 oracle."""
    from astroviper_amd import synth
    from astroviper_amd.measurement import MeasurementOperator
    n = 1024
    uvw, freq, cell = synth.measurement_samples(10_000_000, n)
    rng = np.random.default_rng(3)
    sky = synth.blob_sky(n, 50, rng)
    ops = {dt: MeasurementOperator(torch.from_numpy(uvw).to(gpu), torch.from_numpy(freq).to(gpu),
                                   (n, n), cell, n_pol=1, dtype=dt, device=gpu)
           for dt in (torch.complex64, torch.complex128)}
    assert ops[torch.complex64].n_samples == 10_000_000
    # adjoint dot products
    g = torch.Generator(device=gpu)
    g.manual_seed(31)
    for dt, tol in ((torch.complex64, 1e-6), (torch.complex128, 1e-12)):
        op = ops[dt]
        x = torch.randn(op.shape, dtype=dt, device=gpu, generator=g)
        y = torch.randn((op.n_samples, 1), dtype=dt, device=gpu, generator=g)
        ax, ahy = op.forward(x), op.adjoint(y)
        c128 = torch.complex128
        lhs = torch.vdot(y.flatten().to(c128), ax.flatten().to(c128))
        rhs = torch.vdot(ahy.flatten().to(c128), x.flatten().to(c128))
        mis = ((lhs - rhs).abs() / (torch.linalg.norm(ax.to(c128)) *
                                    torch.linalg.norm(y.to(c128)))).item()
        print(f"full-size adjoint mismatch {dt}: {mis:.3e}")
        assert mis < tol
    # finite differences, FP64
    op = ops[torch.complex128]
    xs = torch.from_numpy(sky).to(gpu)
    data = op.forward(xs * 1.1 + 0.01)
    xg = xs.clone().requires_grad_(True)
    (op.apply(xg) - data).abs().pow(2).sum().backward()
    grad = xg.grad.cpu().numpy()
    h = 1e-3
    scale = max(1.0, np.abs(grad).max())
    for (i, j) in rng.integers(64, n - 64, size=(32, 2)):
        e = torch.zeros_like(xs)
        e[i, j] = h
        fp = (op.forward(xs + e) - data).abs().pow(2).sum().item()
        fm = (op.forward(xs - e) - data).abs().pow(2).sum().item()
        assert abs((fp - fm) / (2 * h) - grad[i, j]) <= 1e-6 * scale, (i, j)
    # chi-squared gradient, FP32 operator against the FP64 oracle
    truth = synth.blob_sky(n, 50, rng)
    uvw3 = uvw.reshape(1, -1, 3)
    d = oracle.measure(truth[None, None], uvw3, freq, cell)
    op32 = ops[torch.complex64]
    x32 = torch.from_numpy(sky).to(gpu, torch.float32).requires_grad_(True)
    dt = torch.from_numpy(d.reshape(-1, 1)).to(gpu, torch.complex64)
    (op32.apply(x32) - dt).abs().pow(2).sum().backward()
    gg = x32.grad.double().cpu().numpy()
    r = oracle.measure(sky[None, None], uvw3, freq, cell) - d
    gref = 2 * np.real(oracle.measure_adjoint(r, uvw3, freq, (n, n), cell))[0, 0]
    _, corr = oracle.create_prolate_spheroidal_kernel(100, 7, np.array([n, n]))
    inner = corr >= 1e-2
    den = np.max(np.abs(gref))
    err, err_in = np.max(np.abs(gg - gref)) / den, np.max(np.abs(gg - gref)[inner]) / den
    print(f"full-size chi2 gradient FP32 vs FP64 oracle: full {err:.3e}, taper >= 1e-2 {err_in:.3e}")
    assert err_in < 1e-4 and err < 1e-2
