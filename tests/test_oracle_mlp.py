"""The surrogate network's oracle (oracle/mlp.py) and the film side of the surrogate
(orc_surrogate_dataset / orc_surrogate_apply), pinned on the CPU.

  * The FP64 backpropagation equals central finite differences of the FP64 loss, and
    torch autograd (an independent implementation) on the same network.
  * The kernels' arithmetic (BF16 operands, FP32 accumulation) stays within BF16
    rounding of the exact gradient, and padded parameters get zero gradient.
  * One Adam step equals torch.optim.Adam's.
  * The surrogate step with a negative tolerance is the plain integrator, bit for bit;
    the dataset's rates reproduce the plain integrator's state change.
"""
import ctypes as C

import numpy as np
import pytest

import mlp


def _problem(B=37, seed=3):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(B, 11)) * np.array([400, 300, 300, 300, 0.5, 0.5, 2, 1, 1, 1, 1]) + \
        np.array([1000, 400, 500, 500, 0.2, -0.2, 3, 2, 300, 40, 540])
    T = rng.normal(size=(B, 6)) * np.array([1e20, 1e12, 1e13, 1e13, 1e11, 1e11])
    norm = mlp.make_norm(X, T)
    p = mlp.init_params(seed).astype(np.float64)
    p[mlp.OB1:mlp.OW2] = rng.normal(size=64) * 0.1
    p[mlp.OB2:mlp.OW3] = rng.normal(size=64) * 0.1
    p[mlp.OB3:mlp.OB3 + 6] = rng.normal(size=6) * 0.1
    return X.astype(np.float32), T.astype(np.float32), norm, p


def test_norm_marks_constant_features():
    X, T, norm, _ = _problem()
    X[:, 8:] = X[0, 8:]
    n = mlp.make_norm(X, T)
    assert np.all(n[mlp.N_XIS + 8:mlp.N_XIS + 11] == 0)
    assert np.all(n[mlp.N_XIS:mlp.N_XIS + 8] > 0)


def test_grad64_matches_finite_differences():
    X, T, norm, p = _problem()
    g = mlp.grad64(p, norm, X, T)
    rng = np.random.default_rng(0)
    real = [i for i in range(mlp.NPARAM)
            if not (mlp.OW1 <= i < mlp.OB1 and (i - mlp.OW1) % 16 >= 11)
            and not (mlp.OW3 <= i < mlp.OB3 and (i - mlp.OW3) // 64 >= 6)
            and not (mlp.OB3 + 6 <= i < mlp.NPARAM)]
    idx = rng.choice(real, 300, replace=False)
    for i in idx:
        h = 1e-6 * max(1.0, abs(p[i]))
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (mlp.loss64(pp, norm, X, T) - mlp.loss64(pm, norm, X, T)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(abs(g).max(), 1e-12) + 1e-5 * abs(g[i]), i


def test_grad64_matches_torch_autograd():
    torch = pytest.importorskip("torch")
    X, T, norm, p = _problem(B=53)
    W1, b1, W2, b2, W3, b3 = (torch.tensor(a, dtype=torch.float64, requires_grad=True) for a in mlp.unpack(p))
    n = torch.tensor(norm, dtype=torch.float64)
    xn = torch.zeros((X.shape[0], 16), dtype=torch.float64)
    xn[:, :11] = (torch.tensor(X, dtype=torch.float64) - n[0:11]) * n[16:27]
    h1 = torch.nn.functional.silu(xn @ W1.T + b1)
    h2 = torch.nn.functional.silu(h1 @ W2.T + b2)
    y = (h2 @ W3.T + b3)[:, :6]
    tn = (torch.tensor(T, dtype=torch.float64) - n[32:38]) * n[64:70]
    loss = torch.mean((y - tn) ** 2)
    loss.backward()
    want = mlp.pack(W1.grad.numpy(), b1.grad.numpy(), W2.grad.numpy(), b2.grad.numpy(), W3.grad.numpy(),
                    b3.grad.numpy(), np.float64)
    assert abs(float(loss.detach()) - mlp.loss64(p, norm, X, T)) < 1e-12 * float(loss.detach())
    np.testing.assert_allclose(mlp.grad64(p, norm, X, T), want, rtol=1e-10, atol=1e-14)


def test_kernel_arithmetic_within_bf16_of_exact():
    X, T, norm, p = _problem(B=256)
    p32 = p.astype(np.float32)
    loss, g = mlp.loss_and_grad(p32, norm, X, T)
    g64 = mlp.grad64(p32.astype(np.float64), norm, X, T)
    assert abs(loss - mlp.loss64(p32.astype(np.float64), norm, X, T)) < 2e-2 * loss
    assert np.linalg.norm(g - g64) < 2e-2 * np.linalg.norm(g64)
    W1, b1, W2, b2, W3, b3 = mlp.unpack(g)
    assert not np.any(W1[:, 11:]) and not np.any(W3[6:]) and not np.any(b3[6:])


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    x = np.random.default_rng(1).normal(size=5000).astype(np.float32) * np.float32(1e3)
    x[:3] = [1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -(1.0 + 2 ** -8)]
    want = torch.from_numpy(x.copy()).to(torch.bfloat16).float().numpy()
    assert np.array_equal(mlp.bf16(x), want)


def test_adam_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(5)
    p = rng.normal(size=mlp.NPARAM).astype(np.float32)
    m1 = np.zeros(mlp.NPARAM, np.float32)
    m2 = np.zeros(mlp.NPARAM, np.float32)
    tp = torch.tensor(p.copy(), requires_grad=True)
    opt = torch.optim.Adam([tp], lr=2e-3, betas=(0.9, 0.999), eps=1e-8)
    for step in range(1, 4):
        g = rng.normal(size=mlp.NPARAM).astype(np.float32)
        p, m1, m2 = mlp.adam(p, g, m1, m2, step, 2e-3)
        tp.grad = torch.tensor(g)
        opt.step()
    np.testing.assert_allclose(p, tp.detach().numpy(), rtol=1e-5, atol=1e-7)


def _film_state(seed, n=257, dtype="f64"):
    import driver
    from mphip import configs
    spec, _ = configs.coupled_pump(nx=24, ny=20, nz=32, steps=10, fluence_mj_cm2=30.0, npml=4, src_z=20e-9,
                                   n_planes=0, enable_mo=False, dtype=dtype)
    st = driver.OracleState(spec)
    rng = np.random.default_rng(seed)
    N = spec.n_film
    CT = spec.compute_dtype
    st.fa["Te"][:] = rng.uniform(300, 2500, N).astype(CT)
    st.fa["Tl"][:] = rng.uniform(300, 700, N).astype(CT)
    st.fa["Ts_TM"][:] = rng.uniform(300, 900, N).astype(CT)
    st.fa["Ts_RE"][:] = rng.uniform(300, 900, N).astype(CT)
    st.fa["m_TM_x"][:] = rng.uniform(-0.9, 0.9, N).astype(CT)
    st.fa["m_RE_x"][:] = rng.uniform(-0.9, 0.9, N).astype(CT)
    st.fa["window_sum"][:] = rng.uniform(0, 1e22, N) * spec.dt
    return st


@pytest.mark.parametrize("dtype", ["f32", "f64"])
def test_surrogate_apply_negative_tol_is_plain(dtype):
    a = _film_state(1, dtype=dtype)
    b = _film_state(1, dtype=dtype)
    L = a.L
    rates = np.random.default_rng(2).normal(size=(a.spec.n_film, 6)).astype(np.float32)
    fb = C.c_int64(0)
    assert L.orc_surrogate_apply(C.byref(a.film), a.gd_arr.ctypes.data, C.byref(a.lut), rates.ctypes.data,
                                 a.dt, 4, a.spec.brillouin_iters, 1, -1.0, C.byref(fb)) == 0
    assert fb.value == a.spec.n_film
    assert L.orc_multiphysics_step(C.byref(b.film), b.gd_arr.ctypes.data, C.byref(b.lut), b.dt, 4,
                                   b.spec.brillouin_iters, 1) == 0
    fa, fbs = a.film_state(), b.film_state()
    for k in a.FILM_NAMES + ("window_sum",):
        assert np.array_equal(fa[k], fbs[k]), k


def test_dataset_rates_reproduce_plain_step():
    st = _film_state(4)
    ref = _film_state(4)
    N = st.spec.n_film
    X = np.zeros((N, 11), np.float32)
    T = np.zeros((N, 6), np.float32)
    L = st.L
    assert L.orc_surrogate_dataset(C.byref(st.film), st.gd_arr.ctypes.data, C.byref(st.lut), st.dt, 4,
                                   st.spec.brillouin_iters, 1, 1, X.ctypes.data, T.ctypes.data, N) == 0
    pabs = ref.fa["window_sum"] / ref.dt
    before = {k: ref.fa[k].copy() for k in ("Te", "Tl", "Ts_TM", "Ts_RE", "m_TM_x", "m_RE_x")}
    assert L.orc_multiphysics_step(C.byref(ref.film), ref.gd_arr.ctypes.data, C.byref(ref.lut), ref.dt, 4,
                                   ref.spec.brillouin_iters, 1) == 0
    after = ref.film_state()
    assert np.allclose(X[:, 0], before["Te"], rtol=1e-7) and np.allclose(X[:, 4], before["m_TM_x"], rtol=1e-7)
    assert np.allclose(X[:, 6], np.arcsinh(pabs * 1e-18), rtol=1e-6)
    gam = st.spec.gd["gamma_e"]
    ue1 = 0.5 * gam * before["Te"] ** 2 + st.dt * (T[:, 0].astype(np.float64) + pabs)
    np.testing.assert_allclose(np.sqrt(2 * ue1 / gam), after["Te"], rtol=1e-6)
    for c, k in ((1, "Tl"), (2, "Ts_TM"), (3, "Ts_RE"), (4, "m_TM_x"), (5, "m_RE_x")):
        np.testing.assert_allclose(before[k] + st.dt * T[:, c].astype(np.float64), after[k], rtol=1e-6, atol=1e-9)
