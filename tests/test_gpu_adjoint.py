"""GPU gray bounce-back and discrete adjoint against the FP64 oracle adjoint
and central finite differences."""
import numpy as np
import pytest

from oracle.oracle import OracleLBM
from windtunnel_amd.lbm import LBM

pytestmark = pytest.mark.gpu


def gray_field(dims, c, r_in, r_out):
    nx, ny, nz = dims
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    r = np.sqrt((x + 0.5 - c[0]) ** 2 + (y + 0.5 - c[1]) ** 2 + (z + 0.5 - c[2]) ** 2)
    return np.clip((r_out - r) / (r_out - r_in), 0, 1).reshape(-1).astype(np.float32)


def setup(L, dims, w, les, trt, tau=0.6, u0=0.05):
    L.set_params(tau=tau, les=les, trt=trt, csm=0.14, u0=u0, **{"lambda": 3 / 16})
    L.reset()
    solid = (w >= 1).reshape(dims[::-1])
    idx = np.argwhere(solid)
    lo, hi = idx.min(0), idx.max(0)
    L.voxelize_mask([lo[2], lo[1], lo[0], hi[2], hi[1], hi[0]],
                    solid[lo[0]:hi[0] + 1, lo[1]:hi[1] + 1, lo[2]:hi[2] + 1].astype(np.uint8).reshape(-1))
    L.set_gray(np.where(w >= 1, 0, w).astype(np.float32))
    return L


DIMS = (20, 12, 12)
W = gray_field(DIMS, (7, 6, 6), 2.0, 3.6)


def fbar(N):
    F = np.zeros((N, 3))
    F[-10:, 0] = 0.1
    F[-10:, 1] = 0.03
    return F


@pytest.mark.parametrize("les,trt", [(False, False), (False, True), (True, True)])
def test_fp64_adjoint_equals_oracle(les, trt):
    N = 40
    print(f"[adjoint f64] les={les} trt={trt}", flush=True)
    O = setup(OracleLBM(*DIMS, prec=8), DIMS, W, les, trt)
    fo, go = O.adjoint(N, fbar(N))
    for interval in (0, 1, 7):
        G = setup(LBM(*DIMS, storage="f64", arith="f64"), DIMS, W, les, trt)
        fg, gg, _ = G.adjoint(N, fbar(N), interval=interval)
        np.testing.assert_allclose(fg, fo, rtol=1e-11, atol=1e-15)
        assert np.max(np.abs(gg - go)) <= 1e-10 * np.max(np.abs(go))


def test_gray_forward_exact_mode_equals_oracle():
    O = setup(OracleLBM(*DIMS), DIMS, W, True, True)
    G = setup(LBM(*DIMS, arith="f64"), DIMS, W, True, True)
    fo, fg = O.step(30), G.step(30)
    np.testing.assert_allclose(fg, fo, rtol=1e-11, atol=1e-15)
    assert np.array_equal(G.f, O.array("f"))


def test_fp32_adjoint_tracks_fp64():
    N = 60
    O = setup(OracleLBM(*DIMS, prec=8), DIMS, W, True, True)
    _, go = O.adjoint(N, fbar(N))
    G = setup(LBM(*DIMS), DIMS, W, True, True)
    _, gg, _ = G.adjoint(N, fbar(N))
    gray = (W > 0) & (W < 1)
    assert np.linalg.norm(gg[gray] - go[gray]) / np.linalg.norm(go[gray]) < 1e-3


def test_adjoint_against_central_differences():
    """Central differences (eps = 1e-3) of the FP64 GPU forward on seeded gray cells."""
    dims = (48, 24, 24)
    w = gray_field(dims, (16, 12, 12), 3.0, 6.0)
    N = 150
    F = np.zeros((N, 3))
    F[-50:, 0] = 1.0 / 50
    G = setup(LBM(*dims, storage="f64", arith="f64"), dims, w, False, True, tau=0.6, u0=0.04)
    _, g, _ = G.adjoint(N, F)
    gray = np.flatnonzero((w > 2e-3) & (w < 1 - 2e-3))
    cells = np.random.default_rng(4).choice(gray, 8, replace=False)
    eps = 1e-3
    for c in cells:
        J, wv = [], []
        for s in (1, -1):
            wp = w.copy()
            wp[c] = np.float32(w[c] + s * eps)
            wv.append(float(wp[c]))
            L = setup(LBM(*dims, storage="f64", arith="f64"), dims, wp, False, True, tau=0.6, u0=0.04)
            J.append(float(np.sum(L.step(N) * F)))
            L.close()
        fd = (J[0] - J[1]) / (wv[0] - wv[1])
        assert abs(g[c] - fd) <= 1e-3 * abs(fd) + 1e-6 * np.abs(g).max(), (c, g[c], fd)


def test_forward_backward_pair_and_autograd():
    """wt_adjoint_forward / wt_adjoint_backward equal the fused wt_adjoint,
    backward repeats and is linear in Fbar, leaves the lattice's state
    alone, refuses a changed configuration; the torch.autograd.Function
    gives the same gradient."""
    import torch

    from windtunnel_amd.autograd import gray_forces
    from windtunnel_amd.lbm import WindtunnelError

    N = 40
    Fb = fbar(N)
    A = setup(LBM(*DIMS), DIMS, W, True, True)
    fa, ga, _ = A.adjoint(N, Fb, interval=6)
    B = setup(LBM(*DIMS), DIMS, W, True, True)
    fb = B.adjoint_forward(N, interval=6)
    f_state = B.f.copy()
    g1, _ = B.adjoint_backward(Fb)
    g2, _ = B.adjoint_backward(2 * Fb)
    assert np.array_equal(fa, fb)
    assert np.array_equal(ga, g1)
    np.testing.assert_allclose(g2, 2 * g1, rtol=1e-6, atol=1e-12 * np.abs(g1).max())
    assert np.array_equal(B.f, f_state) and B.steps == N
    with pytest.raises(ValueError):
        B.adjoint_backward(Fb[:-1])
    B.tau = 0.61
    with pytest.raises(WindtunnelError) as e:
        B.adjoint_backward(Fb)
    assert e.value.code == -4
    B.adjoint_release()

    C = setup(LBM(*DIMS), DIMS, W, True, True)
    w = torch.tensor(np.where(W >= 1, 0, W), dtype=torch.float32, requires_grad=True)
    F = gray_forces(w, C, N, 6)
    J = (F * torch.from_numpy(Fb)).sum()
    J.backward()
    np.testing.assert_allclose(F.detach().numpy(), fa, rtol=0, atol=0)
    np.testing.assert_allclose(w.grad.numpy(), ga.astype(np.float32), rtol=1e-6, atol=0)


DIMS_B = (128, 16, 16)  # nx a multiple of 128: blocked storage for every precision
W_B = gray_field(DIMS_B, (40, 8, 8), 2.5, 4.5)


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_windowed_reverse_sweep_is_bit_identical(storage, monkeypatch):
    """A replay store shorter than the checkpoint segment (WT_ADJ_WINDOW)
    recomputes more but must give the same gradient bit for bit; so must the
    sweep that replays the next segment on a second stream (the default
    with room for two stores) against one stream (WT_ADJ_OVERLAP=0)."""
    N = 47
    Fb = fbar(N)
    print(f"[adjoint window] storage={storage}", flush=True)
    L = setup(LBM(*DIMS_B, storage=storage), DIMS_B, W_B, True, True)
    L.adjoint_forward(N, interval=10)
    g_full, _ = L.adjoint_backward(Fb)
    monkeypatch.setenv("WT_ADJ_OVERLAP", "0")
    g_one, _ = L.adjoint_backward(Fb)
    monkeypatch.delenv("WT_ADJ_OVERLAP")
    assert np.array_equal(g_one, g_full)
    for cap in (3, 1):
        monkeypatch.setenv("WT_ADJ_WINDOW", str(cap))
        g_win, _ = L.adjoint_backward(Fb)
        assert np.array_equal(g_win, g_full), cap
    monkeypatch.delenv("WT_ADJ_WINDOW")
    assert np.abs(g_full).max() > 0
    L.adjoint_release()


def test_fp16_adjoint_tracks_fp64(monkeypatch):
    """binary16 states against the FP64 oracle adjoint (LES + TRT, blocked
    storage), with the adjoint fields in binary16 (scaled per step, the
    default) and in float (WT_ADJ_FIELD=f32). Measured rel-L2 errors of the
    gradient over the gray cells: float32 storage 1e-5, binary16 storage
    about 1e-3 with either field type."""
    N = 60
    O = setup(OracleLBM(*DIMS_B, prec=8), DIMS_B, W_B, True, True)
    _, go = O.adjoint(N, fbar(N))
    gray = (W_B > 0) & (W_B < 1)
    errs = {}
    for name, storage, field in (("f32", "f32", None), ("f16", "f16", None), ("f16/f32 fields", "f16", "f32")):
        if field:
            monkeypatch.setenv("WT_ADJ_FIELD", field)
        G = setup(LBM(*DIMS_B, storage=storage), DIMS_B, W_B, True, True)
        _, gg, _ = G.adjoint(N, fbar(N))
        errs[name] = np.linalg.norm(gg[gray] - go[gray]) / np.linalg.norm(go[gray])
        G.close()
        monkeypatch.delenv("WT_ADJ_FIELD", raising=False)
    print(f"[adjoint f16] rel-L2 vs FP64: {errs}", flush=True)
    assert errs["f32"] < 1e-3
    assert errs["f16"] < 3e-3
    assert errs["f16/f32 fields"] < 3e-3


def test_binary16_fields_over_a_long_sweep(monkeypatch):
    """400 steps on binary16 states: the gradient with binary16 adjoint
    fields against the one with float fields on the same recorded states
    (the difference is the fields' rounding alone), the per-step scale
    following the field as it grows from zero."""
    N = 400
    F = np.zeros((N, 3))
    F[N // 2:, 0] = 1.0 / (N // 2)
    L = setup(LBM(*DIMS_B, storage="f16"), DIMS_B, W_B, True, True)
    L.adjoint_forward(N)
    g16, _ = L.adjoint_backward(F)
    monkeypatch.setenv("WT_ADJ_FIELD", "f32")
    g32, _ = L.adjoint_backward(F)
    monkeypatch.delenv("WT_ADJ_FIELD")
    L.adjoint_release()
    L.close()
    gray = (W_B > 0) & (W_B < 1)
    rel = np.linalg.norm(g16[gray] - g32[gray]) / np.linalg.norm(g32[gray])
    print(f"[adjoint f16 fields] 400 steps, rel-L2 against float fields: {rel:.3e}", flush=True)
    assert np.all(np.isfinite(g16)) and np.abs(g32[gray]).max() > 0
    assert rel < 5e-3


def test_second_graph_on_one_lattice_refuses_stale_backward():
    """Two gray_forces calls on one lattice before backward: the lattice holds
    only the second recorded pass, so the first graph's backward raises
    instead of returning the second graph's gradient; the second works."""
    import torch

    from windtunnel_amd.autograd import gray_forces

    N = 12
    C = setup(LBM(*DIMS), DIMS, W, True, True)
    w = torch.tensor(np.where(W >= 1, 0, W), dtype=torch.float32, requires_grad=True)
    J1 = (gray_forces(w, C, N)[:, 0]).sum()
    r1 = C.adjoint_record()
    J2 = (gray_forces(w, C, N)[:, 0]).sum()
    assert C.adjoint_record() == r1 + 1
    with pytest.raises(RuntimeError, match="replaced"):
        J1.backward()
    J2.backward()
    assert torch.isfinite(w.grad).all() and w.grad.abs().max() > 0
    C.close()


def test_weight_field_is_freed_with_the_lattice():
    """Creating, gray-weighting and destroying lattices in a loop does not
    leak device memory (the weight field goes with the lattice)."""
    from windtunnel_amd.lbm import device_memory

    dims = (128, 64, 64)
    w = np.full(dims[0] * dims[1] * dims[2], 0.2, np.float32)

    def cycle():
        L = LBM(*dims)
        L.reset()
        L.set_gray(w)
        L.close()

    cycle()
    free0 = device_memory()[0]
    for _ in range(4):
        cycle()
    assert device_memory()[0] >= free0 - (1 << 20)


def test_slab_count_limit_is_an_error():
    """More slabs than the device force buffer holds per step is refused
    before any launch."""
    L = LBM(3, 3, 4200)
    L.reset()
    with pytest.raises(ValueError, match="above the limit"):
        L.step_slabs(1, 4197)
    f = L.step_slabs(1, 4096)
    assert f.shape[0] == 1 and np.all(np.isfinite(f))
    L.close()


def test_central_differences_at_the_differentiation_workload_size():
    """The differentiation workload's lattice (192x96x96 TRT, tau 0.6,
    u0 0.04, the seeded gray-walled ellipsoid) over 600 recorded steps,
    J = mean C_d over the last 100: the binary64 adjoint against central
    differences at eps = 1e-3 and their Richardson extrapolation from eps
    and eps/2, on 4 seeded gray cells. The extrapolation removes the
    differences' own O(eps^2) error, so it must match to 1e-5."""
    from bench import gray_ellipsoid
    dims, N, K, eps = (192, 96, 96), 600, 100, 1e-3
    w = gray_ellipsoid(dims)
    solid = (w >= 1).reshape(dims[::-1])
    idx = np.argwhere(solid)
    lo, hi = idx.min(0), idx.max(0)
    box = [int(lo[2]), int(lo[1]), int(lo[0]), int(hi[2]), int(hi[1]), int(hi[0])]
    mask = solid[lo[0]:hi[0] + 1, lo[1]:hi[1] + 1, lo[2]:hi[2] + 1].astype(np.uint8).reshape(-1)
    wg = np.where(w >= 1, 0, w).astype(np.float32)
    D = LBM(*dims, storage="f64", arith="f64")
    D.set_params(tau=0.6, les=False, trt=True, u0=0.04, **{"lambda": 3 / 16})
    D.reset()
    D.voxelize_mask(box, mask)
    q = 0.5 * 0.04 ** 2 * D.frontalArea
    Fbar = np.zeros((N, 3))
    Fbar[N - K:, 0] = 1.0 / (K * q)
    D.set_gray(wg)
    D.adjoint_forward(N)
    g, _ = D.adjoint_backward(Fbar)
    D.adjoint_release()
    ok = np.flatnonzero((wg > 2 * eps) & (wg < 1 - 2 * eps))
    cells = np.random.default_rng(5).choice(ok, 4, replace=False)

    def central(c, e):
        J, wv = [], []
        for s in (1, -1):
            wp = wg.copy()
            wp[c] = np.float32(wg[c] + s * e)
            wv.append(float(wp[c]))
            D.reset()
            D.set_gray(wp)
            J.append(float(np.sum(D.step(N) * Fbar)))
        return (J[0] - J[1]) / (wv[0] - wv[1]), 0.5 * (wv[0] - wv[1])

    scale = np.abs(g[ok]).max()
    for c in cells:
        d1, h1 = central(c, eps)
        d2, h2 = central(c, eps / 2)
        rich = (h1 * h1 * d2 - h2 * h2 * d1) / (h1 * h1 - h2 * h2)
        print(f"[adjoint fd 192] cell {c}: adjoint {g[c]:.9e} fd {d1:.9e} richardson {rich:.9e}", flush=True)
        assert abs(g[c] - d1) <= 1e-3 * abs(d1) + 1e-4 * scale, (c, g[c], d1)
        assert abs(g[c] - rich) <= 1e-5 * abs(rich) + 1e-7 * scale, (c, g[c], rich)
    D.close()
