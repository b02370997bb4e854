"""GPU parity of the batched form and the least-squares driver against the oracle."""
import numpy as np
import pytest

import oracle as O
import rnp2_amd as R
from metrics import rel_err_colscaled

pytestmark = pytest.mark.gpu


def c3_shapes(count, seed):
    """The batched configuration's distribution: m log-uniform in [16, 256],
    n uniform in [8, m]."""
    rng = np.random.default_rng(seed)
    m = np.rint(np.exp(rng.uniform(np.log(16), np.log(256), count))).astype(int)
    n = np.array([rng.integers(8, mm + 1) for mm in m])
    return list(zip(m.tolist(), n.tolist()))


def _pack(shapes, cplx, seed, nrhs):
    import torch
    desc, na, nt, nb_ = R.make_batch(shapes, nrhs)
    rng = np.random.default_rng(seed)
    dt = np.complex128 if cplx else np.float64
    a = rng.standard_normal(na)
    b = rng.standard_normal(max(nb_, 1))
    if cplx:
        a = a + 1j * rng.standard_normal(na)
        b = b + 1j * rng.standard_normal(max(nb_, 1))
    tdt = torch.complex128 if cplx else torch.float64
    A = torch.from_numpy(a.astype(dt)).cuda()
    B = torch.from_numpy(b.astype(dt)).cuda()
    tau = torch.zeros(max(nt, 1), dtype=tdt, device="cuda")
    return desc, a.astype(dt), b.astype(dt), A, tau, B


@pytest.mark.parametrize("cplx", [False, True])
def test_batched_factor_and_ls(cplx):
    shapes = c3_shapes(60, 7) + [(256, 256), (256, 200), (200, 180), (17, 8), (8, 8), (1, 1), (5, 9)]
    print(f"[case] batched {len(shapes)} problems cplx={cplx}", flush=True)
    desc, a, b, A, tau, B = _pack(shapes, cplx, 3, nrhs=3)
    h = R.Handle()
    h.geqrf_batched(desc, A, tau, B)
    h.synchronize()
    Ag, tg, Bg = A.cpu().numpy(), tau.cpu().numpy(), B.cpu().numpy()
    worst = {"R": 0.0, "tau": 0.0, "x": 0.0}
    for d in desc:
        m, n, r = int(d["m"]), int(d["n"]), int(d["nrhs"])
        k = min(m, n)
        a0 = a[d["a_off"]:d["a_off"] + m * n].reshape(n, m).T
        af_o, tau_o = O.factor(a0)
        af_g = Ag[d["a_off"]:d["a_off"] + m * n].reshape(n, m).T
        worst["R"] = max(worst["R"], rel_err_colscaled(np.triu(af_g[:k]), np.triu(af_o[:k])))
        worst["tau"] = max(worst["tau"], float(np.max(np.abs(tg[d["tau_off"]:d["tau_off"] + k] - tau_o))) if k else 0)
        if r:
            b0 = b[d["b_off"]:d["b_off"] + m * r].reshape(r, m).T
            x_o = O.ls_solve(a0, b0)
            x_g = Bg[d["b_off"]:d["b_off"] + m * r].reshape(r, m).T[:n]
            worst["x"] = max(worst["x"], float(np.max(np.abs(x_g - x_o)) / max(np.max(np.abs(x_o)), 1e-300)))
    print(f"        {worst}", flush=True)
    assert worst["R"] < 1e-11 and worst["tau"] < 1e-11 and worst["x"] < 1e-9


def _edge_small(m, n, cplx, seed):
    """Zero, single-entry, underflowing, subnormal, overflowing and tiny columns."""
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((m, n))
    if cplx:
        a = a + 1j * rng.standard_normal((m, n))
    a[:, 1] = 0.0
    a[:, 2] = 0.0
    a[2, 2] = 2.5j if cplx else 2.5
    a[:, 3] *= 1e-300
    a[:, 4] *= 1e-310
    a[:, 5] *= 1e300
    a[:, 6] *= 1e-160
    return a


@pytest.mark.parametrize("cplx", [False, True])
def test_batched_edge_columns(cplx):
    """The scaled-norm and rescale paths of the reflector in both tiers (LDS
    problems and in-place 256-thread problems): R and tau elementwise."""
    import torch
    shapes = [(40, 20), (120, 64), (256, 200), (16, 8)]
    print(f"[case] batched edge columns cplx={cplx}", flush=True)
    desc, na, nt, _ = R.make_batch(shapes, 0)
    dt = np.complex128 if cplx else np.float64
    a = np.zeros(na, dtype=dt)
    mats = []
    for j, d in enumerate(desc):
        m, n = int(d["m"]), int(d["n"])
        x = _edge_small(m, n, cplx, 20 + j)
        mats.append(x)
        a[d["a_off"]:d["a_off"] + m * n] = x.T.reshape(-1)
    A = torch.from_numpy(a.copy()).cuda()
    tau = torch.zeros(max(nt, 1), dtype=torch.complex128 if cplx else torch.float64, device="cuda")
    h = R.Handle()
    h.geqrf_batched(desc, A, tau, None)
    h.synchronize()
    Ag, tg = A.cpu().numpy(), tau.cpu().numpy()
    for d, x in zip(desc, mats):
        m, n = x.shape
        k = min(m, n)
        af_o, tau_o = O.factor(x)
        af_g = Ag[d["a_off"]:d["a_off"] + m * n].reshape(n, m).T
        e_r = rel_err_colscaled(np.triu(af_g[:k]), np.triu(af_o[:k]))
        e_t = float(np.max(np.abs(tg[d["tau_off"]:d["tau_off"] + k] - tau_o)))
        print(f"        {(m, n)}: R {e_r:.3g} tau {e_t:.3g}", flush=True)
        assert e_r < 1e-11 and e_t < 1e-11, (m, n, e_r, e_t)


def test_batched_alone_vs_mix():
    """Each problem's result is bitwise the same alone and inside a batch."""
    import torch
    shapes = c3_shapes(40, 11)
    print("[case] batched determinism", flush=True)
    desc, a, b, A, tau, B = _pack(shapes, False, 5, nrhs=2)
    h = R.Handle()
    h.geqrf_batched(desc, A, tau, B)
    h.synchronize()
    for i in (0, 7, 23):
        d = desc[i:i + 1].copy()
        m, n, r = int(d["m"][0]), int(d["n"][0]), int(d["nrhs"][0])
        A1 = torch.from_numpy(a[d["a_off"][0]:d["a_off"][0] + m * n].copy()).cuda()
        B1 = torch.from_numpy(b[d["b_off"][0]:d["b_off"][0] + m * r].copy()).cuda()
        t1 = torch.zeros(min(m, n), dtype=torch.float64, device="cuda")
        d["a_off"], d["tau_off"], d["b_off"] = 0, 0, 0
        h.geqrf_batched(d, A1, t1, B1)
        h.synchronize()
        assert torch.equal(A1, A[desc["a_off"][i]:desc["a_off"][i] + m * n])
        assert torch.equal(B1, B[desc["b_off"][i]:desc["b_off"][i] + m * r])


def test_batched_plan_replays():
    """The hipGraph plan gives the eager result on every replay."""
    import torch
    shapes = c3_shapes(50, 13) + [(256, 220)]
    print("[case] batched plan replays", flush=True)
    desc, a, b, A, tau, B = _pack(shapes, False, 9, nrhs=1)
    A0, B0 = A.clone(), B.clone()
    h = R.Handle()
    h.geqrf_batched(desc, A, tau, B)
    h.synchronize()
    ref_A, ref_B, ref_t = A.clone(), B.clone(), tau.clone()
    plan = h.batched_plan(desc, A, tau, B)
    for rep in range(3):
        A.copy_(A0)
        B.copy_(B0)
        tau.zero_()
        torch.cuda.synchronize()
        plan.run()
        h.synchronize()
        assert torch.equal(A, ref_A) and torch.equal(B, ref_B) and torch.equal(tau, ref_t), rep
    plan.close()


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("m,n,nrhs", [(300, 200, 3), (2000, 700, 16), (513, 513, 1)])
def test_gels_dense(m, n, nrhs, cplx):
    import torch
    print(f"[case] gels {m}x{n} nrhs={nrhs} cplx={cplx}", flush=True)
    rng = np.random.default_rng(m + n)
    dt = np.complex128 if cplx else np.float64
    a = rng.standard_normal((m, n)).astype(dt)
    b = rng.standard_normal((m, nrhs)).astype(dt)
    if cplx:
        a += 1j * rng.standard_normal((m, n))
        b += 1j * rng.standard_normal((m, nrhs))
    x_o = O.ls_solve(a, b)
    tdt = torch.complex128 if cplx else torch.float64
    A = R.colmajor_empty(m, n, tdt)
    A.copy_(torch.from_numpy(a))
    B = R.colmajor_empty(m, nrhs, tdt)
    B.copy_(torch.from_numpy(b))
    tau = torch.zeros(n, dtype=tdt, device="cuda")
    torch.cuda.synchronize()  # the copies run on torch's stream, gels on the handle's
    h = R.Handle()
    h.gels(A, tau, B)
    h.synchronize()
    x_g = B.cpu().numpy()[:n]
    err = float(np.max(np.abs(x_g - x_o)) / np.max(np.abs(x_o)))
    print(f"        err={err:.2e}", flush=True)
    assert err < 1e-10
