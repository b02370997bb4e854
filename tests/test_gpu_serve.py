"""GPU tests of the least-squares request server: every result checked
against the FP64 solution, the same bits alone and in a concurrent mix, and
HBM-budget rejection."""
import numpy as np
import pytest

import rnp2_amd as R

pytestmark = pytest.mark.gpu


def conditioned(m, n, cond, rng):
    u, _ = np.linalg.qr(rng.standard_normal((m, n)))
    v, _ = np.linalg.qr(rng.standard_normal((n, n)))
    s = np.logspace(0, -np.log10(cond), n)
    return (u * s) @ v.T


def eta(a, b, x):
    r = b - a @ x
    an = np.linalg.norm(a)
    return float(np.max(np.linalg.norm(a.T @ r, axis=0) / (an * (an * np.linalg.norm(x, axis=0) + np.linalg.norm(b, axis=0)))))


def make_requests(count, seed, mmax=1200):
    rng = np.random.default_rng(seed)
    reqs = []
    for _ in range(count):
        m = int(np.rint(np.exp(rng.uniform(np.log(64), np.log(mmax)))))
        n = int(rng.integers(max(1, m // 8), max(2, m // 2) + 1))
        q = int(rng.integers(1, 17))
        cond = float(np.exp(rng.uniform(np.log(1e2), np.log(1e8))))
        a = conditioned(m, n, cond, rng)
        b = a @ rng.standard_normal((n, q)) + 1e-8 * rng.standard_normal((m, q))
        reqs.append((a, b, cond))
    return reqs


@pytest.mark.parametrize("mode", [0, 1])
def test_server_mix_matches_alone(mode):
    reqs = make_requests(24, 3 + mode)
    print(f"[case] server mode={mode} requests={len(reqs)} sizes={[r[0].shape for r in reqs[:6]]}...", flush=True)
    srv = R.Server(workers=4, mode=mode)
    ids = [srv.submit(a, b)[1] for a, b, _ in reqs]
    mix = [srv.poll(i) for i in ids]
    paths = {}
    for (a, b, cond), (x, info) in zip(reqs, mix):
        assert info.status == 0
        paths[R.PATHS[info.path]] = paths.get(R.PATHS[info.path], 0) + 1
        e = eta(a, b, x)
        assert e <= 1e-13, (a.shape, cond, e, R.PATHS[info.path])
    print(f"        paths={paths} stats={srv.stats()}", flush=True)
    srv.close()
    # alone: one request at a time on a fresh server
    srv1 = R.Server(workers=1, mode=mode)
    for k in (0, 5, 11, 17):
        a, b, _ = reqs[k]
        st, i = srv1.submit(a, b)
        x1, _ = srv1.poll(i)
        assert np.array_equal(x1, mix[k][0]), k
    srv1.close()


def test_server_budget_rejects_oversize():
    srv = R.Server(hbm_budget=1 << 20, workers=2, mode=0)
    a = np.random.default_rng(0).standard_normal((2000, 500))
    b = np.ones((2000, 1))
    st, rid = srv.submit(a, b)
    assert st == R.RNPX_ERR_NOMEM
    x, info = srv.poll(rid)
    assert info.status == R.RNPX_ERR_NOMEM
    small = np.random.default_rng(1).standard_normal((100, 20))
    st, rid = srv.submit(small, np.ones((100, 1)))
    assert st == 0
    x, info = srv.poll(rid)
    assert info.status == 0 and np.allclose(x[:, 0], np.linalg.lstsq(small, np.ones(100), rcond=None)[0])
    assert srv.stats()["rejected"] == 1
    srv.close()
