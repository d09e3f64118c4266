"""The training gradient without a workspace (d3x_sur_train_grad_noalloc, static device scratch)
against d3x_sur_train_grad with a caller-allocated workspace: two calls in a row, the second with
fewer rows than the first (the scratch keeps the first call's fragments beyond the second's
rows), and a batch above the static scratch's rows rejected before any launch."""

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _vp(t):
    return C.c_void_p(t.data_ptr())


def _params(seed):
    import surrogate_oracle as SO
    rng = np.random.default_rng(seed)
    p = SO.init_params(seed, mean=rng.normal(0.5, 0.2, SO.N_FEAT), istd=rng.uniform(0.5, 2, SO.N_FEAT),
                       scale=[0.01, 0.02, 0.03], shift=[0.001, -0.002, 0.003])
    return p.astype(np.float32)


def test_noalloc_train_grad_matches_workspace_entry():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    torch.zeros(1, device="cuda")  # the device is brought up by torch first
    import d3_lib
    lib = d3_lib.load()
    dev = lambda a, dt=torch.float32: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device="cuda")  # noqa: E731
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    P = dev(_params(11))
    pk = torch.empty(lib.d3x_sur_packed_bytes(), dtype=torch.uint8, device="cuda")
    assert lib.d3x_sur_pack(_vp(P), _vp(pk), stream) == 0
    rng = np.random.default_rng(12)
    nmax = 16384
    for call, (c, n) in enumerate(((1, nmax), (2, 3001))):
        print(f"case no-alloc train grad call={call} class={c} rows={n}", flush=True)
        X = dev(rng.normal(0.5, 0.6, (n, 64)).astype(np.float32))
        y = dev(rng.normal(0.0, 0.01, n).astype(np.float32))
        w = dev((rng.uniform(0.5, 1.5, n) / n).astype(np.float32))
        out = []
        for noalloc in (False, True):
            E = torch.zeros(n, device="cuda")
            grad = torch.zeros(d3_lib.SUR_NPARAMS, device="cuda")
            loss = torch.zeros(1, device="cuda")
            if noalloc:
                rc = lib.d3x_sur_train_grad_noalloc(_vp(P), _vp(pk), c, n, _vp(X), _vp(y), _vp(w), _vp(E), _vp(grad),
                                                    _vp(loss), stream)
            else:
                work = torch.empty(lib.d3x_sur_train_bytes(n), dtype=torch.uint8, device="cuda")
                rc = lib.d3x_sur_train_grad(_vp(P), _vp(pk), c, n, _vp(X), _vp(y), _vp(w), _vp(E), _vp(work),
                                            _vp(grad), _vp(loss), stream)
            assert rc == 0, d3_lib.last_error()
            torch.cuda.synchronize()
            out.append((E.cpu().numpy(), grad.cpu().numpy().astype(np.float64), float(loss.item())))
        (E1, g1, l1), (E2, g2, l2) = out
        # per-row energies come from the same kernel on the same rows: equal bit for bit; the
        # gradient and the loss are sums of float atomics whose order varies
        assert np.array_equal(E1, E2)
        assert abs(l2 - l1) <= 1e-5 * abs(l1)
        assert np.max(np.abs(g2 - g1)) <= 1e-5 * np.max(np.abs(g1))
        assert np.max(np.abs(g1)) > 0
    # one row more than the static scratch holds: refused, nothing launched
    n = nmax + 1
    E = torch.zeros(n, device="cuda")
    grad = torch.zeros(d3_lib.SUR_NPARAMS, device="cuda")
    loss = torch.zeros(1, device="cuda")
    X = torch.zeros((n, 64), device="cuda")
    rc = lib.d3x_sur_train_grad_noalloc(_vp(P), _vp(pk), 0, n, _vp(X), _vp(E), _vp(E), _vp(E), _vp(grad), _vp(loss),
                                        stream)
    assert rc & d3_lib.D3X_ERR_INPUT and "16384" in d3_lib.last_error()
    torch.cuda.synchronize()
    assert float(loss.item()) == 0.0 and not torch.any(grad != 0)
