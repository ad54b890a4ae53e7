"""Reverse-mode derivatives of the differentiable run (dem_diff_forward /
dem_diff_backward, wrapped by dem_mbd_cuda_amd.diff.BedObservables) against
central finite differences in FP64: settling beds that start with small
overlaps, several checkpoint intervals per run, gradients of a weighted sum
of the bed heights and plate loads with respect to each scene's
(mu, restitution, E).  The FD step is 1e-6 relative for restitution and E
(agreement required to 1e-5 relative); the friction derivative is orders of
magnitude smaller over these horizons and FD resolves it only to the spread
of its estimates at steps 1e-6 and 1e-5, which is added to the bound."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def beds(count=3, n=150, seed=3):
    from dem_mbd_cuda_amd import scene

    return scene.settle_scenes(count=count, n=n, box=0.01, seed=seed)


def base_params(sc):
    return np.array([[s.pp["mu"], s.pp["epsN"], s.pp["kN"] / np.mean(s.radius)] for s in sc])


def adjoint(sc, params, w, steps, K):
    from dem_mbd_cuda_amd import diff

    p = torch.tensor(params, requires_grad=True)
    o = diff.BedObservables.apply(p, sc, steps, K, 4)
    (o * torch.from_numpy(w)).sum().backward()
    return o.detach().numpy(), p.grad.numpy()


def fd(sc, params, w, steps, K, k, rel):
    from dem_mbd_cuda_amd import diff

    pp, pm = params.copy(), params.copy()
    pp[:, k] *= 1 + rel
    pm[:, k] *= 1 - rel
    jp = (diff.observables(sc, pp, steps, K) * w).sum(axis=1)
    jm = (diff.observables(sc, pm, steps, K) * w).sum(axis=1)
    return (jp - jm) / (2 * rel * params[:, k])


@pytest.mark.parametrize("steps,K", [(40, 15), (200, 50)])
def test_adjoint_matches_central_differences(hip, steps, K):
    sc = beds()
    params = base_params(sc)
    w = np.random.RandomState(1).normal(size=(len(sc), 2))
    obs, g = adjoint(sc, params, w, steps, K)
    assert np.all(np.isfinite(g))
    for k in (1, 2):  # restitution, E
        f = fd(sc, params, w, steps, K, k, 1e-6)
        assert np.all(np.abs(g[:, k] - f) <= 1e-5 * np.abs(f)), (k, g[:, k], f)
    # mu: FD at two steps; their own spread bounds what FD can resolve
    f1 = fd(sc, params, w, steps, K, 0, 1e-6)
    f2 = fd(sc, params, w, steps, K, 0, 1e-5)
    spread = np.abs(f1 - f2)
    assert np.all(np.abs(g[:, 0] - f1) <= 1e-5 * np.abs(f1) + spread), (g[:, 0], f1, f2)


def test_checkpoint_interval_does_not_change_gradients(hip):
    sc = beds(count=2, n=120, seed=9)
    params = base_params(sc)
    w = np.random.RandomState(2).normal(size=(len(sc), 2))
    o1, g1 = adjoint(sc, params, w, 60, 7)
    o2, g2 = adjoint(sc, params, w, 60, 60)
    assert np.array_equal(o1, o2)
    np.testing.assert_allclose(g1, g2, rtol=1e-11, atol=0.0)


def test_forward_observables_match_plain_run(hip):
    """The differentiable forward run is the ordinary step loop: its final
    heights equal those of host.run_pure on the same batch, bit for bit."""
    from dem_mbd_cuda_amd import diff, host, scene

    sc = beds(count=2, n=100, seed=4)
    params = base_params(sc)
    obs = diff.observables(sc, params, 30, 8)
    batch = scene.SceneBatch(diff.apply_params(sc, params))
    import dem_mbd_cuda_amd as dm

    s = dm.CudaDemSolver()
    s.initialize_batch(batch)
    host.run_pure(s, batch.combined, steps=30)
    z = s.state()["pos"][:, 2]
    s.close()
    for k, part in enumerate(batch.split(z)):
        acc = 0.0
        for v in part:  # the run's own summation order (rows in order)
            acc += float(v)
        assert acc / len(part) == obs[k, 0], (k, acc / len(part), obs[k, 0])
