"""The differentiable batch at its full size: 64 settling beds of 5,000
spheres, checkpoints every 100 steps, gradients of a weighted sum of each
scene's bed height and plate load with respect to (mu, restitution, E),
against central finite differences in FP64.

The observables of a bed of thousands of spheres are piecewise smooth in the
parameters: a contact that starts one step earlier (the dashpot force
appears at once), or a contact that switches between sticking and sliding,
makes a small jump or kink.  The adjoint differentiates the smooth piece.
A central difference at relative step h equals that derivative (to O(h^2)
and rounding) only when no such event lies within +-h of the parameter, and
the events grow denser with the number of steps:

* over 200 steps (two checkpoint segments), a scene whose differences at
  h = 1e-6 and 1e-7 (1e-5 and 1e-6 for friction, whose derivative is small
  this early) agree to 1e-6 has no event in the interval, and there the
  adjoint must match the h = 1e-6 difference to 1e-5;
* over the full 2,000 steps (twenty segments) almost every scene has an
  event within +-1e-6; there the adjoint must match the difference at one of
  h = 1e-6, 3e-7, 1e-7, 3e-8 to 1e-5 on a good share of the scenes, and the
  share must grow as h shrinks (fewer events fall in a smaller interval).
  An adjoint that is wrong by more than 1e-5 on a scene cannot match any
  difference there, so a mistake shows up as a share that stays near zero.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def beds():
    from dem_mbd_cuda_amd import scene

    sc = scene.settle_scenes(count=64, n=5000, seed=5)
    params = np.array([[s.pp["mu"], s.pp["epsN"], s.pp["kN"] / np.mean(s.radius)] for s in sc])
    w = np.random.RandomState(1).normal(size=(len(sc), 2))
    return sc, params, w


def adjoint(sc, params, w, steps, K=100):
    from dem_mbd_cuda_amd import diff

    p = torch.tensor(params, requires_grad=True)
    o = diff.BedObservables.apply(p, sc, steps, K, 4)
    (o * torch.from_numpy(w)).sum().backward()
    return p.grad.numpy()


def central(sc, params, w, steps, k, rel, K=100):
    from dem_mbd_cuda_amd import diff

    pp, pm = params.copy(), params.copy()
    pp[:, k] *= 1 + rel
    pm[:, k] *= 1 - rel
    jp = (diff.observables(sc, pp, steps, K) * w).sum(axis=1)
    jm = (diff.observables(sc, pm, steps, K) * w).sum(axis=1)
    return (jp - jm) / (2 * rel * params[:, k])


def test_200_steps_where_the_difference_is_smooth(hip, beds):
    sc, params, w = beds
    g = adjoint(sc, params, w, 200)
    assert np.all(np.isfinite(g))
    results = []
    # friction matters little over 200 steps: its difference at 1e-7 is
    # rounding-limited, so its smoothness is judged between 1e-5 and 1e-6
    for k, pair, least in ((0, (1e-5, 1e-6), 4), (1, (1e-6, 1e-7), 24), (2, (1e-6, 1e-7), 32)):
        fa = central(sc, params, w, 200, k, pair[0])
        fb = central(sc, params, w, 200, k, pair[1])
        smooth = (np.abs(fa - fb) <= 1e-6 * np.abs(fb)) & (np.abs(fb) > 0)
        f6 = fb if pair[1] == 1e-6 else fa
        err = np.abs(g[:, k] - f6) / np.maximum(np.abs(f6), 1e-300)
        print(f"param {k}: {smooth.sum()}/64 scenes smooth over +-{pair[0]:g}, adjoint error there max "
              f"{err[smooth].max() if smooth.any() else 0:.2e}", flush=True)
        results.append((k, smooth, err, least))
    for k, smooth, err, least in results:
        assert smooth.sum() >= least, (k, smooth.sum())
        assert np.all(err[smooth] <= 1e-5), (k, np.nonzero(smooth)[0], err[smooth])


def test_2000_steps_full_configuration(hip, beds):
    sc, params, w = beds
    g = adjoint(sc, params, w, 2000)
    assert np.all(np.isfinite(g))
    hs = (1e-6, 3e-7, 1e-7, 3e-8)
    for k, least in ((0, 6), (1, 16), (2, 10)):
        f = np.array([central(sc, params, w, 2000, k, h) for h in hs])
        match = np.abs(f - g[None, :, k]) <= 1e-5 * np.abs(g[None, :, k])
        share = match.mean(axis=1)
        print(f"param {k}: share of scenes matching at h = {hs}: {share}, at any h {match.any(axis=0).mean():.3f}",
              flush=True)
        assert match.any(axis=0).sum() >= least, (k, share)
        assert share[-1] > share[0], (k, share)
