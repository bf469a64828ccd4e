"""Host-side logic of the multi-label lifting model (no GPU): the label grid,
the per-pixel pair matrices against the oracle's and against the lifted
points they must map to (s, -1), the lift / unlift round trip, and the
batched solver's argument checks, which run before any device call."""
import numpy as np
import pytest

from opymize_amd import lifting
from oracle import lifting as olf


@pytest.mark.parametrize("L", [2, 3, 16, 33])
def test_pair_matrices_map_lifted_points_to_the_piece_pair(L):
    gam = lifting.labels(L, -0.3, 1.4)
    AU, Ar = lifting.pair_matrices(gam)
    AUo, Aro = olf.pair_matrices(gam)
    assert np.array_equal(AU, AUo) and np.array_equal(Ar, Aro)
    rng = np.random.default_rng(L)
    s = rng.uniform(gam[0], gam[-1], 200)
    U, r = lifting.lift(s, gam)
    z = (U @ AU + r @ Ar).reshape(200, L - 1, 2)
    piece = np.argmax(r, axis=1)
    assert np.allclose(z[np.arange(200), piece, 0], s, atol=1e-12)
    assert np.all(z[np.arange(200), piece, 1] == -1.0)
    other = np.ones((200, L - 1), bool)
    other[np.arange(200), piece] = False
    assert not other.any() or np.abs(z[other]).max() < 1e-12
    assert np.allclose(lifting.unlift(np.concatenate([U.ravel(), r.ravel()]), gam, 200), s, atol=1e-12)


def test_labels_and_lift_edges():
    gam = lifting.labels(5, 0.0, 1.0)
    assert np.allclose(gam, [0, 0.25, 0.5, 0.75, 1.0])
    U, r = lifting.lift(np.array([0.0, 1.0, -3.0, 7.0, 0.5]), gam)
    assert np.allclose(U[0], 0) and np.allclose(U[1], 1) and np.allclose(U[2], 0) and np.allclose(U[3], 1)
    assert np.allclose(U[4], [1, 1, 0, 0]) and r.sum(axis=1).tolist() == [1.0] * 5
    with pytest.raises(ValueError):
        lifting.labels(1, 0, 1)
    with pytest.raises(ValueError):
        lifting.labels(4, 1.0, 1.0)


def test_liftbatch_rejects_bad_arguments_before_any_device_call():
    from opymize_amd.solvers.lift import LiftBatch
    img = np.zeros((8, 9))
    for kw, msg in ((dict(L=1), "labels"), (dict(L=34), "labels"), (dict(lam=0.0), "lam"),
                    (dict(omega=-1.0), "omega"), (dict(omega=float("nan")), "omega")):
        args = dict(images=[img], L=16, lam=0.25)
        args.update(kw)
        with pytest.raises(ValueError, match=msg):
            LiftBatch(**args)
    with pytest.raises(ValueError, match="2-D"):
        LiftBatch([np.zeros(5)], 16, 0.25)
    with pytest.raises(ValueError, match="non-empty"):
        LiftBatch([], 16, 0.25)
    with pytest.raises(ValueError, match="finite"):
        LiftBatch([np.full((4, 4), np.nan)], 16, 0.25)
