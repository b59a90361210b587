"""The seeded film states of the surrogate's training family
(mms_surrogate_state, host code): determinism, |M| = Ms on the mask and 0
off it, the same state in every layer, the shapes' geometry and the states'
structure."""
import numpy as np
import pytest

F = np.float32
MS = F(8.6e5)


@pytest.mark.parametrize("dims", [(128, 128, 3), (64, 48, 2), (33, 17, 1)])
def test_states_are_seeded_unit_and_layered(mg, dims):
    xd, yd, zd = dims
    plane = xd * yd
    kinds = set()
    for i in range(24):
        M, mat, kind = mg.surrogate_state(*dims, MS, 2010, i)
        M2, mat2, kind2 = mg.surrogate_state(*dims, MS, 2010, i)
        assert np.array_equal(M, M2) and np.array_equal(mat, mat2) and kind == kind2
        kinds.add(kind)
        m = M.reshape(zd, plane, 3)
        k = mat.reshape(zd, plane)
        for z in range(1, zd):
            assert np.array_equal(m[z], m[0]) and np.array_equal(k[z], k[0])
        nrm = np.linalg.norm(m[0].astype(np.float64), axis=1)
        on = k[0] == 1
        assert on.sum() > 0.05 * plane
        assert np.abs(nrm[on] / MS - 1).max() < 1e-6
        assert np.all(m[0][~on] == 0)
    assert kinds == set(range(6))
    other, _, _ = mg.surrogate_state(*dims, MS, 2011, 0)
    assert not np.array_equal(other, mg.surrogate_state(*dims, MS, 2010, 0)[0])


def test_state_shapes_and_textures(mg):
    """Discs are convex and round (area close to pi r^2 of their extent),
    rings have a hole at their centre, vortices have an out-of-plane core of
    either polarity and circulate, domain states are in-plane."""
    xd = yd = 128
    seen = set()
    for i in range(40):
        M, mat, kind = mg.surrogate_state(xd, yd, 1, MS, 7, i)
        shape, state = kind // 2, kind % 2
        k = mat.reshape(yd, xd).astype(bool)
        m = M.reshape(yd, xd, 3) / MS
        ys, xs = np.nonzero(k)
        cy, cx = ys.mean(), xs.mean()
        if shape == 0:
            r = 0.5 * (xs.max() - xs.min() + 1)
            assert abs(k.sum() / (np.pi * r * r) - 1) < 0.08
            assert k[int(cy), int(cx)]
        if shape == 2:
            assert not k[int(round(cy)), int(round(cx))]
        if state == 1:
            assert np.abs(m[k][:, 2]).max() < 0.02
        else:
            mz = m[k][:, 2]
            assert np.abs(mz).max() > 0.5 or shape == 2   # a ring has no core
            # circulation: m is perpendicular to the radius away from the core
            ry, rx = ys - cy, xs - cx
            far = np.hypot(rx, ry) > 6
            radial = (m[k][:, 0] * rx + m[k][:, 1] * ry)[far] / np.hypot(rx, ry)[far]
            assert np.abs(radial).mean() < 0.05
        seen.add(kind)
    assert len(seen) == 6
