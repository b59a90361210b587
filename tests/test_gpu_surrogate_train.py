"""Training of the learned surrogate on the GPU: the generated (charge,
potential) pairs against the oracle, the backward kernels' gradient against
PyTorch autograd of the fp32 restatement (same BF16 roundings), the Adam step
against torch.optim.Adam, the loss falling over training, and the trained
weights reproducing the trainer's predictions in a film."""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
F = np.float32
sys.path.insert(0, os.path.dirname(__file__))
DX = F(5e-9)
MS = F(8.6e5)


def _say(*a):
    print("[gpu-test]", *a, flush=True)


def _weights(mg, seed, scale=0.08):
    n = int(mg.lib().mms_surrogate_weight_count())
    return (np.random.default_rng(seed).standard_normal(n) * scale).astype(F)


def test_generated_pairs_match_oracle(mg, orc):
    """Pairs of the seeded states: charge bitwise equal to divergence_3d
    (vector_functions.cpp:8-28) of mms_surrogate_state's M, potential of the
    direct sum within 1e-6 and of the FMM (P = 6) within 1e-4 of the FP64
    oracle sum."""
    dims = (32, 32, 3)
    d = (DX, DX, DX)
    with mg.SurrogateTrainer(*dims, DX, MS, 2) as t:
        t.generate(2, 11, P=0)
        t.generate(1, 11, P=6)
        assert t.pairs == 3
        for i, tol in ((0, 1e-6), (1, 1e-6), (2, 1e-4)):
            M, mat, kind = mg.surrogate_state(*dims, MS, 11, i)
            rho, phi = t.pair(i)
            _say(f"pair {i} kind {kind} cells {int(mat.sum())}")
            assert np.array_equal(rho, orc.divergence_3d(M, dims, d))
            ref = orc.potential_exact_f64(rho, dims, d)
            err = np.linalg.norm(phi - ref) / np.linalg.norm(ref)
            _say(f"  potential rel L2 {err:.2e}")
            assert err < tol


def test_trainer_weights_round_trip(mg):
    w = _weights(mg, 3)
    with mg.SurrogateTrainer(16, 16, 2, DX, MS, 1, weights=w) as t:
        assert np.array_equal(t.weights(), w)


@pytest.mark.parametrize("dims,nfilm", [((32, 32, 3), 2), ((48, 40, 2), 3)])
def test_gradient_matches_autograd(mg, dims, nfilm):
    """Every weight block's gradient (conv weight gradients on the matrix
    cores, transposed-conv data gradients, spectral backward, lift and head)
    within 1% relative L2 of PyTorch autograd (measured about 2e-3); the loss
    within 1e-3."""
    import surrogate_ref as R
    w = _weights(mg, 5)
    with mg.SurrogateTrainer(*dims, DX, MS, nfilm, weights=w) as t:
        t.generate(nfilm, 2010, P=0)
        idx = np.arange(nfilm)
        g, loss = t.gradient(idx)
        pairs = [t.pair(i) for i in idx]
    rho = np.concatenate([p[0] for p in pairs])
    phi = np.concatenate([p[1] for p in pairs])
    lref, gref = R.loss_and_grad(w, rho, phi, dims, nfilm, float(DX / MS), float(MS * DX))
    _say(f"dims {dims} x {nfilm}: loss {loss:.6e} ref {lref:.6e}")
    assert abs(loss - lref) <= 1e-3 * abs(lref)
    for name, s in R.block_slices():
        e = np.linalg.norm(g[s] - gref[s]) / np.linalg.norm(gref[s])
        _say(f"  {name:6s} |g| {np.linalg.norm(gref[s]):.3e} rel L2 {e:.2e}")
        assert np.all(np.isfinite(g[s]))
        assert e < 1e-2, name


def test_adam_step_matches_torch(mg):
    """Two Adam steps (lr 1e-3, betas 0.9 / 0.999, eps 1e-8) of the FP32
    parameters equal torch.optim.Adam fed the same gradients (the trainer's
    gradient is deterministic, so the step recomputes the one read back)."""
    import torch
    w = _weights(mg, 9)
    dims = (32, 32, 2)
    idx = np.arange(2)
    got, grads = [], []
    with mg.SurrogateTrainer(*dims, DX, MS, 2, weights=w) as t:
        t.generate(2, 5, P=0)
        for _ in range(2):
            grads.append(t.gradient(idx)[0])
            t.step(idx, 1e-3)
            got.append(t.weights())
    p = torch.nn.Parameter(torch.from_numpy(w.copy()))
    opt = torch.optim.Adam([p], lr=1e-3, betas=(0.9, 0.999), eps=1e-8)
    for k in range(2):
        p.grad = torch.from_numpy(grads[k].copy())
        opt.step()
        ref = p.detach().numpy()
        err = np.abs(got[k] - ref).max()
        _say(f"Adam step {k + 1}: max |w - torch| {err:.2e}, max |step| {np.abs(ref - w).max():.2e}")
        assert err < 2e-6


def test_training_lowers_the_loss(mg):
    """Forty Adam steps on 16 direct-sum pairs of 64 x 64 x 3 films lower the
    mean residual of the training pairs; the run is deterministic."""
    dims = (64, 64, 3)
    res = []
    for rep in range(2):
        with mg.SurrogateTrainer(*dims, DX, MS, 4) as t:
            t.generate(16, 77, P=0)
            batches = [np.arange(4 * b, 4 * b + 4) for b in range(4)]
            before = np.mean([t.eval(b) for b in batches])
            losses = [t.step(batches[s % 4], 1e-3) for s in range(40)]
            after = np.mean([t.eval(b) for b in batches])
            res.append((before, after, losses, t.weights()))
        _say(f"run {rep}: residual {before:.3f} -> {after:.3f}; loss {losses[0]:.3f} -> {losses[-1]:.3f}")
    assert res[0][1] < 0.7 * res[0][0]
    assert res[0][2] == res[1][2] and np.array_equal(res[0][3], res[1][3])


def test_trained_weights_in_a_film(mg):
    """Weights taken from a trainer give bitwise the trainer's prediction
    through mms_surrogate_init / mms_surrogate_potential on a film."""
    dims = (48, 32, 3)
    with mg.SurrogateTrainer(*dims, DX, MS, 2) as t:
        t.generate(4, 3, P=0)
        for s in range(3):
            t.step(np.array([2 * (s % 2), 2 * (s % 2) + 1]), 1e-3)
        w = t.weights()
        idx = np.array([3, 1])
        pred = t.predict(idx)
        rho = [t.pair(i)[0] for i in idx]
    with mg.Film(*dims, DX, DX, DX, Ms=MS) as f:
        f.surrogate_init(w)
        for b in range(2):
            assert np.array_equal(f.surrogate_potential(rho[b]), pred[b])


def test_c4_films_train_and_generalise(mg):
    """The surrogate's own workload at its film size (128 x 128 x 3, FMM P = 6
    pairs of the seeded states): 320 Adam steps of 8 films on 448 pairs bring
    the residual of 64 held-out pairs from about 1 below 0.3."""
    B = 8
    with mg.SurrogateTrainer(128, 128, 3, DX, MS, B) as t:
        t.generate(512, 2010, P=6)
        held = [np.arange(448 + B * k, 448 + B * k + B) for k in range(8)]
        before = np.mean([t.eval(b) for b in held])
        rng = np.random.default_rng(1)
        order = np.concatenate([rng.permutation(448) for _ in range(6)])
        for s in range(320):
            t.step(order[B * s:B * s + B], 1e-3)
        after = np.mean([t.eval(b) for b in held])
    _say(f"C4 films: held-out residual {before:.3f} -> {after:.4f}")
    assert before > 0.9 and after < 0.3
