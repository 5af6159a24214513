"""Learned search radius on the GPU: the network's forward pass, gradient
and Adam step against the CPU oracle, and the learned distance path against
the plain one (identical outputs, fewer evaluated triangle pairs)."""
import numpy as np
import pytest

from oracle import oracle as O
from stl2sdf import Learned, SignedDistances, _lib
from tests import _cases
from tests.test_oracle_mlp import he_params

pytestmark = pytest.mark.gpu


def _points(seed, n):
    rng = np.random.default_rng(seed)
    return rng.uniform(-1.1, 1.1, (n, 3)).astype(np.float32)


def test_forward_matches_oracle():
    P = he_params(10).astype(np.float32)
    Learned.mlp_set_params(P)
    assert np.array_equal(Learned.mlp_get_params(), P)
    u = _points(11, 1000)
    y = Learned.mlp_forward(u).astype(np.float64)
    y_ref = O.mlp_forward(P.astype(np.float64), u, O.MLP_BF16)
    err = np.abs(y - y_ref)
    scale = np.sqrt(np.mean(y_ref ** 2))
    print(f"forward: max err {err.max():.3e}, median {np.median(err):.3e}, rms y {scale:.3e}")
    # FP32 accumulation vs FP64 and rare BF16 rounding flips of activations
    assert err.max() <= 5e-3 * scale
    assert np.median(err) <= 2e-4 * scale


def test_gradient_matches_oracle():
    P = he_params(12).astype(np.float32)
    Learned.mlp_set_params(P)
    u = _points(13, 200)          # not a multiple of the 64-point block
    t = np.random.default_rng(14).uniform(0, 1, 200).astype(np.float32)
    loss, g = Learned.mlp_loss_grad(u, t)
    loss_ref, g_ref = O.mlp_loss_grad(P.astype(np.float64), u, t, O.MLP_BF16)
    print(f"loss {loss:.6e} oracle {loss_ref:.6e}")
    assert abs(loss - loss_ref) <= 1e-4 * loss_ref
    for name, n in (("W1", 16384), ("b1", 128), ("W2", 16384), ("b2", 128), ("W3", 16384),
                    ("b3", 128), ("w4", 128), ("b4", 1)):
        o = O.MLP_OFFSETS[name]
        a, b = g[o:o + n].astype(np.float64), g_ref[o:o + n]
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        print(f"  {name}: rel L2 err {rel:.2e}")
        assert rel <= 1e-2, name
    cos = g @ g_ref / (np.linalg.norm(g) * np.linalg.norm(g_ref))
    assert cos >= 0.9999
    # padded feature columns never receive gradient
    W1 = g[:16384].reshape(128, 128)
    assert np.all(W1[:, 100:] == 0)


def test_adam_step_matches_oracle():
    P = he_params(15).astype(np.float32)
    Learned.mlp_set_params(P)
    u = _points(16, 256)
    t = np.random.default_rng(17).uniform(0, 1, 256).astype(np.float32)
    _, g = Learned.mlp_loss_grad(u, t)
    lr = 1e-3
    # one step, full batch in order: the learning-rate schedule starts at lr
    Learned.mlp_train(u, t, steps=1, batch=256, lr=lr, seed=0)
    P1 = Learned.mlp_get_params()
    P1_ref, _, _ = O.mlp_adam(P, np.zeros_like(P), np.zeros_like(P), g, 1, lr)
    diff = np.abs(P1.astype(np.float64) - P1_ref)
    print(f"adam: max |dp| {diff.max():.3e}")
    assert diff.max() <= 2.5e-7  # FP32 rounding of the update, parameters within +-1.5


def test_training_reduces_loss():
    """A smooth target (scaled distance to a sphere) is learned by Adam."""
    rng = np.random.default_rng(18)
    u = rng.uniform(-1, 1, (8192, 3)).astype(np.float32)
    t = np.clip(np.abs(np.linalg.norm(u, axis=1) - 0.6) / 0.3, 0, 1).astype(np.float32)
    Learned.mlp_set_params(he_params(19).astype(np.float32))
    hist = Learned.mlp_train(u, t, steps=400, batch=1024, lr=2e-3, seed=5)
    first, last = hist[:10].mean(), hist[-20:].mean()
    print(f"training loss {first:.4e} -> {last:.4e}")
    assert np.all(np.isfinite(hist))
    assert last < first / 10


def _learned_vs_plain(mesh, grid, fit_kw):
    lr_path = Learned.LearnedRadius()
    fs = lr_path.fit(mesh, grid, **fit_kw)
    print("fit:", {k: round(v, 5) for k, v in fs.items()})
    plain = SignedDistances.signed_distance(mesh, grid)
    learned = lr_path.signed_distance(mesh, grid)
    st = lr_path.stats()
    print("run:", st)
    return plain, learned, st, fs


def test_learned_path_identical_icosphere():
    mesh, grid = _cases.icosphere_case()
    plain, learned, st, fs = _learned_vs_plain(mesh, grid, dict(n_coarse=24, steps=300,
                                                                batch=2048))
    assert np.array_equal(learned, plain)
    assert np.array_equal(np.signbit(learned), np.signbit(plain))
    assert st["tile_pairs"] > 0 and 0 < st["evals"] <= st["evals_plain"] <= st["tile_pairs"]


def test_learned_path_exact_with_untrained_network():
    """Any network output (here: an untrained one, with the radius scale of a
    different mesh) leaves the result unchanged; only the work varies."""
    mesh, grid = _cases.star_case(60, 45)
    plain = SignedDistances.signed_distance(mesh, grid)
    lr_path = Learned.LearnedRadius()
    lr_path.fit(mesh, grid, n_coarse=12, steps=1, batch=64)
    Learned.mlp_set_params(he_params(20).astype(np.float32))
    learned = lr_path.signed_distance(mesh, grid)
    st = lr_path.stats()
    print("untrained run:", st)
    assert np.array_equal(learned, plain)


def test_learned_radius_gain_star():
    """On a star surface the trained network shrinks the candidate set below
    the plain radius h and fewer tile pairs are evaluated, while the output
    stays identical; training cuts the nodes that need the exact fallback
    several-fold against an untrained network with the same calibration.
    (Measured: node pairs 0.96 of those within h, fallback nodes 7x fewer
    than untrained.  The plain radius is already the clamp distance h, so
    only nodes with d well below h can gain.)"""
    mesh, grid = _cases.star_case(200, 121)
    plain, learned, st, fs = _learned_vs_plain(mesh, grid, dict(n_coarse=40, steps=800,
                                                                batch=4096))
    assert np.array_equal(learned, plain)
    lane_ratio = st["lane_pairs_radius"] / st["lane_pairs_h"]
    pair_ratio = st["evals"] / st["evals_plain"]
    fb = st["fallback_nodes"] / grid.ngp
    print(f"trained: node-pairs radius/h {lane_ratio:.3f}; exact evaluations learned/plain "
          f"{pair_ratio:.3f}; fallback nodes {st['fallback_nodes']:.0f} ({fb:.2e} of grid); "
          f"band MAE {fs['mae_band']:.4f} hc; delta {fs['delta']:.4f} hc")
    Learned.mlp_set_params(he_params(21).astype(np.float32))
    lr_path = Learned.LearnedRadius()
    untrained = lr_path.signed_distance(mesh, grid)
    su = lr_path.stats()
    print("untrained:", su)
    assert np.array_equal(untrained, plain)
    assert lane_ratio < 0.99
    assert pair_ratio <= 1.0
    assert st["fallback_nodes"] * 3 < su["fallback_nodes"]


def test_network_skipped_when_every_radius_is_h():
    """With delta * hc >= h every learned radius is h whatever the network
    outputs (its prediction is clamped at 0), so the call runs the plain
    distance loop without the network; forcing the network gives the same
    output.  A one-step fit leaves delta large."""
    mesh, grid = _cases.star_case(60, 45)
    plain = SignedDistances.signed_distance(mesh, grid)
    lr_path = Learned.LearnedRadius()
    fs = lr_path.fit(mesh, grid, n_coarse=12, steps=1, batch=64)
    print("fit:", fs)
    assert fs["delta"] * fs["hc"] >= grid.cell_size
    skipped = lr_path.signed_distance(mesh, grid)
    st = lr_path.stats()
    assert st["network_skipped"] == 1.0 and st["evals"] == 0.0
    assert np.array_equal(skipped, plain)
    ctx = _lib.ctx(0)
    _lib.check(_lib.lib().sdf_ctx_tune(ctx, b"learned_force", 1))
    try:
        forced = lr_path.signed_distance(mesh, grid)
        sf = lr_path.stats()
    finally:
        _lib.check(_lib.lib().sdf_ctx_tune(ctx, b"learned_force", 0))
    print("forced:", sf)
    assert sf["network_skipped"] == 0.0 and sf["evals"] > 0
    assert np.array_equal(forced, plain)
