"""The learned initial guess on the GPU (smd_mlp_* and smd_induce guess 3
through the C ABI) against the CPU oracle in oracle/learned_guess.py.

Bars:
 * network outputs and gradients: the oracle rounds to BF16 where the kernels
   do (inputs, weights, activations, back-propagated gradients), so the two
   differ by FP32 accumulation order plus the rare activation that the order
   moves across a BF16 rounding boundary: 1e-3 relative RMS on the outputs,
   2e-3 on the loss gradient, ten times below the 0.4% that separates the
   emulated from the plain FP64 network;
 * descriptors: FP32 storage of FP64 sums, 1e-5 of each column's scale;
 * the guess: CG from it reaches the dipoles of CG from the reference's
   start to 1e-10 relative RMS (the same system solved to |r| < 1e-14), in
   fewer iterations than from alpha E0 once the network is trained."""
import numpy as np
import pytest

from tests.helpers import rel_rms

pytestmark = pytest.mark.gpu

CUT = 0.7


def _system(seed):
    from swinmd_amd import systems as S
    return S.water(512, density_mol=33.4, seed=seed, lattice=(8, 8, 8), r_oh=0.09572, hoh_deg=108.5,
                   types=S.AMOEBA_TYPES)


def _ctx(system):
    from swinmd_amd.forces import Context, Induction
    ctx = Context(system)
    ctx.set_pairlist(CUT, 0.05, True)
    ind = Induction(ctx, cutoff_ind=CUT, a_damp=0.39)
    return ctx, ind


def test_forward_matches_oracle():
    from oracle import learned_guess as G
    from swinmd_amd.forces import LearnedGuess
    ctx, _ = _ctx(_system(31))
    net = LearnedGuess(ctx, seed=3)
    rng = np.random.default_rng(4)
    p = G.init_params(5)
    p[G.NPARAMS - 3:] = [0.1, -0.2, 0.05]
    p32 = p.astype(np.float32)
    net.set_params(p32)
    assert np.array_equal(net.params(), p32)
    for rows in (1, 17, 1000):
        print(f"case forward rows {rows}", flush=True)
        x = rng.normal(size=(rows, G.NF)).astype(np.float32)
        y = net.forward(x)
        want, _ = G.forward(p32.astype(np.float64), x.astype(np.float64), emulate=True)
        plain, _ = G.forward(p32.astype(np.float64), x.astype(np.float64), emulate=False)
        assert rel_rms(y, want) <= 1e-3, rel_rms(y, want)
        if rows == 1000:
            assert rel_rms(plain, want) > 10 * rel_rms(y, want)
    ctx.close()


def test_loss_gradient_matches_oracle():
    from oracle import learned_guess as G
    from swinmd_amd.forces import LearnedGuess
    ctx, _ = _ctx(_system(31))
    net = LearnedGuess(ctx, seed=3)
    rng = np.random.default_rng(6)
    p32 = G.init_params(7).astype(np.float32)
    p32[G.NPARAMS - 3:] = [0.3, -0.1, 0.2]
    net.set_params(p32)
    offs = np.cumsum([0] + [int(np.prod(s)) for _, s in G.PUB_SHAPES])
    for rows in (16, 777, 4096):
        print(f"case loss_grad rows {rows}", flush=True)
        x = rng.normal(size=(rows, G.NF)).astype(np.float32)
        t = rng.normal(scale=0.5, size=(rows, G.NO)).astype(np.float32)
        loss, g = net.loss_grad(x, t)
        wl, wg = G.loss_grad(p32.astype(np.float64), x.astype(np.float64), t.astype(np.float64), emulate=True)
        assert abs(loss - wl) <= 1e-4 * wl, (loss, wl)
        assert rel_rms(g, wg) <= 2e-3, rel_rms(g, wg)
        for b, (name, _) in enumerate(G.PUB_SHAPES):      # every block, biases included
            gb, wb = g[offs[b]:offs[b + 1]], wg[offs[b]:offs[b + 1]]
            assert rel_rms(gb, wb) <= 1e-2, (name, rel_rms(gb, wb))
    ctx.close()


def test_descriptors_match_oracle():
    from oracle import learned_guess as G
    from swinmd_amd.forces import LearnedGuess
    s = _system(31)
    ctx, ind = _ctx(s)
    net = LearnedGuess(ctx)
    _, conv, _ = ind.induce()
    assert conv
    _, _, e0, _ = ind.results()
    d = net.descriptors()
    want = G.descriptors(s.pos, s.mol_start, s.mol, s.scaled_charge, np.array([s.types[t].polar for t in s.type]),
                         e0, np.asarray(s.box, dtype=float), CUT, 0.39)
    scale = np.abs(want).max(axis=0) + 1e-300
    err = (np.abs(d - want) / scale).max(axis=0)
    assert err.max() <= 1e-5, err
    ctx.close()


def test_mixed_descriptors_match_oracle():
    """The mixed solve's descriptors: FP32 pair math over the partner table
    (after an alpha E0 start) and the shells summed by the mixed field pass
    (with the learned start), against the oracle on that solve's E0: 1e-4 of
    each column's scale (FP32 displacements and sums; columns 3-5 go through
    the mixed A v)."""
    from oracle import learned_guess as G
    from swinmd_amd import _lib as L
    from swinmd_amd.forces import Context, Induction, LearnedGuess
    s = _system(31)
    ctx = Context(s)
    ctx.set_pairlist(CUT, 0.05, True)
    ind = Induction(ctx, cutoff_ind=CUT, a_damp=0.39, tol=1e-6 * L.DEBYE_E_NM, precond=2, prec=L.PREC_MIXED,
                    guess=2)
    net = LearnedGuess(ctx)
    norm = np.concatenate([np.zeros(G.NF), np.ones(G.NF), np.zeros(G.NO), np.ones(G.NO)]).astype(np.float32)
    net.set_norm(norm)
    polar = np.array([s.types[t].polar for t in s.type])
    for guess in (2, 3):
        print(f"case mixed descriptors guess {guess}", flush=True)
        ind.params.guess = guess
        _, conv, _ = ind.induce()
        assert conv
        _, _, e0, _ = ind.results()
        d = net.descriptors()
        want = G.descriptors(s.pos, s.mol_start, s.mol, s.scaled_charge, polar, e0, np.asarray(s.box, dtype=float),
                             CUT, 0.39)
        scale = np.abs(want).max(axis=0) + 1e-300
        err = (np.abs(d - want) / scale).max(axis=0)
        assert err.max() <= 1e-4, (guess, err)
    ctx.close()


def _frames(net, ctx, ind, seeds):
    for sd in seeds:
        s = _system(sd)
        ctx.set_positions(s.pos)
        ctx.set_pairlist(CUT, 0.05, True)
        ind.params.guess = 2
        _, conv, _ = ind.induce()
        assert conv
        rows = net.add_frame()
    return rows


def test_training_reduces_iterations_and_keeps_the_solution(tmp_path):
    """Ablation: on a frame the network never saw, CG from the learned guess
    and CG from alpha E0 (the best guess without the network) stop at the same
    tolerance on the same dipoles; the learned start needs fewer iterations.
    A saved network loaded into another context starts CG at the same point,
    and training twice from the same seed gives the same network bit for
    bit."""
    from swinmd_amd.forces import LearnedGuess
    ctx, ind = _ctx(_system(31))
    net = LearnedGuess(ctx, seed=1)
    rows = _frames(net, ctx, ind, [41, 42, 43, 44, 45, 46])
    assert rows == 6 * 1536
    rep = net.train(epochs=40, batch=512, lr=2e-3, lr_final=2e-5, seed=9)
    print("training", rep, flush=True)
    assert rep["samples"] == rows and rep["steps"] == 40 * 18
    assert rep["loss_last"] < 0.1 * rep["loss_first"]
    params, norm = net.params(), net.norm()
    # held-out frame
    s = _system(31)
    ctx.set_positions(s.pos)
    ctx.set_pairlist(CUT, 0.05, True)
    out = {}
    for guess in (0, 2, 3):
        ind.params.guess = guess
        it, conv, res = ind.induce()
        assert conv and res < 1e-14
        out[guess] = (it, ind.results()[3])
    print("iterations (reference start, alpha E0, learned):", out[0][0], out[2][0], out[3][0], flush=True)
    assert rel_rms(out[3][1], out[0][1]) <= 1e-10 and rel_rms(out[2][1], out[0][1]) <= 1e-10
    assert out[3][0] < out[2][0] < out[0][0]
    # a saved network loaded into another context gives the same start
    net.save(tmp_path / "net.npz")
    ctx3, ind3 = _ctx(_system(31))
    LearnedGuess(ctx3, seed=5).load(tmp_path / "net.npz")
    ind3.params.guess = 3
    it3, conv3, _ = ind3.induce()
    assert conv3 and it3 == out[3][0]
    assert rel_rms(ind3.results()[3], out[3][1]) <= 1e-13
    ctx3.close()
    # determinism
    net2 = LearnedGuess(ctx, seed=1)
    net2.train(epochs=40, batch=512, lr=2e-3, lr_final=2e-5, seed=9)
    assert np.array_equal(net2.params(), params) and np.array_equal(net2.norm(), norm)
    ctx.close()


def test_learned_guess_errors():
    from swinmd_amd import _lib
    from swinmd_amd.forces import LearnedGuess
    ctx, ind = _ctx(_system(31))
    ind.params.guess = 3
    with pytest.raises(_lib.SmdError) as e:
        ind.induce()                                     # no network
    assert e.value.code == _lib.SMD_ESTATE
    net = LearnedGuess(ctx)
    with pytest.raises(_lib.SmdError) as e:
        ind.induce()                                     # not trained, no normalisation
    assert e.value.code == _lib.SMD_ESTATE
    with pytest.raises(_lib.SmdError) as e:
        net.train()                                      # no frames
    assert e.value.code == _lib.SMD_ESTATE
    with pytest.raises(_lib.SmdError) as e:
        net.set_params(np.zeros(10))
    assert e.value.code == _lib.SMD_EINVAL
    bad = np.ones(_lib.MLP_NNORM, dtype=np.float32)
    bad[40] = 0.0
    with pytest.raises(_lib.SmdError) as e:
        net.set_norm(bad)
    assert e.value.code == _lib.SMD_EINVAL
    ind.params.guess = 5
    with pytest.raises(_lib.SmdError) as e:
        ind.induce()
    assert e.value.code == _lib.SMD_EINVAL
    # a frame is only the converged solve of the current state: a new box
    # or moved positions make the solve's coupling table stale
    import ctypes as C

    def set_box(box):
        b = np.ascontiguousarray(box, dtype=np.float64)
        ctx.check(ctx.lib.smd_set_box(ctx.h, b.ctypes.data_as(C.c_void_p)))

    ind.params.guess = 2
    assert ind.induce()[1]
    net.add_frame()
    set_box(ctx.system.box * 1.001)
    with pytest.raises(_lib.SmdError) as e:
        net.add_frame()
    assert e.value.code == _lib.SMD_ESTATE
    set_box(ctx.system.box)
    ctx.set_pairlist(CUT, 0.05, True)
    assert ind.induce()[1]
    ctx.set_positions(ctx.positions() + 1e-4)
    with pytest.raises(_lib.SmdError) as e:
        net.add_frame()
    assert e.value.code == _lib.SMD_ESTATE
    ctx.close()
