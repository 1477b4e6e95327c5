"""The learned guess's CPU oracle (oracle/learned_guess.py) pinned by what
fixes it independently of its own formulas: finite differences for the
network's gradient, the C oracle's dense A (which follows
src/forces/Induction.cpp:181-455) for the dipole tensor, brute force for the
shell sums, and geometry for the frames.  CPU only."""
import math

import numpy as np
import pytest

from oracle import learned_guess as G


def test_bf16_rounds_to_nearest_even():
    one = np.float32(1.0)
    half_ulp = np.float32(2.0 ** -8)                       # half a BF16 ulp at 1
    vals = np.array([one, one + half_ulp, one + 3 * half_ulp, -(one + 3 * half_ulp), one + 1.5 * half_ulp,
                     np.float32(3.0e38)], dtype=np.float32)
    got = G.bf16(vals)
    want = [1.0, 1.0, 1.0 + 2.0 ** -6, -(1.0 + 2.0 ** -6), 1.0 + 2.0 ** -7, float(np.float32(3.0e38))]
    assert got[:5].tolist() == want[:5]
    assert abs(got[5] - want[5]) / want[5] < 2.0 ** -8
    assert np.all(G.bf16(got) == got)                       # idempotent


def test_network_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = G.init_params(1)
    p[G.NPARAMS - 3:] = rng.normal(scale=0.1, size=3)     # non-zero output biases
    x = rng.normal(size=(40, G.NF))
    t = rng.normal(scale=0.3, size=(40, G.NO))
    _, g = G.loss_grad(p, x, t, emulate=False)
    offs = np.cumsum([0] + [int(np.prod(s)) for _, s in G.PUB_SHAPES])
    picks = []
    for b in range(8):                                      # a few entries of every block
        picks += list(rng.integers(offs[b], offs[b + 1], size=4))
    h = 1e-6
    for k in picks:
        pp = p.copy()
        pp[k] += h
        lp, _ = G.loss_grad(pp, x, t, emulate=False)
        pp[k] -= 2 * h
        lm, _ = G.loss_grad(pp, x, t, emulate=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[k]) <= 1e-7 + 1e-5 * abs(fd), (k, fd, g[k])


def test_emulated_network_differs_from_fp64_only_by_bf16_rounding():
    rng = np.random.default_rng(1)
    p = G.init_params(2)
    x = rng.normal(size=(200, G.NF))
    y64, _ = G.forward(p, x, emulate=False)
    y16, _ = G.forward(p, x, emulate=True)
    d = np.sqrt(np.sum((y16 - y64) ** 2) / np.sum(y64 ** 2))
    assert 1e-4 < d < 3e-2
    # with weights and inputs already on the BF16 grid only the activations round
    pb = G.bf16(p)
    xb = G.bf16(x)
    y64b, _ = G.forward(pb, xb, emulate=False)
    y16b, _ = G.forward(pb, xb, emulate=True)
    assert np.sqrt(np.sum((y16b - y64b) ** 2) / np.sum(y64b ** 2)) < d


def _cluster(nmol=6, seed=3, box=50.0):
    """Isolated waters near the centre of a large box."""
    rng = np.random.default_rng(seed)
    pos = []
    for m in range(nmol):
        c = box / 2 + rng.uniform(-0.35, 0.35, size=3)
        q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        th = math.radians(108.5) / 2
        o = np.zeros(3)
        h1 = 0.09572 * np.array([-math.sin(th), 0.0, math.cos(th)])
        h2 = 0.09572 * np.array([math.sin(th), 0.0, math.cos(th)])
        pos += [c + q @ h1, c + q @ o, c + q @ h2]
    pos = np.array(pos)
    mol = np.repeat(np.arange(nmol), 3)
    ms = np.arange(0, 3 * nmol + 1, 3)
    q = np.tile([0.41 * G.SQRT_COULOMB, -0.82 * G.SQRT_COULOMB, 0.41 * G.SQRT_COULOMB], nmol)
    return pos, mol, ms, q, np.full(3, box)


def test_site_frames_geometry():
    pos, _, ms, _, _ = _cluster()
    F = G.site_frames(pos, ms)
    for i in range(len(pos)):
        assert np.allclose(F[i] @ F[i].T, np.eye(3), atol=1e-13)
    dets = np.linalg.det(F).reshape(-1, 3)
    assert np.allclose(dets, [1.0, 1.0, -1.0], atol=1e-12)
    for m in range(len(ms) - 1):
        h1, o, h2 = pos[3 * m], pos[3 * m + 1], pos[3 * m + 2]
        bis = (h1 - o) / np.linalg.norm(h1 - o) + (h2 - o) / np.linalg.norm(h2 - o)
        assert np.allclose(F[3 * m + 1, 2], bis / np.linalg.norm(bis), atol=1e-13)
        assert F[3 * m + 1, 0] @ (h2 - h1) > 0 and abs(F[3 * m + 1, 0] @ bis) < 1e-13
        # the molecule lies in the local x-z plane
        for a in (h1, h2):
            assert abs(F[3 * m + 1, 1] @ (a - o)) < 1e-13
        # H1 and H2 sit at mirror-image local coordinates
        l1 = F[3 * m] @ (h1 - o)
        l2 = F[3 * m + 2] @ (h2 - o)
        assert np.allclose(l1, l2, atol=1e-13)


def test_charge_shells_against_brute_force_with_periodic_images():
    """Waters placed across the faces of a 1.5 nm box: plain loops over every
    pair and the nearest image decide each shell and kind."""
    rng = np.random.default_rng(7)
    pos, mol, ms, q, _ = _cluster(nmol=10, seed=8, box=1.5)
    box = np.full(3, 1.5)
    pos = pos + rng.uniform(-0.75, 0.75, size=3)          # some molecules straddle the faces
    pos = pos - box * np.floor(pos / box)
    got = G.charge_shells(pos, q, mol, box)
    n = len(pos)
    want = np.zeros((n, 8, 3))
    for i in range(n):
        for j in range(n):
            if mol[i] == mol[j]:
                continue
            d = [pos[i][k] - pos[j][k] for k in range(3)]
            d = [d[k] - box[k] * round(d[k] / box[k]) for k in range(3)]
            r = math.sqrt(sum(v * v for v in d))
            if r >= 0.6:
                continue
            shell = 0 if r < 0.3 else 1 if r < 0.4 else 2 if r < 0.5 else 3
            kind = 0 if q[j] > 0 else 1
            for k in range(3):
                want[i, 2 * shell + kind, k] += (q[j] / G.SQRT_COULOMB) * d[k] / r ** 3
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()
    assert np.count_nonzero(np.abs(want).sum(axis=2)) > 8   # several bins occupied


def test_descriptors_rotate_with_the_molecules():
    """A rigid rotation of an isolated cluster leaves every frame-resolved
    shell descriptor unchanged, and so does any field rotated with it."""
    pos, mol, ms, q, box = _cluster()
    rng = np.random.default_rng(9)
    R, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    if np.linalg.det(R) < 0:
        R[:, 0] *= -1
    c = box / 2
    pos2 = (pos - c) @ R.T + c
    polar = np.tile([0.000496, 0.000837, 0.000496], len(ms) - 1)
    e0 = rng.normal(size=pos.shape)
    d1 = G.descriptors(pos, ms, mol, q, polar, e0, box, cutoff_ind=0.7, a_damp=0.39)
    d2 = G.descriptors(pos2, ms, mol, q, polar, e0 @ R.T, box, cutoff_ind=0.7, a_damp=0.39)
    assert np.allclose(d1, d2, rtol=1e-10, atol=1e-12 * np.abs(d1).max())
    assert np.abs(d1[:, 6:30]).max() > 0


def test_tmu_matches_the_dense_A_of_the_c_oracle(oracle_mod):
    """For the C oracle's converged dipoles A Pi = E0 with A = diag(1/alpha) - T,
    so Pi / alpha - T Pi must give back E0."""
    from swinmd_amd import systems as S
    s = S.water(64, density_mol=33.4, seed=5, lattice=(4, 4, 4), r_oh=0.09572, hoh_deg=108.5,
                types=S.AMOEBA_TYPES)
    o = oracle_mod.Oracle(s)
    d, qd = o.local2global()
    e0, pi, _, res = o.induced_dipoles(d, qd, cutoff_ind=0.7, a_damp=0.39)
    assert res < 1e-13
    polar = o.polarizabilities()
    t = G.tmu(s.pos, polar, pi, np.asarray(s.box, dtype=float), 0.7, 0.39)
    back = pi / polar[:, None] - t
    assert np.abs(back - e0).max() <= 1e-9 * np.abs(e0).max()
    # dropping the damping breaks it
    t0 = G.tmu(s.pos, polar, pi, np.asarray(s.box, dtype=float), 0.7, 0.39, damp=False)
    assert np.abs(pi / polar[:, None] - t0 - e0).max() > 1e-3 * np.abs(e0).max()


def test_guess_inverts_targets():
    rng = np.random.default_rng(11)
    pos, _, ms, _, _ = _cluster()
    F = G.site_frames(pos, ms)
    n = len(pos)
    polar = rng.uniform(0.0004, 0.0009, size=n)
    e0 = rng.normal(size=(n, 3))
    mu = rng.normal(size=(n, 3)) * 1e-3
    t = G.targets(F, polar, e0, mu)
    mean, sd = np.array([0.1, -0.2, 0.3]), np.array([2.0, 0.5, 1.5])
    back = G.guess(F, polar, e0, (t - mean) / sd, mean, sd)
    assert np.allclose(back, mu, atol=1e-15)


def test_tmu_rows_equals_tmu_on_those_rows():
    """The sampled T v used to check large solves is tmu's sum on its rows."""
    import numpy as np
    from oracle import learned_guess as G
    rng = np.random.default_rng(5)
    pos = rng.uniform(-0.8, 0.8, size=(300, 3))
    polar = rng.uniform(0.0004, 0.0009, size=300)
    v = rng.normal(size=(300, 3))
    box = np.array([1.6, 1.6, 1.6])
    rows = rng.choice(300, 40, replace=False)
    full = G.tmu(pos, polar, v, box, 0.6, 0.39)
    np.testing.assert_allclose(G.tmu_rows(pos, polar, v, box, 0.6, 0.39, rows), full[rows], rtol=1e-12, atol=1e-9)
