"""CPU oracle of the learned initial guess for the induced dipoles.

Test infrastructure: only tests/, the smoke check and bench.py's CPU leg use
this module, as the checker; the library never imports it.

The reference has no learned guess: its ConjugateGradient starts from
Pi = 1 (src/forces/Induction.cpp:2404-2410) on A Pi = E0, with A = diag(1/alpha)
- T assembled in Induction::InducedDipoles (:181-455).  What is restated here
is this project's own definition of the guess, which the HIP kernels in
swinmd_amd/csrc/learned_guess.hip implement:

  site frames   one orthonormal frame per molecule (z along the H-O-H
                bisector, x along H1->H2 made perpendicular to z, y = z x x);
                H1 and O use it as is, H2 its mirror image (-x, y, z), so the
                two hydrogens see mirror-equivalent environments.
  descriptors   32 per site, all vectors in the site's frame:
                 0-2  alpha_i E0_i                  (the direct induced dipole)
                 3-5  alpha_i (T alpha E0)_i        (the next Neumann term; T
                      over every pair inside cutoffInd, intramolecular too,
                      as A couples them, :248-256, :296-361)
                 6-29 alpha_i sum_j q_j r_ij / r^3 over the intermolecular
                      partners j in four shells (r < 0.3, 0.4, 0.5, 0.6 nm)
                      times two partner kinds (q_j > 0, q_j <= 0), q in units
                      of e (scaled charge / SQRTCOULOMBCONSTANT, NEMD_defs.h:103)
                 30   alpha_i
                 31   1 for a site of positive charge, else 0
  network       32 -> 128 -> 128 -> 128 -> 3, SiLU after the three hidden
                layers; inputs are the descriptors standardised by per-feature
                mean and standard deviation; outputs are the standardised
                residual F_i (Pi_i - alpha_i E0_i).
  guess         Pi_i = alpha_i E0_i + F_i^T (sd * y_i + mean).
  training      mean squared error over the outputs of the rows, Adam.

The GPU computes the network in BF16 on the matrix cores with FP32
accumulation.  forward() and loss_grad() take emulate=True to round exactly
where the kernels round (inputs, weights, each hidden activation, each
back-propagated gradient fed to a product), so that the two agree to FP32
accumulation order; emulate=False is the plain FP64 network.

Pinning (tests/test_oracle_learned_guess.py):
  * loss_grad(emulate=False) against central finite differences of the loss;
  * tmu() against the C oracle's dense A: for its converged dipoles,
    Pi/alpha - T Pi = E0 to the CG tolerance;
  * site_frames(): orthonormality, handedness (+1, +1, -1), the bisector and
    in-plane axes, and the H1/H2 mirror relation;
  * charge_shells(): a hand-computed two-molecule geometry, and invariance
    of the frame-rotated descriptors under a rigid rotation of a cluster.
"""
import numpy as np

SQRT_COULOMB = 11.78708976        # SQRTCOULOMBCONSTANT, src/NEMD_defs.h:103
POLMIN = 0.00000001               # src/forces/Induction.cpp:195
SHELLS = (0.3, 0.4, 0.5, 0.6)     # nm
NF, NH, NO = 32, 128, 3
PUB_SHAPES = (("W1", (NH, NF)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)),
              ("W3", (NH, NH)), ("b3", (NH,)), ("W4", (NO, NH)), ("b4", (NO,)))
NPARAMS = sum(int(np.prod(s)) for _, s in PUB_SHAPES)   # 37635


# ---- BF16 ------------------------------------------------------------------

def bf16(a):
    """Round to the nearest BF16 value, ties to even (the kernels' conversion),
    returned as float64."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


# ---- network ---------------------------------------------------------------

def unpack(params):
    p = np.asarray(params, dtype=np.float64).ravel()
    if p.size != NPARAMS:
        raise ValueError(f"expected {NPARAMS} parameters, got {p.size}")
    out, o = [], 0
    for _, s in PUB_SHAPES:
        k = int(np.prod(s))
        out.append(p[o:o + k].reshape(s))
        o += k
    return out


def pack(arrays):
    return np.concatenate([np.asarray(a, dtype=np.float64).ravel() for a in arrays])


def init_params(seed, scale=1.0):
    """Uniform +-sqrt(6/fan_in) weights (last layer scaled by 0.1), zero biases."""
    rng = np.random.default_rng(seed)
    out = []
    for name, s in PUB_SHAPES:
        if name[0] == "W":
            lim = scale * np.sqrt(6.0 / s[1]) * (0.1 if name == "W4" else 1.0)
            out.append(rng.uniform(-lim, lim, size=s))
        else:
            out.append(np.zeros(s))
    return pack(out)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(params, x, emulate=True):
    """Outputs [rows, 3] and the cache backward needs."""
    r = bf16 if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(params)
    Ws = [r(W1), r(W2), r(W3), r(W4)]
    bs = [b1, b2, b3, b4]
    if emulate:     # the kernels hold biases in FP32
        bs = [b.astype(np.float32).astype(np.float64) for b in bs]
    h = r(np.asarray(x, dtype=np.float64))
    hs, zs = [h], []
    for l in range(3):
        z = h @ Ws[l].T + bs[l]
        zs.append(z)
        h = r(silu(z))
        hs.append(h)
    y = h @ Ws[3].T + bs[3]
    return y, (Ws, hs, zs)


def loss_grad(params, x, target, emulate=True):
    """Mean squared error over rows x 3 outputs and its gradient in the
    parameter order W1 b1 W2 b2 W3 b3 W4 b4."""
    r = bf16 if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    y, (Ws, hs, zs) = forward(params, x, emulate)
    t = np.asarray(target, dtype=np.float64)
    rows = y.shape[0]
    d = y - t
    loss = float(np.sum(d * d) / (3 * rows))
    dz = r(2.0 * d / (3 * rows))
    grads = [None] * 8
    for l in (3, 2, 1, 0):
        grads[2 * l] = dz.T @ hs[l]
        grads[2 * l + 1] = dz.sum(axis=0)
        if l == 0:
            break
        dh = dz @ Ws[l]
        dz = r(dh * dsilu(zs[l - 1]))
    return loss, pack(grads)


# ---- geometry ----------------------------------------------------------------

def _min_image(d, box):
    return d - box * np.round(d / box)


def site_frames(pos, mol_start):
    """[n, 3, 3]: rows are the site's local x, y, z axes (3-site waters H1, O, H2)."""
    pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
    ms = np.asarray(mol_start)
    if np.any(np.diff(ms) != 3):
        raise ValueError("site_frames: 3-site waters only")
    a0 = ms[:-1]
    h1, o, h2 = pos[a0], pos[a0 + 1], pos[a0 + 2]
    u1 = (h1 - o) / np.linalg.norm(h1 - o, axis=1)[:, None]
    u2 = (h2 - o) / np.linalg.norm(h2 - o, axis=1)[:, None]
    z = u1 + u2
    z /= np.linalg.norm(z, axis=1)[:, None]
    d = h2 - h1
    x = d - np.sum(d * z, axis=1)[:, None] * z
    x /= np.linalg.norm(x, axis=1)[:, None]
    y = np.cross(z, x)
    F = np.zeros((pos.shape[0], 3, 3))
    for s, sx in ((0, 1.0), (1, 1.0), (2, -1.0)):
        F[a0 + s, 0] = sx * x
        F[a0 + s, 1] = y
        F[a0 + s, 2] = z
    return F


def charge_shells(pos, q_scaled, mol, box):
    """[n, 8, 3] lab-frame sums of q_j r_ij / r^3 (q in e) over partners of
    other molecules: bin 2 * shell + kind, kind 0 for q_j > 0."""
    pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
    q = np.asarray(q_scaled, dtype=np.float64) / SQRT_COULOMB
    mol = np.asarray(mol)
    box = np.asarray(box, dtype=np.float64)
    n = pos.shape[0]
    out = np.zeros((n, 8, 3))
    edges2 = np.array(SHELLS) ** 2
    for i0 in range(0, n, 512):
        i1 = min(n, i0 + 512)
        d = _min_image(pos[i0:i1, None, :] - pos[None, :, :], box)
        r2 = np.sum(d * d, axis=2)
        other = mol[i0:i1, None] != mol[None, :]
        shell = np.searchsorted(edges2, r2, side="right")     # 0..4, 4 = outside
        use = other & (shell < 4)
        r3 = np.where(use, r2 * np.sqrt(np.where(use, r2, 1.0)), 1.0)
        e = np.where(use[:, :, None], q[None, :, None] * d / r3[:, :, None], 0.0)
        kind = (q <= 0).astype(int)
        b = 2 * shell + kind[None, :]
        for k in range(8):
            out[i0:i1, k] = np.sum(np.where((b == k)[:, :, None] & use[:, :, None], e, 0.0), axis=1)
    return out


def tmu(pos, polar, v, box, cutoff_ind, a_damp, damp=True):
    """(T v)_i = sum_j T_ij v_j over every pair inside cutoffInd, the Thole
    damped dipole tensor of Induction::InducedDipoles (:296-313, :317-361)."""
    pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
    v = np.asarray(v, dtype=np.float64).reshape(-1, 3)
    p = np.maximum(np.asarray(polar, dtype=np.float64), POLMIN)
    box = np.asarray(box, dtype=np.float64)
    n = pos.shape[0]
    out = np.zeros((n, 3))
    for i0 in range(0, n, 512):
        i1 = min(n, i0 + 512)
        d = _min_image(pos[i0:i1, None, :] - pos[None, :, :], box)
        r2 = np.sum(d * d, axis=2)
        idx = np.arange(i0, i1)
        use = (r2 <= cutoff_ind * cutoff_ind)
        use[np.arange(i1 - i0), idx] = False
        r = np.sqrt(np.where(use, r2, 1.0))
        r3, r5 = 1.0 / r ** 3, 1.0 / r ** 5
        if damp:
            u3 = r2 * r / np.sqrt(p[i0:i1, None] * p[None, :])
            e = np.exp(-a_damp * u3)
            l3 = 1.0 - e
            l5 = 1.0 - (1.0 + a_damp * u3) * e
        else:
            l3 = l5 = np.ones_like(r)
        dv = np.sum(d * v[None, :, :], axis=2)
        t = 3.0 * l5 * r5 * dv
        contrib = t[:, :, None] * d - (l3 * r3)[:, :, None] * v[None, :, :]
        out[i0:i1] = np.sum(np.where(use[:, :, None], contrib, 0.0), axis=1)
    return out


def tmu_rows(pos, polar, v, box, cutoff_ind, a_damp, rows, damp=True):
    """(T v)_i of tmu for the sites `rows` only (the same sum, restricted to a
    sample of rows so that a large system can be checked)."""
    pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
    v = np.asarray(v, dtype=np.float64).reshape(-1, 3)
    p = np.maximum(np.asarray(polar, dtype=np.float64), POLMIN)
    box = np.asarray(box, dtype=np.float64)
    rows = np.asarray(rows)
    out = np.zeros((len(rows), 3))
    for b0 in range(0, len(rows), 256):
        idx = rows[b0:b0 + 256]
        d = _min_image(pos[idx, None, :] - pos[None, :, :], box)
        r2 = np.sum(d * d, axis=2)
        use = (r2 <= cutoff_ind * cutoff_ind)
        use[np.arange(len(idx)), idx] = False
        r = np.sqrt(np.where(use, r2, 1.0))
        r3, r5 = 1.0 / r ** 3, 1.0 / r ** 5
        if damp:
            u3 = r2 * r / np.sqrt(p[idx, None] * p[None, :])
            e = np.exp(-a_damp * u3)
            l3 = 1.0 - e
            l5 = 1.0 - (1.0 + a_damp * u3) * e
        else:
            l3 = l5 = np.ones_like(r)
        dv = np.sum(d * v[None, :, :], axis=2)
        contrib = (3.0 * l5 * r5 * dv)[:, :, None] * d - (l3 * r3)[:, :, None] * v[None, :, :]
        out[b0:b0 + len(idx)] = np.sum(np.where(use[:, :, None], contrib, 0.0), axis=1)
    return out


def descriptors(pos, mol_start, mol, q_scaled, polar, e0, box, cutoff_ind, a_damp):
    """[n, 32] raw descriptors (the order of the module docstring)."""
    F = site_frames(pos, mol_start)
    alpha = np.asarray(polar, dtype=np.float64)
    e0 = np.asarray(e0, dtype=np.float64).reshape(-1, 3)
    ae0 = alpha[:, None] * e0
    tae0 = tmu(pos, polar, ae0, box, cutoff_ind, a_damp)
    sh = charge_shells(pos, q_scaled, mol, box)
    n = F.shape[0]
    out = np.zeros((n, NF))
    out[:, 0:3] = np.einsum("nab,nb->na", F, ae0)
    out[:, 3:6] = np.einsum("nab,nb->na", F, alpha[:, None] * tae0)
    out[:, 6:30] = np.einsum("nab,nkb->nka", F, alpha[:, None, None] * sh).reshape(n, 24)
    out[:, 30] = alpha
    out[:, 31] = (np.asarray(q_scaled) > 0).astype(np.float64)
    return out


def targets(frames, polar, e0, mu):
    """F_i (Pi_i - alpha_i E0_i)."""
    d = np.asarray(mu).reshape(-1, 3) - np.asarray(polar)[:, None] * np.asarray(e0).reshape(-1, 3)
    return np.einsum("nab,nb->na", frames, d)


def guess(frames, polar, e0, y, out_mean, out_sd):
    loc = np.asarray(y) * np.asarray(out_sd)[None, :] + np.asarray(out_mean)[None, :]
    return np.asarray(polar)[:, None] * np.asarray(e0).reshape(-1, 3) + np.einsum("nab,na->nb", frames, loc)
