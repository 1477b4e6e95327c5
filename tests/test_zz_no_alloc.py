"""The allocation-free reciprocal-space entry point
(smd_ewald_compute_long_static: scratch in a static device array) against
smd_ewald_compute_long on the same context state, through the C ABI: two
calls in a row, the second on a smaller system, in both precisions; and the
clear error for a system whose scratch exceeds the array."""
import ctypes as C

import numpy as np
import pytest

from tests.helpers import rel

pytestmark = pytest.mark.gpu


def _long(ctx, static, prec):
    from swinmd_amd import _lib as L
    o = L.EwaldLongOut()
    ctx.zero_forces()
    if static:
        rc = ctx.lib.smd_ewald_compute_long_static(ctx.h, prec, C.byref(o))
    else:
        rc = ctx.lib.smd_ewald_compute_long(ctx.h, prec, None, C.byref(o))
    ctx.check(rc)
    return o, ctx.forces()


@pytest.mark.parametrize("prec", [0, 1])
def test_static_scratch_matches_the_allocating_entry_point(prec):
    from swinmd_amd import systems as S
    from swinmd_amd.forces import Context, EwaldForce
    systems = [S.water(512, density_mol=33.4, seed=5, lattice=(8, 8, 8)),
               S.water(216, box=1.9, seed=3, lattice=(6, 6, 6))]
    for s in systems:
        ctx = Context(s)
        ctx.compute_mol_centers()
        EwaldForce(ctx, alpha=3.2, real_cut=0.7, kcut=S.kcut_for(81, s.box[0]))
        want, fw = _long(ctx, False, prec)
        got, fg = _long(ctx, True, prec)
        assert np.array_equal(fg, fw)
        assert np.any(fw != 0.0)
        for key in ("energy", "deriv", "deriv2"):
            assert rel(getattr(got, key), getattr(want, key)) <= 1e-12, key
        assert np.max(np.abs(np.array(got.virial[:]) - np.array(want.virial[:]))) <= \
            1e-12 * np.max(np.abs(np.array(want.virial[:])))


def test_static_scratch_rejects_a_system_beyond_its_size():
    from swinmd_amd import _lib as L
    from swinmd_amd import systems as S
    from swinmd_amd.forces import Context, EwaldForce
    s = S.water(84000, density_mol=33.4, seed=9)      # 252,000 atoms: phases alone exceed 16 MiB
    ctx = Context(s)
    ctx.compute_mol_centers()
    EwaldForce(ctx, alpha=3.2, real_cut=0.7, kcut=S.kcut_for(9, s.box[0]))
    o = L.EwaldLongOut()
    assert ctx.lib.smd_ewald_compute_long_static(ctx.h, 0, C.byref(o)) == L.SMD_ECAPACITY
    assert b"static scratch" in ctx.lib.smd_last_error(ctx.h)
