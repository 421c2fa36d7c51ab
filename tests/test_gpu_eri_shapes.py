"""Launch shapes of the stored ERI build (lanes per shell quartet x threads
per workgroup, csrc/eri.hip) and their search (myqc_eri_tune): every shape
stores the same integrals, so G from the store is unchanged (to the last
bits the LDS-atomic order allows), and the search returns a valid shape for
every class present.  The integrals themselves are pinned to the
reference's int2e by the parity tests; this checks that splitting a
quartet's primitive quartets over lanes (src/integrals/int2e.f90:192-284
sums them in one loop) gives the same sums."""
import numpy as np
import pytest

from myqc_amd import molecules as M

pytestmark = pytest.mark.gpu

DEFAULT = {c: (1, 256) for c in range(6)}


@pytest.fixture(scope="module")
def w12(ctx):
    Z, x = M.water_cluster(12, seed=12)
    return Z, M.to_bohr(Z, x)


def _G(ctx, D):
    ctx.eri_build(1e-12)
    return ctx.fock_G(D)


@pytest.mark.parametrize("split", [2, 4, 8])
def test_split_shapes_store_the_same_integrals(ctx, w12, split):
    Z, xyz = w12
    ctx.set_system(Z, xyz)
    n = ctx.n
    rng = np.random.default_rng(split)
    D = rng.uniform(-1, 1, (n, n))
    D = D + D.T
    try:
        ctx.eri_set_shapes(DEFAULT)
        G1 = _G(ctx, D)
        for tpb in (64, 128, 256):
            ctx.eri_set_shapes({c: (split, tpb) for c in range(6)})
            G2 = _G(ctx, D)
            err = np.max(np.abs(G2 - G1))
            print(f"[shapes] {split} lanes, {tpb} threads: max|dG| = {err:.2e}", flush=True)
            assert err <= 1e-12 * np.max(np.abs(G1))
    finally:
        ctx.eri_set_shapes(DEFAULT)


def test_tune_keeps_a_valid_shape_per_class(ctx, w12):
    Z, xyz = w12
    ctx.set_system(Z, xyz)
    n = ctx.n
    D = np.eye(n)
    try:
        ctx.eri_set_shapes(DEFAULT)
        G1 = _G(ctx, D)
        shapes, ms = ctx.eri_tune(reps=1)
        print(f"[shapes] tuned: {shapes}; ms {ms}", flush=True)
        for c, (sp, tpb) in shapes.items():
            assert sp in (1, 2, 4, 8) and tpb in (64, 128, 256)
            assert ms[c] > 0.0
        G2 = _G(ctx, D)
        assert np.max(np.abs(G2 - G1)) <= 1e-12 * np.max(np.abs(G1))
        from myqc_amd.hip import MyqcError
        with pytest.raises(MyqcError):
            ctx.eri_set_shapes({0: (3, 256)})
    finally:
        ctx.eri_set_shapes(DEFAULT)
