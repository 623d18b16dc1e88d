"""gd_score_poses_noalloc (static device workspace, no allocation) against
gd_score_poses, through the C ABI: bit-identical FP32 and FP64 energies on
two calls in a row, the second smaller than the first, so that the second
call runs over what the first left in the workspace."""
import numpy as np
import pytest

from gpudock_amd import remc, synth


def _poses(rng, n, n_lig, n_prt):
    lc = rng.integers(0, n_lig, n).astype(np.int32)
    pc = rng.integers(0, n_prt, n).astype(np.int32)
    mv = np.concatenate([rng.normal(0, 0.6, (n, 3)), rng.normal(0, 1.0, (n, 3))], 1)
    return lc, pc, mv.astype(np.float32)


@pytest.mark.gpu
def test_noalloc_equals_allocating_entry_point():
    rng = np.random.default_rng(31)
    big = synth.make_config("c1", n_lig=16, n_prt=4)[0]        # pocket 800, ligand 48
    small = synth.make_config("c0", n_lig=3, n_prt=2, seed=5)[0]   # pocket 300, ligand 24
    for cx, n in ((big, 3000), (small, 200)):
        size = cx.size
        lc, pc, mv = _poses(rng, n, size.n_lig, size.n_prt)
        for term_order in (0, 1):
            e, e64 = remc.score_poses(cx, lc, pc, mv, term_order=term_order)
            f, f64 = remc.score_poses(cx, lc, pc, mv, term_order=term_order, noalloc=True)
            np.testing.assert_array_equal(f, e)
            np.testing.assert_array_equal(f64, e64)
            assert np.all(np.isfinite(f))


@pytest.mark.gpu
def test_noalloc_fp32_only_then_fp64_only():
    cx = synth.make_config("c0", n_lig=2, n_prt=1)[0]
    lc, pc, mv = _poses(np.random.default_rng(2), 64, 2, 1)
    e, e64 = remc.score_poses(cx, lc, pc, mv)
    f = remc.score_poses(cx, lc, pc, mv, fp64=False, noalloc=True)
    np.testing.assert_array_equal(f, e)
    out = np.zeros((64, 10), np.float64)
    remc.check(remc.lib().gd_score_poses_noalloc(
        *cx.args(), cx.size, 64, remc._ip(lc), remc._ip(pc), remc._fp(mv), 0, 0, None,
        out.ctypes.data_as(remc.C.POINTER(remc.C.c_double))))
    np.testing.assert_array_equal(out, e64)


def test_noalloc_refuses_oversize_before_the_gpu():
    """More poses than the 64 MiB workspace holds: a clear error before any
    GPU call (so it runs without a GPU too)."""
    cx = synth.make_config("c0", n_lig=1, n_prt=1)[0]
    n = 500000            # 500k poses x 184 B > 64 MiB
    lc = np.zeros(n, np.int32)
    mv = np.zeros((n, 6), np.float32)
    with pytest.raises(remc.GDError, match="static workspace"):
        remc.score_poses(cx, lc, lc, mv, noalloc=True)
