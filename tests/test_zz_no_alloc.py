"""sp_bin_particles_noalloc (work space in the library's fixed device array) against
sp_bin_particles (work space allocated for the call): bit-identical sorted particles and
permutations, on two calls in a row, the second smaller than the first, and the limit."""
import numpy as np
import pytest

from sprout_amd import _native as N
from test_gpu_bin import keys_of, random_particles

pytestmark = pytest.mark.gpu


def run(entry, parts, W, H, tiled):
    n = len(parts)
    src = N.DeviceBuffer.of(parts)
    dst = N.DeviceBuffer(max(n, 1) * 48)
    perm = N.DeviceBuffer(max(n, 1) * 4)
    try:
        L = N.lib()
        rc = getattr(L, entry)(src.ptr, n, W, H, tiled, dst.ptr, perm.ptr, None)
        assert rc == N.OK, L.sp_error_text().decode()
        assert N.hip().hipDeviceSynchronize() == 0
        return dst.download(N.PARTICLE_DTYPE, n), perm.download(np.uint32, n)
    finally:
        for b in (src, dst, perm):
            b.close()


def test_noalloc_matches_the_allocating_entry_point():
    # two calls in a row on the same device array, the second with a smaller input
    for n, grid, tiled, seed in ((1_000_000, (1024, 576), 1, 3), (70_001, (384, 216), 0, 4)):
        W, H = grid
        parts = random_particles(n, W, H, seed=seed, near_sorted=(tiled == 1))
        want, want_perm = run("sp_bin_particles", parts, W, H, tiled)
        got, got_perm = run("sp_bin_particles_noalloc", parts, W, H, tiled)
        assert np.array_equal(got_perm, want_perm)
        assert got.tobytes() == want.tobytes()
        assert np.array_equal(got_perm, np.argsort(keys_of(parts, W, H, tiled), kind="stable").astype(np.uint32))


def test_noalloc_refuses_more_than_its_array_holds():
    L = N.lib()
    n = N.BIN_NOALLOC_MAX + 1
    buf = N.DeviceBuffer(96)
    try:
        # the size is checked before anything touches the (too small) buffers
        assert L.sp_bin_particles_noalloc(buf.ptr, n, 64, 36, 0, buf.ptr + 48 * n, None, None) == N.E_CAPACITY
        assert b"SP_BIN_NOALLOC_MAX" in L.sp_error_text()
    finally:
        buf.close()
