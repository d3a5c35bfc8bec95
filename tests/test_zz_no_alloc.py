"""GPU: fd_phase_mass_noalloc (the entry point that allocates nothing: its
partial sums live in the library's fixed device scratch) gives exactly what
fd_phase_mass gives, called through the C ABI - on two calls in a row, the
second on a smaller domain, and in every state a handle can be in (before
the first iteration, after fd_set_populations, after fused and single
launches).  A launch shape needing more partial-sum slots than the scratch
holds is refused before any launch."""
import ctypes

import numpy as np
import pytest

import bench
from fingering_dynamics_amd.params import LbmParams
from tests.helpers import rect_domain

pytestmark = pytest.mark.gpu


def _both(cm):
    lib = cm._lib
    a = np.zeros(3)
    b = np.zeros(3)
    rc = lib.fd_phase_mass(cm._h, a.ctypes.data_as(ctypes.c_void_p))
    assert rc == 0, lib.fd_last_error()
    rc = lib.fd_phase_mass_noalloc(cm._h, b.ctypes.data_as(ctypes.c_void_p))
    assert rc == 0, lib.fd_last_error()
    return a, b


def test_noalloc_equals_phase_mass_on_two_calls_in_a_row():
    from fingering_dynamics_amd.lbm import Compute
    mask, bits, params, _ = bench.workload(256, 512)
    big = Compute(mask, bits, params)
    H, W = 34, 70
    m2, b2, _ = rect_domain(H, W, 6, size=4, seed=7, x_min=14)
    small = Compute(m2, b2, LbmParams.periodic(H, W, M=5.0, psi_wall=0.0, n_inv_cols=12))
    # before the first iteration, then after odd and even step counts
    for n in (0, 7, 4):
        big.step(n)
        small.step(n)
        for cm in (big, small):            # large call, then a smaller one
            a, b = _both(cm)
            print("t=%d %dx%d phase_mass %r noalloc %r" % (cm.t, cm.H, cm.W, a, b))
            assert np.array_equal(a, b)
            assert b[2] == 0 and b[0] > 0
    # a state handed over with fd_set_populations
    f, g = small.populations()
    small.set_populations(f, g)
    a, b = _both(small)
    assert np.array_equal(a, b)
    small.close()
    big.close()


def test_noalloc_refuses_more_slots_than_its_scratch():
    from fingering_dynamics_amd.lbm import Compute
    H, W = 2100, 4096                      # one wave per 64 x 1 strip: 134,400 slots > 131,072
    mask = np.ones((H, W), dtype=bool)
    cm = Compute(mask, np.zeros((H, W), np.uint8), LbmParams.periodic(H, W, M=2.0, psi_wall=0.0))
    cm.step(2)
    cm.set_tile(1, 1)
    out = np.zeros(3)
    rc = cm._lib.fd_phase_mass_noalloc(cm._h, out.ctypes.data_as(ctypes.c_void_p))
    assert rc == -1 and "partial-sum slots" in cm._lib.fd_last_error().decode()
    cm.set_tile(4, 16)
    a, b = _both(cm)
    assert np.array_equal(a, b)
    cm.close()
