"""pairs_hip_energy_noalloc (the allocation-free variant of pairs_hip_energy,
partials in a static device array) against pairs_hip_energy through the C
ABI: bit-identical totals on two calls in a row, the second on a smaller
system, on full and half lists, and a clear error for a system beyond the
array's stated maximum."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(sim):
    pe, w = ctypes.c_double(), ctypes.c_double()
    code = sim.L.pairs_hip_energy_noalloc(sim.h, ctypes.byref(pe), ctypes.byref(w))
    return code, pe.value, w.value


def _state(n, half=False, steps=40):
    from pairs_amd import md
    s = md.fcc_system(n, temp=1.44, half=half)
    s.run(0, steps, 1000)
    return s


@pytest.mark.parametrize("half", [False, True])
def test_noalloc_energy_matches_energy(half):
    big, small = _state(8, half), _state(4, half)
    for sim in (big, small):  # the second call has the smaller input
        code, pe, w = _noalloc(sim)
        assert code == 0, sim.L.pairs_hip_last_error(sim.h)
        pe_ref, w_ref = sim.energy()
        print(f"{sim.nlocal} locals: noalloc ({pe!r}, {w!r}) energy ({pe_ref!r}, {w_ref!r})")
        assert pe == pe_ref and w == w_ref
        assert np.isfinite(pe) and pe < 0.0 and w != 0.0
    # and again in the other order, after the larger system wrote more partials
    pe_ref, w_ref = small.energy()
    code, pe, w = _noalloc(small)
    assert code == 0 and pe == pe_ref and w == w_ref
    big.close()
    small.close()


def test_noalloc_energy_rejects_beyond_its_scratch():
    from pairs_amd import md
    s = md.fcc_system(65, temp=1.44)  # 1098500 locals > 1048576
    code, _, _ = _noalloc(s)
    assert code == -2
    msg = s.L.pairs_hip_last_error(s.h).decode()
    assert "exceed" in msg and "1048576" in msg, msg
    s.close()
