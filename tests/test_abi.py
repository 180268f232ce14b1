"""CPU checks of the product boundary: the C-ABI library loads and exports
every function include/gwm_hip.h declares, and the host-only helpers
(terrain builders, stretched levels) behave.  No GPU is touched."""
import numpy as np
import pytest

import gwm_amd


def test_library_exports_every_declared_symbol():
    L = gwm_amd.lib()
    names = gwm_amd.header_functions()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_last_error_is_a_string():
    assert isinstance(gwm_amd.lib().gwm_last_error(), bytes)


def test_invalid_arguments_fail_loudly_without_a_gpu():
    with pytest.raises(gwm_amd.GwmError):
        gwm_amd.cosine_mountain(0, 4, 100.0)
    with pytest.raises(gwm_amd.GwmError):
        gwm_amd.stretched_levels(1, 100.0, 500.0, 16000.0)


def test_cosine_mountain_matches_reference_formula():
    # idealized_domain_builder.cpp:14-35, restated independently here
    nx, ny, h0, hx, hy = 21, 17, 1200.0, 0.25, 0.3
    t = gwm_amd.cosine_mountain(nx, ny, h0, hx=hx, hy=hy)
    i = np.arange(nx) / (nx - 1)
    j = np.arange(ny) / (ny - 1)
    X, Y = np.meshgrid(i, j)
    dx, dy = (X - 0.5) / hx, (Y - 0.5) / hy
    ref = np.where((np.abs(dx) >= 1) | (np.abs(dy) >= 1), 0.0,
                   h0 * (0.5 * (1 + np.cos(np.pi * dx))) * (0.5 * (1 + np.cos(np.pi * dy))))
    np.testing.assert_array_equal(t, ref.astype(np.float32))
    assert t.max() == np.float32(h0)


def test_agnesi_ridge_closed_form():
    t = gwm_amd.agnesi_ridge(256, 4, 2000.0, 256000.0, 10000.0, 400.0)
    x = (np.arange(256) + 0.5) * 2000.0 - 256000.0
    np.testing.assert_allclose(t[2], 400.0 / (1 + (x / 1.0e4) ** 2), rtol=1e-6)
    assert np.all(t[0] == t[3])


@pytest.mark.parametrize("nz,dzb,dzt,top", [(64, 100.0, 500.0, 16000.0), (60, 50.0, 600.0, 20000.0)])
def test_stretched_levels(nz, dzb, dzt, top):
    z = gwm_amd.stretched_levels(nz, dzb, dzt, top)
    dz = np.diff(z)
    assert z[0] == 0.0 and z[-1] == top
    assert np.all(dz > 0) and np.all(np.diff(dz) >= -1e-9)
    assert abs(dz[0] - dzb) < 0.5 and abs(dz[-1] - dzt) < 2.0
