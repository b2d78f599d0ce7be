"""CPU checks of the C-ABI library: it loads, exports every symbol that
include/cmx.h declares, and its host-only helpers (slab mesh, model metadata)
agree with the oracle.  No GPU compute is called here."""
import os
import re

import numpy as np
import pytest

from conftest import ROOT


def declared_symbols():
    src = open(os.path.join(ROOT, 'include', 'cmx.h')).read()
    return sorted(set(re.findall(r'\b(cmx_[a-z0-9_]+)\s*\(', src)))


def test_exports_every_declared_symbol():
    from chaste_amd import cmx
    L = cmx.lib()
    names = declared_symbols()
    assert len(names) >= 14
    for name in names:
        assert hasattr(L, name), name
    assert set(names) == set(cmx.SIGNATURES)


def test_library_is_gfx950_code_object():
    from chaste_amd.build import library_path
    data = open(library_path(), 'rb').read()
    assert b'gfx950' in data


def test_slab_mesh_matches_oracle(oracle):
    from chaste_amd import cmx
    for args in [(1, 7, 0, 0), (2, 5, 4, 0), (3, 4, 3, 2)]:
        c1, e1 = cmx.slab_mesh(*args, h=0.01)
        c2, e2 = oracle.slab(*args, h=0.01)
        assert np.array_equal(c1, c2)
        assert np.array_equal(e1, e2)


def test_model_metadata():
    from chaste_amd import cmx
    assert cmx.model_info(cmx.LR91) == (8, 3)
    # the conductances LuoRudy1991.cellml annotates as modifiable parameters
    assert cmx.param_info(cmx.LR91) == [('fast_sodium_current__g_Na', 23.0),
                                        ('time_dependent_potassium_current__g_Kmax', 0.282),
                                        ('slow_inward_current__P_si', 0.09)]
    assert cmx.model_info(cmx.TT06_EPI)[0] == 19
    names = cmx.state_names(cmx.LR91)
    assert names[0] == 'membrane__V' and names[-1] == 'intracellular_calcium_concentration__Cai'
    y = cmx.initial_state(cmx.LR91)
    assert y[0] == -83.853 and y[7] == 0.0002


def test_entry_points_validate_before_touching_the_device():
    """Argument errors of the batched-cell, adjoint and autotune entry points
    come back as CMX_E_ARG with a message, before any device work."""
    import ctypes

    from chaste_amd import cmx
    for kw in ({'t_end': [1.0025], 'dt': 0.005}, {'t_end': [1.0], 'dt': 0.0}, {'t_end': [-1.0], 'dt': 0.01}):
        with pytest.raises(cmx.CmxError) as e:
            cmx.cells_run(cmx.TT06_EPI, cmx.RUSH_LARSEN, kw['t_end'], kw['dt'])
        assert e.value.code == -1 and cmx.lib().cmx_cells_last_error().decode()
    with pytest.raises(cmx.CmxError):
        cmx.cells_run(cmx.TT06_EPI, cmx.BACKWARD_EULER, [1.0], 0.01)
    with pytest.raises(cmx.CmxError):
        cmx.cells_run(cmx.LR91, cmx.FORWARD_EULER, [1.0], 0.01, params=np.ones((4, 1)))
    with pytest.raises(cmx.CmxError):
        cmx.cells_run(7, cmx.FORWARD_EULER, [1.0], 0.01)
    loss = np.zeros(1)
    assert cmx.lib().cmx_adjoint(None, 0.0, 10, 5, 0, None, 5, None, loss.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                 None, None) == -1
    assert cmx.lib().cmx_autotune(None, 0.0, 10, 0, None, None, None) == -1
    assert cmx.lib().cmx_set_params(None, 0, None) == -1
