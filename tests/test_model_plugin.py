"""Cell models loaded at run time (cmx_load_model; the reference's
DynamicCellModelLoader, heart/src/odes/DynamicCellModelLoader.cpp:55-63).

The test plugin is the reference's LuoRudy1991.cellml generated and built by
chaste_amd.build.build_model_plugin (tests/plugins/liblr91_plugin.so, built
by __graft_entry__.build when the CellML file is present).  The built-in LR91
is the same model from the same generator, so the plugin must give the same
metadata, and on the GPU the same bits, in tissue and batched cells.
"""
import os
import re

import numpy as np
import pytest


def _plugin():
    from chaste_amd.build import test_plugin_path
    path = test_plugin_path()
    if not os.path.exists(path):
        pytest.skip('test plugin not built (LuoRudy1991.cellml absent at build time)')
    from chaste_amd import cmx
    return cmx.load_model(path)


def test_plugin_generated_header_is_the_builtin_model():
    """The plugin's generated header is the built-in LR91 header but for the
    namespace name."""
    from chaste_amd.build import CSRC, TEST_PLUGIN_DIR
    path = os.path.join(TEST_PLUGIN_DIR, 'lr91_plugin.h')
    if not os.path.exists(path):
        pytest.skip('test plugin not built')
    a = open(path).read().replace('lr91_plugin', 'lr91')
    b = open(os.path.join(CSRC, 'models', 'lr91.h')).read()
    assert re.sub(r'//.*\n', '\n', a) == re.sub(r'//.*\n', '\n', b)


def test_plugin_metadata_matches_builtin():
    from chaste_amd import cmx
    mid = _plugin()
    assert mid >= 64
    assert cmx.model_info(mid) == cmx.model_info(cmx.LR91)
    assert cmx.param_info(mid) == cmx.param_info(cmx.LR91)
    assert cmx.state_names(mid) == cmx.state_names(cmx.LR91)
    np.testing.assert_array_equal(cmx.initial_state(mid), cmx.initial_state(cmx.LR91))
    # loading again gives another id for the same model
    assert cmx.load_model(__import__('chaste_amd.build', fromlist=['x']).test_plugin_path()) != mid


def test_load_model_errors():
    from chaste_amd import cmx
    from chaste_amd.build import LIBDIR
    with pytest.raises(cmx.CmxError, match='cannot load'):
        cmx.load_model('/nonexistent/libmodel.so')
    with pytest.raises(cmx.CmxError, match='not a cell-model plugin'):
        cmx.load_model(os.path.join(LIBDIR, 'libcmx.so'))
    with pytest.raises(cmx.CmxError):
        cmx.model_info(63)          # below the plugin ids, above the built-in ones


@pytest.mark.gpu
@pytest.mark.parametrize('precision', [0, 1], ids=['fp64', 'mixed'])
@pytest.mark.parametrize('integ', ['fe', 'rl', 'be'])
def test_plugin_tissue_equals_builtin(precision, integ):
    from chaste_amd import cmx
    mid = _plugin()
    print('case plugin tissue', integ, precision, flush=True)
    ig = {'fe': cmx.FORWARD_EULER, 'rl': cmx.RUSH_LARSEN, 'be': cmx.BACKWARD_EULER}[integ]
    coords, elems = cmx.slab_mesh(2, 30, 10, 0, 0.01)
    stim = np.where(coords[:, 0] < 0.02, 0, -1).astype(np.int32)
    out = []
    for model in (cmx.LR91, mid):
        m = cmx.Monodomain(coords, elems, np.array([1.75, 1.75]), model, ig, stimuli=[(-500000.0, 0.0, 1.0)],
                           stim_index=stim, precision=precision, ksp_rtol=1e-10, refine_sweeps=1 if precision else 0,
                           use_graph=True)
        m.step(0.0, 200)
        out.append((m.voltage(), m.state()))
        m.close()
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])
    assert out[0][0].max() > 0.0          # the wave is up


@pytest.mark.gpu
def test_plugin_batched_cells_equal_builtin():
    from chaste_amd import cmx
    mid = _plugin()
    print('case plugin cells', flush=True)
    n = 64
    t_end = np.full(n, 400.0)
    stimuli = [(-25.5, 0.0, 2.0, 300.0 + 5.0 * c) for c in range(n)]
    a = cmx.cells_run(cmx.LR91, cmx.RUSH_LARSEN, t_end, 0.01, stimuli=stimuli, precision=cmx.FP64)
    b = cmx.cells_run(mid, cmx.RUSH_LARSEN, t_end, 0.01, stimuli=stimuli, precision=cmx.FP64)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert (a[1][0] >= 1).all()
