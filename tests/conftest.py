import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, 'tests', 'golden')


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an MI355X (runs the HIP kernels)')


def load_dat(name):
    """Column data file (header line + rows) -> (header, array)."""
    import numpy as np
    with open(os.path.join(GOLDEN, name)) as f:
        hdr = f.readline().split()
        rows = [l.split() for l in f if l.strip()]
    return hdr, np.array(rows, dtype=float)


@pytest.fixture(scope='session')
def oracle():
    from oracle import pyoracle
    pyoracle.lib()
    return pyoracle


def load_exnode(name):
    """Values of a cmgui .exnode data file (one field, one component), in node order."""
    vals = {}
    lines = open(os.path.join(GOLDEN, name)).read().split('\n')
    for i, line in enumerate(lines):
        if line.startswith('Node:'):
            vals[int(line.split()[1])] = float(lines[i + 1])
    import numpy as np
    return np.array([vals[k] for k in range(1, len(vals) + 1)])


# TestMonodomainProblem.hpp:851-975 (the 3-D slab 0.1 cm, h 0.01, LR91,
# -600*5000 uA/cm^3 for 0.5 ms at x = 0, 2 ms): the cmgui data file of printing
# step 43 (heart/test/data/CmguiData/monodomain/monodomain3dValidData.exnode,
# six significant digits) and V at nodes 0, 665, 1330 read back from its VTK
# output at printing steps 100 and 200 (atol 1e-3).
SLAB3D_VTK = {100: ([0, 665, 1330], [48.0637, 26.5404, -55.3058]),
              200: ([0, 665, 1330], [31.8730, 32.6531, 34.5152])}
SLAB3D_EXNODE = (43, 'monodomain3d_43.exnode')


def check_slab3d_geometry(v_by_step):
    """v_by_step: {printing step: V over all 1331 nodes}."""
    import numpy as np
    step, name = SLAB3D_EXNODE
    ref = load_exnode(name)
    # six significant digits: a half unit in the last place is at most 5e-4 mV
    assert ref.shape == v_by_step[step].shape == (1331,)
    assert np.abs(v_by_step[step] - ref).max() <= 1e-3
    for k, (nodes, vals) in SLAB3D_VTK.items():
        assert np.abs(v_by_step[k][nodes] - np.array(vals)).max() <= 1e-3


# TestMonodomainConductionVelocity.hpp:69-127 (TestMonodomainDg01DWith100elements):
# mesh/test/data/1D_0_to_1_100_elements, LR91, sigma 0.0005, Am = Cm = 1 (set
# after Initialise), the default plane stimulus (-600 uA/cm^3 for 0.5 ms at
# x = 0), 30 ms, printing every 0.01 ms, ActivationOutputModifier at 0 and
# -70 mV; its files heart/test/data/MonoConductionVel/activation_map_*.txt
# (first activation, first recovery, second activation, second recovery).
ACTIVATION_MAPS = {0.0: 'activation/activation_map_0.0.txt', -70.0: 'activation/activation_map_-70.0.txt'}


def load_activation_map(threshold):
    import numpy as np
    rows = [l.replace(',', ' ').split() for l in open(os.path.join(GOLDEN, ACTIVATION_MAPS[threshold])) if l.strip()]
    return np.array(rows, dtype=float)


def check_activation_map(got, threshold, dt=0.01, max_flips=2):
    """got[4][n] against the file.  Times are printing steps: a node whose
    voltage lies within a few uV of the threshold at a printing step may land
    one step either side (the reference's iterates differ from any other
    solver's at the 1e-3 mV level), so at most max_flips nodes may differ, by
    exactly one step.  Returns the nodes that did."""
    import numpy as np
    ref = load_activation_map(threshold).T
    assert got.shape == ref.shape
    assert np.array_equal(got < 0, ref < 0)
    d = np.abs(got - ref)
    flipped = np.where(d.max(axis=0) > 1e-9)[0]
    assert len(flipped) <= max_flips, (threshold, flipped)
    assert np.all(np.abs(d[:, flipped] - dt * (d[:, flipped] > 1e-9)) <= 1e-9)
    return flipped
