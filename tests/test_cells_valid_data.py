"""Batched single cells on the GPU against the reference's own ValidData.

cmx_cells_run with a voltage trace (trace_every) reproduces the protocols of
heart/test/ionicmodels/TestIonicModels.hpp -- the cell, its stimulus, its
ODE step and its sampling -- and its voltages are compared with the
reference's files (tests/golden/*ValidData.dat) at the reference's own
tolerance, 1e-3 mV at every sample (RunAndCheckIonicModels.hpp:196-221), in
FP64.  The FP32-rate mixed path is held to 0.5 mV.  Forward Euler throughout:
the files were written with it (the backward-Euler N98 file is checked on the
oracle, test_more_models.py).
"""
import numpy as np
import pytest

from conftest import load_dat

# name: (cmx model, file, dt, t_end, steps per sample, stimulus (mag, start, dur, period), parameters by name);
# Faber-Rudy's 1000 ms are not a whole number of its 0.007 ms steps: the file's
# last sample is at 997.5 ms, where these runs end
PROTOCOLS = {
    'lr91_delayed': ('LR91', 'Lr91DelayedStimValidData.dat', 0.01, 1000.0, 100, (-25.5, 50.0, 2.0, -1.0), {}),
    'lr91_regular': ('LR91', 'Lr91RegularStimValidData.dat', 0.01, 1000.0, 100, (-25.5, 50.0, 2.0, 500.0), {}),
    'tt06': ('TT06_EPI', 'TenTusscherValidData.dat', 0.001, 40.0, 1000, (-38.0, 5.0, 1.0, -1.0), {}),
    'hh52': ('HH52', 'HH52RegResultValidData.dat', 0.01, 150.0, 100, (-20.0, 10.0, 0.5, -1.0), {}),
    'fox2002': ('FOX2002', 'FoxRegularStimValidData.dat', 0.002, 200.0, 500, (-80.0, 50.0, 1.0, 500.0), {}),
    'mahajan2008': ('MAHAJAN2008', 'Mahajan2008ValidData.dat', 0.001, 80.0, 1000, (-1800.0, 10.0, 0.05, 1000.0), {}),
    'dn85': ('DN85', 'DiFrancescoNobleValidData.dat', 0.01, 1800.0, 1000, None, {}),
    'fr2000_delayed': ('FR2000', 'FR2000DelayedStimValidData.dat', 0.007, 997.5, 500, (-25.5, 10.0, 2.0, -1.0), {}),
    'fr2000_endo': ('FR2000', 'FR2000EndoValidData.dat', 0.007, 997.5, 500, (-25.5, 0.0, 2.0, -1.0),
                    {'ScaleFactorGks': 0.462, 'ScaleFactorIto': 0.0, 'ScaleFactorGkr': 1.0}),
    'fr2000_mid': ('FR2000', 'FR2000MidValidData.dat', 0.007, 997.5, 500, (-25.5, 0.0, 2.0, -1.0),
                   {'ScaleFactorGks': 1.154, 'ScaleFactorIto': 0.85, 'ScaleFactorGkr': 1.0}),
    'fr2000_epi': ('FR2000', 'FR2000EpiValidData.dat', 0.007, 997.5, 500, (-25.5, 0.0, 2.0, -1.0),
                   {'ScaleFactorGks': 1.154, 'ScaleFactorIto': 1.0, 'ScaleFactorGkr': 1.0}),
    'maleckar2008': ('MALECKAR2008', 'Maleckar2009ValidData.dat', 0.001, 25.0, 1000, (-5.6, 4.0, 6.0, 1000.0), {}),
    'n98': ('N98', 'N98RegResultValidData.dat', 0.01, 150.0, 100, (-3.0 / 0.095, 10.0, 3.0, -1.0), {}),
}


def _run(name, precision):
    from chaste_amd import cmx
    mname, fname, dt, t_end, every, stim, pnames = PROTOCOLS[name]
    model = getattr(cmx, mname)
    params = None
    if pnames:
        params = np.array([[d] for _, d in cmx.param_info(model)])
        for k, v in pnames.items():
            params[cmx.find_param(model, k), 0] = v
    _, _, st = cmx.cells_run(model, cmx.FORWARD_EULER, np.array([t_end]), dt,
                             stimuli=None if stim is None else [stim], params=params, precision=precision,
                             trace_every=every)
    hdr, ref = load_dat(fname)
    vcol = [k for k, h in enumerate(hdr) if h.startswith(('V(', 'membrane_voltage', 'membrane__V('))][0]
    return st['trace'][:, 0], ref[:, vcol]


def test_trace_argument_checks():
    from chaste_amd import cmx
    with pytest.raises(cmx.CmxError, match='trace_every'):
        cmx.cells_run(cmx.LR91, cmx.FORWARD_EULER, np.array([1.0]), 0.01, trace_every=-1)


@pytest.mark.gpu
@pytest.mark.parametrize('name', sorted(PROTOCOLS))
def test_cells_reproduce_reference_valid_data(name):
    from chaste_amd import cmx
    print('case valid data', name, flush=True)
    v, ref = _run(name, cmx.FP64)
    assert v.shape[0] == ref.shape[0]
    err = np.abs(v - ref).max()
    print('%s: max |V - ValidData| %.2e mV over %d samples' % (name, err, v.shape[0]))
    assert err < 1e-3
    vm, _ = _run(name, cmx.MIXED)
    errm = np.abs(vm - ref).max()
    print('%s mixed: max |V - ValidData| %.2e mV' % (name, errm))
    assert errm < 0.5
