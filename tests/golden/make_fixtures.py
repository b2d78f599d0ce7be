#!/usr/bin/env python3
"""Copies the reference's golden data files for this path into tests/golden/.

Run in a container where the reference is mounted at /root/reference.  The
files are data (time series written by the reference's own test runs), kept
verbatim:

* heart/test/data/ionicmodels/Lr91DelayedStimValidData.dat  (TestIonicModels.hpp:406-466)
* heart/test/data/ionicmodels/Lr91RegularStimValidData.dat  (TestIonicModels.hpp:468-500)
* heart/test/data/ionicmodels/TenTusscherValidData.dat      (TestIonicModels.hpp:793-835)

The tissue goldens of heart/test/monodomain/TestMonodomainProblem.hpp:784-880
(numbers written in that test's source) are transcribed into tissue_goldens.json.
"""
import json
import os
import shutil

REF = '/root/reference'
HERE = os.path.dirname(os.path.abspath(__file__))
FILES = ['Lr91DelayedStimValidData.dat', 'Lr91RegularStimValidData.dat', 'TenTusscherValidData.dat']


def main():
    for f in FILES:
        shutil.copy(os.path.join(REF, 'heart/test/data/ionicmodels', f), os.path.join(HERE, f))
    goldens = {
        'source': 'heart/test/monodomain/TestMonodomainProblem.hpp:784-880',
        'common': {'cell': 'LuoRudy1991 (FE)', 'stimulus': -3e6, 'stim_duration': 0.5, 'stim_nodes': 'x == 0',
                   'am': 1400, 'cm': 1.0, 'sigma': 1.75, 'dt': 0.01, 'ksp_atol': 2e-4, 'atol_check': 5e-3},
        'cases': [
            {'name': '1d', 'dim': 1, 'ex': 100, 'ey': 0, 'ez': 0, 't_end': 5.0,
             'nodes': [1, 3, 5, 7], 'V': [13.9682, 13.9149, 13.8216, 13.7275]},
            {'name': '2d', 'dim': 2, 'ex': 30, 'ey': 30, 'ez': 0, 't_end': 4.0,
             'nodes': [1, 3, 5, 7], 'V': [17.5728, 17.4562, 17.2559, 17.0440]},
            {'name': '3d', 'dim': 3, 'ex': 10, 'ey': 10, 'ez': 10, 't_end': 2.0,
             'nodes': [1, 3, 5, 7], 'V': [31.8958, 32.1109, 32.7315, 33.7011]},
        ],
    }
    with open(os.path.join(HERE, 'tissue_goldens.json'), 'w') as f:
        json.dump(goldens, f, indent=1)


if __name__ == '__main__':
    main()
