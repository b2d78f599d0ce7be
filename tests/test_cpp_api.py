"""The C++ host API (include/chaste_amd/Monodomain.hpp) driven by a C++
program shaped like heart/test/monodomain/TestMonodomainProblem.hpp."""
import json
import os
import subprocess

import pytest


def _binary():
    from chaste_amd.build import cpp_test_path
    return cpp_test_path('TestMonodomainProblem')


def test_cpp_program_links_libcmx():
    path = _binary()
    assert os.path.exists(path), 'run __graft_entry__.build()'
    out = subprocess.run(['ldd', path], stdout=subprocess.PIPE, text=True).stdout
    assert 'libcmx.so' in out and 'not found' not in out


@pytest.mark.gpu
def test_cpp_monodomain_problem_goldens():
    print('case cpp TestMonodomainProblem', flush=True)
    from conftest import GOLDEN
    r = subprocess.run([_binary(), os.path.join(GOLDEN, 'meshes')], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=120)
    print(r.stdout)
    cases = [json.loads(l) for l in r.stdout.splitlines() if l.startswith('{')]
    assert {c['case'] for c in cases} == {'1d', '2d', '3d', 'errors', 'mesh_file_1d'}
    assert all(c['pass'] for c in cases)
    assert r.returncode == 0
