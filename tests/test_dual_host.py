"""Host check of the adjoint's vector forward mode (chaste_amd/csrc/models/
model_dual.h) on the generated LR91 and TT06 models: four tangents per pass
give bitwise the columns of one tangent per pass, and the columns agree with
central differences of the double-precision model (tests/cpp/test_dual.cpp)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_vector_forward_mode_matches_scalar_and_differences(tmp_path):
    exe = str(tmp_path / 'test_dual')
    subprocess.run(['g++', '-O1', '-std=c++17', '-I' + os.path.join(ROOT, 'chaste_amd', 'csrc'),
                    '-I' + os.path.join(ROOT, 'include'), os.path.join(ROOT, 'tests', 'cpp', 'test_dual.cpp'),
                    '-o', exe], check=True, timeout=300)
    r = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=60)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert r.stdout.count('0 differ') == 6
