"""Build recipe for libcmx.so (gfx950) and the C++ host layer."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, 'chaste_amd')
CSRC = os.path.join(PKG, 'csrc')
LIBDIR = os.path.join(PKG, 'lib')

HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
# Packed FP32 instructions stay off: hipcc -O3 emits v_pk_* from scalar code.
HIP_FLAGS = ['--offload-arch=gfx950', '-O3', '-std=c++17', '-fPIC',
             '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
             '-I' + os.path.join(ROOT, 'include'), '-I' + CSRC]

SOURCES = ['cmx_device.hip', 'cmx_adjoint.hip', 'cmx_cells.hip', 'cmx_host.cpp', 'cmx_io.cpp',
           'cmx_model_lr91.hip', 'cmx_model_tt06.hip', 'cmx_model_ord.hip', 'cmx_model_hh52.hip',
           'cmx_model_fox2002.hip', 'cmx_model_mahajan2008.hip']


CPP_TESTS = ['TestMonodomainProblem']


def cpp_test_path(name):
    return os.path.join(ROOT, 'tests', 'cpp', 'bin', name)


def library_path():
    return os.path.join(LIBDIR, 'libcmx.so')


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError('build failed: %s\n%s' % (' '.join(cmd), r.stdout))
    return r.stdout


def build_library(verbose=False):
    """Compile every HIP/C++ source (in parallel, objects under
    chaste_amd/lib/_build) and link chaste_amd/lib/libcmx.so."""
    obj_dir = os.path.join(LIBDIR, '_build')
    os.makedirs(obj_dir, exist_ok=True)
    out = library_path()
    tmp = out + '.tmp'
    objs, procs = [], []
    for s in SOURCES:
        obj = os.path.join(obj_dir, os.path.splitext(s)[0] + '.o')
        cmd = [HIPCC] + HIP_FLAGS + ['-c', '-o', obj, os.path.join(CSRC, s)]
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        objs.append(obj)
    log = ''
    for cmd, p in procs:
        text = p.communicate()[0]
        if p.returncode != 0:
            raise RuntimeError('build failed: %s\n%s' % (' '.join(cmd), text))
        log += text
    log += _run([HIPCC, '--offload-arch=gfx950', '-shared', '-o', tmp] + objs)
    os.replace(tmp, out)
    if verbose:
        print(log)
    return out


def build_cpp_tests():
    """C++ programs written against include/chaste_amd/Monodomain.hpp (host
    code only, linked to libcmx.so)."""
    os.makedirs(os.path.join(ROOT, 'tests', 'cpp', 'bin'), exist_ok=True)
    for name in CPP_TESTS:
        src = os.path.join(ROOT, 'tests', 'cpp', name + '.cpp')
        _run(['g++', '-O2', '-std=c++17', '-Wall', '-I' + os.path.join(ROOT, 'include'), src,
              '-L' + LIBDIR, '-lcmx', '-Wl,-rpath,$ORIGIN/../../../chaste_amd/lib', '-o', cpp_test_path(name)])
