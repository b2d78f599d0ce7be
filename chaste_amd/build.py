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
           'cmx_model_fox2002.hip', 'cmx_model_mahajan2008.hip', 'cmx_model_dn85.hip', 'cmx_model_fr2000.hip',
           'cmx_model_maleckar2008.hip', 'cmx_model_n98.hip']


CPP_TESTS = ['TestMonodomainProblem']


# The plugin the tests load: the reference's LuoRudy1991.cellml generated
# and built as a plugin (the built-in LR91 is the same model), found at run
# time through a path relative to the plugin itself.
TEST_PLUGIN_DIR = os.path.join(ROOT, 'tests', 'plugins')
TEST_PLUGIN_CELLML = '/root/reference/heart/src/odes/cellml/LuoRudy1991.cellml'


def test_plugin_path():
    return os.path.join(TEST_PLUGIN_DIR, 'liblr91_plugin.so')


def build_test_plugin():
    """Builds tests/plugins/liblr91_plugin.so when the CellML source is present."""
    if not os.path.exists(TEST_PLUGIN_CELLML):
        return None
    return build_model_plugin(TEST_PLUGIN_CELLML, TEST_PLUGIN_DIR, 'lr91_plugin',
                              params=('fast_sodium_current.g_Na', 'time_dependent_potassium_current.g_Kmax',
                                      'slow_inward_current.P_si'),
                              rpath='$ORIGIN/../../chaste_amd/lib')


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


def build_model_plugin(cellml, out_dir, name, params=(), rpath=None):
    """A cell model from a CellML file as a plugin for cmx_load_model: the
    model is generated into <out_dir>/<name>.h (namespace cmx::<name>; params:
    'component.variable' constants exposed as per-node parameters), a
    one-line source instantiates every kernel through cmx_model_plugin.h, and
    hipcc links it to libcmx.so as <out_dir>/lib<name>.so.  rpath: where the
    plugin finds libcmx.so at load time (default: this library's directory).
    Returns the plugin's path."""
    from chaste_amd.codegen import cellml2cpp
    if not name.isidentifier():
        raise ValueError('plugin name must be a C identifier')
    os.makedirs(out_dir, exist_ok=True)
    header = os.path.join(out_dir, name + '.h')
    argv = [cellml, header, name]
    for p in params:
        argv += ['--param', p]
    cellml2cpp.main(argv)
    src = os.path.join(out_dir, name + '_plugin.hip')
    with open(src, 'w') as f:
        f.write('#include "cmx_model_plugin.h"\n#include "%s.h"\n\nCMX_MODEL_PLUGIN(%s)\n' % (name, name))
    out = os.path.join(out_dir, 'lib%s.so' % name)
    lib_dir = rpath or LIBDIR
    _run([HIPCC] + HIP_FLAGS + ['-I' + out_dir, '-I' + os.path.join(CSRC, 'models'), '-shared', '-o', out + '.tmp', src, '-L' + LIBDIR, '-lcmx',
                                '-Wl,-rpath,' + lib_dir])
    os.replace(out + '.tmp', out)
    return out
