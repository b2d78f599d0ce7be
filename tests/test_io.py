"""On-disk inputs: Triangle/TetGen meshes and fibre files read through the C
ABI (TrianglesMeshReader / FibreReader semantics) against the oracle's own
reading of the ASCII files, the binary twins of the same meshes, and the
closed forms of the conductivity tensors.  Host code only (no GPU)."""
import os

import numpy as np
import pytest

from conftest import GOLDEN

MESHES = os.path.join(GOLDEN, 'meshes')


@pytest.mark.parametrize('name', ['1D_0_to_1mm_10_elements', '2D_0_to_1mm_400_elements', '2D_0_to_1mm_800_elements',
                                  '3D_0_to_1mm_6000_elements', '3D_Single_tetrahedron_element'])
def test_ascii_mesh_matches_oracle_reader(oracle, name):
    from chaste_amd import cmx
    base = os.path.join(MESHES, name)
    c, e, _ = cmx.read_mesh(base)
    oc, oe = oracle.read_triangles(base)
    assert np.array_equal(c, oc)
    assert np.array_equal(e, oe)
    assert e.min() == 0 and e.max() == c.shape[0] - 1


@pytest.mark.parametrize('name', ['1D_0_to_1mm_10_elements', '3D_0_to_1mm_6000_elements'])
def test_binary_mesh_matches_ascii(name):
    """The BIN files hold the same mesh (nodes as doubles, element node
    indices as 32-bit integers, attributes as doubles, numbered from 0)."""
    from chaste_amd import cmx
    c, e, a = cmx.read_mesh(os.path.join(MESHES, name))
    cb, eb, ab = cmx.read_mesh(os.path.join(MESHES, name + '_binary'))
    assert np.array_equal(c, cb)
    assert np.array_equal(e, eb)
    assert ab.shape[0] == e.shape[0]


def test_mesh_reader_errors(tmp_path):
    from chaste_amd import cmx
    base = str(tmp_path / 'bad')
    with open(base + '.node', 'w') as f:
        f.write('3 2 0 0\n0 0 0\n1 1 0\n')        # one node missing
    with open(base + '.ele', 'w') as f:
        f.write('1 3 0\n0 0 1 2\n')
    with pytest.raises(RuntimeError, match='unexpected end of file'):
        cmx.read_mesh(base)
    with open(base + '.node', 'w') as f:
        f.write('# comment\n3 2 0 0\n1 0 0\n2 1 0 # trailing comment\n\n3 0 1\n')
    with open(base + '.ele', 'w') as f:
        f.write('1 3 0\n1 1 2 3\n')
    c, e, _ = cmx.read_mesh(base)           # numbered from one
    assert np.array_equal(e, [[0, 1, 2]]) and np.array_equal(c, [[0, 0], [1, 0], [0, 1]])
    with pytest.raises(RuntimeError, match='cannot open'):
        cmx.read_mesh(str(tmp_path / 'missing'))


def test_orthotropic_fibres_and_tensors():
    """2D_0_to_1mm_800_elements.ortho: fibres along x for x < 0.05 and at 45
    degrees elsewhere (the file's own description); the tensors are
    diag(sigma_f, sigma_s) and R diag(sigma_f, sigma_s) R^T for the rotation by
    45 degrees, and every fibre/sheet pair is orthonormal."""
    from chaste_amd import cmx
    c, e, _ = cmx.read_mesh(os.path.join(MESHES, '2D_0_to_1mm_800_elements'))
    f = cmx.read_fibres(os.path.join(MESHES, '2D_0_to_1mm_800_elements.ortho'), 2)
    assert f.shape == (800, 2, 2)
    for a in f:
        assert np.allclose(a @ a.T, np.eye(2), atol=1e-8)
    sig = (1.75, 0.19)
    T = cmx.conductivity_tensors(f, sig)
    cx = c[e, 0].mean(axis=1)
    along = np.isclose(f[:, 0, 0], 1.0)
    assert along.sum() == 400 and (cx[along] < 0.05).all() and (cx[~along] > 0.05).all()
    assert np.allclose(T[along], np.diag(sig))
    r = np.sqrt(0.5)
    R = np.array([[r, -r], [r, r]])
    assert np.allclose(T[~along], R @ np.diag(sig) @ R.T, atol=1e-8)


def test_axisymmetric_fibres_and_tensor():
    """3D_Single_tetrahedron_element.axi holds the fibre (0, 1, 0): the tensor
    is sigma_t I + (sigma_f - sigma_t) f f^T = diag(sigma_t, sigma_f, sigma_t);
    the .ortho twin (fibre y, sheet z, normal x) gives diag(s_n, s_f, s_s)."""
    from chaste_amd import cmx
    f = cmx.read_fibres(os.path.join(MESHES, '3D_Single_tetrahedron_element.axi'), 3, axisymmetric=True)
    assert np.array_equal(f, [[0.0, 1.0, 0.0]])
    T = cmx.conductivity_tensors(f, (1.75, 0.19), axisymmetric=True)
    assert np.allclose(T[0], np.diag([0.19, 1.75, 0.19]))
    o = cmx.read_fibres(os.path.join(MESHES, '3D_Single_tetrahedron_element.ortho'), 3)
    T = cmx.conductivity_tensors(o, (1.75, 0.19, 0.05))
    assert np.allclose(T[0], np.diag([0.05, 1.75, 0.19]))


def test_fibre_file_errors(tmp_path):
    from chaste_amd import cmx
    p = str(tmp_path / 'f.ortho')
    with open(p, 'w') as fh:
        fh.write('2\n1 0 0 1\n1 0 0\n')
    with pytest.raises(RuntimeError, match='incomplete'):
        cmx.read_fibres(p, 2)
    with open(p, 'w') as fh:
        fh.write('1\n1 0 0 1 5\n')
    with pytest.raises(RuntimeError, match='Too many entries'):
        cmx.read_fibres(p, 2)
    with open(p, 'w') as fh:
        fh.write('1 2 3\n')
    with pytest.raises(RuntimeError, match='First \\(non comment\\) line'):
        cmx.read_fibres(p, 2)


def test_hdf5_reader_reads_reference_file():
    """The test reader against a file the real HDF5 library wrote (the
    reference's heart/test/data/Monodomain1d/MonodomainLR91_1d.h5: chunked
    layout, v1 messages) and the text twin the reference keeps beside it."""
    import h5lite
    f = h5lite.File(os.path.join(GOLDEN, 'MonodomainLR91_1d.h5'))
    d = f.dataset('Data')
    assert d.shape == (1001, 101, 1)
    assert list(d.attrs['Variable Details']) == ['V(mV)']
    assert list(f.dataset('Time').attrs['Unit']) == ['msecs']
    t = f.read('Time')
    assert np.allclose(t, np.arange(1001) * 0.01)
    x = f.read('Data')
    assert x[0, 0, 0] == -83.853 and abs(x.max() - 48.90776852039833) < 1e-12


def test_hdf5_writer_round_trip(tmp_path):
    """cmx_h5_* writes the reference's layout: read back by the reader that
    reads libhdf5's files, values bit for bit, attributes, and a file that
    ends where the superblock says."""
    import h5lite
    from chaste_amd import cmx
    rng = np.random.default_rng(3)
    data = rng.standard_normal((9, 17, 2))
    path = str(tmp_path / 'r.h5')
    with cmx.H5Writer(path, 17, [('V', 'mV'), ('Cai', 'mM')]) as w:
        for k in range(9):
            w.write(0.25 * k, data[k])
    f = h5lite.File(path)
    assert sorted(f.root) == ['Data', 'Time']
    assert np.array_equal(f.read('Data'), data)
    assert np.array_equal(f.read('Time'), 0.25 * np.arange(9))
    assert list(f.dataset('Data').attrs['Variable Details']) == ['V(mV)', 'Cai(mM)']
    assert int(f.dataset('Data').attrs['IsDataComplete'][0]) == 1
    assert list(f.dataset('Time').attrs['Unit']) == ['msecs']
    assert f.eof == os.path.getsize(path)
    # an empty file is still well formed
    with cmx.H5Writer(str(tmp_path / 'e.h5'), 3) as w:
        pass
    e = h5lite.File(str(tmp_path / 'e.h5'))
    assert e.dataset('Data').shape == (0, 3, 1)


def test_read_cable_in_3d():
    """mesh/test/data/1D_in_3D_0_to_1mm_10_elements: 11 nodes on the line
    y = 1, z = 0.5 and 10 line elements from its .edge file."""
    import numpy as np
    from chaste_amd import cmx
    base = os.path.join(GOLDEN, 'meshes', '1D_in_3D_0_to_1mm_10_elements')
    coords, elems, attr = cmx.read_mesh(base, elem_dim=1)
    assert coords.shape == (11, 3) and elems.shape == (10, 2)
    assert np.allclose(coords[:, 0], np.arange(11) * 0.01) and np.all(coords[:, 1] == 1.0) and np.all(coords[:, 2] == 0.5)
    assert np.array_equal(elems, np.stack([np.arange(10), np.arange(1, 11)], axis=1))
    with pytest.raises(RuntimeError):
        cmx.read_mesh(base)              # no .ele: volume elements are absent
