"""The ellipsoid-shell meshes of the batched-tissue benchmark: a valid,
conforming tetrahedralisation of the truncated shell (volume against the
closed form), fibres in the wall's tangent plane rotating -60 -> +60 degrees
through it; on the GPU, a shell with its fibre tensors against the oracle and
a disjoint union of shells against the shells run one at a time."""
import numpy as np
import pytest


def shell(h=0.25):
    from chaste_amd.shells import ellipsoid_shell
    return ellipsoid_shell(h)


def test_shell_is_a_conforming_tetrahedralisation():
    coords, elems, fibres, apex = shell(0.2)
    x = coords[elems]
    vol6 = np.einsum('ij,ij->i', np.cross(x[:, 1] - x[:, 0], x[:, 2] - x[:, 0]), x[:, 3] - x[:, 0])
    assert np.abs(vol6).min() > 1e-6 * np.abs(vol6).max()
    faces = np.sort(np.concatenate([elems[:, [0, 1, 2]], elems[:, [0, 1, 3]], elems[:, [0, 2, 3]],
                                    elems[:, [1, 2, 3]]]), axis=1)
    _, counts = np.unique(faces, axis=0, return_counts=True)
    assert counts.max() == 2                       # every face shared by at most two tetrahedra
    assert len(np.unique(elems)) == coords.shape[0]


def test_shell_volume_and_mass_matrix(oracle):
    """Truncated at z = 0 the shell is half of each ellipsoid: (2/3) pi (a_out
    b_out^2 - a_in b_in^2), less the thin apical opening theta < 0.1 and the
    chords' deficit; the P1 mass matrix sums to the mesh volume."""
    coords, elems, _, _ = shell(0.1)
    x = coords[elems]
    vol = np.abs(np.einsum('ij,ij->i', np.cross(x[:, 1] - x[:, 0], x[:, 2] - x[:, 0]), x[:, 3] - x[:, 0])).sum() / 6
    exact = 2.0 / 3.0 * np.pi * (3.5 * 2.0 ** 2 - 2.5 * 1.0 ** 2)
    assert abs(vol - exact) <= 0.02 * exact
    _, _, M, _ = oracle.assemble(coords, elems, np.eye(3))
    assert abs(M.sum() - vol) <= 1e-9 * vol


def test_shell_fibres_rotate_through_the_wall():
    """h = 0.2 gives 6 node layers through the 1 cm wall: element layer l has
    its centroid at s = (l + 1/2) / 5, so its fibre angle is -60 + 120 s:
    -48, -24, 0, 24, 48 degrees from the circumferential direction."""
    coords, elems, fibres, apex = shell(0.2)
    assert np.allclose(np.linalg.norm(fibres, axis=1), 1.0)
    x = coords[elems].mean(axis=1)
    phi = np.arctan2(x[:, 1], x[:, 0])
    e_phi = np.stack([-np.sin(phi), np.cos(phi), np.zeros_like(phi)], axis=1)
    ang = np.degrees(np.arccos(np.clip(np.abs(np.einsum('ij,ij->i', fibres, e_phi)), 0, 1)))
    ang *= np.sign(fibres[:, 2])                    # towards the base (+z) is positive
    # (a tetrahedron's centroid sits up to ~2 degrees of azimuth from its
    # hexahedron's centre, where the fibre was set: 3 degrees of slack)
    levels = np.array([-48.0, -24.0, 0.0, 24.0, 48.0])
    nearest = np.abs(ang[:, None] - levels[None, :]).min(axis=1)
    assert nearest.max() <= 3.0
    assert all((np.abs(ang - v) <= 1.0).any() for v in levels)
    assert apex.sum() > 0
    assert coords[apex, 2].max() <= -2.5 * np.cos(0.1) + 1e-9
    assert coords[:, 2].argmin() in np.where(apex)[0]


def run_shell_gpu(coords, elems, fibres, apex, steps, precision=0, rtol=None):
    from chaste_amd import cmx
    sig = cmx.conductivity_tensors(fibres, (1.75, 0.19), axisymmetric=True)
    m = cmx.Monodomain(coords, elems, np.array([1.75, 0.19, 0.19]), cmx.TT06_EPI, cmx.RUSH_LARSEN,
                       stimuli=[(-100000.0, 0.0, 1.0)], stim_index=np.where(apex, 0, -1).astype(np.int32),
                       precision=precision, ksp_rtol=rtol or (1e-12 if precision == 0 else 1e-8),
                       refine_sweeps=0 if precision == 0 else 1, sigma_elem=sig, use_graph=True)
    m.step(0.0, steps)
    return m, sig


@pytest.mark.gpu
def test_shell_gpu_matches_oracle(oracle):
    print('case shell vs oracle', flush=True)
    coords, elems, fibres, apex = shell(0.25)
    m, sig = run_shell_gpu(coords, elems, fibres, apex, 300)
    V = m.voltage()
    r = oracle.mono_run(3, 0, mesh=(coords, elems), model=oracle.TT06_EPI, rush_larsen=True, sigma=(1.75, 0.19, 0.19),
                        sigma_elem=sig, stim_mag=-100000.0, stim_dur=1.0, t_end=3.0, ksp_rtol=1e-12,
                        stim_mask=apex.astype(np.uint8))
    err = np.abs(V - r['V']).max() / np.abs(r['V']).max()
    print('shell nodes', len(V), 'format', m.stats()['format'], 'rel err', err, 'V max', V.max())
    assert V.max() > 0                              # the apical stimulus launched a wave
    assert err <= 1e-9


@pytest.mark.gpu
def test_shell_union_matches_shells_alone():
    """Two shells of different resolution stepped as one context (block-diagonal
    operators, one launch sequence) against each alone: the same solutions to
    the solves' tolerance (the union's CG stops on the joint residual)."""
    from chaste_amd.shells import disjoint_union, ellipsoid_shell
    print('case shell union', flush=True)
    a, b = ellipsoid_shell(0.3), ellipsoid_shell(0.22)
    coords, elems, fibres, apex, off = disjoint_union([a, b])
    mu, _ = run_shell_gpu(coords, elems, fibres, apex, 200)
    Vu = mu.voltage()
    for k, mesh in enumerate((a, b)):
        ms, _ = run_shell_gpu(*mesh, 200)
        Vs = ms.voltage()
        d = np.abs(Vu[off[k]:off[k + 1]] - Vs).max() / np.abs(Vs).max()
        print('shell', k, 'union vs alone', d)
        assert d <= 1e-9


@pytest.mark.gpu
def test_shell_regional_scaling_gradient():
    """The checkpointed adjoint on a shell (TT06 RL, FP64, fibre tensors, ELL):
    the gradient of a voltage mismatch with respect to regional scalings of
    g_Na and g_Kr (four regions by quadrant of azimuth) is the per-node
    gradient summed over each region; against central differences of the
    device's forward run, to 1e-4."""
    from chaste_amd import cmx
    print('case shell regional gradient', flush=True)
    coords, elems, fibres, apex = shell(0.3)
    n = coords.shape[0]
    sig = cmx.conductivity_tensors(fibres, (1.75, 0.19), axisymmetric=True)
    m = cmx.Monodomain(coords, elems, np.array([1.75, 0.19, 0.19]), cmx.TT06_EPI, cmx.RUSH_LARSEN,
                       stimuli=[(-100000.0, 0.0, 1.0)], stim_index=np.where(apex, 0, -1).astype(np.int32),
                       precision=cmx.FP64, ksp_rtol=1e-12, sigma_elem=sig)
    y0 = m.state()
    defaults = np.array([d for _, d in cmx.param_info(cmx.TT06_EPI)])
    region = ((np.arctan2(coords[:, 1], coords[:, 0]) + np.pi) // (np.pi / 2)).astype(int).clip(0, 3)
    rng = np.random.default_rng(4)
    obs = np.sort(rng.choice(n, 24, replace=False)).astype(np.int32)
    steps, every = 400, 20
    ids = (0, 2)                                     # g_Na, g_Kr

    def params(scal):                                # scal [2, 4]: per region
        p = np.repeat(defaults[:, None], n, axis=1)
        for k, q in enumerate(ids):
            p[q] *= scal[k][region]
        return p

    def observe(scal):
        m.set_state(y0)
        m.set_params(params(scal))
        out = []
        for j in range(steps // every):
            m.step(j * every * 0.01, every)
            out.append(m.voltage()[obs])
        return np.array(out)

    truth = np.array([[0.7, 1.0, 1.3, 1.0], [1.0, 1.4, 0.8, 1.0]])
    data = observe(truth)
    s = np.ones((2, 4))
    m.set_state(y0)
    m.set_params(params(s))
    loss, gp, _ = m.adjoint(0.0, steps, obs, every, data, checkpoint_every=50)
    g = np.array([[(gp[q] * defaults[q])[region == r].sum() for r in range(4)] for q in ids])
    print('loss', loss, 'regional gradient', g)
    assert loss > 0
    for k in range(2):
        for r in range(4):
            d = np.zeros((2, 4))
            d[k, r] = 1.0
            eps = 1e-4
            d1 = (((observe(s + eps * d) - data) ** 2).sum() - ((observe(s - eps * d) - data) ** 2).sum()) / (2 * eps)
            d2 = (((observe(s + eps / 2 * d) - data) ** 2).sum() - ((observe(s - eps / 2 * d) - data) ** 2).sum()) / eps
            fd = (4 * d2 - d1) / 3
            print('param', ids[k], 'region', r, 'adjoint', g[k, r], 'fd', fd)
            assert abs(g[k, r] - fd) <= 1e-4 * max(abs(fd), 1e-3 * np.abs(g).max())


@pytest.mark.gpu
def test_shell_union_at_config_size(oracle):
    """The batched-tissue configuration at its size: 8 shells at h from 0.04
    to 0.02 cm (0.5 M - 4 M nodes each, 13 M in all) stepped as one context
    for 200 steps (FP64, fibre tensors, ELL) against each shell run alone.

    The union's CG stops on the joint preconditioned residual, ||D^-1 r|| <=
    1e-12 ||D^-1 b||, so a shell's slice of it meets that shell's own test
    only at rtol 1e-12 sqrt(N / n_k) <= 5e-12 (its share of ||D^-1 b|| is
    about sqrt(n_k / N), N the union's nodes, n_k the shell's).  The bound
    each slice is held to is therefore measured, not guessed: the distance
    between the shell alone at rtol 1e-12 and at rtol 1e-11 (10x looser than
    the union can be), doubled because CG stops at whole iterations, so the
    distance is not monotone in the tolerance, and floored at 1e-9 relative
    to max |V|.  A slice taken from the wrong nodes, or a union assembled
    wrongly, differs by O(1).  The coarsest shell alone is also checked
    against the oracle after 100 steps (1e-9)."""
    from chaste_amd.shells import disjoint_union, ellipsoid_shell
    print('case shell union at config size', flush=True)
    hs = np.linspace(0.04, 0.02, 8)
    meshes = [ellipsoid_shell(h) for h in hs]
    coords, elems, fibres, apex, off = disjoint_union(meshes)
    n = coords.shape[0]
    print('union: %d nodes, %d tetrahedra' % (n, elems.shape[0]), flush=True)
    assert n > 12_000_000 and all(m[0].shape[0] > 300_000 for m in meshes)
    mu, _ = run_shell_gpu(coords, elems, fibres, apex, 200)
    Vu = mu.voltage()
    mu.close()
    del coords, elems, fibres, apex
    rows = []
    for k, mesh in enumerate(meshes):
        if k == 0:
            ms, sig = run_shell_gpu(*mesh, 100)
            c, e, f, a = mesh
            r = oracle.mono_run(3, 0, mesh=(c, e), model=oracle.TT06_EPI, rush_larsen=True, sigma=(1.75, 0.19, 0.19),
                                sigma_elem=sig, stim_mag=-100000.0, stim_dur=1.0, t_end=1.0, ksp_rtol=1e-12,
                                stim_mask=a.astype(np.uint8))
            err = np.abs(ms.voltage() - r['V']).max() / np.abs(r['V']).max()
            print('shell 0 alone vs oracle at 1 ms: %.3e' % err, flush=True)
            assert err <= 1e-9
            ms.step(1.0, 100)
        else:
            ms, _ = run_shell_gpu(*mesh, 200)
        Vs = ms.voltage()
        ms.close()
        ml, _ = run_shell_gpu(*mesh, 200, rtol=1e-11)
        Vl = ml.voltage()
        ml.close()
        scale = np.abs(Vs).max()
        d = np.abs(Vu[off[k]:off[k + 1]] - Vs).max() / scale
        dl = np.abs(Vl - Vs).max() / scale
        print('shell %d (h %.4f, %d nodes): union vs alone %.3e, alone rtol 1e-11 vs 1e-12 %.3e, V max %.1f' % (
            k, hs[k], Vs.shape[0], d, dl, Vs.max()), flush=True)
        rows.append((k, d, dl, Vs.max()))
    assert all(vmax > 0 for _, _, _, vmax in rows)          # every shell's apical stimulus launched a wave
    assert all(d <= max(1e-9, 2.0 * dl) for _, d, dl, _ in rows), rows
