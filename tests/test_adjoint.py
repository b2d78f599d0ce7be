"""The device adjoint (cmx_adjoint) against finite differences.

The loss is the squared voltage mismatch at observation nodes; its gradient
with respect to per-node conductance scalings (LR91 g_Na and g_Kmax) and to
the initial state is checked against FP64 central differences of
  * the oracle's forward run (an independent restatement of the same step),
  * the device's own forward run,
along random directions, to 1e-4 relative.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ADJ_REL = 1e-4


class Square:
    """A 2-D LR91 square (forward Euler, FP64, KSP rtol 1e-12) with a
    supra-threshold stimulus on the left edge, observed every `every` steps."""

    def __init__(self, ex, steps, n_obs, every, seed, integrator=0, model=0):
        from chaste_amd import cmx
        self.ex, self.steps, self.every = ex, steps, every
        self.coords, self.elems = cmx.slab_mesh(2, ex, ex, 0, 0.01)
        self.n = self.coords.shape[0]
        self.stim_index = np.where(self.coords[:, 0] < 0.02, 0, -1).astype(np.int32)
        self.model, self.integrator = model, integrator
        self.m = cmx.Monodomain(self.coords, self.elems, np.array([1.75, 1.75]), model, integrator,
                                stimuli=[(-400000.0, 0.0, 0.5)], stim_index=self.stim_index, precision=cmx.FP64,
                                ksp_rtol=1e-12, ksp_atol=0.0)
        self.y0 = self.m.state()
        rng = np.random.default_rng(seed)
        self.obs = np.sort(rng.choice(self.n, n_obs, replace=False)).astype(np.int32)
        self.defaults = np.array([d for _, d in cmx.param_info(model)])

    def params(self, scalings):
        p = np.repeat(self.defaults[:, None], self.n, axis=1)
        p[:scalings.shape[0]] *= scalings
        return p

    def observe(self, scalings=None, state0=None):
        """Device forward run: V at the observation nodes every `every` steps."""
        m = self.m
        m.set_state(self.y0 if state0 is None else state0)
        if scalings is not None:
            m.set_params(self.params(scalings))
        out = []
        for j in range(self.steps // self.every):
            m.step(j * self.every * 0.01, self.every)
            out.append(m.voltage()[self.obs])
        return np.array(out)

    def adjoint(self, scalings, data, checkpoint_every=100, state0=None):
        m = self.m
        m.set_state(self.y0 if state0 is None else state0)
        m.set_params(self.params(scalings))
        loss, gp, gs = m.adjoint(0.0, self.steps, self.obs, self.every, data, checkpoint_every)
        return loss, gp[:scalings.shape[0]] * self.defaults[:scalings.shape[0], None], gs


def central_difference(f, eps):
    """Richardson-extrapolated central difference of f at 0 (error O(eps^4))."""
    d1 = (f(eps) - f(-eps)) / (2 * eps)
    d2 = (f(eps / 2) - f(-eps / 2)) / eps
    return (4 * d2 - d1) / 3


def oracle_observe(oracle, sq, scalings):
    r = oracle.mono_run(2, sq.ex, sq.ex, model=oracle.LR91_CELLML, sigma=(1.75, 1.75, 1.75), stim_mag=-400000.0,
                        stim_dur=0.5, stim_xlim=0.02, t_end=sq.steps * 0.01, ksp_rtol=1e-12,
                        params=sq.params(scalings), trace_nodes=sq.obs)
    return r['trace'][sq.every - 1::sq.every]


def test_adjoint_matches_oracle_finite_differences(oracle):
    print('case adjoint vs oracle FD', flush=True)
    sq = Square(ex=20, steps=300, n_obs=12, every=10, seed=3)
    rng = np.random.default_rng(7)
    from chaste_amd.autograd import smooth_random_field
    truth = np.stack([smooth_random_field((21, 21), 0.01, 0.05, rng).ravel() for _ in range(2)])
    data = oracle_observe(oracle, sq, truth)
    s = np.ones((2, sq.n))
    loss, g, _ = sq.adjoint(s, data)
    ref_loss = ((oracle_observe(oracle, sq, s) - data) ** 2).sum()
    print('loss', loss, 'oracle', ref_loss)
    assert abs(loss - ref_loss) <= 1e-6 * ref_loss
    for k in range(4):
        d = rng.standard_normal((2, sq.n))
        fd = central_difference(lambda e: ((oracle_observe(oracle, sq, s + e * d) - data) ** 2).sum(), 1e-3)
        ad = float((g * d).sum())
        print('direction', k, 'adjoint', ad, 'fd', fd, 'rel', abs(ad - fd) / abs(fd))
        assert abs(ad - fd) <= ADJ_REL * abs(fd)


def test_adjoint_checkpoint_interval_invariance():
    print('case adjoint checkpoints', flush=True)
    sq = Square(ex=12, steps=120, n_obs=6, every=10, seed=5)
    data = sq.observe(np.full((2, sq.n), 0.9))
    s = np.ones((2, sq.n))
    ref = sq.adjoint(s, data, checkpoint_every=120)
    for L in (1, 7, 50):
        got = sq.adjoint(s, data, checkpoint_every=L)
        assert got[0] == ref[0]
        assert np.array_equal(got[1], ref[1]) and np.array_equal(got[2], ref[2])
    # the context ends where the forward run ends
    sq.m.set_state(sq.y0)
    sq.m.set_params(sq.params(s))
    sq.m.step(0.0, sq.steps)
    V_end = sq.m.voltage()
    sq.adjoint(s, data, checkpoint_every=7)
    assert np.array_equal(sq.m.voltage(), V_end)


def test_adjoint_two_level_checkpoints_match_one_level():
    """Under a memory bound too small for one level of checkpoints, each
    segment is replayed once more to save checkpoints every q steps and only q
    steps are kept at a time: the loss and both gradients are bitwise those of
    the one-level schedule; a bound below even that plan is refused."""
    from chaste_amd import cmx
    print('case adjoint two-level checkpoints', flush=True)
    sq = Square(ex=12, steps=120, n_obs=6, every=10, seed=5)
    data = sq.observe(np.full((2, sq.n), 0.9))
    s = np.ones((2, sq.n))
    ref = sq.adjoint(s, data, checkpoint_every=60)
    st = sq.m.stats()
    slot = st['n_states'] * ((sq.n + 63) // 64 * 64) * 8
    assert st['adjoint_sub_interval'] == 0 and st['adjoint_slots'] == 2 + 60 + 1
    # 2 checkpoints + 60/q + q + 1 states: q = 8 gives 2 + 8 + 8 + 1 = 19
    sq.m.set_adjoint_memory(20 * slot)
    got = sq.adjoint(s, data, checkpoint_every=60)
    st = sq.m.stats()
    assert st['adjoint_sub_interval'] > 0 and st['adjoint_slots'] <= 20
    assert got[0] == ref[0]
    assert np.array_equal(got[1], ref[1]) and np.array_equal(got[2], ref[2])
    sq.m.set_adjoint_memory(10 * slot)
    with pytest.raises(cmx.CmxError, match='exceed the memory bound'):
        sq.adjoint(s, data, checkpoint_every=60)
    sq.m.set_adjoint_memory(0)


@pytest.mark.parametrize('model,integrator', [(0, 0), (0, 1), (1, 1)], ids=['lr91_fe', 'lr91_rl', 'tt06_rl'])
def test_adjoint_initial_state_gradient(model, integrator):
    """dL/d(initial state) against central differences of the device's own
    forward run, along V and along every other state (one direction each)."""
    print('case adjoint initial state', model, integrator, flush=True)
    sq = Square(ex=10, steps=100, n_obs=8, every=5, seed=11, integrator=integrator, model=model)
    P = len(sq.defaults)
    s = np.ones((P, sq.n)) if P else np.zeros((0, sq.n))
    data = sq.observe(None) + 0.5
    loss, _, gs = sq.adjoint(s, data)
    rng = np.random.default_rng(2)
    scale = np.maximum(np.abs(sq.y0), 1e-6)
    fds, ads = [], []
    for plane in range(sq.y0.shape[0]):
        d = np.zeros_like(sq.y0)
        d[plane] = rng.standard_normal(sq.n) * scale[plane]
        fds.append(central_difference(lambda e: ((sq.observe(None, sq.y0 + e * d) - data) ** 2).sum(), 1e-4))
        ads.append(float((gs * d).sum()))
        print('plane', plane, 'adjoint', ads[-1], 'fd', fds[-1], flush=True)
    # each direction to 1e-4 of itself, or of a hundredth of the largest one
    # (states that barely move the loss are checked to that absolute level)
    big = max(abs(f) for f in fds)
    assert big > 0
    for ad, fd in zip(ads, fds):
        assert abs(ad - fd) <= ADJ_REL * max(abs(fd), 1e-2 * big)


def test_autograd_function_and_descent():
    """The torch.autograd.Function harness: backward() returns the adjoint
    gradient, and a step against it lowers the loss."""
    import torch

    from chaste_amd.autograd import AdjointProblem, ConductanceScalingLoss
    print('case adjoint autograd', flush=True)
    sq = Square(ex=12, steps=150, n_obs=10, every=10, seed=13)
    data = sq.observe(np.full((2, sq.n), 1.2))
    prob = AdjointProblem(sq.m, sq.y0, 0.0, sq.steps, sq.obs, sq.every, data, param_ids=[0, 1])
    s = torch.ones((2, sq.n), dtype=torch.float64, requires_grad=True)
    loss = ConductanceScalingLoss.apply(s, prob)
    loss.backward()
    _, g, _ = sq.adjoint(np.ones((2, sq.n)), data)
    assert np.array_equal(s.grad.numpy(), g)
    step = 0.05 / np.abs(g).max()
    with torch.no_grad():
        s2 = s - step * s.grad
    loss2 = ConductanceScalingLoss.apply(s2, prob)
    print('loss', float(loss.detach()), '->', float(loss2.detach()))
    assert float(loss2.detach()) < float(loss.detach())


def test_adjoint_rejects_unsupported_options():
    from chaste_amd import cmx
    coords, elems = cmx.slab_mesh(2, 4, 4, 0, 0.01)
    obs = np.array([0], dtype=np.int32)
    for kw in ({'precision': cmx.MIXED}, {'operator_splitting': True}, {'extrapolate': 1}):
        m = cmx.Monodomain(coords, elems, np.array([1.75, 1.75]), cmx.LR91, cmx.FORWARD_EULER, **kw)
        with pytest.raises(cmx.CmxError):
            m.adjoint(0.0, 10, obs, 5, np.zeros((2, 1)))
    m = cmx.Monodomain(coords, elems, np.array([1.75, 1.75]), cmx.LR91, cmx.BACKWARD_EULER)
    with pytest.raises(cmx.CmxError):
        m.adjoint(0.0, 10, obs, 5, np.zeros((2, 1)))


def test_adjoint_with_mass_lumping_and_fibre_tensors():
    """The adjoint's solve and M mu follow the context's operators: with a
    lumped mass matrix and per-element conductivity tensors the gradient
    still matches central differences of the device's forward run."""
    from chaste_amd import cmx
    print('case adjoint lumped', flush=True)
    coords, elems = cmx.slab_mesh(2, 14, 14, 0, 0.01)
    n = coords.shape[0]
    ang = np.radians(30.0)
    f = np.array([np.cos(ang), np.sin(ang)])
    sig = np.repeat((0.19 * np.eye(2) + (1.75 - 0.19) * np.outer(f, f))[None], elems.shape[0], axis=0)
    m = cmx.Monodomain(coords, elems, np.array([1.75, 0.19]), cmx.LR91, cmx.RUSH_LARSEN,
                       stimuli=[(-400000.0, 0.0, 0.5)], stim_index=np.where(coords[:, 0] < 0.02, 0, -1).astype(np.int32),
                       precision=cmx.FP64, ksp_rtol=1e-12, mass_lumping=True, sigma_elem=sig)
    y0 = m.state()
    defaults = np.array([d for _, d in cmx.param_info(cmx.LR91)])
    rng = np.random.default_rng(9)
    obs = np.sort(rng.choice(n, 10, replace=False)).astype(np.int32)
    steps, every = 150, 10

    def observe(scal):
        m.set_state(y0)
        m.set_params(defaults[:, None] * scal)
        out = []
        for j in range(steps // every):
            m.step(j * every * 0.01, every)
            out.append(m.voltage()[obs])
        return np.array(out)
    data = observe(np.full((3, n), 1.1))
    s = np.ones((3, n))
    m.set_state(y0)
    m.set_params(defaults[:, None] * s)
    loss, gp, _ = m.adjoint(0.0, steps, obs, every, data, checkpoint_every=40)
    g = gp * defaults[:, None]
    for k in range(3):
        d = rng.standard_normal((3, n))
        fd = central_difference(lambda e: ((observe(s + e * d) - data) ** 2).sum(), 1e-4)
        ad = float((g * d).sum())
        print('direction', k, 'adjoint', ad, 'fd', fd)
        assert abs(ad - fd) <= ADJ_REL * abs(fd)


def test_adjoint_config_full_size():
    """The adjoint workload at its full size: LR91 square of 256 x 256 nodes
    (h 0.01 cm, forward Euler, FP64, KSP rtol 1e-12), per-node scalings of
    g_Na and g_Kmax (2 x 65,536 parameters) with a seeded smooth Gaussian
    random field in [0.6, 1.4] (correlation length 0.1 cm) as the ground
    truth, 64 seeded observation nodes every 0.1 ms over 20 ms (2000 steps),
    checkpoints every 100 steps; the gradient at scalings 1 along 8 seeded
    random directions against Richardson-combined central differences of the
    device forward run within 1e-4.  A direction whose differences at eps and
    eps/2 disagree (LR91's h/j gate rates jump at V = -40 mV, so the loss jumps
    where a perturbation moves a node across it) is differenced at 1e-6."""
    from chaste_amd.autograd import smooth_random_field
    print('case adjoint full size 256x256, 2000 steps', flush=True)
    sq = Square(ex=255, steps=2000, n_obs=64, every=10, seed=2024)
    assert sq.n == 65536
    rng = np.random.default_rng(2024)
    truth = np.stack([smooth_random_field((256, 256), 0.01, 0.1, rng).ravel() for _ in range(2)])
    assert truth.min() >= 0.6 - 1e-12 and truth.max() <= 1.4 + 1e-12
    data = sq.observe(truth)
    s1 = np.ones((2, sq.n))
    loss, g, _ = sq.adjoint(s1, data, checkpoint_every=100)
    assert loss > 0

    def loss_at(s):
        return ((sq.observe(s) - data) ** 2).sum()
    rel = []
    for k in range(8):
        d = rng.standard_normal((2, sq.n))
        ad = float((g * d).sum())
        for eps in (1e-4, 1e-6):
            d1 = (loss_at(s1 + eps * d) - loss_at(s1 - eps * d)) / (2 * eps)
            d2 = (loss_at(s1 + eps / 2 * d) - loss_at(s1 - eps / 2 * d)) / eps
            if abs(d1 - d2) <= 1e-5 * abs(d2):
                break
        fd = (4 * d2 - d1) / 3
        rel.append(abs(ad - fd) / abs(fd))
        print('direction %d: adjoint %.9e  fd %.9e  rel %.2e (eps %g)' % (k, ad, fd, rel[-1], eps), flush=True)
    assert max(rel) <= ADJ_REL
