"""ctypes wrapper of the FP64 CPU oracle -- TEST INFRASTRUCTURE ONLY.

Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg,
always as the checker or the timed CPU baseline, never by chaste_amd.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'liboracle.so')

LR91_HAND, LR91_CELLML, TT06_EPI, ORD_ENDO, HH52, FOX2002, MAHAJAN2008, DN85, FR2000, MALECKAR2008, N98 = range(11)
FE, RL, BE, GRL1, GRL2 = 0, 1, 2, 3, 4     # orc_integrator

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class MonoCfg(ctypes.Structure):
    _fields_ = [('dim', ctypes.c_int), ('ex', ctypes.c_int), ('ey', ctypes.c_int), ('ez', ctypes.c_int),
                ('h', ctypes.c_double), ('model', ctypes.c_int), ('integrator', ctypes.c_int),
                ('sigma', ctypes.c_double * 3), ('am', ctypes.c_double), ('cm', ctypes.c_double),
                ('dt_pde', ctypes.c_double), ('dt_ode', ctypes.c_double), ('stim_mag', ctypes.c_double),
                ('stim_dur', ctypes.c_double), ('stim_start', ctypes.c_double), ('stim_xlim', ctypes.c_double),
                ('t_end', ctypes.c_double), ('ksp_rtol', ctypes.c_double), ('ksp_atol', ctypes.c_double),
                ('max_its', ctypes.c_int), ('threads', ctypes.c_int), ('n_param', ctypes.c_int), ('params', _dp),
                ('lumped', ctypes.c_int), ('sigma_elem', _dp), ('stim_period', ctypes.c_double),
                ('modifier_out', _dp), ('modifier_threshold', ctypes.c_double), ('modifier_every', ctypes.c_int),
                ('split', ctypes.c_int), ('svi', ctypes.c_int), ('stim_mask', ctypes.POINTER(ctypes.c_ubyte)),
                ('mesh_n_nodes', ctypes.c_int), ('mesh_n_elems', ctypes.c_int), ('mesh_coords', _dp),
                ('mesh_elems', _ip), ('node_model', _ip), ('node_integrator', _ip)]


class MonoStats(ctypes.Structure):
    _fields_ = [('t_ode', ctypes.c_double), ('t_rhs', ctypes.c_double), ('t_ksp', ctypes.c_double),
                ('t_total', ctypes.c_double), ('cg_iterations', ctypes.c_longlong), ('steps', ctypes.c_int),
                ('n_nodes', ctypes.c_int), ('nnz', ctypes.c_int)]


_L = None


def build():
    subprocess.run(['make', '-s', '-C', HERE], check=True)


def lib():
    global _L
    if _L is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.orc_model_num_states.argtypes = [ctypes.c_int]
        L.orc_cell_compute.restype = ctypes.c_int
        L.orc_cell_compute.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_double, ctypes.c_int, _dp, ctypes.c_int]
        L.orc_model_initial_state.argtypes = [ctypes.c_int, _dp]
        L.orc_cell_compute_v.restype = ctypes.c_int
        L.orc_cell_compute_v.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp]
        L.orc_cell_compute_except_voltage.restype = ctypes.c_double
        L.orc_cell_compute_except_voltage.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.c_double,
                                                      ctypes.c_double, ctypes.c_double]
        L.orc_cell_iionic.restype = ctypes.c_double
        L.orc_cell_iionic.argtypes = [ctypes.c_int, _dp, _dp]
        L.orc_slab_sizes.argtypes = [ctypes.c_int] * 4 + [_ip, _ip]
        L.orc_slab_fill.argtypes = [ctypes.c_int] * 4 + [ctypes.c_double, _dp, _ip]
        L.orc_assemble.argtypes = [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int, _ip, _dp, ctypes.c_int,
                                   ctypes.c_int, _ip, _ip, _ip, _dp, _dp]
        L.orc_pcg.argtypes = [ctypes.c_int, _ip, _ip, _dp, _dp, _dp, ctypes.c_double, ctypes.c_double, ctypes.c_int]
        L.orc_mono_run.argtypes = [ctypes.POINTER(MonoCfg), _dp, _dp, _dp, ctypes.c_double, ctypes.c_int, _ip, _dp,
                                   ctypes.POINTER(MonoStats)]
        _L = L
    return _L


def _p(a, t=_dp):
    return None if a is None else a.ctypes.data_as(t)


def num_states(model):
    return lib().orc_model_num_states(model)


def _integrator(rush_larsen, backward_euler, integrator):
    if integrator is not None:
        return int(integrator)
    return BE if backward_euler else int(rush_larsen)


def cell_compute(model, y, t0, t_end, dt, stim=(0.0, 0.0, 0.0, -1.0), rush_larsen=False,
                 steps_per_sample=1, params=None, backward_euler=False, integrator=None):
    """AbstractCardiacCell::Compute: returns samples[k, S]; y updated in place."""
    S = num_states(model)
    nsteps = int(round((t_end - t0) / dt))
    cap = nsteps // steps_per_sample + 2
    out = np.zeros((cap, S))
    pa = None if params is None else np.ascontiguousarray(params, dtype=np.float64)
    ns = lib().orc_cell_compute(model, _integrator(rush_larsen, backward_euler, integrator), _p(pa), _p(y), t0, t_end, dt, stim[0], stim[1], stim[2],
                                stim[3], steps_per_sample, _p(out), cap)
    if ns < 0:
        raise RuntimeError('orc_cell_compute failed')
    return out[:ns]


def model_initial_state(model):
    y = np.zeros(64)
    S = lib().orc_model_initial_state(model, _p(y))
    if S < 0:
        raise ValueError('unknown model %d' % model)
    return y[:S].copy()


def cell_trace_v(model, y, t_end, dt, stim=(0.0, 0.0, 0.0, -1.0), integrator=0, params=None):
    """Single cell from t = 0: V at t = 0 and after every step; y updated in place."""
    n = int(round(t_end / dt)) + 2
    out = np.zeros(n)
    pa = None if params is None else np.ascontiguousarray(params, dtype=np.float64)
    ns = lib().orc_cell_compute_v(model, integrator, _p(pa), _p(y), t_end, dt, stim[0], stim[1], stim[2], stim[3],
                                  _p(out))
    if ns < 0:
        raise RuntimeError('orc_cell_compute_v failed')
    return out[:ns]


def cell_properties(v, dt, threshold):
    """CellProperties' action-potential bookkeeping on a uniformly sampled
    trace (heart/src/postprocessing/CellProperties.cpp:54-200): completed
    APs, last onset (linear interpolation), peak of the last completed AP and
    the interpolated time it fell back below the threshold (-1 if none)."""
    v = np.asarray(v)
    ups = np.where((v[:-1] <= threshold) & (v[1:] > threshold))[0]
    downs = np.where((v[:-1] >= threshold) & (v[1:] < threshold))[0]

    def cross(i):     # the interval from sample i to i + 1
        return i * dt + dt / (v[i + 1] - v[i]) * (threshold - v[i])
    count, onset, last_peak, end = 0, -1.0, -1.0, -1.0
    i = -1
    while True:
        u = ups[ups > i]
        if len(u) == 0:
            break
        onset = cross(u[0])
        d = downs[downs > u[0]]
        if len(d) == 0:
            break
        last_peak = v[u[0] + 1:d[0] + 2].max()
        end = cross(d[0])
        count += 1
        i = d[0]
    return np.array([count, onset, last_peak, end], dtype=float)


def cell_compute_except_voltage(model, y, t0, t_end, dt, integrator=0, params=None):
    """AbstractCardiacCell::ComputeExceptVoltage then GetIIonic; y updated in place."""
    pa = None if params is None else np.ascontiguousarray(params, dtype=np.float64)
    r = lib().orc_cell_compute_except_voltage(model, integrator, _p(pa), _p(y), t0, t_end, dt)
    if not np.isfinite(r):
        raise RuntimeError('orc_cell_compute_except_voltage failed')
    return r


def cell_iionic(model, y, params=None):
    pa = None if params is None else np.ascontiguousarray(params, dtype=np.float64)
    return lib().orc_cell_iionic(model, _p(pa), _p(np.ascontiguousarray(y, dtype=np.float64)))


def read_triangles(base):
    """ASCII Triangle/TetGen mesh (<base>.node, <base>.ele): the oracle's own
    reading of the format TrianglesMeshReader reads (header, then one item
    per line numbered from 0 or 1, '#' comments) -> (coords, elems from 0)."""
    def lines(path):
        with open(path) as f:
            for raw in f:
                raw = raw.split('#', 1)[0].split()
                if raw:
                    yield raw
    nl = lines(base + '.node')
    head = next(nl)
    n, dim = int(head[0]), int(head[1])
    coords = np.zeros((n, dim))
    first = None
    for i in range(n):
        row = next(nl)
        if first is None:
            first = int(row[0])
        coords[i] = [float(x) for x in row[1:1 + dim]]
    el = lines(base + '.ele')
    head = next(el)
    ne, npe = int(head[0]), int(head[1])
    elems = np.zeros((ne, npe), dtype=np.int32)
    for e in range(ne):
        row = next(el)
        elems[e] = [int(x) - first for x in row[1:1 + npe]]
    return coords, elems


def slab(dim, ex, ey=0, ez=0, h=0.01):
    n, ne = ctypes.c_int(), ctypes.c_int()
    lib().orc_slab_sizes(dim, ex, ey, ez, ctypes.byref(n), ctypes.byref(ne))
    coords = np.zeros((n.value, dim))
    elems = np.zeros((ne.value, dim + 1), dtype=np.int32)
    lib().orc_slab_fill(dim, ex, ey, ez, h, _p(coords), _p(elems, _ip))
    return coords, elems


def assemble(coords, elems, sigma, lumped=False):
    """(row_ptr, col, M, K) with the reference's quadrature.  sigma: one
    dim x dim tensor, or [n_elems, dim, dim] per element."""
    coords = np.ascontiguousarray(coords, dtype=np.float64)
    elems = np.ascontiguousarray(elems, dtype=np.int32)
    n, dim = coords.shape
    sig = np.asarray(sigma, dtype=np.float64)
    per_elem = int(sig.size == elems.shape[0] * dim * dim and sig.size != dim * dim)
    sig = np.ascontiguousarray(sig.reshape(-1))
    nnz = ctypes.c_int()
    lib().orc_assemble(dim, n, _p(coords), elems.shape[0], _p(elems, _ip), _p(sig), per_elem, int(lumped),
                       ctypes.byref(nnz), None, None, None, None)
    rp = np.zeros(n + 1, dtype=np.int32)
    col = np.zeros(nnz.value, dtype=np.int32)
    M = np.zeros(nnz.value)
    K = np.zeros(nnz.value)
    lib().orc_assemble(dim, n, _p(coords), elems.shape[0], _p(elems, _ip), _p(sig), per_elem, int(lumped),
                       ctypes.byref(nnz), _p(rp, _ip), _p(col, _ip), _p(M), _p(K))
    return rp, col, M, K


def pcg(rp, col, val, b, x0, rtol=1e-12, atol=0.0, max_its=10000):
    x = np.array(x0, dtype=np.float64)
    its = lib().orc_pcg(len(b), _p(rp, _ip), _p(col, _ip), _p(val), _p(np.ascontiguousarray(b)), _p(x), rtol, atol,
                        max_its)
    return x, its


def mono_run(dim, ex, ey=0, ez=0, h=0.01, model=LR91_HAND, rush_larsen=False, sigma=(1.75, 1.75, 1.75),
             am=1400.0, cm=1.0, dt_pde=0.01, dt_ode=0.01, stim_mag=-3e6, stim_dur=0.5, stim_start=0.0,
             stim_xlim=1e-5, t_end=1.0, ksp_rtol=1e-12, ksp_atol=0.0, max_its=10000, threads=0,
             params=None, trace_nodes=None, act_threshold=None, want_state=False, lumped=False, sigma_elem=None,
             stim_period=0.0, modifier_threshold=None, modifier_every=1, split=False, backward_euler=False,
             svi=False, stim_mask=None, integrator=None, mesh=None, node_model=None, node_integrator=None):
    """Full FP64 monodomain run (AbstractCardiacProblem::Solve equivalent)
    on the slab, or on mesh = (coords[n, dim], elems[ne, dim+1]); node_model /
    node_integrator [n] give each node its own cell type."""
    if mesh is not None:
        mc = np.ascontiguousarray(mesh[0], dtype=np.float64)
        me = np.ascontiguousarray(mesh[1], dtype=np.int32)
        n = mc.shape[0]
    else:
        mc = me = None
        n = [ex + 1, (ex + 1) * (ey + 1), (ex + 1) * (ey + 1) * (ez + 1)][dim - 1]
    pa = None
    npar = 0
    if params is not None:
        pa = np.ascontiguousarray(params, dtype=np.float64)
        npar = pa.shape[0]
    se = None if sigma_elem is None else np.ascontiguousarray(sigma_elem, dtype=np.float64).reshape(-1)
    mod = None if modifier_threshold is None else np.zeros((4, n))
    mask = None if stim_mask is None else np.ascontiguousarray(stim_mask, dtype=np.uint8)
    cfg = MonoCfg(dim, ex, ey, ez, h, model, _integrator(rush_larsen, backward_euler, integrator), (ctypes.c_double * 3)(*sigma), am, cm, dt_pde, dt_ode,
                  stim_mag, stim_dur, stim_start, stim_xlim, t_end, ksp_rtol, ksp_atol, max_its, threads, npar,
                  _p(pa), int(lumped), _p(se), float(stim_period), _p(mod),
                  0.0 if modifier_threshold is None else float(modifier_threshold), int(modifier_every),
                  int(split), int(svi), _p(mask, ctypes.POINTER(ctypes.c_ubyte)),
                  0 if mc is None else n, 0 if me is None else me.shape[0], _p(mc), _p(me, _ip))
    nm = None if node_model is None else np.ascontiguousarray(node_model, dtype=np.int32)
    ni = None if node_integrator is None else np.ascontiguousarray(node_integrator, dtype=np.int32)
    cfg.node_model, cfg.node_integrator = _p(nm, _ip), _p(ni, _ip)
    V = np.zeros(n)
    S = num_states(model) if nm is None else max(num_states(int(m)) for m in np.unique(nm))
    state = np.zeros((S, n)) if want_state else None
    act = np.zeros(n) if act_threshold is not None else None
    tn = None
    trace = None
    ntr = 0
    if trace_nodes is not None:
        tn = np.ascontiguousarray(trace_nodes, dtype=np.int32)
        ntr = len(tn)
        nsteps = int(round(t_end / dt_pde))
        trace = np.zeros((nsteps + 1, ntr))
    st = MonoStats()
    rc = lib().orc_mono_run(ctypes.byref(cfg), _p(V), _p(state), _p(act),
                            0.0 if act_threshold is None else act_threshold, ntr, _p(tn, _ip), _p(trace),
                            ctypes.byref(st))
    if rc != 0:
        raise RuntimeError('orc_mono_run failed: %d' % rc)
    out = {'V': V, 'stats': {k: getattr(st, k) for k, _ in MonoStats._fields_}}
    if want_state:
        out['state'] = state
    if act is not None:
        out['act'] = act
    if mod is not None:
        out['modifier'] = mod
    if trace is not None:
        out['trace'] = trace[:st.steps]
    return out
