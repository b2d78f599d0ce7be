"""ctypes binding of the C ABI in include/cmx.h (test and benchmark harness).

Every call goes to libcmx.so; there is no fallback: a missing or unloadable
library raises immediately.
"""
import ctypes
import os

import numpy as np

from .build import library_path

FP64, MIXED = 0, 1
LR91, TT06_EPI, ORD_ENDO, HH52, FOX2002, MAHAJAN2008, DN85, FR2000, MALECKAR2008, N98 = range(10)
FORWARD_EULER, RUSH_LARSEN, BACKWARD_EULER, GRL1, GRL2 = 0, 1, 2, 3, 4
FMT_AUTO, FMT_STENCIL, FMT_ELL, FMT_ELL_OFFSETS = 0, 1, 2, 3
KERNEL_KINDS = ['cell', 'residual', 'spmv', 'update', 'pupdate', 'finish', 'svi']

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)


class CmxMesh(ctypes.Structure):
    _fields_ = [('dim', ctypes.c_int32), ('n_nodes', ctypes.c_int32), ('coords', _dp),
                ('n_elems', ctypes.c_int32), ('elems', _ip), ('sigma_elem', _dp),
                ('sigma', ctypes.c_double * 9), ('elem_dim', ctypes.c_int32)]


class CmxStimulus(ctypes.Structure):
    _fields_ = [('magnitude', ctypes.c_double), ('start', ctypes.c_double),
                ('duration', ctypes.c_double), ('period', ctypes.c_double), ('stop', ctypes.c_double)]


class CmxTissue(ctypes.Structure):
    _fields_ = [('model', ctypes.c_int32), ('integrator', ctypes.c_int32),
                ('n_stimuli', ctypes.c_int32), ('stimuli', ctypes.POINTER(CmxStimulus)),
                ('stim_index', _ip), ('n_params', ctypes.c_int32), ('params', _dp),
                ('n_groups', ctypes.c_int32), ('group_of', _ip), ('group_model', _ip), ('group_integrator', _ip)]


class CmxOpts(ctypes.Structure):
    _fields_ = [('precision', ctypes.c_int32), ('format', ctypes.c_int32),
                ('dt_pde', ctypes.c_double), ('dt_ode', ctypes.c_double),
                ('am', ctypes.c_double), ('cm', ctypes.c_double),
                ('ksp_rtol', ctypes.c_double), ('ksp_atol', ctypes.c_double),
                ('ksp_max_its', ctypes.c_int32), ('refine_sweeps', ctypes.c_int32),
                ('use_graph', ctypes.c_int32), ('mass_lumping', ctypes.c_int32),
                ('extrapolate', ctypes.c_int32), ('operator_splitting', ctypes.c_int32),
                ('state_variable_interpolation', ctypes.c_int32), ('separate_refinement', ctypes.c_int32),
                ('lookup_tables', ctypes.c_int32)]


class CmxCells(ctypes.Structure):
    _fields_ = [('model', ctypes.c_int32), ('integrator', ctypes.c_int32), ('precision', ctypes.c_int32),
                ('n_cells', ctypes.c_int32), ('dt', ctypes.c_double), ('t_end', _dp),
                ('stimuli', ctypes.POINTER(CmxStimulus)), ('n_params', ctypes.c_int32), ('params', _dp),
                ('threshold', ctypes.c_double), ('steps_per_launch', ctypes.c_int32),
                ('lookup_tables', ctypes.c_int32), ('lanes_per_cell', ctypes.c_int32),
                ('trace_every', ctypes.c_int32), ('trace', _dp)]


class CmxCellsStats(ctypes.Structure):
    _fields_ = [('kernel_ms', ctypes.c_double), ('cell_steps', ctypes.c_int64), ('launches', ctypes.c_int32),
                ('block', ctypes.c_int32), ('split_launches', ctypes.c_int32)]


class CmxStats(ctypes.Structure):
    _fields_ = [('steps', ctypes.c_int64), ('cg_iterations', ctypes.c_int64),
                ('refine_sweeps', ctypes.c_int64), ('last_iterations', ctypes.c_int32),
                ('format', ctypes.c_int32), ('n_nodes', ctypes.c_int32), ('n_states', ctypes.c_int32),
                ('row_width', ctypes.c_int32), ('n_patterns', ctypes.c_int32), ('nnz', ctypes.c_int64),
                ('last_residual', ctypes.c_double), ('device_bytes', ctypes.c_double),
                ('graph_launches', ctypes.c_int64), ('graph_captures', ctypes.c_int64),
                ('graph_mode', ctypes.c_int32), ('solver_blocks', ctypes.c_int32),
                ('resumes', ctypes.c_int64), ('folded_refinements', ctypes.c_int64),
                ('adjoint_slots', ctypes.c_int32), ('adjoint_sub_interval', ctypes.c_int32),
                ('cg_rows_per_thread', ctypes.c_int32), ('graph_steps', ctypes.c_int32),
                ('adjoint_iterations', ctypes.c_int64)]


# Every symbol include/cmx.h declares, with its ctypes signature.
SIGNATURES = {
    'cmx_create': (ctypes.c_int, [ctypes.POINTER(CmxMesh), ctypes.POINTER(CmxTissue),
                                  ctypes.POINTER(CmxOpts), ctypes.POINTER(ctypes.c_void_p)]),
    'cmx_destroy': (ctypes.c_int, [ctypes.c_void_p]),
    'cmx_read_mesh': (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32),
                                     ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32), _dp, _ip, _dp]),
    'cmx_read_mesh_elements': (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32),
                                              ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32),
                                              ctypes.POINTER(ctypes.c_int32), _dp, _ip, _dp]),
    'cmx_read_fibres': (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32), _dp]),
    'cmx_conductivity_tensors': (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _dp, _dp, _dp]),
    'cmx_io_last_error': (ctypes.c_char_p, []),
    'cmx_h5_open': (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(ctypes.c_char_p),
                                   ctypes.POINTER(ctypes.c_char_p), ctypes.c_char_p, ctypes.POINTER(ctypes.c_void_p)]),
    'cmx_h5_write_step': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, _dp]),
    'cmx_h5_close': (ctypes.c_int, [ctypes.c_void_p]),
    'cmx_last_error': (ctypes.c_char_p, [ctypes.c_void_p]),
    'cmx_step': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_int32]),
    'cmx_get_voltage': (ctypes.c_int, [ctypes.c_void_p, _dp]),
    'cmx_set_voltage': (ctypes.c_int, [ctypes.c_void_p, _dp]),
    'cmx_get_state': (ctypes.c_int, [ctypes.c_void_p, _dp]),
    'cmx_set_state': (ctypes.c_int, [ctypes.c_void_p, _dp]),
    'cmx_get_iionic': (ctypes.c_int, [ctypes.c_void_p, _dp]),
    'cmx_get_stats': (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(CmxStats)]),
    'cmx_set_activation': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_int32, ctypes.c_double]),
    'cmx_get_activation': (ctypes.c_int, [ctypes.c_void_p, _dp]),
    'cmx_set_profiling': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    'cmx_get_kernel_times': (ctypes.c_int, [ctypes.c_void_p, _dp, ctypes.POINTER(ctypes.c_int64), ctypes.c_int32]),
    'cmx_model_info': (ctypes.c_int, [ctypes.c_int32, _ip, _ip]),
    'cmx_load_model': (ctypes.c_int, [ctypes.c_char_p, _ip]),
    'cmx_model_param_alias': (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(ctypes.c_char_p)]),
    'cmx_model_find_param': (ctypes.c_int, [ctypes.c_int32, ctypes.c_char_p, _ip]),
    'cmx_model_state_name': (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(ctypes.c_char_p)]),
    'cmx_model_initial_state': (ctypes.c_int, [ctypes.c_int32, _dp]),
    'cmx_model_param': (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(ctypes.c_char_p), _dp]),
    'cmx_set_params': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, _dp]),
    'cmx_gate_table': (ctypes.c_int, [ctypes.c_int32, ctypes.c_double, ctypes.POINTER(ctypes.c_float), _ip, _ip,
                                      _dp, _dp]),
    'cmx_autotune': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_int32, ctypes.c_int32, _ip,
                                    ctypes.POINTER(ctypes.c_int32), _dp]),
    'cmx_cells_run': (ctypes.c_int, [ctypes.POINTER(CmxCells), _dp, _dp, _dp, ctypes.POINTER(CmxCellsStats)]),
    'cmx_cells_last_error': (ctypes.c_char_p, []),
    'cmx_cells_run_noalloc': (ctypes.c_int, [ctypes.POINTER(CmxCells), ctypes.c_void_p, ctypes.c_void_p,
                                             ctypes.c_void_p, ctypes.c_void_p]),
    'cmx_adjoint': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                   _ip, ctypes.c_int32, _dp, _dp, _dp, _dp]),
    'cmx_set_adjoint_memory': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double]),
    'cmx_set_launch_config': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]),
    'cmx_slab_mesh': (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                     ctypes.c_double, _ip, _ip, _dp, _ip]),
}

_LIB = None


def lib():
    """Load libcmx.so (raises if it is missing: there is no CPU fallback)."""
    global _LIB
    if _LIB is None:
        path = library_path()
        if not os.path.exists(path):
            raise RuntimeError('libcmx.so not built (%s): run __graft_entry__.build()' % path)
        L = ctypes.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


class CmxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__('cmx error %d: %s' % (code, msg))
        self.code = code


def _check(rc, ctx=None):
    if rc != 0:
        raise CmxError(rc, lib().cmx_last_error(ctx).decode())


def _p(a, t=_dp):
    return a.ctypes.data_as(t)


def load_model(path):
    """Loads a cell-model plugin (build.build_model_plugin) and returns its
    model id (cmx_load_model)."""
    mid = ctypes.c_int32()
    _check(lib().cmx_load_model(os.fsencode(os.path.abspath(path)), ctypes.byref(mid)))
    return mid.value


def model_info(model):
    s, p = ctypes.c_int32(), ctypes.c_int32()
    _check(lib().cmx_model_info(model, ctypes.byref(s), ctypes.byref(p)))
    return s.value, p.value


def state_names(model):
    S, _ = model_info(model)
    out = []
    for i in range(S):
        c = ctypes.c_char_p()
        _check(lib().cmx_model_state_name(model, i, ctypes.byref(c)))
        out.append(c.value.decode())
    return out


def initial_state(model):
    S, _ = model_info(model)
    y = np.zeros(S)
    _check(lib().cmx_model_initial_state(model, _p(y)))
    return y


def param_info(model):
    """[(name, default)] of the model's per-node parameters."""
    _, P = model_info(model)
    out = []
    for q in range(P):
        c, d = ctypes.c_char_p(), np.zeros(1)
        _check(lib().cmx_model_param(model, q, ctypes.byref(c), _p(d)))
        out.append((c.value.decode(), float(d[0])))
    return out


def param_aliases(model):
    """The reference's names of the model's parameters ('' where none)."""
    _, P = model_info(model)
    out = []
    for q in range(P):
        c = ctypes.c_char_p()
        _check(lib().cmx_model_param_alias(model, q, ctypes.byref(c)))
        out.append(c.value.decode())
    return out


def find_param(model, name):
    """Index of a parameter by the library's name or the reference's."""
    i = ctypes.c_int32()
    _check(lib().cmx_model_find_param(model, name.encode(), ctypes.byref(i)))
    return i.value


def cells_run(model, integrator, t_end, dt, stimuli=None, params=None, y0=None, precision=FP64,
              threshold=-30.0, steps_per_launch=0, lookup_tables=False, lanes_per_cell=0, trace_every=0):
    """Batched single cells (cmx_cells_run).  t_end [n]; stimuli: one
    (magnitude, start, duration[, period[, stop]]) per cell or None; params
    [n_params, n].  Returns (final states [S, n], props [4, n]: completed APs,
    last onset, last peak, last end; stats dict).  trace_every > 0: V at t = 0
    and every trace_every steps, as stats['trace'] [n_samples, n]."""
    t_end = np.ascontiguousarray(t_end, dtype=np.float64)
    n = t_end.shape[0]
    S, _ = model_info(model)
    c = CmxCells()
    c.model, c.integrator, c.precision, c.n_cells, c.dt = model, integrator, precision, n, dt
    c.t_end = _p(t_end)
    keep = [t_end]
    if stimuli is not None:
        st = (CmxStimulus * n)()
        for k, row in enumerate(stimuli):
            st[k] = CmxStimulus(row[0], row[1], row[2], row[3] if len(row) > 3 else -1.0,
                                row[4] if len(row) > 4 else 0.0)
        c.stimuli = st
        keep.append(st)
    if params is not None:
        pa = np.ascontiguousarray(np.atleast_2d(params), dtype=np.float64)
        c.n_params, c.params = pa.shape[0], _p(pa)
        keep.append(pa)
    c.threshold, c.steps_per_launch = threshold, steps_per_launch
    c.lookup_tables = 1 if lookup_tables else 0
    c.lanes_per_cell = int(lanes_per_cell)
    trace = None
    c.trace_every = int(trace_every)
    if trace_every > 0:
        longest = int(np.round(t_end / dt).max()) if n else 0
        trace = np.zeros((longest // trace_every + 1, n))
        c.trace = _p(trace)
    yi = None if y0 is None else np.ascontiguousarray(y0, dtype=np.float64)
    y = np.zeros((S, n))
    props = np.zeros((4, n))
    stats = CmxCellsStats()
    rc = lib().cmx_cells_run(ctypes.byref(c), _p(yi) if yi is not None else None, _p(y), _p(props),
                             ctypes.byref(stats))
    if rc != 0:
        raise CmxError(rc, lib().cmx_cells_last_error().decode())
    out = {k: getattr(stats, k) for k, _ in CmxCellsStats._fields_}
    if trace is not None:
        out['trace'] = trace
    return y, props, out


def gate_table(model, dt):
    """The gate lookup table of a model for Rush-Larsen steps of length dt:
    (table [rows, width] float32, vmin, dv); row r is V = vmin + r dv, gate j
    has its steady state in column 2j and its decay factor in 2j + 1."""
    L = lib()
    rows, width = ctypes.c_int32(), ctypes.c_int32()
    vmin, dv = ctypes.c_double(), ctypes.c_double()
    _check(L.cmx_gate_table(model, dt, None, ctypes.byref(rows), ctypes.byref(width), ctypes.byref(vmin),
                            ctypes.byref(dv)))
    t = np.zeros((rows.value, width.value), dtype=np.float32)
    _check(L.cmx_gate_table(model, dt, t.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), None, None, None, None))
    return t, vmin.value, dv.value


def slab_mesh(dim, ex, ey=0, ez=0, h=0.01):
    """ConstructRegularSlabMesh-numbered slab: (coords[n, dim], elems[ne, dim+1])."""
    n, ne = ctypes.c_int32(), ctypes.c_int32()
    _check(lib().cmx_slab_mesh(dim, ex, ey, ez, h, ctypes.byref(n), ctypes.byref(ne), None, None))
    coords = np.zeros((n.value, dim))
    elems = np.zeros((ne.value, dim + 1), dtype=np.int32)
    _check(lib().cmx_slab_mesh(dim, ex, ey, ez, h, None, None, _p(coords), _p(elems, _ip)))
    return coords, elems


def _io_check(rc):
    if rc != 0:
        raise RuntimeError(lib().cmx_io_last_error().decode())


def read_mesh(base, elem_dim=0):
    """TrianglesMeshReader: <base>.node/.ele (ASCII or binary), or with
    elem_dim = 1 the line elements of <base>.edge (a cable in 2-D or 3-D
    space) -> (coords[n, dim], elems[ne, elem_dim+1] from zero, elem_attr[ne, n_attr])."""
    d, n, ne, na = (ctypes.c_int32() for _ in range(4))
    b = os.fsencode(base)
    _io_check(lib().cmx_read_mesh_elements(b, elem_dim, ctypes.byref(d), ctypes.byref(n), ctypes.byref(ne),
                                           ctypes.byref(na), None, None, None))
    ed = elem_dim if elem_dim > 0 else d.value
    coords = np.zeros((n.value, d.value))
    elems = np.zeros((ne.value, ed + 1), dtype=np.int32)
    attr = np.zeros((ne.value, na.value))
    _io_check(lib().cmx_read_mesh_elements(b, elem_dim, ctypes.byref(d), ctypes.byref(n), ctypes.byref(ne),
                                           ctypes.byref(na), _p(coords), _p(elems, _ip), _p(attr) if na.value else None))
    return coords, elems, attr


def read_fibres(path, dim, axisymmetric=False):
    """FibreReader: [n, dim, dim] (.ortho: fibre, sheet, normal rows) or
    [n, dim] (.axi)."""
    per = dim if axisymmetric else dim * dim
    n = ctypes.c_int32()
    b = os.fsencode(path)
    _io_check(lib().cmx_read_fibres(b, per, ctypes.byref(n), None))
    out = np.zeros((n.value, per))
    _io_check(lib().cmx_read_fibres(b, per, ctypes.byref(n), _p(out)))
    return out.reshape((n.value, dim) if axisymmetric else (n.value, dim, dim))


def conductivity_tensors(fibres, sigma, axisymmetric=False):
    """Per-element conductivity tensors [n, dim, dim] from fibre directions."""
    f = np.ascontiguousarray(fibres, dtype=np.float64)
    n = f.shape[0]
    dim = f.shape[1]
    sg = np.ascontiguousarray(np.resize(np.asarray(sigma, dtype=np.float64), 3))
    out = np.zeros((n, dim, dim))
    _io_check(lib().cmx_conductivity_tensors(dim, n, 1 if axisymmetric else 0, _p(f), _p(sg), _p(out)))
    return out


class H5Writer:
    """Hdf5DataWriter-layout results file ("Data" [t][node][var], "Time"),
    written one time step at a time:

        with cmx.H5Writer('results.h5', n_nodes, [('V', 'mV')]) as w:
            w.write(t, V)
    """

    def __init__(self, path, n_nodes, variables=(('V', 'mV'),), time_unit='msecs'):
        self.n_nodes, self.n_vars = int(n_nodes), len(variables)
        names = (ctypes.c_char_p * self.n_vars)(*[v[0].encode() for v in variables])
        units = (ctypes.c_char_p * self.n_vars)(*[v[1].encode() for v in variables])
        h = ctypes.c_void_p()
        _io_check(lib().cmx_h5_open(os.fsencode(path), self.n_nodes, self.n_vars, names, units, time_unit.encode(),
                                    ctypes.byref(h)))
        self.h = h

    def write(self, t, values):
        v = np.ascontiguousarray(values, dtype=np.float64).reshape(self.n_nodes, self.n_vars)
        _io_check(lib().cmx_h5_write_step(self.h, float(t), _p(v)))

    def close(self):
        if self.h:
            h, self.h = self.h, None
            _io_check(lib().cmx_h5_close(h))

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Monodomain:
    """One cmx context: a tissue on a mesh with its options."""

    def __init__(self, coords, elems, sigma, model, integrator=FORWARD_EULER,
                 stimuli=(), stim_index=None, params=None, precision=FP64,
                 dt_pde=0.01, dt_ode=0.01, am=1400.0, cm=1.0, ksp_rtol=1e-8, ksp_atol=0.0,
                 ksp_max_its=1000, refine_sweeps=0, fmt=FMT_AUTO, mass_lumping=False,
                 sigma_elem=None, use_graph=False, extrapolate=False, operator_splitting=False, svi=False,
                 cell_types=None, separate_refinement=False, lookup_tables=False):
        """cell_types = (group_of [n], [(model, integrator), ...]) gives node i
        the cell type group_of[i] (model and integrator are then ignored).
        separate_refinement: mixed precision's FP64 true-residual check always
        in a kernel of its own instead of in the last CG iteration.
        lookup_tables: mixed precision with Rush-Larsen cells reads each
        gate's steady state and decay factor from a table in V."""
        L = lib()
        self._keep = []
        coords = np.ascontiguousarray(coords, dtype=np.float64)
        elems = np.ascontiguousarray(elems, dtype=np.int32)
        n, dim = coords.shape
        self.n, self.dim, self.model = n, dim, model
        self.n_states, self.n_params = model_info(model)
        if cell_types is not None:
            self.n_states = max(model_info(mo)[0] for mo, _ in cell_types[1])
        m = CmxMesh()
        m.dim, m.n_nodes, m.coords = dim, n, _p(coords)
        m.n_elems, m.elems = elems.shape[0], _p(elems, _ip)
        m.elem_dim = elems.shape[1] - 1      # line elements (cables) when below dim
        sig = np.zeros(9)
        s = np.asarray(sigma, dtype=np.float64)
        if s.ndim == 1:
            s = np.diag(s[:dim])
        sig[:dim * dim] = s.reshape(-1)
        m.sigma = (ctypes.c_double * 9)(*sig)
        if sigma_elem is not None:
            se = np.ascontiguousarray(sigma_elem, dtype=np.float64)
            self._keep.append(se)
            m.sigma_elem = _p(se)
        t = CmxTissue()
        t.model, t.integrator = model, integrator
        st = (CmxStimulus * max(len(stimuli), 1))()
        for k, row in enumerate(stimuli):
            mag, start, dur = row[0], row[1], row[2]
            period = row[3] if len(row) > 3 else -1.0
            stop = row[4] if len(row) > 4 else 0.0
            st[k] = CmxStimulus(mag, start, dur, period, stop)
        t.n_stimuli, t.stimuli = len(stimuli), st
        if stim_index is not None:
            si = np.ascontiguousarray(stim_index, dtype=np.int32)
            self._keep.append(si)
            t.stim_index = _p(si, _ip)
        if params is not None:
            pa = np.ascontiguousarray(params, dtype=np.float64)
            self._keep.append(pa)
            t.n_params, t.params = pa.shape[0], _p(pa)
        if cell_types is not None:
            go = np.ascontiguousarray(cell_types[0], dtype=np.int32)
            gm = np.array([mo for mo, _ in cell_types[1]], dtype=np.int32)
            gi = np.array([ig for _, ig in cell_types[1]], dtype=np.int32)
            self._keep += [go, gm, gi]
            t.n_groups, t.group_of, t.group_model, t.group_integrator = len(gm), _p(go, _ip), _p(gm, _ip), _p(gi, _ip)
        o = CmxOpts(precision, fmt, dt_pde, dt_ode, am, cm, ksp_rtol, ksp_atol, ksp_max_its,
                    refine_sweeps, 1 if use_graph else 0, 1 if mass_lumping else 0, int(extrapolate), 1 if operator_splitting else 0,
                    1 if svi else 0, 1 if separate_refinement else 0, 1 if lookup_tables else 0)
        self.dt = dt_pde
        ctx = ctypes.c_void_p()
        rc = L.cmx_create(ctypes.byref(m), ctypes.byref(t), ctypes.byref(o), ctypes.byref(ctx))
        _check(rc, None)
        self.ctx = ctx

    def close(self):
        if getattr(self, 'ctx', None):
            lib().cmx_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, t0, n_steps):
        _check(lib().cmx_step(self.ctx, t0, n_steps), self.ctx)

    def voltage(self):
        V = np.zeros(self.n)
        _check(lib().cmx_get_voltage(self.ctx, _p(V)), self.ctx)
        return V

    def set_voltage(self, V):
        V = np.ascontiguousarray(V, dtype=np.float64)
        _check(lib().cmx_set_voltage(self.ctx, _p(V)), self.ctx)

    def state(self):
        S = np.zeros((self.n_states, self.n))
        _check(lib().cmx_get_state(self.ctx, _p(S)), self.ctx)
        return S

    def set_state(self, S):
        S = np.ascontiguousarray(S, dtype=np.float64)
        _check(lib().cmx_set_state(self.ctx, _p(S)), self.ctx)

    def iionic(self):
        out = np.zeros(self.n)
        _check(lib().cmx_get_iionic(self.ctx, _p(out)), self.ctx)
        return out

    def autotune(self, t0=0.0, n_steps=20, candidates=None):
        """Time n_steps from the current state for each candidate solver block
        count, then the CG rows per thread and the steps per captured graph at
        the fastest; keep the fastest configuration (launch_config()); the state
        is restored.  Returns (chosen blocks, {blocks: ms per step})."""
        cand = np.ascontiguousarray([128, 192, 256, 320, 384, 512] if candidates is None else candidates,
                                    dtype=np.int32)
        ms = np.zeros(len(cand))
        chosen = ctypes.c_int32()
        _check(lib().cmx_autotune(self.ctx, float(t0), int(n_steps), len(cand), _p(cand, _ip), ctypes.byref(chosen),
                                  _p(ms)), self.ctx)
        return chosen.value, dict(zip(cand.tolist(), ms.tolist()))

    def set_params(self, params):
        """Per-node parameters [n_params, n] (AbstractParameterisedSystem::SetParameter
        on every node); fewer planes leave the rest at the model defaults."""
        pa = np.ascontiguousarray(np.atleast_2d(params), dtype=np.float64)
        _check(lib().cmx_set_params(self.ctx, pa.shape[0], _p(pa)), self.ctx)

    def adjoint(self, t0, n_steps, obs_nodes, obs_every, obs_values, checkpoint_every=100):
        """Loss sum_j sum_o (V_o(t0 + j obs_every dt) - obs_values[j-1, o])^2 over
        the n_steps steps from the current state, and its gradient with respect
        to the per-node parameters [n_params, n] and the initial state
        [n_states, n].  The context ends at t0 + n_steps dt."""
        nodes = np.ascontiguousarray(obs_nodes, dtype=np.int32)
        vals = np.ascontiguousarray(obs_values, dtype=np.float64).reshape(n_steps // obs_every, len(nodes))
        loss = np.zeros(1)
        gp = np.zeros((max(self.n_params, 1), self.n))
        gs = np.zeros((self.n_states, self.n))
        _check(lib().cmx_adjoint(self.ctx, float(t0), int(n_steps), int(checkpoint_every), len(nodes),
                                 _p(nodes, _ip), int(obs_every), _p(vals), _p(loss),
                                 _p(gp) if self.n_params else None, _p(gs)), self.ctx)
        return float(loss[0]), gp[:self.n_params], gs

    def set_launch_config(self, solver_blocks=0, cg_rows=0, graph_steps=0):
        """Solver blocks, CG rows per thread, steps per captured graph (0 keeps)."""
        _check(lib().cmx_set_launch_config(self.ctx, int(solver_blocks), int(cg_rows), int(graph_steps)), self.ctx)

    def launch_config(self):
        st = self.stats()
        return {'solver_blocks': st['solver_blocks'], 'cg_rows': st['cg_rows_per_thread'],
                'graph_steps': st['graph_steps']}

    def set_adjoint_memory(self, nbytes):
        """Bound on the forward states adjoint() stores (0: the device's free
        memory); past one level of checkpoints it nests a second."""
        _check(lib().cmx_set_adjoint_memory(self.ctx, float(nbytes)), self.ctx)

    def set_activation(self, threshold, every_steps=1, t_now=0.0):
        """ActivationOutputModifier on the device: processes the current voltage
        at t_now, then every every_steps-th PDE step of each step() call."""
        _check(lib().cmx_set_activation(self.ctx, float(threshold), int(every_steps), float(t_now)), self.ctx)

    def activation(self):
        """[5, n]: first activation, first recovery, second activation, second
        recovery (ms, -1 if not yet seen), interpolated first upward crossing."""
        out = np.zeros((5, self.n))
        _check(lib().cmx_get_activation(self.ctx, _p(out)), self.ctx)
        return out

    def set_profiling(self, kinds):
        mask = 0
        for k in kinds:
            mask |= 1 << KERNEL_KINDS.index(k)
        _check(lib().cmx_set_profiling(self.ctx, mask), self.ctx)

    def kernel_times(self, reset=True):
        ms = np.zeros(len(KERNEL_KINDS))
        cnt = np.zeros(len(KERNEL_KINDS), dtype=np.int64)
        _check(lib().cmx_get_kernel_times(self.ctx, _p(ms), cnt.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                          1 if reset else 0), self.ctx)
        return {k: (float(ms[i]), int(cnt[i])) for i, k in enumerate(KERNEL_KINDS)}

    def stats(self):
        s = CmxStats()
        _check(lib().cmx_get_stats(self.ctx, ctypes.byref(s)), self.ctx)
        return {k: getattr(s, k) for k, _ in CmxStats._fields_}
